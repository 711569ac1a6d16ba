// Parameter layout and host engine of the learned preconditioner (gnn.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace alien_hip {
namespace gnn {

constexpr int kHidden = 32;
// node features f_i = [1, s_i, 1 - s_i, max_j e_ij, sum_{j>i} e_ij^2, sum_{j<i} e_ij^2]
// (s_i = sum_j e_ij): the squared couplings to later and to earlier rows carry what the
// factorised sparse approximate inverse on A's lower pattern needs (its column scales)
constexpr int kFeat = 6;
// flat FP32 parameter vector: W_in[6x32] b_in[32] W1[64x32] b1[32] W2[64x32] b2[32]
// W3[64x32] b3[32] w_d[32] c_d w_a[32] w_b[32] c_e   (row-major matrices)
constexpr int kOffWin = 0;
constexpr int kOffBin = kOffWin + kFeat * kHidden;
constexpr int kOffW1 = kOffBin + kHidden;
constexpr int kLayerSize = 2 * kHidden * kHidden + kHidden; // W_l and b_l
constexpr int kOffW2 = kOffW1 + kLayerSize;
constexpr int kOffW3 = kOffW2 + kLayerSize;
constexpr int kOffWd = kOffW3 + kLayerSize;
constexpr int kOffCd = kOffWd + kHidden;
constexpr int kOffWa = kOffCd + 1;
constexpr int kOffWb = kOffWa + kHidden;
constexpr int kOffCe = kOffWb + kHidden;
constexpr int kNParam = kOffCe + 1; // 6562
constexpr int kHeadGrads = kNParam - kOffWd; // w_d, c_d, w_a, w_b, c_e = 98
constexpr int kMaxProbes = 8;
constexpr int kMaxCgSteps = 200;

} // namespace gnn
} // namespace alien_hip

#include "../alien/hip/HIPContext.h"
#include "krylov_state.h"

namespace alien_hip {
namespace gnn {

// A system read in place: host CSR with the diagonal positions, the same
// arrays on the device and, when the system is stored as RB-DIA, its host
// bundle offsets and device offsets / lane masks (the factors' structures are
// then the system's split by the sign of the offset).
struct SystemView {
  int n = 0;
  const int* ro = nullptr;
  const int* cols = nullptr;
  const int* dcol = nullptr;
  const double* vals = nullptr; // not read by the engine (its checks run on the device)
  const int* d_ro = nullptr;
  const int* d_cols = nullptr;
  const int* d_dcol = nullptr;
  const double* d_vals = nullptr;
  const int* rb_bo = nullptr;
  const int* d_rb_doff = nullptr;
  const unsigned long long* d_rb_dmask = nullptr;
};

// Device state of the network applied to one matrix (SPD, symmetric pattern,
// stored positive diagonal): scaled couplings, features, activations, Ghat.
class Engine {
 public:
  // copies the host CSR
  Engine(alien_hip_ctx* ctx, int n, const int* ro, const int* cols, const double* vals);
  // reads the system in place: its arrays must outlive the engine
  Engine(alien_hip_ctx* ctx, const SystemView& v);
  int n() const { return m_n; }
  int nnz() const { return m_nnz; }
  alien_hip_ctx* context() const { return m_ctx; }
  // A kept engine for a system with the same structure (new values, arrays possibly moved):
  // re-reads the view, re-runs the checks of the values and re-allocates the per-system buffers;
  // the factors' structures built earlier are kept
  void reset_system(const SystemView& v) { init(v); }
  // activations of every layer and Ghat (per CSR entry; upper entries 0) for
  // the parameters P (device), as loss_grad needs them
  void forward(const float* P);
  // Ghat only, through two activation buffers (the solver's set-up)
  void infer(const float* P);
  // loss = L1 + beta * probe term on 4 probes V (device, N x 4); L1 = Kaporin's condition number
  // (cg_steps = 0) or the residual reduction after cg_steps PCG steps with V as right-hand sides;
  // grad (device, kNParam); parts = {L1, probe term}
  double loss_grad(const float* P, const float* V, int K, float beta, int cg_steps, float* grad, double* parts);
  const float* ghat() const { return d_ghat.data(); }
  // Row-bundled layouts (RB-DIA or RB-CSR) of G and G^T and their values (after forward)
  void build_factor_layout();
  template <typename T> void fill_factor(T* g_vals, T* gt_vals);
  // frees everything but the factors' structures (after fill_factor)
  void release_network_buffers();
  size_t g_slots() const { return m_g.slots; }
  size_t gt_slots() const { return m_gt.slots; }
  bool factor_is_dia() const { return m_g.dia && m_gt.dia; }
  template <typename T, typename V = T> RBMat<T, V> g_mat(const V* v) const { return m_g.mat<T, V>(v); }
  template <typename T, typename V = T> RBMat<T, V> gt_mat(const V* v) const { return m_gt.mat<T, V>(v); }

 private:
  double cg_term(const float* B, int k, float* gg);
  void init(const SystemView& v);
  void prep_features();
  void ensure_activations(bool all_layers);
  alien_hip_ctx* m_ctx;
  int m_n, m_nnz = 0;
  SystemView m_v;
  std::vector<int> h_ro, h_cols, h_dcol, m_gbo, m_gtbo;
  DeviceArray<int> d_ro, d_cols, d_dcol, d_tpos;
  DeviceArray<double> d_vals;
  DeviceArray<float> d_dis, d_e, d_f, d_ghat;
  DeviceArray<float> d_h[4], d_z[4], d_m[4];
  struct Factor {
    bool dia = false;
    size_t slots = 0;
    DeviceArray<int> bo, bc, doff;
    DeviceArray<unsigned long long> dmask;
    // uniform RB-DIA (every bundle holds the same uni slots at offsets uoff; RBMat::uni)
    int uni = 0, nrows = 0;
    int uoff[kUniMaxSlots] = {};
    template <typename T, typename V = T> RBMat<T, V> mat(const V* v) const {
      RBMat<T, V> a{bo.data(), dia ? nullptr : bc.data(), v, dia ? doff.data() : nullptr,
                    dia ? dmask.data() : nullptr};
      if (uni) {
        a.uni = uni;
        a.nrows = nrows;
        for (int j = 0; j < uni; ++j) a.uoff[j] = uoff[j];
      }
      return a;
    }
  };
  Factor m_g, m_gt;
};

void launch_adam(hipStream_t s, int np, float* P, const float* g, float* m, float* v, float lr, int step);

} // namespace gnn
} // namespace alien_hip
