// Synthetic system generators for the Krylov workloads (host code, no GPU).
//
// All grids are lexicographic, idx = i + nx*(j + ny*k), with columns stored in
// ascending order inside each row (the sorted-column contract of Alien's
// MatrixProfilerT, src/core/alien/handlers/profiler/MatrixProfilerT.h:94-114).
// Neighbour order in a row is therefore: bottom (k-1), south (j-1), west
// (i-1), diagonal, east (i+1), north (j+1), top (k+1).
//
// Random streams: splitmix64 (Steele, Lea & Flood, "Fast splittable
// pseudorandom number generators", OOPSLA 2014), U[0,1) = (u >> 11) * 2^-53,
// standard normals by Box-Muller on consecutive uniform pairs.
#include "../alien_hip_internal.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace alien_hip {

uint64_t splitmix64_next(uint64_t& s) {
  s += 0x9e3779b97f4a7c15ull;
  uint64_t z = s;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

double splitmix64_u01(uint64_t& s) { return (double)(splitmix64_next(s) >> 11) * 0x1.0p-53; }

void fill_uniform01(uint64_t seed, int64_t n, double* out) {
  uint64_t s = seed;
  for (int64_t i = 0; i < n; ++i) out[i] = splitmix64_u01(s);
}

// n standard normals; pair (u1,u2) -> r = sqrt(-2 ln(1-u1)), z0 = r cos(2 pi u2), z1 = r sin(2 pi u2)
void fill_normal(uint64_t seed, int64_t n, double* out) {
  uint64_t s = seed;
  const double two_pi = 6.283185307179586476925286766559;
  for (int64_t i = 0; i < n; i += 2) {
    double u1 = splitmix64_u01(s);
    double u2 = splitmix64_u01(s);
    double r = std::sqrt(-2.0 * std::log(1.0 - u1));
    out[i] = r * std::cos(two_pi * u2);
    if (i + 1 < n) out[i + 1] = r * std::sin(two_pi * u2);
  }
}

// Stream used for the cell coefficients of the heterogeneous generators, kept
// apart from the x_true stream of the same seed.
static inline uint64_t field_seed(uint64_t seed) { return seed ^ 0xa5a5a5a5a5a5a5a5ull; }

int64_t gen_nnz(int kind, int nx, int ny, int nz) {
  (void)kind;
  // 7-point pattern minus the neighbours that fall outside the grid
  int64_t N = (int64_t)nx * ny * nz;
  int64_t missing = 2ll * ny * nz + 2ll * nx * nz + 2ll * nx * ny;
  return 7 * N - missing;
}

namespace {

struct Grid {
  int nx, ny, nz;
  int64_t N() const { return (int64_t)nx * ny * nz; }
};

// Per-face off-diagonal coefficient callback style: the caller gives, for each
// cell and each of the 6 directions, the coefficient (or marks it absent).
template <typename CoefFn, typename DiagFn>
int build7(const Grid& g, CoefFn coef, DiagFn diag, int* ro, int* cols, double* vals) {
  const int64_t nxy = (int64_t)g.nx * g.ny;
  int64_t k_out = 0;
  ro[0] = 0;
  int64_t row = 0;
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i, ++row) {
        // direction ids: 0 bottom, 1 south, 2 west, 3 east, 4 north, 5 top
        if (k > 0) { cols[k_out] = (int)(row - nxy); vals[k_out++] = coef(i, j, k, 0); }
        if (j > 0) { cols[k_out] = (int)(row - g.nx); vals[k_out++] = coef(i, j, k, 1); }
        if (i > 0) { cols[k_out] = (int)(row - 1); vals[k_out++] = coef(i, j, k, 2); }
        cols[k_out] = (int)row; vals[k_out++] = diag(i, j, k);
        if (i < g.nx - 1) { cols[k_out] = (int)(row + 1); vals[k_out++] = coef(i, j, k, 3); }
        if (j < g.ny - 1) { cols[k_out] = (int)(row + g.nx); vals[k_out++] = coef(i, j, k, 4); }
        if (k < g.nz - 1) { cols[k_out] = (int)(row + nxy); vals[k_out++] = coef(i, j, k, 5); }
        if (k_out > INT32_MAX) return -1;
        ro[row + 1] = (int)k_out;
      }
  return 0;
}

// Harmonic-mean face transmissibility, always evaluated with the lower cell
// index first so that a_ij == a_ji bit for bit.
static inline double harmonic(double ka, double kb) { return (2.0 * ka * kb) / (ka + kb); }

} // namespace

int gen_csr(int kind, int nx, int ny, int nz, const double* params, uint64_t seed, int* ro,
            int* cols, double* vals) {
  Grid g{nx, ny, nz};
  if (nx < 1 || ny < 1 || nz < 1) return -1;
  if (g.N() > INT32_MAX) return -1;
  switch (kind) {
  case ALIEN_HIP_GEN_POISSON7: {
    // 3-D Laplacian, stencil [6, -1 x 6], Dirichlet neighbours eliminated
    return build7(g, [](int, int, int, int) { return -1.0; }, [](int, int, int) { return 6.0; },
                  ro, cols, vals);
  }
  case ALIEN_HIP_GEN_CONVDIFF7: {
    // first-order upwind convection-diffusion, h = 1:
    // params = {D, Peclet, vx, vy, vz}; velocity v = Peclet * (vx, vy, vz) >= 0
    double D = params ? params[0] : 1.0;
    double pe = params ? params[1] : 50.0;
    double v[3] = {params ? pe * params[2] : 50.0, params ? pe * params[3] : 25.0,
                   params ? pe * params[4] : 12.5};
    double dg = 6.0 * D + v[0] + v[1] + v[2];
    double c[6] = {-(D + v[2]), -(D + v[1]), -(D + v[0]), -D, -D, -D};
    return build7(g, [&](int, int, int, int d) { return c[d]; }, [&](int, int, int) { return dg; },
                  ro, cols, vals);
  }
  case ALIEN_HIP_GEN_LOGNORMAL7: {
    // heterogeneous diffusion: k = exp(sigma * n_c), n_c ~ N(0,1) i.i.d. per cell,
    // face T = aniso[axis] * 2 k_a k_b / (k_a + k_b); Dirichlet x = 0 on the face
    // i = 0 adds aniso[0] * 2 k to the diagonal; the other faces are no-flux.
    // params = {sigma, ax, ay, az}
    double sigma = params ? params[0] : 2.0;
    double an[3] = {params ? params[1] : 1.0, params ? params[2] : 1.0, params ? params[3] : 1.0};
    std::vector<double> kc(g.N());
    fill_normal(field_seed(seed), g.N(), kc.data());
    for (auto& v : kc) v = std::exp(sigma * v);
    const int64_t nxy = (int64_t)nx * ny;
    auto cell = [&](int i, int j, int k) { return (int64_t)i + (int64_t)nx * (j + (int64_t)ny * k); };
    auto T = [&](int64_t a, int64_t b, int axis) {
      return an[axis] * (a < b ? harmonic(kc[a], kc[b]) : harmonic(kc[b], kc[a]));
    };
    const int64_t dn[6] = {-nxy, -nx, -1, 1, nx, nxy};
    const int ax[6] = {2, 1, 0, 0, 1, 2};
    auto coef = [&](int i, int j, int k, int d) {
      int64_t a = cell(i, j, k);
      return -T(a, a + dn[d], ax[d]);
    };
    auto diag = [&](int i, int j, int k) {
      int64_t a = cell(i, j, k);
      double s = 0.0;
      // same order as the stored row: bottom, south, west, (dirichlet), east, north, top
      if (k > 0) s += T(a, a - nxy, 2);
      if (j > 0) s += T(a, a - nx, 1);
      if (i > 0) s += T(a, a - 1, 0);
      else s += an[0] * (2.0 * kc[a]);
      if (i < nx - 1) s += T(a, a + 1, 0);
      if (j < ny - 1) s += T(a, a + nx, 1);
      if (k < nz - 1) s += T(a, a + nxy, 2);
      return s;
    };
    return build7(g, coef, diag, ro, cols, vals);
  }
  default:
    return -1;
  }
}

// 2 unknowns per cell, 2x2 blocks in Alien's column-major block layout
// (m[ieq + bs*iu], SimpleCSRMatrixMultT.h:337). Per unknown u an independent
// log-normal field k^u (normals 2c and 2c+1 of the field stream), face block
// F = diag(T^0, T^1) * [[1, c],[c, 1]] with c = coupling: off-diagonal block
// -F, diagonal block = sum of the face blocks (+ Dirichlet on i = 0).
// params = {sigma, coupling}
int gen_bsr2(int nx, int ny, int nz, const double* params, uint64_t seed, int* ro, int* cols,
             double* vals) {
  Grid g{nx, ny, nz};
  if (nx < 1 || ny < 1 || nz < 1 || g.N() > INT32_MAX / 2) return -1;
  double sigma = params ? params[0] : 2.0;
  double cpl = params ? params[1] : 0.3;
  const int64_t N = g.N();
  std::vector<double> kc(2 * N);
  fill_normal(field_seed(seed), 2 * N, kc.data());
  for (auto& v : kc) v = std::exp(sigma * v);
  const int64_t nxy = (int64_t)nx * ny;
  auto T = [&](int64_t a, int64_t b, int u) {
    return a < b ? harmonic(kc[2 * a + u], kc[2 * b + u]) : harmonic(kc[2 * b + u], kc[2 * a + u]);
  };
  int64_t kout = 0;
  ro[0] = 0;
  int64_t a = 0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i, ++a) {
        int64_t nb[6];
        int nn = 0;
        if (k > 0) nb[nn++] = a - nxy;
        if (j > 0) nb[nn++] = a - nx;
        if (i > 0) nb[nn++] = a - 1;
        int64_t lower = nn;
        if (i < nx - 1) nb[nn++] = a + 1;
        if (j < ny - 1) nb[nn++] = a + nx;
        if (k < nz - 1) nb[nn++] = a + nxy;
        double d0 = 0.0, d1 = 0.0;
        for (int q = 0; q < nn; ++q) { d0 += T(a, nb[q], 0); d1 += T(a, nb[q], 1); }
        if (i == 0) { d0 += 2.0 * kc[2 * a]; d1 += 2.0 * kc[2 * a + 1]; }
        auto put = [&](int64_t col, double t0, double t1, double sgn) {
          cols[kout] = (int)col;
          double* m = vals + 4 * kout;
          m[0] = sgn * t0;       // (0,0)
          m[1] = sgn * cpl * t1; // (1,0)
          m[2] = sgn * cpl * t0; // (0,1)
          m[3] = sgn * t1;       // (1,1)
          ++kout;
        };
        for (int q = 0; q < lower; ++q) put(nb[q], T(a, nb[q], 0), T(a, nb[q], 1), -1.0);
        put(a, d0, d1, 1.0);
        for (int q = (int)lower; q < nn; ++q) put(nb[q], T(a, nb[q], 0), T(a, nb[q], 1), -1.0);
        ro[a + 1] = (int)kout;
      }
  return 0;
}

} // namespace alien_hip
