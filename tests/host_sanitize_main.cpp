// Driver for the sanitizer builds of the host structure builders (tests/test_host_sanitizers.py):
// runs the colouring (symmetric pattern: threaded symmetry check, no transpose; nonsymmetric:
// transpose), level sets, BEllPack, RB-CSR and RB-DIA builders on a 7-point grid large enough
// for the host threads, checks a few invariants and prints "ok".
#include <cstdio>
#include <vector>

#include "alien_hip_internal.h"

using namespace alien_hip;

static int check(bool c, const char* what) {
  if (!c) std::printf("FAILED: %s\n", what);
  return c ? 0 : 1;
}

int main() {
  const int N = 64, n = N * N * N;
  std::vector<int> ro(n + 1, 0), cols;
  std::vector<double> vals;
  for (int i = 0; i < n; ++i) {
    const int x = i % N, y = (i / N) % N, z = i / (N * N);
    const int c[7] = {i - N * N, i - N, i - 1, i, i + 1, i + N, i + N * N};
    const bool ok[7] = {z > 0, y > 0, x > 0, true, x < N - 1, y < N - 1, z < N - 1};
    for (int k = 0; k < 7; ++k)
      if (ok[k]) {
        cols.push_back(c[k]);
        vals.push_back(c[k] == i ? 6.0 : -1.0);
      }
    ro[i + 1] = (int)cols.size();
  }
  int bad = 0;
  std::vector<int> color(n), level(n), dcol(n);
  const int nc = host_color(n, ro.data(), cols.data(), color.data());
  bad += check(nc == 2, "red-black colouring of the 7-point grid");
  for (int i = 0; i < n; ++i)
    for (int k = ro[i]; k < ro[i + 1]; ++k)
      if (cols[k] != i && color[cols[k]] == color[i]) bad += check(false, "coupled rows share a colour");
  // a nonsymmetric pattern: drop the x-1 coupling of every row
  std::vector<int> ro2(n + 1, 0), cols2;
  for (int i = 0; i < n; ++i) {
    for (int k = ro[i]; k < ro[i + 1]; ++k)
      if (cols[k] != i - 1) cols2.push_back(cols[k]);
    ro2[i + 1] = (int)cols2.size();
  }
  const int nc2 = host_color(n, ro2.data(), cols2.data(), color.data());
  bad += check(nc2 == 2, "colouring of the nonsymmetric pattern (x+1 couplings keep 2 colours)");
  host_dcol(n, ro.data(), cols.data(), dcol.data());
  const int nl = host_levels(n, ro.data(), cols.data(), dcol.data(), 0, level.data());
  bad += check(nl == 3 * (N - 1) + 1, "level count of the lower DAG");
  std::vector<int> ptr(nl + 1), rows(n);
  group_by_level(n, level.data(), nl, ptr.data(), rows.data());
  const int slice = 1024, ns = bellpack_nslices(n, slice);
  std::vector<int> bro(ns + 1);
  bellpack_bro(n, slice, ro.data(), bro.data());
  std::vector<int> bc((size_t)bro[ns] * slice);
  std::vector<double> bv(bc.size());
  std::vector<uint8_t> lm(bc.size()), um(bc.size());
  bellpack_fill(n, slice, ro.data(), cols.data(), vals.data(), bro.data(), bc.data(), bv.data(), lm.data(),
                um.data());
  std::vector<int> bo((n + kBundle - 1) / kBundle + 1);
  bad += check(rbcsr_bundle_offsets(n, ro.data(), bo.data()) == 0, "RB-CSR bundle offsets");
  std::vector<int> dbo, doff;
  std::vector<uint64_t> dmask;
  const long slots = rbdia_structure(n, ro.data(), cols.data(), nullptr, nullptr, kRbdiaMaxWidth, dbo, doff, dmask);
  bad += check(slots > 0 && slots <= 7L * (long)((n + kBundle - 1) / kBundle), "RB-DIA structure");
  if (!bad) std::printf("ok\n");
  return bad;
}
