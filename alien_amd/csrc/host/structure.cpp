// Host-side structure builders: diagonal offsets, Alien's BEllPack layout,
// the row-bundled CSR layout and the ILU0 schedules (integer work only).
#include "../alien_hip_internal.h"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <iterator>
#include <vector>

namespace alien_hip {

// dcol[i] = position of the first entry of row i whose column is >= i
// (CSRStructInfo::computeUpperDiagOffset, simple_csr/CSRStructInfo.h:291-309).
// Columns are ascending inside a row, so this is a count of the leading
// entries with column < i.
void host_dcol(int nrows, const int* ro, const int* cols, int* dcol) {
  for (int i = 0; i < nrows; ++i) {
    int k = ro[i];
    const int e = ro[i + 1];
    while (k < e && cols[k] < i) ++k;
    dcol[i] = k;
  }
}

int bellpack_nslices(int nrows, int slice) { return (nrows + slice - 1) / slice; }

// Slice s covers rows [s*slice, s*slice+slice); its width is the longest row
// of the slice and bro[s] is the sum of the widths of the earlier slices
// (computeBlockRowOffset, sycl/data/SYCLBEllPackMatrix.cc:69-88).
void bellpack_bro(int nrows, int slice, const int* ro, int* bro) {
  const int ns = bellpack_nslices(nrows, slice);
  int acc = 0;
  for (int s = 0; s < ns; ++s) {
    bro[s] = acc;
    const int r0 = s * slice, r1 = std::min(nrows, r0 + slice);
    int w = 0;
    for (int r = r0; r < r1; ++r) w = std::max(w, ro[r + 1] - ro[r]);
    acc += w;
  }
  bro[ns] = acc;
}

// Entry j of row r sits at bro[s]*slice + j*slice + (r % slice); slots past the
// end of a row hold column 0 / value 0 (SYCLBEllPackMatrix.cc:127-170, 422-460).
// Masks mark strictly-lower / strictly-upper entries (SYCLBEllPackMatrix.cc:235-315).
void bellpack_fill(int nrows, int slice, const int* ro, const int* cols, const double* vals,
                   const int* bro, int* bcols, double* bvals, uint8_t* lmask, uint8_t* umask) {
  const int ns = bellpack_nslices(nrows, slice);
  const size_t total = (size_t)bro[ns] * (size_t)slice;
  if (total == 0) return;
  std::memset(bcols, 0, total * sizeof(int));
  if (bvals) std::fill(bvals, bvals + total, 0.0);
  if (lmask) std::memset(lmask, 0, total);
  if (umask) std::memset(umask, 0, total);
  std::vector<int> dcol;
  if (lmask || umask) {
    dcol.resize(nrows);
    host_dcol(nrows, ro, cols, dcol.data());
  }
  for (int s = 0; s < ns; ++s) {
    const size_t base = (size_t)bro[s] * slice;
    const int r0 = s * slice, r1 = std::min(nrows, r0 + slice);
    for (int r = r0; r < r1; ++r) {
      const int lane = r - r0;
      const int len = ro[r + 1] - ro[r];
      for (int j = 0; j < len; ++j) {
        const size_t e = base + (size_t)j * slice + lane;
        bcols[e] = cols[ro[r] + j];
        if (bvals) bvals[e] = vals[ro[r] + j];
        if (lmask) lmask[e] = (ro[r] + j < dcol[r]) ? 1 : 0;
        if (umask) umask[e] = (ro[r] + j > dcol[r]) ? 1 : 0;
      }
    }
  }
}

// Level of a row in the dependency DAG of a triangular sweep. Lower sweep:
// row i needs rows cols[k] for k in [ro[i], dcol[i]); upper sweep: rows
// cols[k] for k in (dcol[i], ro[i+1]). Levels are longest-path depths, so rows
// of one level are independent and levels run in increasing order.
int host_levels(int nrows, const int* ro, const int* cols, const int* dcol, int upper, int* level) {
  int nl = 0;
  if (!upper) {
    for (int i = 0; i < nrows; ++i) {
      int lv = 0;
      for (int k = ro[i]; k < dcol[i]; ++k) lv = std::max(lv, level[cols[k]] + 1);
      level[i] = lv;
      nl = std::max(nl, lv + 1);
    }
  } else {
    for (int i = nrows - 1; i >= 0; --i) {
      int lv = 0;
      for (int k = dcol[i] + 1; k < ro[i + 1]; ++k) lv = std::max(lv, level[cols[k]] + 1);
      level[i] = lv;
      nl = std::max(nl, lv + 1);
    }
  }
  return nl;
}

// True when every stored (i, j) has a stored (j, i) (rows sorted ascending;
// an unsorted row can only make the answer false).
static bool pattern_symmetric(int nrows, const int* ro, const int* cols) {
  std::vector<char> ok(16, 1);
  std::atomic<int> slot{0};
  parallel_ranges(nrows, [&](long a, long b) {
    const int me = slot++;
    for (long i = a; i < b && ok[me]; ++i)
      for (int k = ro[i]; k < ro[i + 1]; ++k) {
        const int j = cols[k];
        if (j == i) continue;
        if (!std::binary_search(cols + ro[j], cols + ro[j + 1], (int)i)) {
          ok[me] = 0;
          break;
        }
      }
  });
  return std::all_of(ok.begin(), ok.end(), [](char c) { return c != 0; });
}

// First-fit greedy colouring, rows in natural order, over the union of the
// stored pattern and its transpose. At row i only rows j < i are coloured, so
// only those neighbours count: with a symmetric pattern they are row i's own
// entries of column < i and the transpose is not built.
int host_color(int nrows, const int* ro, const int* cols, int* color) {
  if (pattern_symmetric(nrows, ro, cols)) {
    std::vector<int> mark; // mark[c] == i + 1 <=> colour c is taken by a neighbour of i
    int nc = 0;
    for (int i = 0; i < nrows; ++i) {
      if ((int)mark.size() < nc + 1) mark.resize(nc + 1, 0);
      for (int k = ro[i]; k < ro[i + 1]; ++k) {
        const int j = cols[k];
        if (j < i) mark[color[j]] = i + 1;
      }
      int c = 0;
      while (c < nc && mark[c] == i + 1) ++c;
      color[i] = c;
      nc = std::max(nc, c + 1);
    }
    return nc;
  }
  // transpose pattern (CSC of the strictly off-diagonal part)
  std::vector<int> tcount(nrows + 1, 0);
  for (int i = 0; i < nrows; ++i)
    for (int k = ro[i]; k < ro[i + 1]; ++k)
      if (cols[k] != i) ++tcount[cols[k] + 1];
  for (int i = 0; i < nrows; ++i) tcount[i + 1] += tcount[i];
  std::vector<int> trow(tcount[nrows]);
  std::vector<int> fillp(tcount.begin(), tcount.end() - 1);
  for (int i = 0; i < nrows; ++i)
    for (int k = ro[i]; k < ro[i + 1]; ++k)
      if (cols[k] != i) trow[fillp[cols[k]]++] = i;

  std::vector<int> mark; // mark[c] == i + 1 <=> colour c is taken by a neighbour of i
  int nc = 0;
  for (int i = 0; i < nrows; ++i) color[i] = -1;
  for (int i = 0; i < nrows; ++i) {
    if ((int)mark.size() < nc + 1) mark.resize(nc + 1, 0);
    for (int k = ro[i]; k < ro[i + 1]; ++k) {
      int j = cols[k];
      if (j != i && color[j] >= 0) mark[color[j]] = i + 1;
    }
    for (int k = tcount[i]; k < tcount[i + 1]; ++k) {
      int j = trow[k];
      if (color[j] >= 0) mark[color[j]] = i + 1;
    }
    int c = 0;
    while (c < nc && mark[c] == i + 1) ++c;
    color[i] = c;
    nc = std::max(nc, c + 1);
  }
  return nc;
}

void group_by_level(int nrows, const int* level, int nlevels, int* ptr, int* rows) {
  std::fill(ptr, ptr + nlevels + 1, 0);
  for (int i = 0; i < nrows; ++i) ++ptr[level[i] + 1];
  for (int l = 0; l < nlevels; ++l) ptr[l + 1] += ptr[l];
  std::vector<int> pos(ptr, ptr + nlevels);
  for (int i = 0; i < nrows; ++i) rows[pos[level[i]]++] = i;
}

// RB-CSR: bundle b holds rows [64b, 64b+64); its width is the longest row in
// it; bundle_off[b] counts slots of 64 entries before bundle b.
int rbcsr_bundle_offsets(int nrows, const int* ro, int* bundle_off) {
  const int nb = (nrows + kBundle - 1) / kBundle;
  long acc = 0;
  for (int b = 0; b < nb; ++b) {
    bundle_off[b] = (int)acc;
    const int r0 = b * kBundle, r1 = std::min(nrows, r0 + kBundle);
    int w = 0;
    for (int r = r0; r < r1; ++r) w = std::max(w, ro[r + 1] - ro[r]);
    acc += w;
    if (acc * kBundle > 0x7fffffffL) return -1;
  }
  bundle_off[nb] = (int)acc;
  return 0;
}

// RB-DIA: for each bundle of 64 rows the ascending distinct column offsets
// c - r of the entries k in [kb(r), ke(r)) of its rows (kb = ro[r], ke =
// ro[r+1] unless given), and for each (bundle, offset) the 64-bit mask of the
// rows that hold it. Columns must ascend strictly inside a row (-1 otherwise),
// so a row's offsets ascend too and each bundle's set is a merge of sorted
// lists; consecutive rows of a stencil matrix repeat the same list, which is
// detected and skipped.
long rbdia_structure(int nrows, const int* ro, const int* cols, const int* kb, const int* ke, int max_w,
                     std::vector<int>& bo, std::vector<int>& doff, std::vector<uint64_t>& dmask) {
  const int nb = (nrows + kBundle - 1) / kBundle;
  bo.assign(nb + 1, 0);
  doff.clear();
  dmask.clear();
  std::vector<int> set, tmp, row_off, prev;
  long acc = 0;
  for (int b = 0; b < nb; ++b) {
    bo[b] = (int)acc;
    const int r0 = b * kBundle, r1 = std::min(nrows, r0 + kBundle);
    set.clear();
    prev.clear();
    for (int r = r0; r < r1; ++r) {
      const int k0 = kb ? kb[r] : ro[r], k1 = ke ? ke[r] : ro[r + 1];
      row_off.clear();
      for (int k = k0; k < k1; ++k) {
        if (k > k0 && cols[k] <= cols[k - 1]) return -1; // columns not strictly ascending
        row_off.push_back(cols[k] - r);
      }
      if (row_off == prev) continue;
      tmp.clear();
      std::set_union(set.begin(), set.end(), row_off.begin(), row_off.end(), std::back_inserter(tmp));
      set.swap(tmp);
      prev.swap(row_off);
      if ((int)set.size() > max_w) return -1;
    }
    const size_t s0 = doff.size();
    doff.insert(doff.end(), set.begin(), set.end());
    dmask.resize(doff.size(), 0);
    for (int r = r0; r < r1; ++r) {
      const int k0 = kb ? kb[r] : ro[r], k1 = ke ? ke[r] : ro[r + 1];
      size_t j = s0;
      for (int k = k0; k < k1; ++k) {
        const int d = cols[k] - r;
        while (j < doff.size() && doff[j] != d) ++j;
        if (j == doff.size()) return -1;
        dmask[j++] |= 1ull << (r - r0);
      }
    }
    acc += (long)set.size();
    if (acc * kBundle > 0x7fffffffL) return -1;
  }
  bo[nb] = (int)acc;
  return acc;
}

// Variable-block CSR -> scalar CSR, the order of _seqMultVariableBlock
// (SimpleCSRMatrixMultT.h:439-476): for output row (i, kr) the blocks of row i in
// order, inside each its columns kc of row kr.
long vblock_scalar_nnz(int nbrows, const int* ro, const int* cols, const int* block_size) {
  long nnz = 0;
  for (int i = 0; i < nbrows; ++i)
    for (int j = ro[i]; j < ro[i + 1]; ++j) nnz += (long)block_size[i] * block_size[cols[j]];
  return nnz;
}

int vblock_to_csr(int nbrows, const int* ro, const int* cols, const int* block_cols, const int* block_size,
                  const int* block_offset, const double* vals, int* sro, int* scols, double* svals) {
  long off = 0;
  for (int i = 0; i < nbrows; ++i) {
    if (block_size[i] < 1 || block_offset[i] != off) return -1;
    off += block_size[i];
  }
  for (int j = 0; j < (nbrows ? ro[nbrows] : 0); ++j)
    if (cols[j] < 0 || cols[j] >= nbrows) return -1;
  long pos = 0;
  sro[0] = 0;
  for (int i = 0; i < nbrows; ++i) {
    const int bsr = block_size[i];
    for (int kr = 0; kr < bsr; ++kr) {
      for (int j = ro[i]; j < ro[i + 1]; ++j) {
        const int col = cols[j], bsc = block_size[col];
        for (int kc = 0; kc < bsc; ++kc) {
          scols[pos] = block_offset[col] + kc;
          if (vals) svals[pos] = vals[(long)block_cols[j] + (long)kr * bsc + kc];
          ++pos;
        }
      }
      sro[block_offset[i] + kr + 1] = (int)pos;
    }
  }
  return 0;
}

} // namespace alien_hip
