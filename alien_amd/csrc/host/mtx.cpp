// MatrixMarket system I/O (host). Reading follows Alien's MatrixMarketReader
// (src/movesemantic/alien/move/impl/MatrixMarketReader.cc): the banner must be
// "coordinate real" for a matrix; "general" entries are taken as given, any other
// symmetry mirrors each off-diagonal entry; 1-based indices; repeated entries are
// summed (DoK addNNZ); the result is CSR with ascending columns. Vectors are
// "array real" with one column, one value per line. Writing follows
// MatrixMarketSystemWriter::dump (src/refsemantic/alien/ref/import_export/
// MatrixMarketSystemWriter.cc:34-140) for the layout; values are written with 17
// significant digits so that a write/read round trip is exact.
#include "../alien_hip_internal.h"

#include <algorithm>
#include <cctype>
#include <climits>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace alien_hip {

namespace {
std::string lower(std::string s) {
  std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return (char)std::tolower(c); });
  return s;
}
struct Banner {
  std::string format, scalar, symmetry;
};
// reads up to and including the size line
bool read_banner(std::istream& in, Banner& b, std::string& size_line) {
  std::string line;
  bool have = false;
  while (std::getline(in, line)) {
    if (!have) {
      std::stringstream ss(line);
      std::string mm, obj;
      ss >> mm;
      if (mm != "%%MatrixMarket") continue;
      ss >> obj >> b.format >> b.scalar >> b.symmetry;
      b.format = lower(b.format);
      b.scalar = lower(b.scalar);
      b.symmetry = lower(b.symmetry);
      have = true;
      continue;
    }
    if (line.empty() || line[0] == '%') continue;
    size_line = line;
    return true;
  }
  return false;
}
} // namespace

int mtx_read_csr(const char* path, int& n, std::vector<int>& ro, std::vector<int>& cols, std::vector<double>& vals,
                 std::string& err) {
  std::ifstream in(path);
  if (!in) return err = std::string("cannot open ") + path, -1;
  Banner b;
  std::string size_line;
  if (!read_banner(in, b, size_line)) return err = "no MatrixMarket banner / size line", -1;
  if (b.format != "coordinate" || b.scalar != "real") return err = "matrix must be 'coordinate real'", -1;
  const bool sym = b.symmetry != "general";
  long nr = 0, nc = 0, nz = 0;
  {
    std::stringstream ss(size_line);
    ss >> nr >> nc >> nz;
  }
  if (nr < 0 || nr != nc || nz < 0 || nr > INT32_MAX) return err = "matrix must be square", -1;
  std::vector<std::pair<long, double>> e; // key = row * n + col
  e.reserve(sym ? 2 * nz : nz);
  std::string line;
  long count = 0;
  while (count < nz && std::getline(in, line)) {
    if (line.empty() || line[0] == '%') continue;
    std::stringstream ss(line);
    long r = 0, c = 0;
    double v = 0.0;
    if (!(ss >> r >> c >> v)) return err = "bad entry line: " + line, -1;
    if (r < 1 || r > nr || c < 1 || c > nc) return err = "entry out of range: " + line, -1;
    ++count;
    e.emplace_back((r - 1) * nr + (c - 1), v);
    if (sym && r != c) e.emplace_back((c - 1) * nr + (r - 1), v);
  }
  if (count != nz) return err = "fewer entries than announced", -1;
  // stable sort keeps the file order of repeated entries, which are then summed in that order
  std::stable_sort(e.begin(), e.end(), [](const auto& a, const auto& b2) { return a.first < b2.first; });
  n = (int)nr;
  ro.assign(n + 1, 0);
  cols.clear();
  vals.clear();
  for (size_t k = 0; k < e.size(); ++k) {
    if (k > 0 && e[k].first == e[k - 1].first) {
      vals.back() += e[k].second;
      continue;
    }
    const int r = (int)(e[k].first / nr);
    cols.push_back((int)(e[k].first % nr));
    vals.push_back(e[k].second);
    ro[r + 1]++;
  }
  for (int i = 0; i < n; ++i) ro[i + 1] += ro[i];
  return 0;
}

int mtx_read_vector(const char* path, std::vector<double>& v, std::string& err) {
  std::ifstream in(path);
  if (!in) return err = std::string("cannot open ") + path, -1;
  Banner b;
  std::string size_line;
  if (!read_banner(in, b, size_line)) return err = "no MatrixMarket banner / size line", -1;
  if (b.format != "array" || b.scalar != "real") return err = "vector must be 'array real'", -1;
  long rows = 0, ncol = 0;
  {
    std::stringstream ss(size_line);
    ss >> rows >> ncol;
  }
  if (ncol != 1 || rows < 0) return err = "vector must have one column", -1;
  v.clear();
  std::string line;
  while ((long)v.size() < rows && std::getline(in, line)) {
    if (line.empty() || line[0] == '%') continue;
    std::stringstream ss(line);
    double x = 0.0;
    if (!(ss >> x)) return err = "bad value line: " + line, -1;
    v.push_back(x);
  }
  if ((long)v.size() != rows) return err = "fewer values than announced", -1;
  return 0;
}

int mtx_write_csr(const char* path, int n, const int* ro, const int* cols, const double* vals, const char* desc,
                  std::string& err) {
  FILE* f = std::fopen(path, "w");
  if (!f) return err = std::string("cannot open ") + path, -1;
  std::fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n%%\n%% %s\n%%\n", desc ? desc : "");
  std::fprintf(f, "%% #rows #cols #nonzeros\n%% #vertices #hyperedges #pins\n");
  std::fprintf(f, "%d %d %d\n", n, n, ro[n]);
  for (int i = 0; i < n; ++i)
    for (int k = ro[i]; k < ro[i + 1]; ++k) std::fprintf(f, "%d %d %.17g\n", i + 1, cols[k] + 1, vals[k]);
  return std::fclose(f) == 0 ? 0 : (err = "write failed", -1);
}

// one value per line (the layout the reader expects)
int mtx_write_vector(const char* path, int n, const double* v, const char* desc, std::string& err) {
  FILE* f = std::fopen(path, "w");
  if (!f) return err = std::string("cannot open ") + path, -1;
  std::fprintf(f, "%%%%MatrixMarket matrix array real general\n%%\n%% %s\n%%\n%d 1\n", desc ? desc : "", n);
  for (int i = 0; i < n; ++i) std::fprintf(f, "%.17g\n", v[i]);
  return std::fclose(f) == 0 ? 0 : (err = "write failed", -1);
}

} // namespace alien_hip
