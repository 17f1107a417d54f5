// Writing an LP out: free-format MPS that read_mps (and the reference's reader,
// src/io/mps_reader.cpp) reads back to the same arrays, and a binary instance cache.
//
// The reference library writes no MPS; its benchmark generator writes MPS from Python for the
// solver's CLI (bench/runners/generate_large_lp.py:49-206). write_mps plays that part for the
// generated workloads here, so the reference's own binary can be run on identical instances.
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <set>
#include <string>
#include <type_traits>
#include <vector>

#include "lp.hpp"
#include "lpfile.hpp"
#include "mps.hpp"

namespace sk {
namespace {

// 17 significant digits: the reader recovers the exact double.
std::string num(double v) {
  char buf[40];
  std::snprintf(buf, sizeof(buf), "%.17g", v);
  return buf;
}

bool usable(const std::vector<std::string>& names, size_t count) {
  if (names.size() != count) return false;
  std::set<std::string> seen;
  for (const std::string& s : names) {
    if (s.empty() || s.find_first_of(" \t\r\n") != std::string::npos || !seen.insert(s).second) return false;
    if (s[0] == '*' || s == "MARKER") return false;
  }
  return true;
}

std::vector<std::string> generated(char prefix, size_t count) {
  std::vector<std::string> out(count);
  for (size_t k = 0; k < count; ++k) out[k] = prefix + std::to_string(k);
  return out;
}

class File {
 public:
  explicit File(const std::string& path) : f_(std::fopen(path.c_str(), "wb")) {}
  ~File() {
    if (f_) std::fclose(f_);
  }
  bool ok() const { return f_ != nullptr && !failed_; }
  void line(const std::string& s) {
    if (!f_) return;
    if (std::fwrite(s.data(), 1, s.size(), f_) != s.size() || std::fputc('\n', f_) == EOF) failed_ = true;
  }
  void bytes(const void* p, size_t n) {
    if (!f_ || n == 0) return;
    if (std::fwrite(p, 1, n, f_) != n) failed_ = true;
  }
  bool close() {
    if (!f_) return false;
    const bool good = !failed_ && std::fclose(f_) == 0;
    f_ = nullptr;
    return good;
  }

 private:
  std::FILE* f_;
  bool failed_ = false;
};

// FNV-1a, 64 bit
uint64_t fnv1a(const unsigned char* p, size_t n, uint64_t h = 1469598103934665603ULL) {
  for (size_t k = 0; k < n; ++k) {
    h ^= p[k];
    h *= 1099511628211ULL;
  }
  return h;
}

constexpr char kMagic[8] = {'S', 'K', 'L', 'P', 'B', 'I', 'N', '1'};

struct Header {
  int32_t num_rows;
  int32_t num_cols;
  int32_t sense;
  int32_t reserved;
  double objective_offset;
  int64_t nnz;
};

}  // namespace

bool write_mps(const HostLp& lp, const MpsInfo* info, const std::string& path, std::string* error) {
  const std::string problem = lp.validate();
  if (!problem.empty()) {
    *error = "write_mps: " + problem;
    return false;
  }
  const auto m = static_cast<size_t>(lp.num_rows), n = static_cast<size_t>(lp.num_cols);
  const std::vector<std::string> rows =
      info && usable(info->row_names, m) ? info->row_names : generated('R', m);
  const std::vector<std::string> cols =
      info && usable(info->col_names, n) ? info->col_names : generated('C', n);
  std::set<std::string> row_set(rows.begin(), rows.end());
  std::string obj = "OBJ";
  for (int k = 0; row_set.count(obj); ++k) obj = "OBJ_" + std::to_string(k);
  std::string name = info && !info->name.empty() ? info->name : "SK_LP";
  if (name.find_first_of(" \t\r\n") != std::string::npos) name = "SK_LP";

  File f(path);
  if (!f.ok()) {
    *error = "write_mps: cannot open '" + path + "' for writing";
    return false;
  }
  f.line("NAME " + name);
  if (lp.sense < 0) {
    f.line("OBJSENSE");
    f.line("    MAX");
  }
  // row types; a two-sided row is written as G or L with a range, whichever gives back both
  // bounds exactly (one of them does unless the difference itself rounds)
  std::vector<char> type(m);
  std::vector<double> rhs(m, 0.0), range(m, 0.0);
  std::vector<uint8_t> ranged(m, 0);
  for (size_t i = 0; i < m; ++i) {
    const double lo = lp.row_lower[i], hi = lp.row_upper[i];
    const bool lf = is_finite_bound(lo), hf = is_finite_bound(hi);
    if (lf && hf && lo == hi) {
      type[i] = 'E';
      rhs[i] = lo;
    } else if (lf && hf) {
      const double r = hi - lo;
      ranged[i] = 1;
      range[i] = r;
      if (lo + r == hi || !(hi - r == lo)) {
        type[i] = 'G';
        rhs[i] = lo;
      } else {
        type[i] = 'L';
        rhs[i] = hi;
      }
    } else if (lf) {
      type[i] = 'G';
      rhs[i] = lo;
    } else if (hf) {
      type[i] = 'L';
      rhs[i] = hi;
    } else {  // a free row: an L row whose right-hand side reads back as +infinity
      type[i] = 'L';
      rhs[i] = 1e30;
    }
  }
  f.line("ROWS");
  f.line(" N  " + obj);
  for (size_t i = 0; i < m; ++i) f.line(std::string(" ") + type[i] + "  " + rows[i]);
  f.line("COLUMNS");
  for (size_t j = 0; j < n; ++j) {
    const int32_t p0 = lp.col_starts[j], p1 = lp.col_starts[j + 1];
    // a column with neither entries nor cost still needs a record, or the reader loses it
    if (lp.cost[j] != 0.0 || p0 == p1) f.line("    " + cols[j] + "  " + obj + "  " + num(lp.cost[j]));
    for (int32_t p = p0; p < p1; ++p)
      f.line("    " + cols[j] + "  " + rows[static_cast<size_t>(lp.row_indices[static_cast<size_t>(p)])] + "  " +
             num(lp.values[static_cast<size_t>(p)]));
  }
  f.line("RHS");
  if (lp.objective_offset != 0.0) f.line("    RHS  " + obj + "  " + num(-lp.objective_offset));
  for (size_t i = 0; i < m; ++i)
    if (rhs[i] != 0.0) f.line("    RHS  " + rows[i] + "  " + num(rhs[i]));
  bool any_range = false;
  for (size_t i = 0; i < m; ++i) any_range = any_range || ranged[i];
  if (any_range) {
    f.line("RANGES");
    for (size_t i = 0; i < m; ++i)
      if (ranged[i]) f.line("    RNG  " + rows[i] + "  " + num(range[i]));
  }
  f.line("BOUNDS");
  for (size_t j = 0; j < n; ++j) {
    const double lo = lp.col_lower[j], hi = lp.col_upper[j];
    const bool lf = is_finite_bound(lo), hf = is_finite_bound(hi);
    const std::string& c = cols[j];
    if (lf && hf && lo == hi) {
      f.line(" FX BND  " + c + "  " + num(lo));
      continue;
    }
    if (!lf && !hf) {
      f.line(" FR BND  " + c);
      continue;
    }
    // the lower bound first, so an upper bound below zero is never read as "lower -inf"
    if (!lf) f.line(" MI BND  " + c);
    else if (lo != 0.0 || (hf && hi < 0.0)) f.line(" LO BND  " + c + "  " + num(lo));
    if (hf) f.line(" UP BND  " + c + "  " + num(hi));
  }
  f.line("ENDATA");
  if (!f.close()) {
    *error = "write_mps: writing '" + path + "' failed";
    return false;
  }
  return true;
}

bool save_lp(const HostLp& lp, const std::string& path, std::string* error) {
  const std::string problem = lp.validate();
  if (!problem.empty()) {
    *error = "save_lp: " + problem;
    return false;
  }
  Header h{lp.num_rows, lp.num_cols, lp.sense, 0, lp.objective_offset, lp.nnz()};
  std::vector<unsigned char> body;
  auto put = [&](const void* p, size_t bytes) {
    const auto* c = static_cast<const unsigned char*>(p);
    body.insert(body.end(), c, c + bytes);
  };
  put(&h, sizeof(h));
  put(lp.col_starts.data(), lp.col_starts.size() * sizeof(int32_t));
  put(lp.row_indices.data(), lp.row_indices.size() * sizeof(int32_t));
  put(lp.values.data(), lp.values.size() * sizeof(double));
  for (const std::vector<double>* v : {&lp.cost, &lp.col_lower, &lp.col_upper, &lp.row_lower, &lp.row_upper})
    put(v->data(), v->size() * sizeof(double));
  const uint64_t hash = fnv1a(body.data(), body.size());
  File f(path);
  if (!f.ok()) {
    *error = "save_lp: cannot open '" + path + "' for writing";
    return false;
  }
  f.bytes(kMagic, sizeof(kMagic));
  f.bytes(body.data(), body.size());
  f.bytes(&hash, sizeof(hash));
  if (!f.close()) {
    *error = "save_lp: writing '" + path + "' failed";
    return false;
  }
  return true;
}

bool load_lp(const std::string& path, HostLp* lp, std::string* error) {
  std::FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) {
    *error = "load_lp: cannot open '" + path + "'";
    return false;
  }
  std::vector<unsigned char> data;
  unsigned char chunk[1 << 16];
  size_t got;
  while ((got = std::fread(chunk, 1, sizeof(chunk), f)) > 0) data.insert(data.end(), chunk, chunk + got);
  const bool read_error = std::ferror(f) != 0;
  std::fclose(f);
  auto fail = [&](const std::string& why) {
    *error = "load_lp: '" + path + "': " + why;
    return false;
  };
  if (read_error) return fail("read error");
  if (data.size() < sizeof(kMagic) + sizeof(Header) + sizeof(uint64_t) ||
      std::memcmp(data.data(), kMagic, sizeof(kMagic)) != 0)
    return fail("not an instance cache file");
  Header h;
  std::memcpy(&h, data.data() + sizeof(kMagic), sizeof(h));
  if (h.num_rows < 0 || h.num_cols < 0 || h.nnz < 0 || h.nnz > INT32_MAX)
    return fail("corrupt header");
  const auto m = static_cast<uint64_t>(h.num_rows), n = static_cast<uint64_t>(h.num_cols),
             z = static_cast<uint64_t>(h.nnz);
  const uint64_t body = sizeof(Header) + (n + 1 + z) * sizeof(int32_t) + (z + 3 * n + 2 * m) * sizeof(double);
  if (data.size() != sizeof(kMagic) + body + sizeof(uint64_t)) return fail("size does not match its header");
  uint64_t stored;
  std::memcpy(&stored, data.data() + sizeof(kMagic) + body, sizeof(stored));
  if (fnv1a(data.data() + sizeof(kMagic), body) != stored) return fail("checksum mismatch");
  const unsigned char* p = data.data() + sizeof(kMagic) + sizeof(Header);
  auto take = [&](auto* v, uint64_t count) {
    using E = typename std::remove_pointer<decltype(v)>::type::value_type;
    v->resize(count);
    std::memcpy(v->data(), p, count * sizeof(E));
    p += count * sizeof(E);
  };
  HostLp out;
  out.num_rows = h.num_rows;
  out.num_cols = h.num_cols;
  out.sense = h.sense;
  out.objective_offset = h.objective_offset;
  take(&out.col_starts, n + 1);
  take(&out.row_indices, z);
  take(&out.values, z);
  take(&out.cost, n);
  take(&out.col_lower, n);
  take(&out.col_upper, n);
  take(&out.row_lower, m);
  take(&out.row_upper, m);
  const std::string problem = out.validate();
  if (!problem.empty()) return fail(problem);
  *lp = std::move(out);
  return true;
}

}  // namespace sk
