// MPS reader (reference src/io/mps_reader.cpp, src/io/mps_bounds.cpp, io.hpp:35-46): free
// format, and fixed columns when the free reading of a file is impossible (names with blanks).
// Sections NAME, OBJSENSE (value on the header line or the next), ROWS, COLUMNS (with integer
// MARKER records), RHS, RANGES, BOUNDS, ENDATA; the first vector of RHS, RANGES and BOUNDS is
// used and later ones are ignored; quadratic sections are refused (this is an LP engine).
// Every rejection names the file and line.
#include "mps.hpp"

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <numeric>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace sk {
namespace {

constexpr double kZeroDrop = 1e-11;  // tolerances.hpp:115, SparseMatrix::finalize
// fixed-format fields 1-6 as (0-based offset, width) (mps_reader.cpp:66)
constexpr int kFixed[6][2] = {{1, 2}, {4, 8}, {14, 8}, {24, 12}, {39, 8}, {49, 12}};

enum class Section { kNone, kObjsenseValue, kRows, kColumns, kRhs, kRanges, kBounds };

std::string upper(std::string s) {
  for (char& c : s) c = static_cast<char>(std::toupper(static_cast<unsigned char>(c)));
  return s;
}

std::string trim(const std::string& s) {
  size_t b = 0, e = s.size();
  while (b < e && std::isspace(static_cast<unsigned char>(s[b]))) ++b;
  while (e > b && std::isspace(static_cast<unsigned char>(s[e - 1]))) --e;
  return s.substr(b, e - b);
}

// token.hpp:63-95: the whole field must be a number; Fortran D exponents accepted; NaN refused
bool parse_number(const std::string& text, double* out) {
  std::string t = trim(text);
  if (t.empty()) return false;
  for (char& c : t)
    if (c == 'D' || c == 'd') c = 'E';
  errno = 0;
  char* end = nullptr;
  const double v = std::strtod(t.c_str(), &end);
  if (end == t.c_str() || *end != '\0' || std::isnan(v)) return false;
  *out = v;
  return true;
}

double normalize_infinity(double v) {
  if (v >= 1e30) return INFINITY;
  if (v <= -1e30) return -INFINITY;
  return v;
}

class Parser {
 public:
  Parser(const std::string& path, bool fixed) : path_(path), fixed_(fixed) {}
  bool run(HostLp* lp, MpsInfo* info, std::string* error);

 private:
  bool fail(const std::string& why) {
    error_ = path_ + ":" + std::to_string(line_no_) + ": " + why;
    return false;
  }
  void split(const std::string& line);
  bool rows();
  bool columns();
  bool rhs();
  bool ranges();
  bool bounds();
  int32_t find_row(const std::string& name) const {
    const auto it = row_index_.find(name);
    return it == row_index_.end() ? kUnknown : it->second;
  }
  int32_t find_col(const std::string& name) const {
    const auto it = col_index_.find(name);
    return it == col_index_.end() ? -1 : it->second;
  }

  static constexpr int32_t kObjective = -1, kIgnored = -2, kUnknown = -3;
  std::string path_;
  bool fixed_;
  int64_t line_no_ = 0;
  std::string error_;
  std::vector<std::string> tok_;
  bool have_objective_ = false;
  bool integer_marker_ = false;
  std::unordered_map<std::string, int32_t> row_index_, col_index_;
  std::vector<char> row_type_;
  std::vector<double> row_rhs_, row_range_;
  std::vector<char> row_has_range_;
  std::vector<std::string> row_names_, col_names_;
  std::vector<double> cost_, lower_, upper_;
  std::vector<char> lower_explicit_, integer_;
  std::unordered_set<int32_t> objective_seen_;
  std::unordered_set<uint64_t> entry_seen_;
  std::vector<int32_t> tri_row_, tri_col_;
  std::vector<double> tri_val_;
  double objective_rhs_ = 0.0;
  std::string rhs_vector_, range_vector_, bound_vector_;
  int32_t sense_ = 1;
  std::string name_;
};

void Parser::split(const std::string& line) {
  tok_.clear();
  const bool header = !line.empty() && !std::isspace(static_cast<unsigned char>(line[0]));
  if (fixed_ && !header) {
    for (const auto& f : kFixed) {
      if (static_cast<size_t>(f[0]) >= line.size()) break;
      const std::string field = trim(line.substr(static_cast<size_t>(f[0]), static_cast<size_t>(f[1])));
      if (!field.empty()) tok_.push_back(field);
    }
    return;
  }
  size_t i = 0;
  while (i < line.size()) {
    while (i < line.size() && std::isspace(static_cast<unsigned char>(line[i]))) ++i;
    size_t j = i;
    while (j < line.size() && !std::isspace(static_cast<unsigned char>(line[j]))) ++j;
    if (j > i) tok_.push_back(line.substr(i, j - i));
    i = j;
  }
  if (fixed_) return;
  // '$' opens a comment in free format
  for (size_t k = 0; k < tok_.size(); ++k)
    if (tok_[k][0] == '$') {
      tok_.resize(k);
      break;
    }
}

// mps_reader.cpp:206-251
bool Parser::rows() {
  if (tok_.size() != 2)
    return fail("ROWS entry has " + std::to_string(tok_.size()) + " fields, expected exactly 2 (type and name)");
  const std::string type = upper(tok_[0]);
  if (type != "N" && type != "L" && type != "G" && type != "E")
    return fail("unknown row type '" + tok_[0] + "', expected N, L, G or E");
  int32_t assigned;
  if (type == "N") {
    assigned = have_objective_ ? kIgnored : kObjective;  // later N rows are free rows: dropped
    have_objective_ = true;
  } else {
    assigned = static_cast<int32_t>(row_type_.size());
    row_type_.push_back(type[0]);
    row_rhs_.push_back(0.0);
    row_range_.push_back(0.0);
    row_has_range_.push_back(0);
    row_names_.push_back(tok_[1]);
  }
  if (!row_index_.emplace(tok_[1], assigned).second) return fail("duplicate row name '" + tok_[1] + "'");
  return true;
}

// mps_reader.cpp:158-204, 253-319
bool Parser::columns() {
  bool marker = false, intorg = false, intend = false;
  for (const std::string& raw : tok_) {
    std::string t = raw;
    if (t.size() >= 2 && t.front() == '\'' && t.back() == '\'') t = t.substr(1, t.size() - 2);
    t = upper(t);
    marker = marker || t == "MARKER";
    intorg = intorg || t == "INTORG";
    intend = intend || t == "INTEND";
  }
  if (marker || intorg || intend) {
    if (intorg) integer_marker_ = true;
    else if (intend) integer_marker_ = false;
    else return fail("MARKER record names neither INTORG nor INTEND");
    return true;
  }
  if (tok_.size() != 3 && tok_.size() != 5)
    return fail("COLUMNS entry has " + std::to_string(tok_.size()) +
                " fields, expected 3 or 5 (column, row, value[, row, value])");
  int32_t col = find_col(tok_[0]);
  if (col < 0) {
    col = static_cast<int32_t>(cost_.size());
    col_index_.emplace(tok_[0], col);
    col_names_.push_back(tok_[0]);
    cost_.push_back(0.0);
    lower_.push_back(0.0);  // the MPS default is [0, +inf)
    upper_.push_back(INFINITY);
    lower_explicit_.push_back(0);
    integer_.push_back(integer_marker_ ? 1 : 0);
  }
  for (size_t k = 1; k + 1 < tok_.size(); k += 2) {
    const int32_t row = find_row(tok_[k]);
    if (row == kUnknown) return fail("COLUMNS entry names unknown row '" + tok_[k] + "'");
    double v = 0.0;
    if (!parse_number(tok_[k + 1], &v)) return fail("'" + tok_[k + 1] + "' is not a number");
    if (row == kIgnored) continue;
    if (!std::isfinite(v))
      return fail("coefficient for column '" + col_names_[static_cast<size_t>(col)] +
                  "' is not finite; a constraint or objective coefficient must be finite");
    if (row == kObjective) {
      if (!objective_seen_.insert(col).second)
        return fail("duplicate objective coefficient for column '" + col_names_[static_cast<size_t>(col)] + "'");
      cost_[static_cast<size_t>(col)] = v;
      continue;
    }
    const uint64_t key = (static_cast<uint64_t>(static_cast<uint32_t>(row)) << 32) | static_cast<uint32_t>(col);
    if (!entry_seen_.insert(key).second)
      return fail("duplicate entry for row '" + row_names_[static_cast<size_t>(row)] + "' and column '" +
                  col_names_[static_cast<size_t>(col)] + "'; MPS has no accumulate semantics");
    if (v != 0.0) {
      tri_row_.push_back(row);
      tri_col_.push_back(col);
      tri_val_.push_back(v);
    }
  }
  return true;
}

// mps_reader.cpp:321-371
bool Parser::rhs() {
  size_t k = 0;
  if (tok_.size() % 2 == 1) {
    if (rhs_vector_.empty()) rhs_vector_ = tok_[0];
    if (tok_[0] != rhs_vector_) return true;  // a later RHS vector: ignored
    k = 1;
  }
  if (tok_.size() <= k) return fail("RHS entry has no row/value pair");
  for (; k + 1 < tok_.size(); k += 2) {
    const int32_t row = find_row(tok_[k]);
    if (row == kUnknown) return fail("RHS entry names unknown row '" + tok_[k] + "'");
    double v = 0.0;
    if (!parse_number(tok_[k + 1], &v)) return fail("'" + tok_[k + 1] + "' is not a number");
    if (row == kIgnored) continue;
    if (row == kObjective) objective_rhs_ = v;
    else row_rhs_[static_cast<size_t>(row)] = v;
  }
  return true;
}

// mps_bounds.cpp (RANGES)
bool Parser::ranges() {
  size_t k = 0;
  if (tok_.size() % 2 == 1) {
    if (range_vector_.empty()) range_vector_ = tok_[0];
    if (tok_[0] != range_vector_) return true;
    k = 1;
  }
  if (tok_.size() <= k) return fail("RANGES entry has no row/value pair");
  for (; k + 1 < tok_.size(); k += 2) {
    const int32_t row = find_row(tok_[k]);
    if (row == kUnknown) return fail("RANGES entry names unknown row '" + tok_[k] + "'");
    double v = 0.0;
    if (!parse_number(tok_[k + 1], &v)) return fail("'" + tok_[k + 1] + "' is not a number");
    if (row == kIgnored) continue;
    if (row == kObjective) return fail("RANGES entry on the objective row is not meaningful");
    row_range_[static_cast<size_t>(row)] = v;
    row_has_range_[static_cast<size_t>(row)] = 1;
  }
  return true;
}

// mps_bounds.cpp (BOUNDS)
bool Parser::bounds() {
  if (tok_.size() < 2) return fail("BOUNDS entry needs a type and a column");
  const std::string type = upper(tok_[0]);
  const bool takes_value = type == "UP" || type == "LO" || type == "FX" || type == "LI" ||
                           type == "UI" || type == "SC";
  const bool valueless = type == "FR" || type == "MI" || type == "PL" || type == "BV";
  if (!takes_value && !valueless)
    return fail("unknown bound type '" + tok_[0] + "'; expected UP LO FX FR MI PL BV LI UI");
  size_t field = 1;
  if (takes_value) {
    if (tok_.size() == 4) field = 2;
    else if (tok_.size() == 3) field = 1;
    else return fail("bound type " + type + " needs a column and a value");
  } else if (tok_.size() >= 3) {
    field = 2;  // a dummy value in field 4 is tolerated and ignored
  }
  if (field == 2) {
    if (bound_vector_.empty()) bound_vector_ = tok_[1];
    if (tok_[1] != bound_vector_) return true;
  }
  const int32_t col = find_col(tok_[field]);
  if (col < 0) return fail("BOUNDS entry names column '" + tok_[field] + "', which has no COLUMNS entry");
  const auto u = static_cast<size_t>(col);
  double value = 0.0;
  if (takes_value) {
    if (!parse_number(tok_[field + 1], &value)) return fail("'" + tok_[field + 1] + "' is not a number");
    value = normalize_infinity(value);
  }
  if (type == "UP") {
    upper_[u] = value;
    // the negative-upper convention: with no explicit lower bound, the lower bound is -inf
    if (value < 0.0 && !lower_explicit_[u]) lower_[u] = -INFINITY;
  } else if (type == "LO") {
    lower_[u] = value;
    lower_explicit_[u] = 1;
  } else if (type == "FX") {
    lower_[u] = upper_[u] = value;
    lower_explicit_[u] = 1;
  } else if (type == "FR") {
    lower_[u] = -INFINITY;
    upper_[u] = INFINITY;
    lower_explicit_[u] = 1;
  } else if (type == "MI") {
    lower_[u] = -INFINITY;
    lower_explicit_[u] = 1;
  } else if (type == "PL") {
    upper_[u] = INFINITY;
  } else if (type == "BV") {
    integer_[u] = 1;
    lower_[u] = 0.0;
    upper_[u] = 1.0;
    lower_explicit_[u] = 1;
  } else if (type == "LI") {
    integer_[u] = 1;
    lower_[u] = value;
    lower_explicit_[u] = 1;
  } else if (type == "UI") {
    integer_[u] = 1;
    upper_[u] = value;
  } else {
    return fail("semi-continuous bounds (SC) are not supported");
  }
  return true;
}

// mps_reader.cpp:437-535, 537-643
bool Parser::run(HostLp* lp, MpsInfo* info, std::string* error) {
  if (path_.size() > 3 && path_.compare(path_.size() - 3, 3, ".gz") == 0) {
    *error = path_ + ": gzip input is not supported by this reader";
    return false;
  }
  std::ifstream in(path_);
  if (!in) {
    *error = path_ + ": cannot open";
    return false;
  }
  Section section = Section::kNone;
  bool saw_end = false;
  std::string line;
  while (std::getline(in, line)) {
    ++line_no_;
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line.empty() || line[0] == '*') continue;
    const bool indented = std::isspace(static_cast<unsigned char>(line[0])) != 0;
    split(line);
    if (tok_.empty()) continue;
    if (!indented) {
      const std::string key = upper(tok_[0]);
      if (key == "ENDATA") {
        saw_end = true;
        break;
      }
      if (key == "NAME") {
        name_ = tok_.size() >= 2 ? tok_[1] : std::string();
        section = Section::kNone;
        continue;
      }
      if (key == "OBJSENSE" || key == "OBJSENS") {
        if (tok_.size() >= 2) {
          const std::string v = upper(tok_[1]);
          sense_ = (v == "MAX" || v == "MAXIMIZE") ? -1 : 1;
          section = Section::kNone;
        } else {
          section = Section::kObjsenseValue;
        }
        continue;
      }
      if (key == "QUADOBJ" || key == "QMATRIX" || key == "QSECTION" || key == "QUADS") {
        *error = path_ + ":" + std::to_string(line_no_) +
                 ": quadratic objective section; this is an LP reader";
        return false;
      }
      Section next = Section::kNone;
      if (key == "ROWS") next = Section::kRows;
      else if (key == "COLUMNS") next = Section::kColumns;
      else if (key == "RHS") next = Section::kRhs;
      else if (key == "RANGES") next = Section::kRanges;
      else if (key == "BOUNDS") next = Section::kBounds;
      if (next != Section::kNone) {
        section = next;
        continue;
      }
    }
    bool ok = true;
    switch (section) {
      case Section::kObjsenseValue: {
        const std::string v = upper(tok_[0]);
        sense_ = (v == "MAX" || v == "MAXIMIZE") ? -1 : 1;
        section = Section::kNone;
        break;
      }
      case Section::kRows: ok = rows(); break;
      case Section::kColumns: ok = columns(); break;
      case Section::kRhs: ok = rhs(); break;
      case Section::kRanges: ok = ranges(); break;
      case Section::kBounds: ok = bounds(); break;
      case Section::kNone: ok = fail("data line appears before any recognised section header"); break;
    }
    if (!ok) {
      *error = error_;
      return false;
    }
  }
  if (!saw_end) {
    *error = path_ + ": file ends without an ENDATA record";
    return false;
  }
  // row bounds from type, right-hand side and range
  const size_t m = row_type_.size(), n = cost_.size();
  lp->num_rows = static_cast<int32_t>(m);
  lp->num_cols = static_cast<int32_t>(n);
  lp->sense = sense_;
  lp->row_lower.assign(m, -INFINITY);
  lp->row_upper.assign(m, INFINITY);
  for (size_t i = 0; i < m; ++i) {
    const double b = normalize_infinity(row_rhs_[i]);
    double lo = -INFINITY, hi = INFINITY;
    const char t = row_type_[i];
    if (!row_has_range_[i]) {
      if (t == 'L') hi = b;
      else if (t == 'G') lo = b;
      else lo = hi = b;
    } else {
      const double r = row_range_[i], mag = std::fabs(r);
      if (t == 'G') {
        lo = b;
        hi = b + mag;
      } else if (t == 'L') {
        lo = b - mag;
        hi = b;
      } else if (r >= 0.0) {
        lo = b;
        hi = b + r;
      } else {
        lo = b + r;
        hi = b;
      }
    }
    if (lo > hi) {
      *error = path_ + ": row '" + row_names_[i] + "' has an empty range";
      return false;
    }
    lp->row_lower[i] = lo;
    lp->row_upper[i] = hi;
  }
  lp->cost = cost_;
  lp->col_lower.resize(n);
  lp->col_upper.resize(n);
  for (size_t j = 0; j < n; ++j) {
    lp->col_lower[j] = normalize_infinity(lower_[j]);
    lp->col_upper[j] = normalize_infinity(upper_[j]);
  }
  const double offset = -objective_rhs_;
  lp->objective_offset = offset == 0.0 ? 0.0 : offset;
  // CSC: counting sort by column, rows ascending within a column, |v| < 1e-11 dropped
  // (SparseMatrix::finalize, sparse.cpp:55-136)
  std::vector<int32_t> start(n + 1, 0);
  for (int32_t c : tri_col_) ++start[static_cast<size_t>(c) + 1];
  for (size_t j = 0; j < n; ++j) start[j + 1] += start[j];
  std::vector<int32_t> next(start.begin(), start.end() - 1), rows(tri_row_.size());
  std::vector<double> vals(tri_row_.size());
  for (size_t k = 0; k < tri_row_.size(); ++k) {
    const auto at = static_cast<size_t>(next[static_cast<size_t>(tri_col_[k])]++);
    rows[at] = tri_row_[k];
    vals[at] = tri_val_[k];
  }
  lp->col_starts.assign(n + 1, 0);
  lp->row_indices.clear();
  lp->values.clear();
  std::vector<size_t> order;
  for (size_t j = 0; j < n; ++j) {
    const auto b = static_cast<size_t>(start[j]), e = static_cast<size_t>(start[j + 1]);
    order.resize(e - b);
    std::iota(order.begin(), order.end(), b);
    std::sort(order.begin(), order.end(), [&](size_t x, size_t y) { return rows[x] < rows[y]; });
    for (size_t k : order) {
      if (std::fabs(vals[k]) < kZeroDrop) continue;
      lp->row_indices.push_back(rows[k]);
      lp->values.push_back(vals[k]);
    }
    lp->col_starts[j + 1] = static_cast<int32_t>(lp->row_indices.size());
  }
  const std::string problem = lp->validate();
  if (!problem.empty()) {
    *error = path_ + ": model failed validation: " + problem;
    return false;
  }
  info->name = name_;
  info->row_names = std::move(row_names_);
  info->col_names = std::move(col_names_);
  info->integer_columns = static_cast<int32_t>(std::count(integer_.begin(), integer_.end(), 1));
  info->fixed_format = fixed_;
  return true;
}

}  // namespace

// read_mps (mps_reader.cpp:662-690): free format first; fixed columns only when that fails
bool read_mps(const std::string& path, HostLp* lp, MpsInfo* info, std::string* error) {
  HostLp free_lp;
  MpsInfo free_info;
  std::string free_error;
  if (Parser(path, false).run(&free_lp, &free_info, &free_error)) {
    *lp = std::move(free_lp);
    *info = std::move(free_info);
    return true;
  }
  HostLp fixed_lp;
  MpsInfo fixed_info;
  std::string fixed_error;
  if (Parser(path, true).run(&fixed_lp, &fixed_info, &fixed_error)) {
    *lp = std::move(fixed_lp);
    *info = std::move(fixed_info);
    return true;
  }
  *error = free_error + " (the fixed-format reader also failed: " + fixed_error + ")";
  return false;
}

}  // namespace sk
