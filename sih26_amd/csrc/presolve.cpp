// Presolve reductions and postsolve; see presolve.hpp. Line references are to the reference's
// src/presolve/presolve.cpp unless another file is named.
#include "presolve.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <limits>
#include <thread>
#include <unordered_map>
#include <utility>

#include "threads.hpp"

namespace sk::presolve {
namespace {

constexpr double kInf = std::numeric_limits<double>::infinity();
// include/sankhya/tolerances.hpp:19, 22, 25, 115
constexpr double kPrimalFeasibility = 1e-7;
constexpr double kDualFeasibility = 1e-7;
constexpr double kIntegrality = 1e-6;
constexpr double kZeroDrop = 1e-11;
constexpr int kMaxPasses = 20;     // :244
constexpr int kMaxDualPasses = 8;  // :1374

// The reference's core holds true infinities only (types.hpp:27-45 normalises MPS-scale
// ones on input); the same normalisation is applied to the working bounds here.
double normalized(double v) { return is_finite_bound(v) ? v : (v > 0.0 ? kInf : -kInf); }
bool finite(double v) { return std::fabs(v) < kInf; }  // :84-86

// :103-106: on the bound to the precision asked of a solver at that scale.
bool at_bound(double value, double bound) {
  return finite(bound) && std::fabs(value - bound) <= kPrimalFeasibility * std::max(1.0, std::fabs(bound));
}
double round_integer_lower(double v) { return std::ceil(v - kIntegrality); }   // :115-117
double round_integer_upper(double v) { return std::floor(v + kIntegrality); }  // :118-120

uint64_t key(int32_t a, int32_t b) {
  return (static_cast<uint64_t>(static_cast<uint32_t>(a)) << 32) | static_cast<uint32_t>(b);
}

// f(begin, end) over [0, n) split into contiguous ranges, one host thread each (at most 16);
// small ranges run inline. Every element is handled by exactly one call, in order, so the
// arithmetic per element is the same whatever the split.
template <class F>
void parallel_ranges(int64_t n, F&& f) {
  const unsigned hw = std::thread::hardware_concurrency();
  const int64_t threads = std::min<int64_t>({16, hw ? static_cast<int64_t>(hw) : 1, n / 65536 + 1});
  if (threads <= 1) {
    f(int64_t{0}, n);
    return;
  }
  std::vector<std::thread> pool;
  JoinAll joiner{pool};
  for (int64_t t = 0; t < threads; ++t)
    pool.emplace_back([&, t]() { f(n * t / threads, n * (t + 1) / threads); });
}

// The rows of the working LP: the original entries of each row in ascending column order (the
// order :230-236 builds them in), then the entries doubleton fill-in appended, in that order.
struct RowStore {
  std::vector<int64_t> start;
  std::vector<int32_t> col;
  std::vector<double> val;
  std::unordered_map<int32_t, std::vector<std::pair<int32_t, double>>> fill;

  // The transpose of the CSC in parallel over column chunks: each chunk counts its entries per
  // row, a row's entries from chunk t go after those of the chunks before it, and each chunk
  // then places its own, so every row keeps ascending column order.
  void build(const HostLp& lp) {
    const int64_t m = lp.num_rows, n = lp.num_cols, z = lp.nnz();
    start.assign(static_cast<size_t>(m) + 1, 0);
    col.resize(static_cast<size_t>(z));
    val.resize(static_cast<size_t>(z));
    const int32_t* ri = lp.row_indices.data();
    const int32_t* cs = lp.col_starts.data();
    const unsigned hw = std::thread::hardware_concurrency();
    const int64_t chunks = std::min<int64_t>({16, hw ? static_cast<int64_t>(hw) : 1, z / 262144 + 1});
    std::vector<int32_t> count(static_cast<size_t>(chunks * m), 0);  // [chunk][row]
    auto col_of_chunk = [&](int64_t t) {  // chunk boundaries by nonzeros, on column starts
      const int64_t target = z * t / chunks;
      return static_cast<int64_t>(std::lower_bound(cs, cs + n, static_cast<int32_t>(target)) - cs);
    };
    std::vector<int64_t> c0(static_cast<size_t>(chunks) + 1);
    for (int64_t t = 0; t <= chunks; ++t) c0[static_cast<size_t>(t)] = t == chunks ? n : col_of_chunk(t);
    auto each_chunk = [&](auto&& f) {
      std::vector<std::thread> pool;
      JoinAll joiner{pool};
      for (int64_t t = 1; t < chunks; ++t) pool.emplace_back([&, t]() { f(t); });
      f(int64_t{0});
    };
    each_chunk([&](int64_t t) {
      int32_t* c = count.data() + t * m;
      for (int64_t p = cs[c0[static_cast<size_t>(t)]]; p < cs[c0[static_cast<size_t>(t) + 1]]; ++p) ++c[ri[p]];
    });
    // per row: each chunk's first slot within the row, and the row's length; then the starts
    parallel_ranges(m, [&](int64_t r0, int64_t r1) {
      for (int64_t i = r0; i < r1; ++i) {
        int32_t at = 0;
        for (int64_t t = 0; t < chunks; ++t) {
          int32_t& c = count[static_cast<size_t>(t * m + i)];
          const int32_t k = c;
          c = at;
          at += k;
        }
        start[static_cast<size_t>(i) + 1] = at;
      }
    });
    for (int64_t i = 0; i < m; ++i) start[static_cast<size_t>(i) + 1] += start[static_cast<size_t>(i)];
    each_chunk([&](int64_t t) {
      int32_t* next = count.data() + t * m;
      for (int64_t j = c0[static_cast<size_t>(t)]; j < c0[static_cast<size_t>(t) + 1]; ++j)
        for (int32_t p = cs[j]; p < cs[j + 1]; ++p) {
          const int32_t r = ri[p];
          const auto s = static_cast<size_t>(start[static_cast<size_t>(r)] + next[r]++);
          col[s] = static_cast<int32_t>(j);
          val[s] = lp.values[static_cast<size_t>(p)];
        }
    });
  }
  // f(column, coefficient&) over every entry of row r, dead columns included
  template <class F>
  void each(int32_t r, F&& f) {
    for (int64_t p = start[static_cast<size_t>(r)]; p < start[static_cast<size_t>(r) + 1]; ++p)
      f(col[static_cast<size_t>(p)], val[static_cast<size_t>(p)]);
    const auto it = fill.find(r);
    if (it != fill.end())
      for (auto& e : it->second) f(e.first, e.second);
  }
};

struct Activity {  // :24-29
  double lower = 0.0, upper = 0.0;
  bool lower_finite = true, upper_finite = true;
};

// :34-82
struct Work {
  const HostLp* lp = nullptr;
  std::vector<double> col_lower, col_upper, row_lower, row_upper, cost;
  std::vector<uint8_t> col_dead, row_dead, doubleton_touched, singleton_row_touched;
  std::vector<int32_t> col_count, row_count;
  RowStore rows;
  double offset_delta = 0.0;
  // the accumulated fill-in delta of (column, row), and the rows fill-in added a column to
  std::unordered_map<uint64_t, double> delta;
  std::unordered_map<int32_t, std::vector<int32_t>> new_rows;
  // A row examination reads its bounds and count, its coefficients, and its columns' bounds,
  // liveness and touched flags. `dirty` marks a row one of those changed for since it was last
  // examined: an examination of a row that is not dirty would repeat the last one, which
  // removed nothing, so it is skipped. `stale` marks a row whose precomputed activity bounds
  // (taken at the start of the row pass) no longer hold. `filled` marks a column fill-in
  // changed or created a coefficient of.
  std::vector<uint8_t> dirty, stale, filled;
  std::vector<Activity> activity;

  void touch_column(int32_t c) {
    for (int32_t p = lp->col_starts[static_cast<size_t>(c)]; p < lp->col_starts[static_cast<size_t>(c) + 1]; ++p) {
      const auto r = static_cast<size_t>(lp->row_indices[static_cast<size_t>(p)]);
      dirty[r] = stale[r] = 1;
    }
    if (const std::vector<int32_t>* extra = new_rows_of(c))
      for (const int32_t r : *extra) dirty[static_cast<size_t>(r)] = stale[static_cast<size_t>(r)] = 1;
  }

  const double* find_delta(int32_t column, int32_t row) const {
    const auto it = delta.find(key(column, row));
    return it == delta.end() ? nullptr : &it->second;
  }
  const std::vector<int32_t>* new_rows_of(int32_t column) const {
    const auto it = new_rows.find(column);
    return it == new_rows.end() ? nullptr : &it->second;
  }
};

// :122-145: the activity limits the live columns' bounds allow, summed in row order.
Activity activity_bounds(Work& w, int32_t row) {
  Activity b;
  w.rows.each(row, [&](int32_t c, double& a) {
    const auto u = static_cast<size_t>(c);
    if (w.col_dead[u]) return;
    const double low = a > 0.0 ? w.col_lower[u] : w.col_upper[u];
    const double high = a > 0.0 ? w.col_upper[u] : w.col_lower[u];
    if (finite(low)) b.lower += a * low;
    else b.lower_finite = false;
    if (finite(high)) b.upper += a * high;
    else b.upper_finite = false;
  });
  return b;
}

// :161-188: drop a fixed column from its rows, its value moved into their bounds. The
// coefficients are the original ones patched by fill-in, then the rows fill-in added it to.
void fold_fixed_column(Work& w, int32_t column, double value) {
  const HostLp& lp = *w.lp;
  w.touch_column(column);
  for (int32_t p = lp.col_starts[static_cast<size_t>(column)]; p < lp.col_starts[static_cast<size_t>(column) + 1]; ++p) {
    const int32_t row = lp.row_indices[static_cast<size_t>(p)];
    const auto r = static_cast<size_t>(row);
    if (w.row_dead[r]) continue;
    double coefficient = lp.values[static_cast<size_t>(p)];
    if (const double* d = w.find_delta(column, row)) coefficient += *d;
    const double shift = coefficient * value;
    if (finite(w.row_lower[r])) w.row_lower[r] -= shift;
    if (finite(w.row_upper[r])) w.row_upper[r] -= shift;
    --w.row_count[r];
  }
  if (const std::vector<int32_t>* extra = w.new_rows_of(column)) {
    for (const int32_t row : *extra) {
      const auto r = static_cast<size_t>(row);
      if (w.row_dead[r]) continue;
      const double* d = w.find_delta(column, row);
      if (!d) continue;
      const double shift = *d * value;
      if (finite(w.row_lower[r])) w.row_lower[r] -= shift;
      if (finite(w.row_upper[r])) w.row_upper[r] -= shift;
      --w.row_count[r];
    }
  }
  w.col_dead[static_cast<size_t>(column)] = 1;
}

void kill_row(Work& w, int32_t row) {  // :190-198
  const auto r = static_cast<size_t>(row);
  if (w.row_dead[r]) return;
  w.row_dead[r] = 1;
  w.rows.each(row, [&](int32_t c, double&) {
    if (!w.col_dead[static_cast<size_t>(c)]) --w.col_count[static_cast<size_t>(c)];
  });
}

std::string format(const char* fmt, int32_t index, double a, double b) {
  char buf[256];
  std::snprintf(buf, sizeof buf, fmt, index, a, b);
  return buf;
}

}  // namespace

Result run(const HostLp& lp, const std::vector<uint8_t>& is_integer, double feasibility) {
  Result result;
  const int32_t m = lp.num_rows, n = lp.num_cols;
  result.original_rows = m;
  result.original_cols = n;
  result.original_nonzeros = lp.nnz();
  const double sense = static_cast<double>(lp.sense);
  auto integer = [&](int32_t j) { return !is_integer.empty() && is_integer[static_cast<size_t>(j)] != 0; };

  Work w;
  w.lp = &lp;
  auto norm_all = [](const std::vector<double>& v) {
    std::vector<double> out(v.size());
    std::transform(v.begin(), v.end(), out.begin(), normalized);
    return out;
  };
  w.col_lower = norm_all(lp.col_lower);
  w.col_upper = norm_all(lp.col_upper);
  w.row_lower = norm_all(lp.row_lower);
  w.row_upper = norm_all(lp.row_upper);
  w.cost = lp.cost;
  w.col_dead.assign(static_cast<size_t>(n), 0);
  w.row_dead.assign(static_cast<size_t>(m), 0);
  w.doubleton_touched.assign(static_cast<size_t>(n), 0);
  w.singleton_row_touched.assign(static_cast<size_t>(n), 0);
  w.col_count.resize(static_cast<size_t>(n));
  for (int32_t j = 0; j < n; ++j)
    w.col_count[static_cast<size_t>(j)] = lp.col_starts[static_cast<size_t>(j) + 1] - lp.col_starts[static_cast<size_t>(j)];
  w.rows.build(lp);
  w.row_count.resize(static_cast<size_t>(m));
  for (int32_t i = 0; i < m; ++i)
    w.row_count[static_cast<size_t>(i)] = static_cast<int32_t>(w.rows.start[static_cast<size_t>(i) + 1] - w.rows.start[static_cast<size_t>(i)]);
  w.dirty.assign(static_cast<size_t>(m), 1);
  w.stale.assign(static_cast<size_t>(m), 0);
  w.filled.assign(static_cast<size_t>(n), 0);
  w.activity.resize(static_cast<size_t>(m));

  auto infeasible = [&](std::string why) {
    result.proved_infeasible = true;
    result.message = std::move(why);
  };
  auto push = [&](const Record& rec) { result.records.push_back(rec); };

  int pass = 0;
  for (; pass < kMaxPasses && !result.proved_infeasible; ++pass) {
    bool changed = false;

    // ---- columns (:250-410)
    for (int32_t j = 0; j < n && !result.proved_infeasible; ++j) {
      const auto u = static_cast<size_t>(j);
      if (w.col_dead[u]) continue;
      if (w.col_lower[u] > w.col_upper[u] + feasibility) {
        infeasible(format("column %d has crossed bounds after presolve: [%.6g, %.6g]", j,
                          w.col_lower[u], w.col_upper[u]));
        break;
      }
      // fixed column (:262-274)
      if (w.col_upper[u] - w.col_lower[u] <= feasibility && finite(w.col_lower[u])) {
        const double value = w.col_lower[u];
        Record rec;
        rec.kind = Kind::kFixedColumn;
        rec.index = j;
        rec.value = value;
        push(rec);
        fold_fixed_column(w, j, value);
        changed = true;
        continue;
      }
      // empty column (:276-310): parked where the cost prefers; left alone when that is an
      // infinite bound (presolve cannot conclude unboundedness)
      if (w.col_count[u] == 0) {
        const double c = sense * w.cost[u];
        double value;
        if (c > 0.0) value = w.col_lower[u];
        else if (c < 0.0) value = w.col_upper[u];
        else value = finite(w.col_lower[u]) ? w.col_lower[u] : (finite(w.col_upper[u]) ? w.col_upper[u] : 0.0);
        if (!finite(value)) continue;
        Record rec;
        rec.kind = Kind::kEmptyColumn;
        rec.index = j;
        rec.value = value;
        push(rec);
        w.col_dead[u] = 1;
        changed = true;
        continue;
      }
      // free column singleton (:312-409)
      if (w.col_count[u] == 1 && !finite(w.col_lower[u]) && !finite(w.col_upper[u])) {
        int32_t row = -1;
        double a = 0.0;
        for (int32_t p = lp.col_starts[u]; p < lp.col_starts[u + 1]; ++p) {
          const int32_t rr = lp.row_indices[static_cast<size_t>(p)];
          if (w.row_dead[static_cast<size_t>(rr)]) continue;
          row = rr;
          a = lp.values[static_cast<size_t>(p)];
          if (const double* d = w.find_delta(j, rr)) a += *d;
          break;
        }
        if (row < 0) {
          if (const std::vector<int32_t>* extra = w.new_rows_of(j)) {
            for (const int32_t cand : *extra) {
              if (w.row_dead[static_cast<size_t>(cand)]) continue;
              row = cand;
              a = *w.find_delta(j, cand);
              break;
            }
          }
        }
        if (row < 0 || std::fabs(a) < kZeroDrop) continue;
        const auto r = static_cast<size_t>(row);
        const double cj = w.cost[u];
        const double rate = sense * cj / a;
        double target;
        if (rate > kZeroDrop) {
          if (!finite(w.row_lower[r])) continue;
          target = w.row_lower[r];
        } else if (rate < -kZeroDrop) {
          if (!finite(w.row_upper[r])) continue;
          target = w.row_upper[r];
        } else {
          target = finite(w.row_lower[r]) ? w.row_lower[r] : (finite(w.row_upper[r]) ? w.row_upper[r] : 0.0);
        }
        Record rec;
        rec.kind = Kind::kFreeColumnSingleton;
        rec.index = row;
        rec.column = j;
        rec.coefficient = a;
        rec.substituted_rhs = target;
        rec.eliminated_cost = cj;
        push(rec);
        // the column's cost moves onto every other live column of the row (:398-406)
        w.rows.each(row, [&](int32_t other, double& coef) {
          if (other == j || w.col_dead[static_cast<size_t>(other)]) return;
          w.cost[static_cast<size_t>(other)] -= cj * (coef / a);
        });
        w.offset_delta += cj * target / a;
        w.col_dead[u] = 1;
        kill_row(w, row);
        changed = true;
        continue;
      }
    }

    // ---- rows (:412-714). The activity bounds of every dirty row, from the state at the
    // start of the row pass, in parallel; a row a reduction of this pass touches before its
    // turn is recomputed then, so each row sees the state the reference's sequential pass has.
    parallel_ranges(m, [&](int64_t r0, int64_t r1) {
      for (int64_t i = r0; i < r1; ++i) {
        const auto r = static_cast<size_t>(i);
        if (w.row_dead[r] || !w.dirty[r]) continue;
        w.activity[r] = activity_bounds(w, static_cast<int32_t>(i));
        w.stale[r] = 0;
      }
    });
    for (int32_t i = 0; i < m && !result.proved_infeasible; ++i) {
      const auto r = static_cast<size_t>(i);
      if (w.row_dead[r] || !w.dirty[r]) continue;
      w.dirty[r] = 0;
      if (w.row_lower[r] > w.row_upper[r] + feasibility) {
        infeasible(format("row %d has crossed bounds after presolve: [%.6g, %.6g]", i,
                          w.row_lower[r], w.row_upper[r]));
        break;
      }
      // empty row (:423-442)
      if (w.row_count[r] == 0) {
        const bool violates_lower = finite(w.row_lower[r]) && w.row_lower[r] > feasibility;
        const bool violates_upper = finite(w.row_upper[r]) && w.row_upper[r] < -feasibility;
        if (violates_lower || violates_upper) {
          infeasible(format("row %d has no entries but requires activity in [%.6g, %.6g], which "
                            "excludes zero", i, w.row_lower[r], w.row_upper[r]));
          break;
        }
        Record rec;
        rec.kind = Kind::kEmptyRow;
        rec.index = i;
        push(rec);
        kill_row(w, i);
        changed = true;
        continue;
      }
      const Activity b = w.stale[r] ? activity_bounds(w, i) : w.activity[r];
      w.stale[r] = 0;
      // redundant row (:446-460)
      const bool lower_slack = !finite(w.row_lower[r]) || (b.lower_finite && b.lower >= w.row_lower[r] - feasibility);
      const bool upper_slack = !finite(w.row_upper[r]) || (b.upper_finite && b.upper <= w.row_upper[r] + feasibility);
      if (lower_slack && upper_slack) {
        Record rec;
        rec.kind = Kind::kRedundantRow;
        rec.index = i;
        push(rec);
        kill_row(w, i);
        changed = true;
        continue;
      }
      // infeasible by activity (:462-477)
      if (b.upper_finite && finite(w.row_lower[r]) && b.upper < w.row_lower[r] - feasibility) {
        infeasible(format("row %d needs activity of at least %.6g but the column bounds cap it at %.6g",
                          i, w.row_lower[r], b.upper));
        break;
      }
      if (b.lower_finite && finite(w.row_upper[r]) && b.lower > w.row_upper[r] + feasibility) {
        infeasible(format("row %d allows activity of at most %.6g but the column bounds force at least %.6g",
                          i, w.row_upper[r], b.lower));
        break;
      }
      // singleton row (:479-547): a bound on its one live column
      if (w.row_count[r] == 1) {
        int32_t column = -1;
        double coefficient = 0.0;
        w.rows.each(i, [&](int32_t c, double& a) {
          if (column < 0 && !w.col_dead[static_cast<size_t>(c)]) {
            column = c;
            coefficient = a;
          }
        });
        if (column < 0 || std::fabs(coefficient) < kZeroDrop) continue;
        const auto c = static_cast<size_t>(column);
        if (w.doubleton_touched[c]) continue;
        double lo = -kInf, hi = kInf;
        if (coefficient > 0.0) {
          if (finite(w.row_lower[r])) lo = w.row_lower[r] / coefficient;
          if (finite(w.row_upper[r])) hi = w.row_upper[r] / coefficient;
        } else {
          if (finite(w.row_upper[r])) lo = w.row_upper[r] / coefficient;
          if (finite(w.row_lower[r])) hi = w.row_lower[r] / coefficient;
        }
        if (integer(column)) {
          if (finite(lo)) lo = round_integer_lower(lo);
          if (finite(hi)) hi = round_integer_upper(hi);
        }
        Record rec;
        rec.kind = Kind::kSingletonRow;
        rec.index = i;
        rec.column = column;
        rec.coefficient = coefficient;
        rec.row_lower = w.row_lower[r];
        rec.row_upper = w.row_upper[r];
        push(rec);
        w.singleton_row_touched[c] = 1;
        w.touch_column(column);
        if (finite(lo)) w.col_lower[c] = std::max(w.col_lower[c], lo);
        if (finite(hi)) w.col_upper[c] = std::min(w.col_upper[c], hi);
        kill_row(w, i);
        changed = true;
        continue;
      }
      // doubleton equation (:549-711): one column solved for in terms of the other
      if (w.row_count[r] == 2 && w.row_upper[r] - w.row_lower[r] <= feasibility && finite(w.row_lower[r])) {
        int32_t first = -1, second = -1;
        double first_coeff = 0.0, second_coeff = 0.0;
        w.rows.each(i, [&](int32_t c, double& a) {
          if (w.col_dead[static_cast<size_t>(c)]) return;
          if (first < 0) {
            first = c;
            first_coeff = a;
          } else {
            second = c;
            second_coeff = a;
          }
        });
        if (first < 0 || second < 0) continue;
        int32_t elim = first, keep = second;
        double a = first_coeff, bk = second_coeff;
        if (integer(first)) {
          if (integer(second)) continue;
          elim = second;
          keep = first;
          a = second_coeff;
          bk = first_coeff;
        }
        if (std::fabs(a) < kZeroDrop || std::fabs(bk) < kZeroDrop) continue;
        const auto ue = static_cast<size_t>(elim), uk = static_cast<size_t>(keep);
        if (w.doubleton_touched[ue] || w.doubleton_touched[uk]) continue;
        if (w.singleton_row_touched[ue] || w.singleton_row_touched[uk]) continue;
        const double rhs = w.row_lower[r];
        // the range of a x_elim, then of b x_keep = rhs - a x_elim (:608-649)
        const double lo_e = w.col_lower[ue], hi_e = w.col_upper[ue];
        double term_lo = 0.0, term_hi = 0.0;
        bool term_lo_finite = false, term_hi_finite = false;
        if (a > 0.0) {
          if (finite(lo_e)) { term_lo = a * lo_e; term_lo_finite = true; }
          if (finite(hi_e)) { term_hi = a * hi_e; term_hi_finite = true; }
        } else {
          if (finite(hi_e)) { term_lo = a * hi_e; term_lo_finite = true; }
          if (finite(lo_e)) { term_hi = a * lo_e; term_hi_finite = true; }
        }
        const double bxk_lo = term_hi_finite ? rhs - term_hi : -kInf;
        const double bxk_hi = term_lo_finite ? rhs - term_lo : kInf;
        double k_lo = -kInf, k_hi = kInf;
        if (bk > 0.0) {
          if (term_hi_finite) k_lo = bxk_lo / bk;
          if (term_lo_finite) k_hi = bxk_hi / bk;
        } else {
          if (term_lo_finite) k_lo = bxk_hi / bk;
          if (term_hi_finite) k_hi = bxk_lo / bk;
        }
        if (integer(keep)) {
          if (finite(k_lo)) k_lo = round_integer_lower(k_lo);
          if (finite(k_hi)) k_hi = round_integer_upper(k_hi);
        }
        Record rec;
        rec.kind = Kind::kDoubletonEquation;
        rec.index = i;
        rec.column = elim;
        rec.coefficient = a;
        rec.partner_column = keep;
        rec.partner_coefficient = bk;
        rec.substituted_rhs = rhs;
        rec.eliminated_cost = w.cost[ue];
        push(rec);
        w.doubleton_touched[ue] = w.doubleton_touched[uk] = 1;
        w.touch_column(elim);
        w.touch_column(keep);
        if (finite(k_lo)) w.col_lower[uk] = std::max(w.col_lower[uk], k_lo);
        if (finite(k_hi)) w.col_upper[uk] = std::min(w.col_upper[uk], k_hi);
        const double ce = w.cost[ue];
        w.cost[uk] -= ce * (bk / a);
        w.offset_delta += ce * rhs / a;
        // fill-in: every other live row of the eliminated column loses it and gains (or
        // adjusts) a term on the kept column (:674-704)
        for (int32_t p = lp.col_starts[ue]; p < lp.col_starts[ue + 1]; ++p) {
          const int32_t other = lp.row_indices[static_cast<size_t>(p)];
          if (other == i) continue;
          const auto o = static_cast<size_t>(other);
          if (w.row_dead[o]) continue;
          const double a_other = lp.values[static_cast<size_t>(p)];
          if (std::fabs(a_other) < kZeroDrop) continue;
          const double factor = a_other / a;
          if (finite(w.row_lower[o])) w.row_lower[o] -= factor * rhs;
          if (finite(w.row_upper[o])) w.row_upper[o] -= factor * rhs;
          const double delta_keep = -factor * bk;
          bool found = false;
          w.rows.each(other, [&](int32_t c, double& coef) {
            if (!found && c == keep) {
              coef += delta_keep;
              found = true;
            }
          });
          w.delta[key(keep, other)] += delta_keep;
          w.filled[uk] = 1;
          if (!found) {
            w.rows.fill[other].emplace_back(keep, delta_keep);
            w.new_rows[keep].push_back(other);
            ++w.row_count[o];
            ++w.col_count[uk];
          }
          --w.row_count[o];
        }
        w.col_dead[ue] = 1;
        kill_row(w, i);
        changed = true;
      }
    }
    if (!changed) {
      ++pass;
      break;
    }
  }
  result.passes = pass;
  if (result.proved_infeasible) return result;

  // ---- compaction (:727-803)
  std::vector<int32_t> new_row(static_cast<size_t>(m), -1);
  for (int32_t j = 0; j < n; ++j)
    if (!w.col_dead[static_cast<size_t>(j)]) result.col_to_original.push_back(j);
  for (int32_t i = 0; i < m; ++i)
    if (!w.row_dead[static_cast<size_t>(i)]) {
      new_row[static_cast<size_t>(i)] = static_cast<int32_t>(result.row_to_original.size());
      result.row_to_original.push_back(i);
    }
  HostLp& red = result.reduced;
  red.sense = lp.sense;
  red.num_cols = static_cast<int32_t>(result.col_to_original.size());
  red.num_rows = static_cast<int32_t>(result.row_to_original.size());
  red.cost.resize(result.col_to_original.size());
  red.col_lower.resize(result.col_to_original.size());
  red.col_upper.resize(result.col_to_original.size());
  for (size_t jr = 0; jr < result.col_to_original.size(); ++jr) {
    const auto u = static_cast<size_t>(result.col_to_original[jr]);
    red.cost[jr] = w.cost[u];
    red.col_lower[jr] = w.col_lower[u];
    red.col_upper[jr] = w.col_upper[u];
  }
  red.row_lower.resize(result.row_to_original.size());
  red.row_upper.resize(result.row_to_original.size());
  for (size_t ir = 0; ir < result.row_to_original.size(); ++ir) {
    red.row_lower[ir] = w.row_lower[static_cast<size_t>(result.row_to_original[ir])];
    red.row_upper[ir] = w.row_upper[static_cast<size_t>(result.row_to_original[ir])];
  }
  // the objective constant of the removed columns, at their post-fold costs (:773-786)
  red.objective_offset = lp.objective_offset;
  for (const Record& rec : result.records)
    if (rec.kind == Kind::kFixedColumn || rec.kind == Kind::kEmptyColumn)
      red.objective_offset += w.cost[static_cast<size_t>(rec.index)] * rec.value;
  red.objective_offset += w.offset_delta;

  // The reduced matrix as SparseMatrix::finalize (sparse.cpp:55-130) lays out the working
  // rows (fill-in included): columns in order, rows ascending, |a| < 1e-11 dropped. A column
  // fill-in never changed has exactly its original coefficients, so it is copied from the
  // CSC with its dead rows dropped (in parallel over columns); a column fill-in changed takes
  // its coefficients from the working rows.
  const auto rn = static_cast<size_t>(red.num_cols);
  std::unordered_map<int64_t, std::vector<std::pair<int32_t, double>>> filled_cols;
  for (size_t jr = 0; jr < rn; ++jr)
    if (w.filled[static_cast<size_t>(result.col_to_original[jr])]) filled_cols[static_cast<int64_t>(jr)];
  std::vector<int32_t> cnt(rn + 1, 0);
  parallel_ranges(static_cast<int64_t>(rn), [&](int64_t c0, int64_t c1) {
    for (int64_t jr = c0; jr < c1; ++jr) {
      const int32_t j = result.col_to_original[static_cast<size_t>(jr)];
      const auto u = static_cast<size_t>(j);
      int32_t k = 0;
      if (!w.filled[u]) {
        for (int32_t p = lp.col_starts[u]; p < lp.col_starts[u + 1]; ++p)
          if (!w.row_dead[static_cast<size_t>(lp.row_indices[static_cast<size_t>(p)])] &&
              !(std::fabs(lp.values[static_cast<size_t>(p)]) < kZeroDrop))
            ++k;
      } else {
        auto& out = filled_cols.at(jr);
        auto take = [&](int32_t row) {
          if (w.row_dead[static_cast<size_t>(row)]) return;
          bool seen = false;
          double a = 0.0;
          w.rows.each(row, [&](int32_t c, double& v) {
            if (!seen && c == j) {
              seen = true;
              a = v;
            }
          });
          if (seen && !(std::fabs(a) < kZeroDrop)) out.emplace_back(new_row[static_cast<size_t>(row)], a);
        };
        for (int32_t p = lp.col_starts[u]; p < lp.col_starts[u + 1]; ++p) take(lp.row_indices[static_cast<size_t>(p)]);
        if (const std::vector<int32_t>* extra = w.new_rows_of(j))
          for (const int32_t row : *extra) take(row);
        std::sort(out.begin(), out.end());
        k = static_cast<int32_t>(out.size());
      }
      cnt[static_cast<size_t>(jr) + 1] = k;
    }
  });
  for (size_t j = 0; j < rn; ++j) cnt[j + 1] += cnt[j];
  red.col_starts = cnt;
  red.row_indices.resize(static_cast<size_t>(cnt.back()));
  red.values.resize(static_cast<size_t>(cnt.back()));
  parallel_ranges(static_cast<int64_t>(rn), [&](int64_t c0, int64_t c1) {
    for (int64_t jr = c0; jr < c1; ++jr) {
      const auto u = static_cast<size_t>(result.col_to_original[static_cast<size_t>(jr)]);
      auto s = static_cast<size_t>(cnt[static_cast<size_t>(jr)]);
      if (!w.filled[u]) {
        for (int32_t p = lp.col_starts[u]; p < lp.col_starts[u + 1]; ++p) {
          const int32_t row = lp.row_indices[static_cast<size_t>(p)];
          const double a = lp.values[static_cast<size_t>(p)];
          if (w.row_dead[static_cast<size_t>(row)] || std::fabs(a) < kZeroDrop) continue;
          red.row_indices[s] = new_row[static_cast<size_t>(row)];
          red.values[s++] = a;
        }
      } else {
        for (const auto& [row, a] : filled_cols.at(jr)) {
          red.row_indices[s] = row;
          red.values[s++] = a;
        }
      }
    }
  });
  return result;
}

// ============================================================================================

void postsolve(const Result& R, const HostLp& orig, const double* xr, const double* yr,
               const double* dr, double* x, double* y, double* d) {
  const int32_t n = orig.num_cols, m = orig.num_rows;
  const size_t N = static_cast<size_t>(n), M = static_cast<size_t>(m), NR = R.records.size();
  const double sense = static_cast<double>(orig.sense);
  std::fill(x, x + N, 0.0);
  std::fill(d, d + N, 0.0);
  std::fill(y, y + M, 0.0);
  // the reduced point, scattered into the original numbering (:824-856)
  for (size_t j = 0; j < R.col_to_original.size(); ++j) {
    const auto t = static_cast<size_t>(R.col_to_original[j]);
    x[t] = xr[j];
    d[t] = dr[j];
  }
  for (size_t i = 0; i < R.row_to_original.size(); ++i) y[static_cast<size_t>(R.row_to_original[i])] = yr[i];

  // which record removed each column and each row (:866-904)
  std::vector<size_t> column_removed_at(N, NR), row_removed_at(M, NR);
  for (size_t p = 0; p < NR; ++p) {
    const Record& rec = R.records[p];
    switch (rec.kind) {
      case Kind::kFixedColumn:
      case Kind::kEmptyColumn: column_removed_at[static_cast<size_t>(rec.index)] = p; break;
      case Kind::kFreeColumnSingleton:
      case Kind::kDoubletonEquation: column_removed_at[static_cast<size_t>(rec.column)] = p; break;
      default: break;
    }
    switch (rec.kind) {
      case Kind::kFixedColumn:
      case Kind::kEmptyColumn: break;
      default: row_removed_at[static_cast<size_t>(rec.index)] = p; break;
    }
  }

  // The rows of the original matrix postsolve reads, the removed ones (as CsrView::build
  // gives them: columns ascending), gathered in one sweep of the CSC.
  std::vector<int32_t> slot(M, -1);
  int32_t needed = 0;
  for (size_t i = 0; i < M; ++i)
    if (row_removed_at[i] < NR) slot[i] = needed++;
  std::vector<std::vector<std::pair<int32_t, double>>> rowv(static_cast<size_t>(needed));
  if (needed > 0)
    for (int32_t j = 0; j < n; ++j)
      for (int32_t p = orig.col_starts[static_cast<size_t>(j)]; p < orig.col_starts[static_cast<size_t>(j) + 1]; ++p) {
        const int32_t sl = slot[static_cast<size_t>(orig.row_indices[static_cast<size_t>(p)])];
        if (sl >= 0) rowv[static_cast<size_t>(sl)].emplace_back(j, orig.values[static_cast<size_t>(p)]);
      }
  auto original_row = [&](int32_t row) -> const std::vector<std::pair<int32_t, double>>& {
    return rowv[static_cast<size_t>(slot[static_cast<size_t>(row)])];
  };
  auto original_at = [&](int32_t row, int32_t col) {  // SparseMatrix::at
    const int32_t* b = orig.row_indices.data() + orig.col_starts[static_cast<size_t>(col)];
    const int32_t* e = orig.row_indices.data() + orig.col_starts[static_cast<size_t>(col) + 1];
    const int32_t* it = std::lower_bound(b, e, row);
    return (it != e && *it == row) ? orig.values[static_cast<size_t>(it - orig.row_indices.data())] : 0.0;
  };

  // The folds replayed forwards (:906-1051): the post-fold costs, the folded rows and the
  // columns that received each fold, and the coefficients fill-in changed or created.
  std::vector<double> adjusted_cost = orig.cost;
  std::vector<uint8_t> row_is_folded(M, 0);
  std::unordered_map<int32_t, std::vector<int32_t>> recipients;  // folded row -> columns
  std::unordered_map<uint64_t, double> overrides;                // (row, col) -> fill-in delta
  std::unordered_map<int32_t, std::vector<int32_t>> extra_rows_for_column, extra_columns_for_row;
  auto find_override = [&](int32_t row, int32_t col) -> const double* {
    const auto it = overrides.find(key(row, col));
    return it == overrides.end() ? nullptr : &it->second;
  };
  auto list_of = [](const std::unordered_map<int32_t, std::vector<int32_t>>& mp, int32_t k) -> const std::vector<int32_t>* {
    const auto it = mp.find(k);
    return it == mp.end() ? nullptr : &it->second;
  };
  for (size_t p = 0; p < NR; ++p) {
    const Record& rec = R.records[p];
    if (rec.kind != Kind::kFreeColumnSingleton && rec.kind != Kind::kDoubletonEquation) continue;
    row_is_folded[static_cast<size_t>(rec.index)] = 1;
    std::vector<int32_t>& got = recipients[rec.index];
    if (rec.kind == Kind::kDoubletonEquation) {
      adjusted_cost[static_cast<size_t>(rec.partner_column)] -= rec.eliminated_cost * (rec.partner_coefficient / rec.coefficient);
      got.push_back(rec.partner_column);
      const auto ce = static_cast<size_t>(rec.column);
      for (int32_t q = orig.col_starts[ce]; q < orig.col_starts[ce + 1]; ++q) {
        const int32_t other = orig.row_indices[static_cast<size_t>(q)];
        if (other == rec.index) continue;
        if (row_removed_at[static_cast<size_t>(other)] < p) continue;
        const double a_other = orig.values[static_cast<size_t>(q)];
        if (std::fabs(a_other) < kZeroDrop) continue;
        const double delta = -(a_other / rec.coefficient) * rec.partner_coefficient;
        const uint64_t k = key(other, rec.partner_column);
        const bool is_new_key = overrides.find(k) == overrides.end();
        overrides[k] += delta;
        if (is_new_key && original_at(other, rec.partner_column) == 0.0) {
          extra_rows_for_column[rec.partner_column].push_back(other);
          extra_columns_for_row[other].push_back(rec.partner_column);
        }
      }
    } else {
      for (const auto& [col, value] : original_row(rec.index)) {
        if (col == rec.column) continue;
        if (column_removed_at[static_cast<size_t>(col)] < p) continue;
        double coefficient = value;
        if (const double* o = find_override(rec.index, col)) coefficient += *o;
        adjusted_cost[static_cast<size_t>(col)] -= rec.eliminated_cost * (coefficient / rec.coefficient);
        got.push_back(col);
      }
      if (const std::vector<int32_t>* extra = list_of(extra_columns_for_row, rec.index)) {
        for (const int32_t col : *extra) {
          if (col == rec.column) continue;
          if (column_removed_at[static_cast<size_t>(col)] < p) continue;
          const double* o = find_override(rec.index, col);
          if (!o) continue;
          adjusted_cost[static_cast<size_t>(col)] -= rec.eliminated_cost * (*o / rec.coefficient);
          got.push_back(col);
        }
      }
    }
  }
  auto is_recipient = [&](int32_t row, int32_t col) {
    const std::vector<int32_t>* g = list_of(recipients, row);
    return g && std::find(g->begin(), g->end(), col) != g->end();
  };

  // The primal values, records replayed backwards (:1053-1120). Row duals of removed rows
  // start at zero; the dual passes below price them.
  for (size_t idx = NR; idx-- > 0;) {
    const Record& rec = R.records[idx];
    switch (rec.kind) {
      case Kind::kFixedColumn:
      case Kind::kEmptyColumn: x[static_cast<size_t>(rec.index)] = rec.value; break;
      case Kind::kEmptyRow:
      case Kind::kRedundantRow:
      case Kind::kForcingRow:
      case Kind::kSingletonRow: y[static_cast<size_t>(rec.index)] = 0.0; break;
      case Kind::kFreeColumnSingleton: {
        // x_j = (target - the other columns still live when it fired) / a
        double sum = 0.0;
        for (const auto& [col, value] : original_row(rec.index)) {
          if (col == rec.column) continue;
          if (column_removed_at[static_cast<size_t>(col)] < idx) continue;
          double coefficient = value;
          if (const double* o = find_override(rec.index, col)) coefficient += *o;
          sum += coefficient * x[static_cast<size_t>(col)];
        }
        if (const std::vector<int32_t>* extra = list_of(extra_columns_for_row, rec.index)) {
          for (const int32_t col : *extra) {
            if (col == rec.column) continue;
            if (column_removed_at[static_cast<size_t>(col)] < idx) continue;
            const double* o = find_override(rec.index, col);
            if (!o) continue;
            sum += *o * x[static_cast<size_t>(col)];
          }
        }
        x[static_cast<size_t>(rec.column)] = (rec.substituted_rhs - sum) / rec.coefficient;
        y[static_cast<size_t>(rec.index)] = 0.0;
        break;
      }
      case Kind::kDoubletonEquation: {
        const double xk = x[static_cast<size_t>(rec.partner_column)];
        x[static_cast<size_t>(rec.column)] = (rec.substituted_rhs - rec.partner_coefficient * xk) / rec.coefficient;
        y[static_cast<size_t>(rec.index)] = 0.0;
        break;
      }
    }
  }

  // ---- the duals (:1122-1637)
  auto effective_coefficient = [&](int32_t row, int32_t col, double value) {
    const double* o = find_override(row, col);
    return o ? value + *o : value;
  };
  // c - A^T y with the post-fold cost, skipping the folded rows whose fold the column received
  auto reduced_cost_of = [&](int32_t col) {
    const auto u = static_cast<size_t>(col);
    double dc = adjusted_cost[u];
    for (int32_t k = orig.col_starts[u]; k < orig.col_starts[u + 1]; ++k) {
      const int32_t row = orig.row_indices[static_cast<size_t>(k)];
      if (row_is_folded[static_cast<size_t>(row)] && is_recipient(row, col)) continue;
      const double coefficient = effective_coefficient(row, col, orig.values[static_cast<size_t>(k)]);
      dc -= coefficient * y[static_cast<size_t>(row)];
    }
    if (const std::vector<int32_t>* extra = list_of(extra_rows_for_column, col)) {
      for (const int32_t row : *extra) {
        if (row_is_folded[static_cast<size_t>(row)]) continue;
        const double coefficient = effective_coefficient(row, col, 0.0);
        dc -= coefficient * y[static_cast<size_t>(row)];
      }
    }
    return dc;
  };
  std::vector<uint8_t> dual_finalized(N, 0), is_doubleton_participant(N, 0);
  for (const Record& rec : R.records)
    if (rec.kind == Kind::kDoubletonEquation) {
      is_doubleton_participant[static_cast<size_t>(rec.column)] = 1;
      is_doubleton_participant[static_cast<size_t>(rec.partner_column)] = 1;
    }
  // a column with an original entry in a folded row whose fold it did not receive
  auto depends_on_unresolved_fold = [&](int32_t col) {
    const auto u = static_cast<size_t>(col);
    for (int32_t k = orig.col_starts[u]; k < orig.col_starts[u + 1]; ++k) {
      const int32_t row = orig.row_indices[static_cast<size_t>(k)];
      if (row_is_folded[static_cast<size_t>(row)] && !is_recipient(row, col)) return true;
    }
    return false;
  };
  auto settle = [&](size_t c, double dc) {  // the column's reduced cost stands as computed
    if (!dual_finalized[c] && !is_doubleton_participant[c]) {
      d[c] = dc;
      dual_finalized[c] = 1;
    }
  };
  // :1232-1322: price a singleton row only when its column needs it, the row is active and
  // the price has the sign the row's active side admits
  auto process_singleton_row = [&](const Record& it) {
    const auto c = static_cast<size_t>(it.column);
    y[static_cast<size_t>(it.index)] = 0.0;
    const double xv = x[c];
    const bool at_lower = at_bound(xv, orig.col_lower[c]);
    const bool at_upper = at_bound(xv, orig.col_upper[c]);
    const double dc = reduced_cost_of(it.column);
    const double signed_d = sense * dc;
    const bool needs_price = (!at_lower && !at_upper) || (at_lower && !at_upper && signed_d < -kDualFeasibility) ||
                             (at_upper && !at_lower && signed_d > kDualFeasibility);
    if (!needs_price) return settle(c, dc);
    const double activity = it.coefficient * xv;
    const bool row_at_lower = at_bound(activity, it.row_lower);
    const bool row_at_upper = at_bound(activity, it.row_upper);
    if (!row_at_lower && !row_at_upper) return settle(c, dc);
    const double candidate = dc / it.coefficient;
    const double signed_candidate = sense * candidate;
    if (row_at_lower && !row_at_upper && signed_candidate < -kDualFeasibility) return settle(c, dc);
    if (row_at_upper && !row_at_lower && signed_candidate > kDualFeasibility) return settle(c, dc);
    y[static_cast<size_t>(it.index)] = candidate;
    dual_finalized[c] = 1;
    d[c] = 0.0;
  };

  // The three stages repeated to a fixed point (:1324-1637): singleton rows backwards (those
  // depending on an unresolved fold deferred), the folds backwards, the deferred rows.
  std::vector<const Record*> deferred;
  std::vector<double> before;
  for (int pass = 0; pass < kMaxDualPasses; ++pass) {
    before.assign(y, y + M);
    std::fill(dual_finalized.begin(), dual_finalized.end(), 0);
    deferred.clear();
    for (size_t idx = NR; idx-- > 0;) {
      const Record& it = R.records[idx];
      if (it.kind != Kind::kSingletonRow) continue;
      if (std::fabs(it.coefficient) <= kZeroDrop) continue;
      if (depends_on_unresolved_fold(it.column)) {
        deferred.push_back(&it);
        continue;
      }
      process_singleton_row(it);
    }
    for (size_t idx = NR; idx-- > 0;) {
      const Record& rec = R.records[idx];
      if (rec.kind != Kind::kFreeColumnSingleton && rec.kind != Kind::kDoubletonEquation) continue;
      if (std::fabs(rec.coefficient) <= kZeroDrop) continue;
      y[static_cast<size_t>(rec.index)] = 0.0;
      const auto c = static_cast<size_t>(rec.column);
      if (rec.kind == Kind::kDoubletonEquation && !dual_finalized[static_cast<size_t>(rec.partner_column)]) {
        // :1407-1550: the row's price Y from whichever column pins it, or the admissible
        // interval both at-bound columns leave
        const auto pc = static_cast<size_t>(rec.partner_column);
        const double a = rec.coefficient, b = rec.partner_coefficient;
        const bool elim_at_lo = at_bound(x[c], orig.col_lower[c]), elim_at_hi = at_bound(x[c], orig.col_upper[c]);
        const bool keep_at_lo = at_bound(x[pc], orig.col_lower[pc]), keep_at_hi = at_bound(x[pc], orig.col_upper[pc]);
        const bool elim_interior = !elim_at_lo && !elim_at_hi;
        const bool keep_interior = !keep_at_lo && !keep_at_hi;
        const double rco_elim = reduced_cost_of(rec.column);
        const double rco_keep = reduced_cost_of(rec.partner_column);
        const double y_from_elim = rco_elim / a;
        const double y_from_keep = y_from_elim + rco_keep / b;
        auto admissible_at_bound = [&](double dv, bool lo, bool hi) {
          const double s = sense * dv;
          if (lo && !hi) return s >= -kDualFeasibility;
          if (hi && !lo) return s <= kDualFeasibility;
          return true;
        };
        double yv = 0.0;
        if (elim_interior) {
          yv = y_from_elim;
        } else if (keep_interior) {
          yv = y_from_keep;
        } else {
          double y_min = -kInf, y_max = kInf;
          auto constrain = [&](double rv0, double g, double anchor, bool lo, bool hi) {
            if (lo == hi) return;
            const double k = -sense * g;
            const double c0 = sense * (rv0 + g * anchor);
            if (std::fabs(k) <= kZeroDrop) return;
            const double bound = -c0 / k;
            if (lo) {
              if (k > 0.0) y_min = std::max(y_min, bound);
              else y_max = std::min(y_max, bound);
            } else {
              if (k > 0.0) y_max = std::min(y_max, bound);
              else y_min = std::max(y_min, bound);
            }
          };
          constrain(rco_elim, a, 0.0, elim_at_lo, elim_at_hi);
          constrain(rco_keep, b, y_from_elim, keep_at_lo, keep_at_hi);
          if (y_min <= y_max + kDualFeasibility) yv = std::min(std::max(0.0, y_min), y_max);
          else if (!admissible_at_bound(rco_elim, elim_at_lo, elim_at_hi)) yv = y_from_elim;
          else yv = y_from_keep;
        }
        y[static_cast<size_t>(rec.index)] = yv;
        d[pc] = rco_keep - b * (yv - y_from_elim);
        dual_finalized[pc] = 1;
        d[c] = rco_elim - a * yv;
        dual_finalized[c] = 1;
        continue;
      }
      // a free column singleton, or a doubleton whose kept column something else priced
      const bool at_lower = at_bound(x[c], orig.col_lower[c]);
      const bool at_upper = at_bound(x[c], orig.col_upper[c]);
      const double dc = reduced_cost_of(rec.column);
      const double signed_d = sense * dc;
      const bool needs_price = (!at_lower && !at_upper) || (at_lower && !at_upper && signed_d < -kDualFeasibility) ||
                               (at_upper && !at_lower && signed_d > kDualFeasibility);
      if (!needs_price) {
        y[static_cast<size_t>(rec.index)] = 0.0;
        d[c] = dc;
        dual_finalized[c] = 1;
        continue;
      }
      y[static_cast<size_t>(rec.index)] = dc / rec.coefficient;
      d[c] = 0.0;
      dual_finalized[c] = 1;
    }
    for (const Record* it : deferred) process_singleton_row(*it);
    double largest_change = 0.0;
    for (size_t r = 0; r < M; ++r) largest_change = std::max(largest_change, std::fabs(y[r] - before[r]));
    if (largest_change == 0.0) break;
  }

  // d = c - A^T y for every column presolve could have changed (:1639-1685): removed
  // columns, and the columns of a removed row that was folded or ended with a price
  std::vector<uint8_t> recompute(N, 0);
  for (size_t u = 0; u < N; ++u) recompute[u] = column_removed_at[u] < NR;
  for (size_t i = 0; i < M; ++i) {
    if (row_removed_at[i] >= NR) continue;
    if (!row_is_folded[i] && y[i] == 0.0) continue;
    for (const auto& entry : original_row(static_cast<int32_t>(i))) recompute[static_cast<size_t>(entry.first)] = 1;
  }
  for (size_t u = 0; u < N; ++u) {
    if (!recompute[u]) continue;
    double dc = orig.cost[u];
    for (int32_t k = orig.col_starts[u]; k < orig.col_starts[u + 1]; ++k)
      dc -= orig.values[static_cast<size_t>(k)] * y[static_cast<size_t>(orig.row_indices[static_cast<size_t>(k)])];
    d[u] = dc;
  }
}

}  // namespace sk::presolve
