// Host-side LP container and seeded instance generator.
//
// An LP is held in the reference's Model layout (include/sankhya/model.hpp:131-225): a frozen
// CSC matrix with ascending row indices inside each column, costs, column bounds and
// two-sided row bounds. A bound counts as absent when it is infinite or |v| >= 1e30
// (include/sankhya/types.hpp:32-39).
#pragma once

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

namespace sk {

inline bool is_finite_bound(double v) { return !(std::isinf(v) || std::fabs(v) >= 1e30); }

struct HostLp {
  int32_t num_rows = 0;
  int32_t num_cols = 0;
  int32_t sense = 1;  // +1 minimise, -1 maximise
  double objective_offset = 0.0;
  std::vector<int32_t> col_starts;   // n + 1
  std::vector<int32_t> row_indices;  // nnz
  std::vector<double> values;        // nnz
  std::vector<double> cost, col_lower, col_upper, row_lower, row_upper;

  int64_t nnz() const { return col_starts.empty() ? 0 : col_starts.back(); }
  // The checks of Model::validate (src/core/model.cpp:125-211) that apply to an LP; returns
  // an empty string when the LP is well formed.
  std::string validate() const;
};

// Portable seeded generator: SplitMix64 (Steele, Lea & Flood, OOPSLA 2014), so this host
// and any other produce identical instances from the same seed.
class SplitMix64 {
 public:
  explicit SplitMix64(uint64_t seed) : state_(seed) {}
  uint64_t next() {
    uint64_t z = (state_ += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  }
  double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }  // [0, 1)
  double uniform(double lo, double hi) { return lo + (hi - lo) * uniform(); }
  // Uniform integer in [0, bound) by rejection on the top bits (no modulo bias).
  uint64_t below(uint64_t bound) {
    if (bound <= 1) return 0;
    const uint64_t limit = UINT64_MAX - UINT64_MAX % bound;
    uint64_t r;
    do {
      r = next();
    } while (r >= limit);
    return r % bound;
  }
  int64_t integer(int64_t lo, int64_t hi) {  // inclusive
    return lo + static_cast<int64_t>(below(static_cast<uint64_t>(hi - lo + 1)));
  }

 private:
  uint64_t state_;
};

// Parameters of the blending-style generator. The caller (or a preset) fixes every choice.
struct GenParams {
  uint64_t seed = 1;
  int32_t num_rows = 500;
  int32_t num_cols = 1000;
  int32_t num_equality_rows = 300;  // rows [0, k) are equalities, the rest are <=
  int32_t degree_kind = 0;          // 0 fixed degree_lo, 1 uniform [lo, hi], 2 power law
  int32_t degree_lo = 5;
  int32_t degree_hi = 5;
  double power_exponent = 2.1;
  int64_t target_nnz = 0;  // power law: trim or pad by single entries to exactly this
  int32_t value_kind = 0;  // 0 uniform |a| in [lo, hi], 1 log-uniform
  double value_lo = 0.1;
  double value_hi = 2.0;
  double negative_probability = 0.2;
  double upper_lo = 1.0, upper_hi = 100.0;
  double slack_lo = 0.0, slack_hi = 5.0;
  double cost_lo = -1.0, cost_hi = 1.0;
  int32_t fix_empty_rows = 1;
};

// Fill `params` for one of the five workloads:
//   1: one 500 x 1,000 blending LP (300 E / 200 <=), 5 nnz per column; instance = seed
//   2: 200k x 500k, 2.5M nnz, power-law column degrees (exponent 2.1, max 2,000),
//      |a| log-uniform on [1e-3, 1e3]; instance = seed
//   3: member `instance` of the 256-LP batch: m ~ U{200..5000}, n = round(m * U[2, 3]),
//      4-8 nnz per column, seed 1000 + instance
//   4: serving request `instance`: m log-uniform on [500, 50,000], n = 2m, 5 nnz per column,
//      seed 100000 + instance
//   5: 2M x 5M, 25M nnz, power-law column degrees, |a| ~ U[0.1, 10]; instance = seed
// Returns false for an unknown workload.
bool preset_params(int config, uint64_t instance, GenParams* params);

// Build the LP described by `params`. Feasible by construction: x0 is drawn inside the box
// and b = A x0, with <= rows given extra slack.
HostLp generate_lp(const GenParams& params);

}  // namespace sk
