// Interior-point finish (reference src/ipm/ipm.cpp): Mehrotra's predictor-corrector on the
// bounded form [A | -I] [x; s] = 0, l <= x <= u, row_lower <= s <= row_upper, with the normal
// equations (A Theta_x A^T + Theta_s + delta I) dy = r assembled and factorised on the device
// (normal.hpp). The O(nnz) vector work of an iteration stays on the host; the O(m^3)
// factorisation is the GPU's. Constants, starting point, warm start, step rule and stopping
// tests follow the reference (line references below).
#include "ipm.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <limits>

#include "device_types.hpp"
#include "engine.hpp"
#include "normal.hpp"
#include "status.hpp"

namespace sk {
namespace {

// ipm.cpp:50-110
constexpr double kTolerance = 1e-8;            // relative primal and dual infeasibility
constexpr double kComplementarity = 1e-8;      // largest s z / (1 + |x|)
constexpr double kGap = 1e-8;                  // relative duality gap
constexpr double kPrimalRegularization = 1e-8; // added to every Theta^-1
constexpr double kDualRegularization = 1e-10;  // diagonal shift of M and pivot floor
constexpr double kStepToBoundary = 0.995;
constexpr int64_t kMaxIterations = 300;
constexpr int kRefinementSteps = 2;
constexpr int32_t kRuizIterations = 10;
// a jump in floored pivots this large, within this factor of every tolerance, ends the solve
constexpr double kExhaustedSlack = 10.0;
constexpr int64_t kExhaustedPivots = 64;
constexpr double kExhaustedFraction = 0.01;

using Clock = std::chrono::steady_clock;

std::string format(const char* f, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, f);
  std::vsnprintf(buf, sizeof(buf), f, ap);
  va_end(ap);
  return buf;
}

// The model the method runs on: scaled (or not), with both orientations of A.
struct Problem {
  int32_t m = 0, n = 0;
  double sense = 1.0;
  std::vector<int32_t> cs, ri;  // CSC
  std::vector<double> cv;
  std::vector<int32_t> rs, ci;  // CSR
  std::vector<double> rv;
  std::vector<double> cost;          // n, minimisation sense
  std::vector<double> lower, upper;  // n + m: structurals, then logicals
  std::vector<double> dr, dc;        // multipliers mapping back to the original units
};

Problem make_problem(const HostLp& lp, bool scale) {
  Problem p;
  p.m = lp.num_rows;
  p.n = lp.num_cols;
  p.sense = static_cast<double>(lp.sense);
  p.cs = lp.col_starts;
  p.ri = lp.row_indices;
  p.dr.assign(static_cast<size_t>(p.m), 1.0);
  p.dc.assign(static_cast<size_t>(p.n), 1.0);
  if (scale) equilibrate(lp, kRuizIterations, &p.dr, &p.dc, &p.cv);
  else p.cv = lp.values;
  const size_t n = static_cast<size_t>(p.n), m = static_cast<size_t>(p.m);
  p.cost.resize(n);
  p.lower.resize(n + m);
  p.upper.resize(n + m);
  // ipm.cpp:911-939: cost * Dc, column bounds / Dc, row bounds * Dr (infinite ones stay)
  for (size_t j = 0; j < n; ++j) {
    p.cost[j] = p.sense * (lp.cost[j] * p.dc[j]);
    p.lower[j] = is_finite_bound(lp.col_lower[j]) ? lp.col_lower[j] / p.dc[j] : lp.col_lower[j];
    p.upper[j] = is_finite_bound(lp.col_upper[j]) ? lp.col_upper[j] / p.dc[j] : lp.col_upper[j];
  }
  for (size_t i = 0; i < m; ++i) {
    p.lower[n + i] = is_finite_bound(lp.row_lower[i]) ? lp.row_lower[i] * p.dr[i] : lp.row_lower[i];
    p.upper[n + i] = is_finite_bound(lp.row_upper[i]) ? lp.row_upper[i] * p.dr[i] : lp.row_upper[i];
  }
  // CSR by a counting pass over the CSC (rows keep ascending column order)
  p.rs.assign(m + 1, 0);
  for (int32_t r : p.ri) ++p.rs[static_cast<size_t>(r) + 1];
  for (size_t i = 0; i < m; ++i) p.rs[i + 1] += p.rs[i];
  p.ci.resize(p.ri.size());
  p.rv.resize(p.ri.size());
  std::vector<int32_t> next(p.rs.begin(), p.rs.end() - 1);
  for (int32_t j = 0; j < p.n; ++j)
    for (int32_t q = p.cs[static_cast<size_t>(j)]; q < p.cs[static_cast<size_t>(j) + 1]; ++q) {
      const auto at = static_cast<size_t>(next[static_cast<size_t>(p.ri[static_cast<size_t>(q)])]++);
      p.ci[at] = j;
      p.rv[at] = p.cv[static_cast<size_t>(q)];
    }
  return p;
}

class Method {
 public:
  Method(const Problem& p, const IpmOptions& o, const WarmPoint* warm)
      : p_(p), opt_(o), warm_(warm), n_(p.n), m_(p.m), total_(p.n + p.m) {}

  int run(IpmOutcome* out);

 private:
  size_t N() const { return static_cast<size_t>(total_); }
  void build();
  void apply_warm_start();
  void residuals();
  int factor(int64_t* floored_now);
  int direction();
  double step_length(const std::vector<double>& s, const std::vector<double>& ds,
                     const std::vector<double>& t, const std::vector<double>& dt) const;
  void finish(int32_t status, const std::string& message, IpmOutcome* out);
  void remember(double merit);
  void restore();

  // Abar v = A v_x - v_s (m); Abar^T w = [A^T w; -w] (n + m)
  void abar(const std::vector<double>& v, std::vector<double>* out) const;
  void abar_t(const std::vector<double>& w, std::vector<double>* out) const;
  // M v with the matrix the device assembled (before its pivots were floored)
  void normal_times(const std::vector<double>& v, std::vector<double>* out) const;

  const Problem& p_;
  const IpmOptions& opt_;
  const WarmPoint* warm_;
  int32_t n_, m_, total_;
  int64_t bound_count_ = 0;
  std::vector<char> has_l_, has_u_, fixed_;
  std::vector<double> cost_;  // n + m, zero on the logicals
  std::vector<double> x_, y_, sl_, zl_, su_, zu_;
  std::vector<double> rb_, rc_, rl_, ru_;
  std::vector<double> theta_, shift_, theta_x_;
  std::vector<double> dx_, dy_, dsl_, dzl_, dsu_, dzu_, rmu_l_, rmu_u_;
  double primal_inf_ = 0.0, dual_inf_ = 0.0, mu_ = 0.0, max_product_ = 0.0, objective_ = 0.0;
  int64_t factorizations_ = 0, floored_ = 0;
  NormalEquations ne_;
  struct Best {
    double merit = std::numeric_limits<double>::infinity();
    std::vector<double> x, y, zl, zu, sl, su;
  } best_;
};

void Method::abar(const std::vector<double>& v, std::vector<double>* out) const {
  out->assign(static_cast<size_t>(m_), 0.0);
  for (int32_t j = 0; j < n_; ++j) {
    const double vj = v[static_cast<size_t>(j)];
    if (vj == 0.0) continue;
    for (int32_t q = p_.cs[static_cast<size_t>(j)]; q < p_.cs[static_cast<size_t>(j) + 1]; ++q)
      (*out)[static_cast<size_t>(p_.ri[static_cast<size_t>(q)])] += p_.cv[static_cast<size_t>(q)] * vj;
  }
  for (int32_t i = 0; i < m_; ++i) (*out)[static_cast<size_t>(i)] -= v[static_cast<size_t>(n_ + i)];
}

void Method::abar_t(const std::vector<double>& w, std::vector<double>* out) const {
  out->assign(N(), 0.0);
  for (int32_t j = 0; j < n_; ++j) {
    double dot = 0.0;
    for (int32_t q = p_.cs[static_cast<size_t>(j)]; q < p_.cs[static_cast<size_t>(j) + 1]; ++q)
      dot += p_.cv[static_cast<size_t>(q)] * w[static_cast<size_t>(p_.ri[static_cast<size_t>(q)])];
    (*out)[static_cast<size_t>(j)] = dot;
  }
  for (int32_t i = 0; i < m_; ++i) (*out)[static_cast<size_t>(n_ + i)] = -w[static_cast<size_t>(i)];
}

void Method::normal_times(const std::vector<double>& v, std::vector<double>* out) const {
  std::vector<double> t(static_cast<size_t>(n_), 0.0);
  for (int32_t j = 0; j < n_; ++j) {
    double dot = 0.0;
    for (int32_t q = p_.cs[static_cast<size_t>(j)]; q < p_.cs[static_cast<size_t>(j) + 1]; ++q)
      dot += p_.cv[static_cast<size_t>(q)] * v[static_cast<size_t>(p_.ri[static_cast<size_t>(q)])];
    t[static_cast<size_t>(j)] = theta_x_[static_cast<size_t>(j)] * dot;
  }
  out->assign(static_cast<size_t>(m_), 0.0);
  for (int32_t j = 0; j < n_; ++j) {
    const double tj = t[static_cast<size_t>(j)];
    for (int32_t q = p_.cs[static_cast<size_t>(j)]; q < p_.cs[static_cast<size_t>(j) + 1]; ++q)
      (*out)[static_cast<size_t>(p_.ri[static_cast<size_t>(q)])] += p_.cv[static_cast<size_t>(q)] * tj;
  }
  for (int32_t i = 0; i < m_; ++i)
    (*out)[static_cast<size_t>(i)] += shift_[static_cast<size_t>(i)] * v[static_cast<size_t>(i)];
}

// ipm.cpp:233-337: bounds, the fixed variables, and the cold starting point
void Method::build() {
  const size_t T = N();
  cost_.assign(T, 0.0);
  for (int32_t j = 0; j < n_; ++j) cost_[static_cast<size_t>(j)] = p_.cost[static_cast<size_t>(j)];
  has_l_.assign(T, 0);
  has_u_.assign(T, 0);
  fixed_.assign(T, 0);
  bound_count_ = 0;
  for (size_t k = 0; k < T; ++k) {
    const double l = p_.lower[k], u = p_.upper[k];
    if (l == u && is_finite_bound(l)) {  // a constant: no slacks, no multipliers
      fixed_[k] = 1;
      continue;
    }
    has_l_[k] = is_finite_bound(l);
    has_u_[k] = is_finite_bound(u);
    bound_count_ += has_l_[k] + has_u_[k];
  }
  x_.assign(T, 0.0);
  sl_.assign(T, 0.0);
  su_.assign(T, 0.0);
  zl_.assign(T, 0.0);
  zu_.assign(T, 0.0);
  for (size_t k = 0; k < T; ++k) {
    const double l = p_.lower[k], u = p_.upper[k];
    if (fixed_[k]) x_[k] = l;
    else if (has_l_[k] && has_u_[k]) x_[k] = 0.5 * (l + u);
    else if (has_l_[k]) x_[k] = l + 1.0;
    else if (has_u_[k]) x_[k] = u - 1.0;
  }
  // logicals consistent with the structurals: s = A x
  std::vector<double> ax;
  abar(x_, &ax);  // A x - s with s = x_[n..], so A x = ax + s
  for (int32_t i = 0; i < m_; ++i) {
    const size_t k = static_cast<size_t>(n_ + i);
    if (!fixed_[k]) x_[k] = ax[static_cast<size_t>(i)] + x_[k];
  }
  for (size_t k = 0; k < T; ++k) {
    if (fixed_[k]) continue;
    if (has_l_[k]) {
      sl_[k] = std::max(x_[k] - p_.lower[k], 1.0);
      zl_[k] = std::max(cost_[k], 1.0);
    }
    if (has_u_[k]) {
      su_[k] = std::max(p_.upper[k] - x_[k], 1.0);
      zu_[k] = std::max(-cost_[k], 1.0);
    }
  }
  y_.assign(static_cast<size_t>(m_), 0.0);
  if (warm_) apply_warm_start();
  for (auto* v : {&rc_, &rl_, &ru_, &theta_, &dx_, &dsl_, &dzl_, &dsu_, &dzu_, &rmu_l_, &rmu_u_})
    v->assign(T, 0.0);
  rb_.assign(static_cast<size_t>(m_), 0.0);
  dy_.assign(static_cast<size_t>(m_), 0.0);
  shift_.assign(static_cast<size_t>(m_), 0.0);
  theta_x_.assign(static_cast<size_t>(n_), 0.0);
}

// ipm.cpp:388-434: the caller's point projected into its box, logicals = A x, y the caller's
// row duals in minimisation sense, multiplier pairs read off the reduced cost, slacks and
// multipliers floored at the start margin (warm point already in the scaled units)
void Method::apply_warm_start() {
  const WarmPoint& w = *warm_;
  const double floor = opt_.start_margin;
  for (int32_t j = 0; j < n_; ++j) {
    const size_t u = static_cast<size_t>(j);
    if (fixed_[u]) continue;
    double x = w.col_value[u];
    if (has_l_[u]) x = std::max(x, p_.lower[u]);
    if (has_u_[u]) x = std::min(x, p_.upper[u]);
    x_[u] = x;
  }
  std::vector<double> ax;
  abar(x_, &ax);
  for (int32_t i = 0; i < m_; ++i) {
    const size_t k = static_cast<size_t>(n_ + i);
    if (!fixed_[k]) x_[k] = ax[static_cast<size_t>(i)] + x_[k];
  }
  for (int32_t i = 0; i < m_; ++i) y_[static_cast<size_t>(i)] = p_.sense * w.row_dual[static_cast<size_t>(i)];
  for (size_t k = 0; k < N(); ++k) {
    if (fixed_[k]) continue;
    // a logical's column of [A -I] is -e_i, so its reduced cost is y_i
    const double d = k < static_cast<size_t>(n_) ? p_.sense * w.col_dual[k] : y_[k - static_cast<size_t>(n_)];
    if (has_l_[k]) {
      sl_[k] = std::max(x_[k] - p_.lower[k], floor);
      zl_[k] = std::max(d, floor);
    }
    if (has_u_[k]) {
      su_[k] = std::max(p_.upper[k] - x_[k], floor);
      zu_[k] = std::max(-d, floor);
    }
  }
}

// ipm.cpp:436-480
void Method::residuals() {
  abar(x_, &rb_);
  for (double& v : rb_) v = -v;
  abar_t(y_, &rc_);
  double x_norm = 0.0, c_norm = 0.0, complementarity = 0.0;
  max_product_ = 0.0;
  objective_ = 0.0;
  primal_inf_ = 0.0;
  dual_inf_ = 0.0;
  for (size_t k = 0; k < N(); ++k) {
    rc_[k] = fixed_[k] ? 0.0 : cost_[k] - rc_[k] - zl_[k] + zu_[k];
    objective_ += cost_[k] * x_[k];
    c_norm = std::max(c_norm, std::fabs(cost_[k]));
    x_norm = std::max(x_norm, std::fabs(x_[k]));
    dual_inf_ = std::max(dual_inf_, std::fabs(rc_[k]));
    if (has_l_[k]) {
      rl_[k] = x_[k] - p_.lower[k] - sl_[k];
      complementarity += sl_[k] * zl_[k];
      max_product_ = std::max(max_product_, sl_[k] * zl_[k] / (1.0 + std::fabs(x_[k])));
      primal_inf_ = std::max(primal_inf_, std::fabs(rl_[k]));
    }
    if (has_u_[k]) {
      ru_[k] = p_.upper[k] - x_[k] - su_[k];
      complementarity += su_[k] * zu_[k];
      max_product_ = std::max(max_product_, su_[k] * zu_[k] / (1.0 + std::fabs(x_[k])));
      primal_inf_ = std::max(primal_inf_, std::fabs(ru_[k]));
    }
  }
  for (double v : rb_) primal_inf_ = std::max(primal_inf_, std::fabs(v));
  primal_inf_ /= 1.0 + x_norm;
  dual_inf_ /= 1.0 + c_norm;
  mu_ = bound_count_ > 0 ? complementarity / static_cast<double>(bound_count_) : 0.0;
}

// ipm.cpp:482-528: Theta = (z_l/s_l + z_u/s_u + rho)^-1; M = A Theta_x A^T + Theta_s + delta I
int Method::factor(int64_t* floored_now) {
  for (size_t k = 0; k < N(); ++k) {
    double inverse = kPrimalRegularization;
    if (has_l_[k]) inverse += zl_[k] / sl_[k];
    if (has_u_[k]) inverse += zu_[k] / su_[k];
    theta_[k] = fixed_[k] ? 0.0 : 1.0 / inverse;
  }
  for (int32_t j = 0; j < n_; ++j) theta_x_[static_cast<size_t>(j)] = theta_[static_cast<size_t>(j)];
  for (int32_t i = 0; i < m_; ++i)
    shift_[static_cast<size_t>(i)] = theta_[static_cast<size_t>(n_ + i)] + kDualRegularization;
  const int st = ne_.factor(theta_x_.data(), shift_.data(), kDualRegularization, floored_now);
  if (st) return st;
  ++factorizations_;
  floored_ += *floored_now;
  return kStatusSuccess;
}

// ipm.cpp:532-595: one Newton direction for the current r_mu terms
int Method::direction() {
  const size_t T = N();
  std::vector<double> g(T), theta_g(T);
  for (size_t k = 0; k < T; ++k) {
    double v = rc_[k];
    if (has_l_[k]) v += (-rmu_l_[k] + zl_[k] * rl_[k]) / sl_[k];
    if (has_u_[k]) v += (rmu_u_[k] - zu_[k] * ru_[k]) / su_[k];
    g[k] = v;
    theta_g[k] = theta_[k] * v;
  }
  std::vector<double> rhs;
  abar(theta_g, &rhs);
  for (int32_t i = 0; i < m_; ++i) rhs[static_cast<size_t>(i)] += rb_[static_cast<size_t>(i)];
  dy_ = rhs;
  int st = ne_.solve(dy_.data());
  if (st) return st;
  std::vector<double> res;
  for (int step = 0; step < kRefinementSteps; ++step) {
    normal_times(dy_, &res);
    for (int32_t i = 0; i < m_; ++i) res[static_cast<size_t>(i)] = rhs[static_cast<size_t>(i)] - res[static_cast<size_t>(i)];
    if ((st = ne_.solve(res.data()))) return st;
    for (int32_t i = 0; i < m_; ++i) dy_[static_cast<size_t>(i)] += res[static_cast<size_t>(i)];
  }
  abar_t(dy_, &dx_);
  for (size_t k = 0; k < T; ++k) {
    dx_[k] = theta_[k] * (dx_[k] - g[k]);
    if (has_l_[k]) {
      dsl_[k] = dx_[k] + rl_[k];
      dzl_[k] = (rmu_l_[k] - zl_[k] * dsl_[k]) / sl_[k];
    }
    if (has_u_[k]) {
      dsu_[k] = -dx_[k] + ru_[k];
      dzu_[k] = (rmu_u_[k] - zu_[k] * dsu_[k]) / su_[k];
    }
  }
  return kStatusSuccess;
}

double Method::step_length(const std::vector<double>& s, const std::vector<double>& ds,
                           const std::vector<double>& t, const std::vector<double>& dt) const {
  double alpha = 1.0;
  for (size_t k = 0; k < N(); ++k) {
    if (has_l_[k] && ds[k] < 0.0) alpha = std::min(alpha, -s[k] / ds[k]);
    if (has_u_[k] && dt[k] < 0.0) alpha = std::min(alpha, -t[k] / dt[k]);
  }
  return alpha;
}

// ipm.cpp:189-229: the best finite iterate by the convergence measures
void Method::remember(double merit) {
  if (!std::isfinite(merit) || !std::isfinite(mu_) || !std::isfinite(objective_)) return;
  if (!(merit < best_.merit)) return;
  best_.merit = merit;
  best_.x = x_;
  best_.y = y_;
  best_.zl = zl_;
  best_.zu = zu_;
  best_.sl = sl_;
  best_.su = su_;
}

void Method::restore() {
  if (best_.x.empty()) return;
  x_ = best_.x;
  y_ = best_.y;
  zl_ = best_.zl;
  zu_ = best_.zu;
  sl_ = best_.sl;
  su_ = best_.su;
}

// ipm.cpp:597-677: the point in the method's units; solve_ipm maps it back
void Method::finish(int32_t status, const std::string& message, IpmOutcome* out) {
  out->status = status;
  out->message = message;
  out->factorizations = factorizations_;
  out->floored_pivots = floored_;
  bool have_point = status == kStOptimal || status == kStFeasible || status == kStIterationLimit ||
                    status == kStTimeLimit;
  if (have_point) {
    for (int32_t j = 0; j < n_; ++j) {
      if (!std::isfinite(x_[static_cast<size_t>(j)])) {
        have_point = false;
        out->status = kStNumericalError;
        out->message = message + (message.empty() ? "" : "; ") +
                       "the iterate is not finite, so it is reported as a numerical failure";
        break;
      }
    }
  }
  out->col_value.clear();
  out->row_dual.clear();
  out->col_dual.clear();
  if (!have_point) return;
  const size_t n = static_cast<size_t>(n_), m = static_cast<size_t>(m_);
  out->col_value.assign(x_.begin(), x_.begin() + static_cast<std::ptrdiff_t>(n));
  out->col_dual.resize(n);
  for (size_t j = 0; j < n; ++j) {
    double d = zl_[j] - zu_[j];
    if (fixed_[j]) {  // c_j - a_j^T y: no multipliers exist for a fixed column
      double dot = 0.0;
      for (int32_t q = p_.cs[j]; q < p_.cs[j + 1]; ++q)
        dot += p_.cv[static_cast<size_t>(q)] * y_[static_cast<size_t>(p_.ri[static_cast<size_t>(q)])];
      d = cost_[j] - dot;
    }
    out->col_dual[j] = p_.sense * d;
  }
  out->row_dual.resize(m);
  for (size_t i = 0; i < m; ++i) out->row_dual[i] = p_.sense * y_[i];
}

// ipm.cpp:679-892
int Method::run(IpmOutcome* out) {
  const auto t0 = Clock::now();
  auto elapsed = [&] { return std::chrono::duration<double>(Clock::now() - t0).count(); };
  build();
  int st = ne_.init(m_, n_, p_.rs, p_.ci, p_.rv, p_.cs, p_.ri, p_.cv);
  if (st) return st;
  int64_t iterations = 0;
  double previous_mu = std::numeric_limits<double>::infinity();
  int stalled = 0;
  for (;; ++iterations) {
    out->iterations = iterations;
    residuals();
    const double relative_gap = mu_ * static_cast<double>(bound_count_) / (1.0 + std::fabs(objective_));
    if (!std::isfinite(mu_) || !std::isfinite(primal_inf_) || !std::isfinite(dual_inf_) ||
        !std::isfinite(objective_)) {
      // a non-finite iterate ends the solve; the best finite one is the answer if feasible
      restore();
      residuals();
      const bool usable = std::isfinite(best_.merit) && std::isfinite(objective_) &&
                          primal_inf_ <= 1e-6 && dual_inf_ <= 1e-6;
      finish(usable ? kStFeasible : kStNumericalError,
             format("the iterate stopped being finite at iteration %lld; %s",
                    static_cast<long long>(iterations),
                    usable ? "the best earlier iterate is reported as a feasible point"
                           : "the best earlier iterate does not meet the feasibility tolerances"),
             out);
      break;
    }
    remember(std::max({primal_inf_, dual_inf_, relative_gap, max_product_}));
    if (primal_inf_ <= kTolerance && dual_inf_ <= kTolerance && relative_gap <= kGap &&
        max_product_ <= kComplementarity) {
      finish(kStOptimal, "", out);
      break;
    }
    if (iterations >= kMaxIterations || (opt_.iteration_limit >= 0 && iterations >= opt_.iteration_limit)) {
      restore();
      finish(kStIterationLimit,
             format("iteration limit reached after %lld iterations; relative infeasibility "
                    "%.1e / %.1e, gap %.1e",
                    static_cast<long long>(iterations), primal_inf_, dual_inf_, relative_gap),
             out);
      break;
    }
    if (elapsed() > opt_.max_seconds) {
      restore();
      finish(kStTimeLimit, format("time limit %gs reached", opt_.max_seconds), out);
      break;
    }
    int64_t floored_now = 0;
    if ((st = factor(&floored_now))) return st;
    // the barrier's last word (ipm.cpp:818-845): a spike of floored pivots at an iterate within
    // a decade of every tolerance means the normal equations have lost rank at working
    // precision; the iterate measured above is the answer
    const int64_t spike = std::max<int64_t>(kExhaustedPivots,
                                            static_cast<int64_t>(kExhaustedFraction * m_));
    const bool nearly = primal_inf_ <= kExhaustedSlack * kTolerance &&
                        dual_inf_ <= kExhaustedSlack * kTolerance &&
                        relative_gap <= kExhaustedSlack * kGap &&
                        max_product_ <= kExhaustedSlack * kComplementarity;
    if (floored_now >= spike && nearly) {
      finish(kStOptimal,
             format("converged at a relative gap of %.1e when the barrier vanished: %lld of %d "
                    "pivots floored in one factorization",
                    relative_gap, static_cast<long long>(floored_now), m_),
             out);
      break;
    }
    // predictor: the affine-scaling direction
    for (size_t k = 0; k < N(); ++k) {
      rmu_l_[k] = has_l_[k] ? -sl_[k] * zl_[k] : 0.0;
      rmu_u_[k] = has_u_[k] ? -su_[k] * zu_[k] : 0.0;
    }
    if ((st = direction())) return st;
    const double ap_aff = step_length(sl_, dsl_, su_, dsu_);
    const double ad_aff = step_length(zl_, dzl_, zu_, dzu_);
    double mu_aff = 0.0;
    for (size_t k = 0; k < N(); ++k) {
      if (has_l_[k]) mu_aff += (sl_[k] + ap_aff * dsl_[k]) * (zl_[k] + ad_aff * dzl_[k]);
      if (has_u_[k]) mu_aff += (su_[k] + ap_aff * dsu_[k]) * (zu_[k] + ad_aff * dzu_[k]);
    }
    mu_aff = bound_count_ > 0 ? mu_aff / static_cast<double>(bound_count_) : 0.0;
    const double ratio = mu_ > 0.0 ? mu_aff / mu_ : 0.0;
    const double sigma = std::min(1.0, ratio * ratio * ratio);
    // corrector: centring plus the predictor's second-order term
    for (size_t k = 0; k < N(); ++k) {
      if (has_l_[k]) rmu_l_[k] = sigma * mu_ - sl_[k] * zl_[k] - dsl_[k] * dzl_[k];
      if (has_u_[k]) rmu_u_[k] = sigma * mu_ - su_[k] * zu_[k] - dsu_[k] * dzu_[k];
    }
    if ((st = direction())) return st;
    const double ap = std::min(1.0, kStepToBoundary * step_length(sl_, dsl_, su_, dsu_));
    const double ad = std::min(1.0, kStepToBoundary * step_length(zl_, dzl_, zu_, dzu_));
    for (size_t k = 0; k < N(); ++k) {
      x_[k] += ap * dx_[k];
      if (has_l_[k]) {
        sl_[k] += ap * dsl_[k];
        zl_[k] += ad * dzl_[k];
      }
      if (has_u_[k]) {
        su_[k] += ap * dsu_[k];
        zu_[k] += ad * dzu_[k];
      }
    }
    for (int32_t i = 0; i < m_; ++i) y_[static_cast<size_t>(i)] += ad * dy_[static_cast<size_t>(i)];
    // a method that stops moving is not converging (ipm.cpp:865-889)
    if (mu_ >= 0.999 * previous_mu && ap < 1e-6 && ad < 1e-6) {
      if (++stalled >= 5) {
        restore();
        residuals();
        const bool usable = std::isfinite(best_.merit) && primal_inf_ <= 1e-6 && dual_inf_ <= 1e-6;
        finish(usable ? kStFeasible : kStNumericalError,
               format("the interior-point iteration stalled after %lld iterations (steps %.1e / %.1e)",
                      static_cast<long long>(iterations), ap, ad),
               out);
        break;
      }
    } else {
      stalled = 0;
    }
    previous_mu = mu_;
  }
  out->iterations = iterations;
  out->seconds = elapsed();
  return kStatusSuccess;
}

}  // namespace

// scaling.cpp:14-159 (LP part): ruiz_iterations rounds of r_i = 1/sqrt(max_j |a_ij|),
// s_j = 1/sqrt(max_i |a_ij|), then one round with the sums; both multipliers of a round come
// from the same matrix and are applied together; an empty row or column keeps 1.
void equilibrate(const HostLp& lp, int32_t ruiz_iterations, std::vector<double>* row,
                 std::vector<double>* column, std::vector<double>* values) {
  const int32_t m = lp.num_rows, n = lp.num_cols;
  row->assign(static_cast<size_t>(m), 1.0);
  column->assign(static_cast<size_t>(n), 1.0);
  *values = lp.values;
  std::vector<double> racc(static_cast<size_t>(m)), cacc(static_cast<size_t>(n));
  std::vector<double> rpass(static_cast<size_t>(m)), cpass(static_cast<size_t>(n));
  for (int32_t round = 0; round <= ruiz_iterations; ++round) {
    const bool use_sum = round == ruiz_iterations;
    std::fill(racc.begin(), racc.end(), 0.0);
    std::fill(cacc.begin(), cacc.end(), 0.0);
    for (int32_t j = 0; j < n; ++j) {
      for (int32_t q = lp.col_starts[static_cast<size_t>(j)]; q < lp.col_starts[static_cast<size_t>(j) + 1]; ++q) {
        const double mag = std::fabs((*values)[static_cast<size_t>(q)]);
        double& r = racc[static_cast<size_t>(lp.row_indices[static_cast<size_t>(q)])];
        double& c = cacc[static_cast<size_t>(j)];
        if (use_sum) {
          r += mag;
          c += mag;
        } else {
          r = std::max(r, mag);
          c = std::max(c, mag);
        }
      }
    }
    for (int32_t i = 0; i < m; ++i) {
      const double a = racc[static_cast<size_t>(i)];
      rpass[static_cast<size_t>(i)] = a > 0.0 ? 1.0 / std::sqrt(a) : 1.0;
    }
    for (int32_t j = 0; j < n; ++j) {
      const double a = cacc[static_cast<size_t>(j)];
      cpass[static_cast<size_t>(j)] = a > 0.0 ? 1.0 / std::sqrt(a) : 1.0;
    }
    for (int32_t j = 0; j < n; ++j) {
      const double cj = cpass[static_cast<size_t>(j)];
      for (int32_t q = lp.col_starts[static_cast<size_t>(j)]; q < lp.col_starts[static_cast<size_t>(j) + 1]; ++q) {
        double& v = (*values)[static_cast<size_t>(q)];
        v = v * rpass[static_cast<size_t>(lp.row_indices[static_cast<size_t>(q)])] * cj;
      }
    }
    for (int32_t i = 0; i < m; ++i) (*row)[static_cast<size_t>(i)] *= rpass[static_cast<size_t>(i)];
    for (int32_t j = 0; j < n; ++j) (*column)[static_cast<size_t>(j)] *= cpass[static_cast<size_t>(j)];
  }
}

// ipm.cpp:897-988 (solve_scaled): scale, map the warm start in (x / Dc, y / Dr, d * Dc), run,
// map the answer back (x * Dc, y * Dr, d / Dc)
int solve_ipm(const HostLp& lp, const IpmOptions& options, const WarmPoint* warm, IpmOutcome* out) {
  const size_t n = static_cast<size_t>(lp.num_cols), m = static_cast<size_t>(lp.num_rows);
  // a non-finite start is ignored and the solve is cold (ipm.cpp:388-402)
  bool finite = warm && warm->col_value && warm->row_dual && warm->col_dual;
  for (size_t j = 0; j < n && finite; ++j)
    finite = std::isfinite(warm->col_value[j]) && std::isfinite(warm->col_dual[j]);
  for (size_t i = 0; i < m && finite; ++i) finite = std::isfinite(warm->row_dual[i]);
  // the polish declines a factor above its cap before anything is built, scaling included
  // (ipm.cpp:509-518); the dense factor holds m (m + 1) / 2 nonzeros
  const int64_t factor_nonzeros = static_cast<int64_t>(m) * static_cast<int64_t>(m + 1) / 2;
  if (finite && options.max_factor_nonzeros >= 0 && factor_nonzeros > options.max_factor_nonzeros) {
    *out = IpmOutcome{};
    out->status = kStNotSolved;
    out->message = format("declined: the factor would hold %lld nonzeros, above polish_max_factor_nonzeros = %lld",
                          static_cast<long long>(factor_nonzeros),
                          static_cast<long long>(options.max_factor_nonzeros));
    return kStatusSuccess;
  }
  const Problem p = make_problem(lp, options.scaling != 0);
  std::vector<double> wx, wy, wd;
  WarmPoint scaled_warm;
  const WarmPoint* use = nullptr;
  {
    if (finite) {
      wx.resize(n);
      wd.resize(n);
      wy.resize(m);
      for (size_t j = 0; j < n; ++j) {
        wx[j] = warm->col_value[j] / p.dc[j];
        wd[j] = warm->col_dual[j] * p.dc[j];
      }
      for (size_t i = 0; i < m; ++i) wy[i] = warm->row_dual[i] / p.dr[i];
      scaled_warm.col_value = wx.data();
      scaled_warm.row_dual = wy.data();
      scaled_warm.col_dual = wd.data();
      use = &scaled_warm;
    }
  }
  Method method(p, options, use);
  const int st = method.run(out);
  if (st) return st;
  for (size_t j = 0; j < out->col_value.size(); ++j) {
    out->col_value[j] *= p.dc[j];
    out->col_dual[j] /= p.dc[j];
  }
  for (size_t i = 0; i < out->row_dual.size(); ++i) out->row_dual[i] *= p.dr[i];
  return kStatusSuccess;
}

}  // namespace sk
