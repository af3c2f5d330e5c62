#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <vector>

#include "cmi/model.hpp"

// Seeded synthetic QCQP/MIQCQP generator. Self-contained RNG (SplitMix64 for
// uniforms, Box-Muller for normals) so an instance is the same on every host
// and in every process: the device engine, the CPU oracle and the benches all
// build the identical DAG from the same (spec, seed).
//
// Each function f (the objective, then each constraint) is
//   sum_{p in pool, kept} coef_{f,p} * p  +  sum_{i kept} lin_{f,i} * x_i
// where the product pool is every square x_i^2 plus each pair x_i*x_j (i<j)
// kept with probability pair_density. A pool product enters f with
// probability term_prob; coef ~ N(0,1), doubled for pairs when
// symmetric_pairs is set (x^T Q x with Q symmetric); lin ~ U[-1,1] kept with
// probability lin_prob. Terms are added left to right in pool order, then the
// linear terms in variable order. Product nodes are shared between functions
// and emitted at first use. Constraints are `f <= rhs`, rhs ~ U[rhs_lo, rhs_hi].
namespace cmi::qcqp
{

class Rng
{
public:
  explicit Rng(std::uint64_t seed)
      : state_(seed)
  {
  }

  std::uint64_t next()
  {
    std::uint64_t z = (state_ += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  }

  /// Uniform on [0, 1) with 53 random bits.
  double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }

  double uniform(double lo, double hi) { return lo + (hi - lo) * uniform(); }

  /// Standard normal by Box-Muller (both outputs used).
  double normal()
  {
    if (has_spare_) {
      has_spare_ = false;
      return spare_;
    }
    double const u1 = 1.0 - uniform();  // (0, 1]
    double const u2 = uniform();
    double const r = std::sqrt(-2.0 * std::log(u1));
    double const theta = 6.283185307179586 * u2;
    spare_ = r * std::sin(theta);
    has_spare_ = true;
    return r * std::cos(theta);
  }

  bool bernoulli(double p) { return p >= 1.0 || uniform() < p; }

private:
  std::uint64_t state_;
  double spare_ = 0.0;
  bool has_spare_ = false;
};

struct Spec
{
  std::size_t n = 4;  // variables
  std::size_t m = 2;  // constraints
  std::size_t n_int = 0;  // the first n_int variables are integer
  double cont_lb = -10.0, cont_ub = 10.0;
  double int_lb = 0.0, int_ub = 10.0;
  double pair_density = 1.0;
  double term_prob = 1.0;
  double lin_prob = 1.0;
  bool symmetric_pairs = false;
  double rhs_lo = 1.0, rhs_hi = 10.0;
};

inline model::Problem<double> generate(const Spec& s, std::uint64_t seed)
{
  using model::Expr;
  Rng rng(seed);
  model::Problem<double> p;
  std::vector<Expr<double>> x;
  x.reserve(s.n);
  for (std::size_t i = 0; i < s.n; ++i) {
    if (i < s.n_int) {
      x.push_back(p.int_var(s.int_lb, s.int_ub));
    } else {
      x.push_back(p.var(s.cont_lb, s.cont_ub));
    }
  }

  struct Product
  {
    std::size_t i, j;
  };
  std::vector<Product> pool;
  for (std::size_t i = 0; i < s.n; ++i) {
    pool.push_back({i, i});
  }
  for (std::size_t i = 0; i < s.n; ++i) {
    for (std::size_t j = i + 1; j < s.n; ++j) {
      if (rng.bernoulli(s.pair_density)) {
        pool.push_back({i, j});
      }
    }
  }
  std::vector<std::size_t> product_node(pool.size(), SIZE_MAX);

  auto function = [&]() -> Expr<double> {
    std::vector<std::pair<double, std::size_t>> prod_terms;  // (coef, pool index)
    for (std::size_t k = 0; k < pool.size(); ++k) {
      if (!rng.bernoulli(s.term_prob)) {
        continue;
      }
      double c = rng.normal();
      if (s.symmetric_pairs && pool[k].i != pool[k].j) {
        c *= 2.0;
      }
      prod_terms.push_back({c, k});
    }
    std::vector<std::pair<double, std::size_t>> lin_terms;  // (coef, var)
    for (std::size_t i = 0; i < s.n; ++i) {
      if (!rng.bernoulli(s.lin_prob)) {
        continue;
      }
      lin_terms.push_back({rng.uniform(-1.0, 1.0), i});
    }
    if (prod_terms.empty() && lin_terms.empty()) {
      prod_terms.push_back({rng.normal(), 0});
    }
    bool have = false;
    Expr<double> acc(&p.graph, 0);
    auto add_term = [&](Expr<double> term) {
      if (!have) {
        acc = term;
        have = true;
      } else {
        acc = acc + term;
      }
    };
    for (auto const& [c, k] : prod_terms) {
      if (product_node[k] == SIZE_MAX) {
        Expr<double> const prod = x[pool[k].i] * x[pool[k].j];  // Sqr when i == j
        product_node[k] = prod.id();
      }
      add_term(c * p.expr(product_node[k]));
    }
    for (auto const& [c, i] : lin_terms) {
      add_term(c * x[i]);
    }
    return acc;
  };

  p.set_objective(function());
  for (std::size_t k = 0; k < s.m; ++k) {
    Expr<double> const f = function();
    p.add_constraint(f, model::Cmp::LE, rng.uniform(s.rhs_lo, s.rhs_hi));
  }
  return p;
}

/// Number of non-Const nodes reachable from the objective and constraint roots.
inline std::size_t reachable_nodes(const model::Problem<double>& p)
{
  std::size_t const n = p.graph.nodes.size();
  std::vector<bool> reach(n, false);
  if (p.objective_root < n) {
    reach[p.objective_root] = true;
  }
  for (auto const& c : p.constraints) {
    reach[c.root_id] = true;
  }
  std::size_t count = 0;
  for (std::size_t i = n; i-- > 0;) {
    if (!reach[i]) {
      continue;
    }
    for (std::size_t in : p.graph.nodes[i].in) {
      reach[in] = true;
    }
    if (p.graph.nodes[i].op != model::Op::Const) {
      ++count;
    }
  }
  return count;
}

/// Spec with term_prob and lin_prob (set equal) chosen by bisection so the
/// reachable non-Const node count is as close as possible to `target` for
/// this seed.
inline Spec tune_term_prob(Spec s, std::uint64_t seed, std::size_t target)
{
  double lo = 0.0, hi = 1.0;
  Spec best = s;
  std::size_t best_err = SIZE_MAX;
  for (int it = 0; it < 40; ++it) {
    double const mid = 0.5 * (lo + hi);
    s.term_prob = mid;
    s.lin_prob = mid;
    std::size_t const got = reachable_nodes(generate(s, seed));
    std::size_t const err = got > target ? got - target : target - got;
    if (err < best_err) {
      best_err = err;
      best = s;
    }
    if (got < target) {
      lo = mid;
    } else {
      hi = mid;
    }
  }
  return best;
}

}  // namespace cmi::qcqp
