// Request-stream serving: continuous batching of searches on one device, and
// the one-at-a-time baseline it is measured against.

#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <new>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "batch.hpp"
#include "cmi.h"
#include "cmi/config.hpp"
#include "cmi/policy.hpp"
#include "cmi/search.hpp"
#include "cmi/tape.hpp"
#include "handles.hpp"
#include "hip_backend.hpp"

namespace cmi::capi
{
std::shared_ptr<const cmi::policy::CompositionPolicy> make_policy(const cmi_solve_options& o);
}  // namespace cmi::capi

namespace
{

using Clock = std::chrono::steady_clock;

/// Device bytes one resident solve can come to hold: its tape and variable
/// kinds, plus the roles' bytes for the worst composition its search can
/// reach at this fan-out and slot cap (the memory-fit resolver's footprint,
/// so admission and shape fitting price a solve alike). Priced from a host
/// compile of the problem's tape, before any device memory is taken.
std::size_t resident_bytes(const cmi::model::Problem<double>& p, const cmi_solve_options& o)
{
  p.validate();
  cmi::tape::Tape const t = cmi::tape::compile(p, {});
  std::size_t const n_vars = t.n_vars;
  std::size_t const cap = std::min<std::size_t>(o.max_cycle_size == 0 ? cmi::config::kMaxSlots : o.max_cycle_size,
                                                n_vars);
  auto const cost = cmi::hip::cost_model_for(t, std::max<std::size_t>(o.sample_points, 1));
  return t.code.size() * sizeof(cmi::tape::Instr) + n_vars
      + cmi::config::worst_footprint(cap, cmi::config::profile_of(p), cost, o.partition_num, o.enumerate_cap);
}

struct Active
{
  std::size_t idx;
  std::unique_ptr<cmi::hip::DeviceTape> tape;
  std::unique_ptr<cmi::search::SolveState> state;
  std::size_t bytes;
};

void wait_until(Clock::time_point t0, double at)
{
  auto const target = t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(at));
  std::this_thread::sleep_until(target);
}

}  // namespace

extern "C" cmi_status cmi_serve_stream(cmi_ctx* ctx, size_t n, const cmi_problem* const* problems,
                                       const double* arrival, const cmi_solve_options* o,
                                       const cmi_server_options* so, cmi_solve_outcome* outs, double* finish,
                                       double* latency, cmi_serve_stats* stats)
{
  struct Args
  {
    cmi_ctx* ctx;
    size_t n;
    const cmi_problem* const* problems;
    const double* arrival;
    const cmi_solve_options* o;
    const cmi_server_options* so;
    cmi_solve_outcome* outs;
    double* finish;
    double* latency;
    cmi_serve_stats* stats;
  } a {ctx, n, problems, arrival, o, so, outs, finish, latency, stats};
  if (ctx == nullptr || o == nullptr || so == nullptr || (n > 0 && (problems == nullptr || outs == nullptr))) {
    return CMI_INVALID_HANDLE;
  }
  return cmi::capi::run_guarded(
      [](void* p) {
        Args& a = *static_cast<Args*>(p);
        if (a.so->max_active < 1) {
          throw cmi::InvalidConfiguration("server max_active must be at least 1");
        }
        auto const pol = cmi::capi::make_policy(*a.o);
        // Each resident solve's host budget: as given, or an equal share of
        // half of MemAvailable among the max_active slots.
        cmi::search::SolveOptions const sopt =
            cmi::batch::search_options(*a.o, cmi::batch::shared_host_budget(*a.o, a.so->max_active));
        std::vector<std::size_t> order(a.n);
        std::iota(order.begin(), order.end(), 0);
        auto arrive = [&](std::size_t i) { return (a.so->realtime && a.arrival != nullptr) ? a.arrival[i] : 0.0; };
        std::stable_sort(order.begin(), order.end(), [&](std::size_t x, std::size_t y) { return arrive(x) < arrive(y); });

        cmi::batch::Runner runner(a.ctx->ctx);
        std::vector<Active> active;
        std::vector<cmi::batch::Member> members;
        std::size_t next = 0, completed = 0, used_bytes = 0, peak = 0, peak_bytes = 0;
        std::uint64_t steps = 0;
        // Each request is priced once, however many steps it waits.
        std::vector<std::size_t> price(a.n, 0);
        std::vector<bool> priced(a.n, false);
        std::vector<std::size_t> cand, admit;
        std::vector<Active> fresh;
        auto const t0 = Clock::now();
        auto elapsed = [&] { return std::chrono::duration<double>(Clock::now() - t0).count(); };

        while (completed < a.n) {
          double const now = elapsed();
          // Admission: arrived, a slot free, and the device budget fits. The
          // candidates are priced, and the admitted ones' tapes compiled and
          // uploaded, on the runner's host workers (a long tape takes tens of
          // milliseconds to compile); admission itself stays in arrival order.
          cand.clear();
          while (next + cand.size() < a.n && active.size() + cand.size() < a.so->max_active
                 && arrive(order[next + cand.size()]) <= now) {
            cand.push_back(order[next + cand.size()]);
          }
          if (!cand.empty()) {
            auto price_one = [&](std::size_t k) {
              if (!priced[cand[k]]) {
                price[cand[k]] = resident_bytes(a.problems[cand[k]]->problem, *a.o);
              }
            };
            if (cand.size() == 1) {
              price_one(0);
            } else {
              runner.workers().run(cand.size(), price_one);
            }
            admit.clear();
            for (std::size_t i : cand) {
              priced[i] = true;
              if (a.so->hbm_budget_bytes != 0 && !(active.empty() && admit.empty())
                  && used_bytes + price[i] > a.so->hbm_budget_bytes) {
                break;  // wait for a resident solve to finish
              }
              used_bytes += price[i];
              admit.push_back(i);
            }
            fresh.clear();
            fresh.resize(admit.size());
            auto build_one = [&](std::size_t k) {
              std::size_t const i = admit[k];
              // A tape that cannot be built stops its own solve (DeviceMemory at
              // its first iteration, as a plain solve's first role build would).
              std::unique_ptr<cmi::hip::DeviceTape> tape;
              try {
                tape = std::make_unique<cmi::hip::DeviceTape>(a.ctx->ctx, a.problems[i]->problem);
              } catch (const cmi::ResourceExhausted&) {
                tape.reset();
              } catch (const std::bad_alloc&) {
                tape.reset();
              }
              fresh[k] = Active {i, std::move(tape),
                                 std::make_unique<cmi::search::SolveState>(a.problems[i]->problem, pol, sopt),
                                 price[i]};
            };
            if (admit.size() == 1) {
              build_one(0);
            } else if (!admit.empty()) {
              runner.workers().run(admit.size(), build_one);
            }
            for (Active& f : fresh) {
              active.push_back(std::move(f));
            }
            next += admit.size();
          }
          peak = std::max(peak, active.size());
          peak_bytes = std::max(peak_bytes, used_bytes);
          if (active.empty()) {
            if (next < a.n) {
              wait_until(t0, arrive(order[next]));
            }
            continue;
          }
          members.clear();
          for (Active& r : active) {
            members.push_back({r.state.get(), r.tape.get()});
          }
          if (cmi::batch::step(runner, members, *a.o, *a.ctx->ctx) > 0) {
            ++steps;
          }
          // Retire finished solves; their slots are free for the next admission.
          double const t_done = elapsed();
          for (std::size_t s = active.size(); s-- > 0;) {
            if (!active[s].state->finished()) {
              continue;
            }
            std::size_t const i = active[s].idx;
            cmi::capi::fill_outcome_public(active[s].state->finish(), t_done - arrive(i), &a.outs[i], nullptr,
                                           a.problems[i]->problem.box_bounds.size());
            if (a.finish != nullptr) {
              a.finish[i] = t_done;
            }
            if (a.latency != nullptr) {
              a.latency[i] = t_done - arrive(i);
            }
            used_bytes -= active[s].bytes;
            active.erase(active.begin() + static_cast<std::ptrdiff_t>(s));
            ++completed;
          }
        }
        if (a.stats != nullptr) {
          a.stats->steps = steps;
          a.stats->launches = runner.launches();
          a.stats->peak_active = peak;
          a.stats->wall_seconds = elapsed();
          a.stats->peak_resident_bytes = peak_bytes;
          a.stats->kernel_seconds = runner.kernel_seconds();
          a.stats->kernel_fp64_ops = runner.fp64_ops();
        }
      },
      &a);
}

extern "C" cmi_status cmi_serve_plain(cmi_ctx* ctx, size_t n, const cmi_problem* const* problems,
                                      const double* arrival, const cmi_solve_options* o, cmi_solve_outcome* outs,
                                      double* finish, double* latency, cmi_serve_stats* stats)
{
  struct Args
  {
    cmi_ctx* ctx;
    size_t n;
    const cmi_problem* const* problems;
    const double* arrival;
    const cmi_solve_options* o;
    cmi_solve_outcome* outs;
    double* finish;
    double* latency;
    cmi_serve_stats* stats;
  } a {ctx, n, problems, arrival, o, outs, finish, latency, stats};
  if (ctx == nullptr || o == nullptr || (n > 0 && (problems == nullptr || outs == nullptr))) {
    return CMI_INVALID_HANDLE;
  }
  return cmi::capi::run_guarded(
      [](void* p) {
        Args& a = *static_cast<Args*>(p);
        auto const pol = cmi::capi::make_policy(*a.o);
        std::vector<std::size_t> order(a.n);
        std::iota(order.begin(), order.end(), 0);
        auto arrive = [&](std::size_t i) { return a.arrival != nullptr ? a.arrival[i] : 0.0; };
        std::stable_sort(order.begin(), order.end(), [&](std::size_t x, std::size_t y) { return arrive(x) < arrive(y); });
        auto const t0 = Clock::now();
        auto elapsed = [&] { return std::chrono::duration<double>(Clock::now() - t0).count(); };
        std::uint64_t iters = 0;
        for (std::size_t i : order) {
          if (elapsed() < arrive(i)) {
            wait_until(t0, arrive(i));
          }
          auto factory = std::make_shared<cmi::hip::HipBackendFactory>(a.ctx->ctx);
          std::optional<double> known;
          if (a.o->has_known_primal) {
            known = a.o->known_primal;
          }
          cmi::search::SearchDriver drv(pol, factory, a.o->iter_limit, a.o->tolerance, a.o->sample_points,
                                        a.o->budget_bytes, a.o->host_budget_bytes,
                                        a.o->compact_frontier ? cmi::search::FrontierPolicy::Compact
                                                              : cmi::search::FrontierPolicy::StopAtBudget,
                                        known);
          auto const res = drv.solve(a.problems[i]->problem);
          double const t_done = elapsed();
          iters += res.counters.iterations;
          cmi::capi::fill_outcome_public(res, t_done - arrive(i), &a.outs[i], nullptr,
                                         a.problems[i]->problem.box_bounds.size());
          if (a.finish != nullptr) {
            a.finish[i] = t_done;
          }
          if (a.latency != nullptr) {
            a.latency[i] = t_done - arrive(i);
          }
        }
        if (a.stats != nullptr) {
          a.stats->steps = iters;
          a.stats->launches = 0;
          a.stats->peak_active = a.n > 0 ? 1 : 0;
          a.stats->wall_seconds = elapsed();
          a.stats->peak_resident_bytes = 0;
        }
      },
      &a);
}
