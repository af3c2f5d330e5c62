#include "aggregate_solve.hpp"

#include <algorithm>
#include <new>
#include <string>
#include <utility>

#include "aggregate_backend.hpp"
#include "cmi/qcqp.hpp"

namespace cmi::hip
{

namespace
{

/// One role of each kind, kept while successive nodes launch the same
/// composition and replaced when it changes.
class AggregateRoles
{
public:
  AggregateRoles(std::shared_ptr<const DeviceTape> tape, std::size_t n, std::size_t k, std::size_t sample_fan_out,
                 std::size_t samples, std::size_t budget_bytes)
      : tape_(std::move(tape))
      , n_(n)
      , k_(k)
      , sample_fan_out_(sample_fan_out)
      , samples_(samples)
      , bounder_budget_(budget_bytes - budget_bytes / 10)
      , sampler_budget_(budget_bytes / 10)
  {
  }

  AggregateBounderImpl& bounder(const region::Composition& c)
  {
    if (!bounder_ || bounder_comp_ != c) {
      rebuilt_ += bounder_ ? 1 : 0;
      bounder_.reset();
      bounder_ = std::make_unique<AggregateBounderImpl>(tape_, c, n_, k_, 1, bounder_budget_);
      bounder_comp_ = c;
      ++built_;
    }
    return *bounder_;
  }

  RoleImpl& sampler(const region::Composition& c)
  {
    if (!sampler_ || sampler_comp_ != c) {
      rebuilt_ += sampler_ ? 1 : 0;
      sampler_.reset();
      sampler_ = std::make_unique<RoleImpl>(tape_, dev::Mode::Sample, c, sample_fan_out_, samples_, sampler_budget_);
      sampler_comp_ = c;
      ++built_;
    }
    return *sampler_;
  }

  std::uint64_t built() const { return built_; }
  std::uint64_t rebuilt() const { return rebuilt_; }

private:
  std::shared_ptr<const DeviceTape> tape_;
  std::size_t n_, k_, sample_fan_out_, samples_;
  std::size_t bounder_budget_, sampler_budget_;
  std::unique_ptr<AggregateBounderImpl> bounder_;
  std::unique_ptr<RoleImpl> sampler_;
  region::Composition bounder_comp_ {}, sampler_comp_ {};
  std::uint64_t built_ = 0, rebuilt_ = 0;
};

aggregate::SelectionRule rule_of(int s)
{
  switch (s) {
    case CMI_SELECT_LEAST_LOWER_BOUND: return aggregate::SelectionRule::LeastLowerBound;
    case CMI_SELECT_REJECT_INDEX: return aggregate::SelectionRule::RejectIndex;
    case CMI_SELECT_REJECT_INDEX_STALE: return aggregate::SelectionRule::RejectIndexStale;
    default: break;
  }
  throw InvalidConfiguration("unknown selection rule " + std::to_string(s));
}

std::size_t largest_power_of_two_within(std::size_t element_bytes, std::size_t budget)
{
  if (element_bytes == 0 || budget < element_bytes) {
    return 0;
  }
  std::size_t n = 1;
  for (std::size_t q = budget / element_bytes; q > 1; q >>= 1) {
    n <<= 1;
  }
  return n;
}

}  // namespace

void aggregate_fit(std::size_t buffer_nodes, std::size_t free_device_bytes, std::size_t branch_fan_out,
                   std::size_t sample_points, std::size_t& partition_budget, std::size_t& sample_fan_out)
{
  auto const total = static_cast<std::size_t>(static_cast<double>(free_device_bytes) * 0.8);
  auto const sampler = static_cast<std::size_t>(static_cast<double>(total) * 0.10);
  std::size_t const bounder = total - sampler;
  partition_budget = largest_power_of_two_within(buffer_nodes * 16 + 33, bounder);
  std::size_t const s = std::max<std::size_t>(sample_points, 1);
  sample_fan_out = std::max<std::size_t>(2, largest_power_of_two_within((buffer_nodes * 8 + 25) * s, sampler));
  sample_fan_out = std::max(sample_fan_out, 2 * branch_fan_out);
}

AggregateRun aggregate_solve(std::shared_ptr<Context> ctx, const model::Problem<double>& problem,
                             const cmi_aggregate_options& o)
{
  if (o.sample_points < 1) {
    throw InvalidConfiguration("sample_points must be at least 1");
  }
  aggregate::WidestBranchPolicy const bound_policy(o.branch_fan_out, o.partition_budget, {}, o.relative_width != 0);
  aggregate::WidestBranchPolicy const sample_policy(o.branch_fan_out, o.sample_fan_out, {}, o.relative_width != 0);
  aggregate::AggregateOptions opt;
  opt.iter_limit = o.iter_limit;
  opt.tolerance = o.tolerance;
  opt.host_budget_bytes = o.host_budget_bytes;
  opt.rule = rule_of(o.selection);
  if (o.has_known_primal) {
    opt.known_primal_bound = o.known_primal;
  }
  problem.validate();
  if (o.kernel < CMI_BOX_AUTO || o.kernel > CMI_BOX_COMPILED) {
    throw InvalidConfiguration("unknown kernel " + std::to_string(o.kernel));
  }
  auto built = std::make_shared<DeviceTape>(ctx, problem);
  built->set_items_engine(static_cast<DeviceTape::BoxEngine>(o.kernel));
  std::shared_ptr<const DeviceTape> const tape = built;
  AggregateRoles roles(tape, o.partition_budget, o.branch_fan_out, o.sample_fan_out, o.sample_points, o.budget_bytes);
  aggregate::AggregateSolveState st(problem, bound_policy, sample_policy, opt);
  try {
    while (const auto* req = st.next()) {
      backend::CandidateResult drawn;
      std::size_t n_samples = 0;
      if (req->sample != nullptr) {
        const region::SlotAssignment& a = req->sample->launch;
        RoleImpl& s = roles.sampler(a.composition);
        s.run(req->box, a.var_ids, a.widths, req->salt);
        drawn = {s.has_candidate(), s.cand_value(), s.witness(), s.cand_index()};
        n_samples = s.n_elems();
      }
      const region::SlotAssignment& b = req->bound->launch;
      AggregateBounderImpl& bounder = roles.bounder(b.composition);
      AggregateBounderImpl::RegionView const view {req->box, b.var_ids, b.widths};
      auto const children = bounder.bound_children({&view, 1});
      st.on_results(drawn, n_samples, children);
    }
  } catch (const backend::OverBudgetError& e) {
    st.stop_over_budget(e.facts());
  } catch (const ResourceExhausted&) {
    st.stop(search::StopReason::DeviceMemory);
  } catch (const std::bad_alloc&) {
    st.stop(search::StopReason::AllocationFailure);
  }
  st.set_graphs_built(roles.built());
  AggregateRun run;
  run.outcome = st.finish();
  run.rule_engaged = st.selection_rule_engaged();
  run.roles_rebuilt = roles.rebuilt();
  return run;
}

}  // namespace cmi::hip
