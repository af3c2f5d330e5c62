#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <thread>
#include <mutex>
#include <functional>
#include <exception>
#include <condition_variable>
#include <atomic>
#include <span>
#include <vector>

#include "cmi.h"
#include "cmi/search.hpp"
#include "hip_backend.hpp"

// Multi-solve execution: many searches advanced in lock step, their device
// work for one step packed into one launch per role kind.
namespace cmi::batch
{

/// One role evaluation inside a batched step.
struct Request
{
  const hip::DeviceTape* tape;
  dev::Mode mode;
  std::span<const interval> box;
  const region::SlotAssignment* assignment;
  std::uint64_t salt;
  std::uint32_t samples;
};

/// Packs any number of role evaluations (any mix of problems and of the three
/// role kinds) into one host->device copy, at most three kernel launches (one
/// per role kind, each covering every request of that kind) and one
/// device->host copy. Buffers grow on demand and are reused across steps.
/// Host threads started once and woken per call: fn(i) for i in [0, n) on
/// them and the caller together. The first exception is rethrown after every
/// thread has stopped working on the call.
class Workers
{
public:
  explicit Workers(std::size_t threads);
  ~Workers();
  Workers(const Workers&) = delete;
  Workers& operator=(const Workers&) = delete;

  void run(std::size_t n, const std::function<void(std::size_t)>& fn);
  std::size_t size() const { return threads_.size() + 1; }

private:
  void loop();
  void work();

  std::mutex mu_;
  std::condition_variable wake_, idle_;
  std::vector<std::thread> threads_;
  const std::function<void(std::size_t)>* fn_ = nullptr;
  std::size_t n_ = 0;
  std::atomic<std::size_t> next_ {0};
  std::uint64_t generation_ = 0;
  std::size_t busy_ = 0;
  bool stop_ = false;
  std::exception_ptr error_;
};

class Runner
{
public:
  explicit Runner(std::shared_ptr<hip::Context> ctx);

  /// One launch per role kind over every request's items, with the waves
  /// per workgroup of each launch from DeviceTape::slice_waves
  /// (CMI_ITEM_WAVES fixes them).
  void run(std::span<const Request> reqs);

  std::span<const unsigned char> feasible(std::size_t i) const;
  std::span<const double> obj_lb(std::size_t i) const;
  backend::CandidateResult candidate(std::size_t i) const;
  std::size_t n_elems(std::size_t i) const { return slots_[i].n_elems; }

  std::uint64_t launches() const { return launches_; }
  /// Algorithmic FP64 work of every run's launches: the whole tape's
  /// interval operations per region of a bound item, its DAG operations per
  /// point of a sampler or enumerator item; and their device time, from HIP
  /// events on the context's stream around every kTimedEvery-th run's
  /// launches, scaled to all runs by their FP64 work (a pair of event records
  /// on every step cost 2 % of the served rate of small requests).
  double kernel_seconds() const;
  double fp64_ops() const { return fp64_ops_; }
  static constexpr std::uint64_t kTimedEvery = 8;
  /// The device's free memory, read on this runner's first call and every
  /// kFreeBytesEvery-th call after it.
  std::size_t free_bytes(hip::Context& ctx);
  static constexpr std::uint64_t kFreeBytesEvery = 32;

  Runner(const Runner&) = delete;
  Runner& operator=(const Runner&) = delete;
  ~Runner();

  /// Host threads for the per-solve halves of a large step (started on first use).
  Workers& workers();

private:
  struct Placement
  {
    std::size_t item;         // index into the items array
    std::size_t region_off;   // bound: offset into feasible/obj_lb
    std::size_t cand_off;     // sample/enumerate: candidate index
    std::size_t witness_off;  // sample/enumerate: offset into witness doubles
    std::size_t n_regions;
    std::size_t n_elems;
    std::size_t n_vars;
  };

  void ensure(hip::DeviceBuffer& d, std::size_t bytes);
  template<typename T>
  void ensure_host(hip::PinnedBuffer<T>& h, std::size_t n);

  std::shared_ptr<hip::Context> ctx_;
  std::vector<Placement> slots_;
  std::vector<hip::DeviceTape::ItemPlan> plans_;  // per launch-ordered item
  hip::DeviceBuffer staging_d_, results_d_, scratch_d_, done_d_;
  hip::DeviceBuffer slice_d_;  // sliced items' per-tile scratch, zero between launches
  std::size_t slice_capacity_ = 0;  // tiles slice_d_ holds
  hip::PinnedBuffer<unsigned char> staging_h_, results_h_;
  std::size_t done_capacity_ = 0;
  std::size_t off_obj_lb_ = 0, off_cand_ = 0, off_witness_ = 0, results_bytes_ = 0;
  std::uint64_t launches_ = 0;
  std::uint64_t runs_ = 0;
  std::uint64_t free_steps_ = 0;
  std::size_t free_bytes_ = 0;
  double fp64_ops_ = 0.0;
  double timed_seconds_ = 0.0, timed_ops_ = 0.0;  // the timed runs
  hipEvent_t ev_[2] = {nullptr, nullptr};  // around a run's launches
  /// A step as one HIP graph (copy in, the launches, copy out), rebuilt when
  /// the step's shape changes and otherwise given each step's parameters.
  struct StepGraph
  {
    std::uint32_t key = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipGraphNode_t h2d = nullptr, d2h = nullptr, ev0 = nullptr, ev1 = nullptr;
    hipGraphNode_t kernel[3] = {nullptr, nullptr, nullptr};
    void reset();
  };
  StepGraph graph_;
  std::unique_ptr<Workers> workers_;
};

/// One solve advanced by the lock-step loops: its search and its device
/// tape (null when the tape could not be built).
struct Member
{
  search::SolveState* state = nullptr;
  const hip::DeviceTape* tape = nullptr;
};

/// One lock-step iteration of every unfinished member: the host part of its
/// iteration, the plain driver's role-build checks for its request (a solve
/// whose roles would not fit, or whose host allocation fails, stops exactly as
/// a plain solve stops, and the others go on), one Runner::run over all the
/// requests, and the results handed back. Returns the requests run.
std::size_t step(Runner& runner, std::span<Member> members, const cmi_solve_options& o, hip::Context& ctx);

/// The host budget of each resident solve: host_budget_bytes as given (the
/// per-solve budget a plain solve has), or, when it is 0, half of MemAvailable
/// read once and shared equally among `residents` solves.
std::size_t shared_host_budget(const cmi_solve_options& o, std::size_t residents);

/// SearchDriver settings from the C options, with the given host budget.
search::SolveOptions search_options(const cmi_solve_options& o, std::size_t host_budget);

struct Result
{
  search::SolveOutcome outcome;
  double seconds = 0.0;  // from the batch start to this solve's completion
};

/// What a batch did: lock-step steps, launches, and the Runner's device time
/// and FP64 work.
struct BatchStats
{
  std::uint64_t steps = 0;
  std::uint64_t launches = 0;
  double wall_seconds = 0.0;
  double kernel_seconds = 0.0;
  double fp64_ops = 0.0;
};

/// Solve every problem with the same options, advancing all unfinished
/// searches one iteration per step.
std::vector<Result> solve_batch(std::shared_ptr<hip::Context> ctx,
                                std::span<const model::Problem<double>* const> problems,
                                const cmi_solve_options& options, BatchStats* stats = nullptr);

}  // namespace cmi::batch
