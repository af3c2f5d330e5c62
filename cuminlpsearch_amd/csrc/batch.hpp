#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
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
class Runner
{
public:
  explicit Runner(std::shared_ptr<hip::Context> ctx);

  void run(std::span<const Request> reqs);

  std::span<const unsigned char> feasible(std::size_t i) const;
  std::span<const double> obj_lb(std::size_t i) const;
  backend::CandidateResult candidate(std::size_t i) const;
  std::size_t n_elems(std::size_t i) const { return slots_[i].n_elems; }

  std::uint64_t launches() const { return launches_; }

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
  hip::DeviceBuffer staging_d_, results_d_, scratch_d_, done_d_;
  hip::PinnedBuffer<unsigned char> staging_h_, results_h_;
  std::size_t done_capacity_ = 0;
  std::size_t off_obj_lb_ = 0, off_cand_ = 0, off_witness_ = 0, results_bytes_ = 0;
  std::uint64_t launches_ = 0;
};

struct Result
{
  search::SolveOutcome outcome;
  double seconds = 0.0;  // from the batch start to this solve's completion
};

/// Solve every problem with the same options, advancing all unfinished
/// searches one iteration per step.
std::vector<Result> solve_batch(std::shared_ptr<hip::Context> ctx,
                                std::span<const model::Problem<double>* const> problems,
                                const cmi_solve_options& options);

}  // namespace cmi::batch
