#pragma once

#include <cstddef>
#include <limits>
#include <cstdint>
#include <memory>
#include <optional>
#include <span>
#include <string>

#include "cmi/errors.hpp"
#include "cmi/interval.hpp"
#include "cmi/model.hpp"
#include "cmi/region.hpp"

// The three backend roles the search asks for -- sampler, bounder,
// enumerator -- and the factory that builds them per Composition. Names,
// result views and ownership follow the reference's plugin boundary
// (backend/backend.hpp:39-287): results are views into role-owned host
// staging, valid until the next call on the same role instance.
namespace cmi::backend
{

struct Region
{
  std::span<const interval> box;
  const region::SlotAssignment& assignment;
};

struct CandidateResult
{
  bool found = false;
  double value {};
  std::span<const double> point;
  std::int64_t index = -1;
};

struct BoundResult
{
  std::span<const unsigned char> feasible;
  std::span<const double> obj_lb;
  std::size_t n_regions = 0;
};

class RegionSampler
{
public:
  virtual ~RegionSampler() = default;
  virtual CandidateResult sample(const Region& region, std::uint64_t salt) = 0;
  virtual std::size_t n_samples() const = 0;
};

class RegionBounder
{
public:
  virtual ~RegionBounder() = default;
  virtual BoundResult bound(const Region& region) = 0;
  virtual std::size_t n_regions() const = 0;
};

class RegionEnumerator
{
public:
  virtual ~RegionEnumerator() = default;
  virtual CandidateResult enumerate(const Region& region) = 0;
  virtual std::size_t n_points() const = 0;
};

struct BackendCapabilities
{
  bool exact_enumeration = true;
  bool deterministic_sampling = true;
};

struct SubdivisionBundle
{
  std::unique_ptr<RegionSampler> sampler;
  std::unique_ptr<RegionBounder> bounder;
  std::unique_ptr<RegionEnumerator> enumerator;
};

struct BuildBudget
{
  std::size_t bytes = 0;  // 0 = what the device reports free
  std::size_t samples_per_region = 1;
  bool report_build = false;
};

struct RoleRequest
{
  bool sampler = false;
  bool bounder = false;
  bool enumerator = false;

  static RoleRequest subdividing() { return {true, true, false}; }
  static RoleRequest enumerating() { return {false, false, true}; }
  static RoleRequest all() { return {true, true, true}; }
};

/// Affine device cost of one role: fixed + per_region * n_regions bytes
/// (reference backend/cost_model.hpp:26-76).
struct RegionCostModel
{
  std::size_t sampler_fixed = 0;
  std::size_t sampler_per_region = 0;
  std::size_t bounder_fixed = 0;
  std::size_t bounder_per_region = 0;
  std::size_t enumerator_fixed = 0;
  std::size_t enumerator_per_region = 0;
  /// Most regions one role launch can index (the kernels address
  /// regions x samples with 32-bit lanes); a wider composition never fits,
  /// whatever the bytes.
  std::size_t max_regions = std::numeric_limits<std::size_t>::max();
};

/// Facts of a build that did not fit, member for member the reference's
/// OverBudget (backend.hpp:236-262): thrown before anything is allocated, so
/// a caller can explain the failure or retry with a smaller composition.
struct OverBudget
{
  std::string role;
  std::size_t needed_bytes = 0;
  std::size_t budget_bytes = 0;
  std::size_t n_regions = 0;
  std::size_t elements_per_region = 1;  // > 1 only for the sampler
  std::size_t bytes_per_element = 0;
  std::string element_breakdown;  // what one element's bytes pay for
  std::string element_inventory;  // what the per-element cost grows with
  RegionCostModel cost;
  region::Composition composition;
  std::optional<region::FanOutSpec> fan_out;
  std::size_t solve_samples_per_region = 1;
  std::size_t sampler_bytes = 0;
};

class OverBudgetError : public ResourceExhausted
{
public:
  explicit OverBudgetError(OverBudget facts)
      : ResourceExhausted(message(facts))
      , facts_(std::move(facts))
  {
  }

  const OverBudget& facts() const { return facts_; }

private:
  static std::string message(const OverBudget& f)
  {
    return "out of device memory for the " + f.role + ": needs " + std::to_string(f.needed_bytes)
        + " bytes against a budget of " + std::to_string(f.budget_bytes) + " (" + std::to_string(f.n_regions)
        + " regions, composition " + region::spell(f.composition) + ")";
  }

  OverBudget facts_;
};

class RegionBackendFactory
{
public:
  virtual ~RegionBackendFactory() = default;
  virtual BackendCapabilities capabilities() const = 0;
  virtual RegionCostModel cost_model(const model::Problem<double>& problem) const = 0;
  virtual SubdivisionBundle build_subdivision(const model::Problem<double>& problem,
                                              const region::Composition& composition, std::size_t n_regions,
                                              const BuildBudget& budget, RoleRequest roles) const = 0;
};

}  // namespace cmi::backend
