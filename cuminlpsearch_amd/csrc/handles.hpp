#pragma once

#include <memory>

#include "cmi.h"
#include "cmi/gams.hpp"
#include "cmi/model.hpp"
#include "cmi/region.hpp"
#include "cmi/search.hpp"
#include "aggregate_backend.hpp"
#include "hip_backend.hpp"

// Definitions of the C-ABI's opaque handle types, shared by the C-ABI
// translation units.
struct cmi_problem
{
  cmi::model::Problem<double> problem;
};

struct cmi_gams_model
{
  cmi::gams::ParsedModel model;  // its problem has been moved into a cmi_problem
};

struct cmi_ctx
{
  std::shared_ptr<cmi::hip::Context> ctx;
};

struct cmi_tape
{
  std::shared_ptr<cmi::hip::DeviceTape> tape;
};

struct cmi_role
{
  std::unique_ptr<cmi::hip::RoleImpl> impl;
  cmi::region::Composition composition;
  std::unique_ptr<cmi::hip::AggregateBounderImpl> aggregate;
};

namespace cmi::capi
{
/// Run fn(arg), mapping an escaping exception onto a status and the
/// thread-local error message.
cmi_status run_guarded(void (*fn)(void*), void* arg);
void fill_outcome_public(const cmi::search::SolveOutcome& o, double seconds, cmi_solve_outcome* out, double* best,
                         std::size_t n_vars);
}  // namespace cmi::capi
