#pragma once

// A closed interval [lb, ub]. Same field names and 16-byte layout as the
// reference's cu::interval<double> (its fields are read as .lb/.ub at
// kernels.cuh:426-427 and problem.hpp:329), so a box can cross the C-ABI as a
// plain array of these.
namespace cmi
{

template<typename T>
struct basic_interval
{
  T lb;
  T ub;
};

using interval = basic_interval<double>;

static_assert(sizeof(interval) == 16, "interval must be two packed doubles");

}  // namespace cmi
