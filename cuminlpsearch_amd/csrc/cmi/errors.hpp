#pragma once

#include <stdexcept>
#include <string>

// Error taxonomy of the engine. The class names and their meaning follow the
// reference's errors.hpp:11-57 so that a caller catching the reference's types
// finds the same hierarchy here; the C-ABI maps each class onto one status code
// (include/cmi.h, cmi_status).
namespace cmi
{

class error : public std::runtime_error
{
  using std::runtime_error::runtime_error;
};

/// The expression graph is structurally broken (reference errors.hpp:19-22).
class InvalidDAG : public error
{
  using error::error;
};

/// Variables, bounds, kinds, objective or constraints are malformed
/// (reference errors.hpp:25-29).
class InvalidProblem : public error
{
  using error::error;
};

/// A runtime argument does not match an expected size (reference
/// errors.hpp:32-35).
class ShapeMismatch : public error
{
  using error::error;
};

/// A tuning value is outside the range that makes sense (reference
/// errors.hpp:38-48).
class InvalidConfiguration : public error
{
  using error::error;
};

/// A legal request needs more device memory than the budget allows
/// (reference errors.hpp:51-55).
class ResourceExhausted : public error
{
  using error::error;
};

/// A HIP runtime call failed; carries the hipError_t value.
class DeviceError : public error
{
public:
  DeviceError(int code, const std::string& what)
      : error(what)
      , code_(code)
  {
  }

  int code() const { return code_; }

private:
  int code_;
};

}  // namespace cmi
