#pragma once

// GAMS scalar-format models (the form GAMS Convert writes, as in MINLPLib) read
// into a Problem. Host only.
//
// Same accepted language, diagnostics and lowering rules as the reference
// frontend (include/cuminlp/gams.hpp:1-136, source/gams/frontend.cpp:1-1691,
// source/gams/ast.hpp:1-453): the objective variable is solved for and
// substituted away where an =E= definition, or an inequality tight at every
// optimum, allows it; `x.fx` variables become literals; free variables get a
// +-default_bound box with a warning; `x**0.5` is Sqrt, integer exponents are
// PowN, other exponents exp(c*log(x)); maximisation negates the objective.

#include <cstddef>
#include <filesystem>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

#include "cmi/model.hpp"

namespace cmi::gams
{

/// Direction of the file's Solve statement (reference gams.hpp:14-19).
enum class Sense
{
  Minimise,
  Maximise
};

/// Reference gams.hpp:21-50.
struct ParseOptions
{
  /// Half-width of the box given to a variable with no finite bound.
  double default_bound = 1e6;
  /// `x.fx = v` variables become the literal v (no search dimension).
  bool fold_fixed_to_const = true;
  /// Reject files that declare Binary or Integer variables.
  bool reject_discrete = false;
};

/// A non-fatal observation (reference gams.hpp:52-57). Line 0 = whole model.
struct Diagnostic
{
  int line = 0;
  std::string message;
};

/// Why a file was rejected (reference gams.hpp:59-70); values are the C-ABI's.
enum class ErrorKind
{
  Syntax = 0,
  NonScalar = 1,
  Discrete = 2,
  UnsupportedFunction = 3,
  Unrepresentable = 4,
  Io = 5
};

/// Reference gams.hpp:72-83: what() reads "line N: message".
class ParseError : public std::runtime_error
{
public:
  ParseError(int line, const std::string& message, ErrorKind kind = ErrorKind::Syntax)
      : std::runtime_error("line " + std::to_string(line) + ": " + message), line(line), kind(kind)
  {
  }
  int line;
  ErrorKind kind;
};

/// A parsed model plus what a harness needs to report on it (reference
/// gams.hpp:85-116). If sense is Maximise the problem's objective is already
/// negated; var_names[i] names box_bounds[i]; constraint_names[i] names
/// constraints[i]; initial_point is empty unless the file gave `x.l` levels.
struct ParsedModel
{
  model::Problem<double> problem;
  Sense sense = Sense::Minimise;
  std::vector<std::string> var_names;
  std::vector<std::string> constraint_names;
  std::vector<double> initial_point;
  /// The objective variable could not be solved for and is a search dimension.
  bool objvar_kept = false;
  std::vector<Diagnostic> warnings;
};

/// Parse source text (reference frontend.cpp:1592-1667). Throws ParseError.
ParsedModel parse(std::string_view source, const ParseOptions& options = {});

/// Parse a file (reference frontend.cpp:1669-1681); ErrorKind::Io if unreadable.
ParsedModel parse_file(const std::filesystem::path& path, const ParseOptions& options = {});

/// The solver's machine-readable result line (reference source/gams/solve.cu:
/// 836-867): "RESULT\tsense=min|max\tprimal=<v|none>\tdual=<v|none>", values
/// with 17 significant digits. `lower`/`upper` are the minimisation bounds the
/// search reports; under Maximise both are negated back to the file's sense.
/// primal is "none" while upper is still the no-incumbent sentinel (>= DBL_MAX),
/// dual is "none" while lower is still at lowest().
std::string result_line(Sense sense, double lower, double upper);

}  // namespace cmi::gams
