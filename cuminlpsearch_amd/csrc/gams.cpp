// GAMS scalar format (what GAMS Convert writes, as in MINLPLib) into a Problem.
//
// Pipeline:
//   1. lex     the whole text once into lexemes, honouring the column-1 rules
//              (`*` comment lines, `$` control lines, `$ontext`..`$offtext`);
//   2. split   the lexemes into statements at `;`;
//   3. record  each statement through a handler picked by its first word:
//              declarations, bounds and levels, equations (parsed by a Pratt
//              parser into an expression arena), the Solve statement;
//   4. fold    `x.fx` variables into literals;
//   5. solve   for the objective variable where an equation allows it;
//   6. emit    the DAG: variables in declaration order, the objective, the
//              constraints, then bounds left on an eliminated objective.
// The language accepted and the lowering rules are the reference frontend's
// (design/GAMS_FRONTEND.md, and its tests in test/source/gams_test.cpp); they
// decide which Problem a file becomes, node for node.

#include "cmi/gams.hpp"

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <functional>
#include <limits>
#include <optional>
#include <sstream>
#include <unordered_map>
#include <unordered_set>
#include <utility>

#include "cmi/errors.hpp"

namespace cmi::gams
{
namespace
{

constexpr double kInfinity = std::numeric_limits<double>::infinity();

/// The comparison every identity below uses: neither side below the other.
/// (It also holds when a side is NaN; the rules are written for that.)
bool equal(double a, double b) { return !(a < b) && !(b < a); }

char lower(char c) { return (c >= 'A' && c <= 'Z') ? static_cast<char>(c - 'A' + 'a') : c; }

std::string lowered(std::string_view s)
{
  std::string out(s.size(), ' ');
  std::transform(s.begin(), s.end(), out.begin(), lower);
  return out;
}

std::string number_text(double v)
{
  char buf[40];
  std::snprintf(buf, sizeof buf, "%.15g", v);
  return buf;
}

// ===========================================================================
// 1. Lexing
// ===========================================================================

enum class Lex : std::uint8_t
{
  Word,    // identifier (lower-cased), or a %macro% as one opaque word
  Number,  // numeric literal
  Plus,
  Minus,
  Times,
  Divide,
  Power,   // **
  Open,
  Close,
  Comma,
  Semicolon,
  Assign,    // =
  Relation,  // =E= =L= =G= =N=
  Range,     // ..
  Dot,
  Eof
};

struct Lexeme
{
  Lex kind = Lex::Eof;
  int line = 0;
  char relation = 0;  // 'e', 'l', 'g' or 'n'
  double number = 0.0;
  std::string word;
};

bool is_letter(char c) { return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z'); }
bool word_start(char c) { return is_letter(c) || c == '_'; }
bool word_part(char c) { return word_start(c) || (c >= '0' && c <= '9'); }
bool is_digit(char c) { return c >= '0' && c <= '9'; }

/// Starts-with, ignoring case.
bool opens_with(std::string_view line, std::string_view prefix)
{
  if (line.size() < prefix.size()) {
    return false;
  }
  for (std::size_t i = 0; i < prefix.size(); ++i) {
    if (lower(line[i]) != prefix[i]) {
      return false;
    }
  }
  return true;
}

class Lexer
{
public:
  explicit Lexer(std::string_view text)
      : text_(text)
  {
  }

  std::vector<Lexeme> run()
  {
    bool skipping_block = false;
    std::size_t start = 0;
    int line = 0;
    while (start <= text_.size()) {
      std::size_t stop = text_.find('\n', start);
      if (stop == std::string_view::npos) {
        stop = text_.size();
      }
      ++line;
      std::string_view const row = text_.substr(start, stop - start);
      start = stop + 1;
      if (skipping_block) {
        skipping_block = !opens_with(row, "$offtext");
      } else if (!row.empty() && row.front() == '$') {
        skipping_block = opens_with(row, "$ontext");
      } else if (row.empty() || row.front() != '*') {
        lex_row(row, line);
      }
    }
    Lexeme eof;
    eof.line = line;
    out_.push_back(eof);
    return std::move(out_);
  }

private:
  /// One line's lexemes; no lexeme spans a line break.
  void lex_row(std::string_view row, int line)
  {
    std::size_t i = 0;
    while (i < row.size()) {
      char const c = row[i];
      if (c == ' ' || c == '\t' || c == '\r' || c == '\f' || c == '\v') {
        ++i;
        continue;
      }
      Lexeme t;
      t.line = line;
      if (word_start(c)) {
        std::size_t j = i + 1;
        while (j < row.size() && word_part(row[j])) {
          ++j;
        }
        t.kind = Lex::Word;
        t.word = lowered(row.substr(i, j - i));
        i = j;
      } else if (is_digit(c) || (c == '.' && i + 1 < row.size() && is_digit(row[i + 1]))) {
        // strtod needs a terminated string: copy the run of characters a
        // decimal literal can contain, and let strtod decide where it ends.
        std::size_t j = i;
        while (j < row.size() && (is_digit(row[j]) || row[j] == '.' || row[j] == 'e' || row[j] == 'E'
                                  || ((row[j] == '+' || row[j] == '-') && (row[j - 1] == 'e' || row[j - 1] == 'E')))) {
          ++j;
        }
        std::string const digits(row.substr(i, j - i));
        char* stop = nullptr;
        t.number = std::strtod(digits.c_str(), &stop);
        if (stop == digits.c_str()) {
          throw ParseError(line, "cannot read the number here");
        }
        t.kind = Lex::Number;
        i += static_cast<std::size_t>(stop - digits.c_str());
      } else if (c == '%') {
        // A compile-time variable such as %NLP%: kept as one word.
        std::size_t const close = row.find('%', i + 1);
        std::size_t const j = close == std::string_view::npos ? row.size() : close + 1;
        t.kind = Lex::Word;
        t.word = lowered(row.substr(i, j - i));
        i = j;
      } else {
        i += punctuation(row, i, t);
      }
      out_.push_back(std::move(t));
    }
  }

  /// The punctuation lexeme at row[i]; returns its length.
  static std::size_t punctuation(std::string_view row, std::size_t i, Lexeme& t)
  {
    char const c = row[i];
    char const next = i + 1 < row.size() ? row[i + 1] : '\0';
    switch (c) {
      case '+': t.kind = Lex::Plus; return 1;
      case '-': t.kind = Lex::Minus; return 1;
      case '/': t.kind = Lex::Divide; return 1;
      case '(': t.kind = Lex::Open; return 1;
      case ')': t.kind = Lex::Close; return 1;
      case ',': t.kind = Lex::Comma; return 1;
      case ';': t.kind = Lex::Semicolon; return 1;
      case '*':
        t.kind = next == '*' ? Lex::Power : Lex::Times;
        return next == '*' ? 2 : 1;
      case '.':
        t.kind = next == '.' ? Lex::Range : Lex::Dot;
        return next == '.' ? 2 : 1;
      case '=':
        if (is_letter(next) && i + 2 < row.size() && row[i + 2] == '=') {
          char const r = lower(next);
          if (r != 'e' && r != 'l' && r != 'g' && r != 'n') {
            throw ParseError(t.line, std::string("=") + next + "= is not a GAMS relation (=E=, =L=, =G=, =N=)");
          }
          t.kind = Lex::Relation;
          t.relation = r;
          return 3;
        }
        t.kind = Lex::Assign;
        return 1;
      case '\'':
      case '"':
        throw ParseError(t.line,
                         "a quoted label or set element: this is indexed GAMS, not the scalar format GAMS "
                         "Convert writes",
                         ErrorKind::NonScalar);
      default: throw ParseError(t.line, std::string("character '") + c + "' has no meaning here");
    }
  }

  std::string_view text_;
  std::vector<Lexeme> out_;
};

// ===========================================================================
// 3a. The expression arena. Children precede parents. Terms are built only
// through the constructors below, which fold literals and apply identities
// that hold for intervals as well as for numbers: x+0, 0+x, x-0, 0-x = -x,
// x-(-y) = x+y, x*1, 1*x, x*(-1), (-1)*x, x/1, x/(-1), -(-x). Never 0*x or
// 0/x, whose interval versions are not 0 on an unbounded box.
// ===========================================================================

enum class Fn : std::uint8_t
{
  Sqr,
  Sqrt,
  Exp,
  Log,
  Log10,
  Log2,
  Abs,
  Sin,
  Cos,
  Tanh,
  Min,
  Max,
  Power
};

enum class Kind : std::uint8_t
{
  Literal,
  Symbol,
  Add,
  Sub,
  Mul,
  Div,
  Neg,
  Call
};

struct Term
{
  Kind kind = Kind::Literal;
  Fn fn = Fn::Sqr;
  int a = -1;  // operands (-1: none)
  int b = -1;
  double value = 0.0;  // Literal
  int symbol = -1;     // Symbol
  int line = 0;

  int operands() const { return a < 0 ? 0 : b < 0 ? 1 : 2; }
};

/// A function of constants, or nothing outside its domain.
std::optional<double> evaluate_call(Fn f, double x, double y)
{
  auto domain = [](bool ok, double v) { return ok ? std::optional<double>(v) : std::nullopt; };
  switch (f) {
    case Fn::Sqr: return x * x;
    case Fn::Sqrt: return domain(x >= 0.0, std::sqrt(x));
    case Fn::Exp: return std::exp(x);
    case Fn::Log: return domain(x > 0.0, std::log(x));
    case Fn::Log10: return domain(x > 0.0, std::log10(x));
    case Fn::Log2: return domain(x > 0.0, std::log2(x));
    case Fn::Abs: return std::fabs(x);
    case Fn::Sin: return std::sin(x);
    case Fn::Cos: return std::cos(x);
    case Fn::Tanh: return std::tanh(x);
    case Fn::Min: return std::fmin(x, y);
    case Fn::Max: return std::fmax(x, y);
    case Fn::Power: return domain(!(std::floor(y) < y && x <= 0.0), std::pow(x, y));
  }
  return std::nullopt;
}

class Arena
{
public:
  const Term& at(int i) const { return terms_[static_cast<std::size_t>(i)]; }
  int count() const { return static_cast<int>(terms_.size()); }

  /// The value of a term with no symbol below it.
  std::optional<double> fixed_value(int i) const
  {
    std::size_t const k = static_cast<std::size_t>(i);
    return known_[k] ? std::optional<double>(values_[k]) : std::nullopt;
  }

  bool is_literal(int i) const { return at(i).kind == Kind::Literal; }
  double literal_of(int i) const { return at(i).value; }

  int literal(double v, int line)
  {
    Term t;
    t.value = v;
    t.line = line;
    return add(t);
  }

  int symbol(int sym, int line)
  {
    Term t;
    t.kind = Kind::Symbol;
    t.symbol = sym;
    t.line = line;
    return add(t);
  }

  int call(Fn f, int x, int y, int line)
  {
    Term t;
    t.kind = Kind::Call;
    t.fn = f;
    t.a = x;
    t.b = y;
    t.line = line;
    return add(t);
  }

  int neg(int x)
  {
    if (is_literal(x)) {
      return literal(-literal_of(x), at(x).line);
    }
    if (at(x).kind == Kind::Neg) {
      return at(x).a;
    }
    return op(Kind::Neg, x, -1);
  }

  int add(int x, int y)
  {
    if (is_literal(x) && is_literal(y)) {
      return literal(literal_of(x) + literal_of(y), at(x).line);
    }
    if (is_literal(y) && equal(literal_of(y), 0.0)) {
      return x;
    }
    if (is_literal(x) && equal(literal_of(x), 0.0)) {
      return y;
    }
    return op(Kind::Add, x, y);
  }

  int sub(int x, int y)
  {
    if (is_literal(x) && is_literal(y)) {
      return literal(literal_of(x) - literal_of(y), at(x).line);
    }
    if (is_literal(y) && equal(literal_of(y), 0.0)) {
      return x;
    }
    if (is_literal(x) && equal(literal_of(x), 0.0)) {
      return neg(y);
    }
    if (at(y).kind == Kind::Neg) {
      return add(x, at(y).a);
    }
    return op(Kind::Sub, x, y);
  }

  int mul(int x, int y)
  {
    if (is_literal(x) && is_literal(y)) {
      return literal(literal_of(x) * literal_of(y), at(x).line);
    }
    for (double unit : {1.0, -1.0}) {
      if (is_literal(y) && equal(literal_of(y), unit)) {
        return unit > 0.0 ? x : neg(x);
      }
      if (is_literal(x) && equal(literal_of(x), unit)) {
        return unit > 0.0 ? y : neg(y);
      }
    }
    return op(Kind::Mul, x, y);
  }

  int div(int x, int y)
  {
    if (is_literal(x) && is_literal(y) && !equal(literal_of(y), 0.0)) {
      return literal(literal_of(x) / literal_of(y), at(x).line);
    }
    if (is_literal(y) && equal(literal_of(y), 1.0)) {
      return x;
    }
    if (is_literal(y) && equal(literal_of(y), -1.0)) {
      return neg(x);
    }
    return op(Kind::Div, x, y);
  }

  /// Every Symbol term whose symbol has a value in `values` becomes that
  /// literal, in place; values of constant subtrees are then recomputed. The
  /// terms above keep the shape they were built with.
  void substitute(const std::vector<std::optional<double>>& values)
  {
    for (Term& t : terms_) {
      if (t.kind != Kind::Symbol) {
        continue;
      }
      auto const s = static_cast<std::size_t>(t.symbol);
      if (s < values.size() && values[s]) {
        t.kind = Kind::Literal;
        t.value = *values[s];
        t.symbol = -1;
      }
    }
    for (std::size_t i = 0; i < terms_.size(); ++i) {
      settle(i);
    }
    mentions_for_ = -1;
  }

  /// Whether term i has symbol `sym` below it. One forward sweep per symbol
  /// asked about, extended as the arena grows.
  bool mentions(int i, int sym) const
  {
    if (mentions_for_ != sym) {
      mentions_for_ = sym;
      mention_.clear();
    }
    for (std::size_t k = mention_.size(); k < terms_.size(); ++k) {
      const Term& t = terms_[k];
      bool m = t.kind == Kind::Symbol && t.symbol == sym;
      if (!m && t.a >= 0) {
        m = mention_[static_cast<std::size_t>(t.a)] != 0;
      }
      if (!m && t.b >= 0) {
        m = mention_[static_cast<std::size_t>(t.b)] != 0;
      }
      mention_.push_back(m ? 1 : 0);
    }
    return mention_[static_cast<std::size_t>(i)] != 0;
  }

private:
  int op(Kind k, int x, int y)
  {
    Term t;
    t.kind = k;
    t.a = x;
    t.b = y;
    t.line = at(x).line;
    return add(t);
  }

  int add(const Term& t)
  {
    terms_.push_back(t);
    values_.push_back(0.0);
    known_.push_back(0);
    settle(terms_.size() - 1);
    return static_cast<int>(terms_.size()) - 1;
  }

  /// Constant propagation for term k from its operands' values.
  void settle(std::size_t k)
  {
    const Term& t = terms_[k];
    std::optional<double> v;
    if (t.kind == Kind::Literal) {
      v = t.value;
    } else if (t.kind != Kind::Symbol) {
      bool const have_a = known_[static_cast<std::size_t>(t.a)] != 0;
      bool const have_b = t.b < 0 || known_[static_cast<std::size_t>(t.b)] != 0;
      if (have_a && have_b) {
        double const x = values_[static_cast<std::size_t>(t.a)];
        double const y = t.b < 0 ? 0.0 : values_[static_cast<std::size_t>(t.b)];
        switch (t.kind) {
          case Kind::Add: v = x + y; break;
          case Kind::Sub: v = x - y; break;
          case Kind::Mul: v = x * y; break;
          case Kind::Div:
            if (!equal(y, 0.0)) {
              v = x / y;
            }
            break;
          case Kind::Neg: v = -x; break;
          case Kind::Call: v = evaluate_call(t.fn, x, y); break;
          default: break;
        }
      }
    }
    known_[k] = v.has_value() ? 1 : 0;
    values_[k] = v.value_or(0.0);
  }

  std::vector<Term> terms_;
  std::vector<double> values_;
  std::vector<char> known_;
  mutable std::vector<char> mention_;
  mutable int mentions_for_ = -1;
};

// ===========================================================================
// 3b. What the file says.
// ===========================================================================

struct Variable
{
  std::string name;
  int declared_at = 0;
  double lo = -kInfinity;
  double up = kInfinity;
  std::optional<double> fixed;
  std::optional<double> level;
  model::VarKind kind = model::VarKind::Continuous;
  int discrete_at = 0;  // the declaration that made it Integer or Binary
  bool substituted = false;
};

struct EquationDef
{
  std::string name;
  int line = 0;
  char relation = 'e';
  int lhs = -1;
  int rhs = -1;
  bool used_for_objective = false;
};

struct Model
{
  Arena arena;
  std::vector<Variable> vars;
  std::unordered_map<std::string, int> index;
  std::vector<EquationDef> equations;
  std::optional<int> objective_var;
  bool maximise = false;
  int solve_line = 0;
};

// ===========================================================================
// 2 + 3c. Statements.
// ===========================================================================

struct FunctionInfo
{
  Fn fn;
  int arity;  // 0: two or more, folded left (min, max)
};

const std::unordered_map<std::string, FunctionInfo>& known_functions()
{
  static const std::unordered_map<std::string, FunctionInfo> table {
      {"sqr", {Fn::Sqr, 1}},     {"sqrt", {Fn::Sqrt, 1}},    {"exp", {Fn::Exp, 1}},       {"log", {Fn::Log, 1}},
      {"log10", {Fn::Log10, 1}}, {"log2", {Fn::Log2, 1}},    {"abs", {Fn::Abs, 1}},       {"sin", {Fn::Sin, 1}},
      {"cos", {Fn::Cos, 1}},     {"tanh", {Fn::Tanh, 1}},    {"min", {Fn::Min, 0}},       {"max", {Fn::Max, 0}},
      {"power", {Fn::Power, 2}}, {"rpower", {Fn::Power, 2}},
  };
  return table;
}

/// GAMS functions recognised but refused: no exact interval form is built for
/// them, and a looser rewrite would silently weaken bounds.
bool refused(const std::string& name)
{
  static const std::unordered_set<std::string> names {
      "tan",   "arctan", "arcsin", "arccos",  "errorf", "sign",     "signpower", "ceil",   "floor",   "round",
      "trunc", "mod",    "div",    "centropy", "sinh",  "cosh",     "gamma",     "loggamma", "beta",  "entropy",
      "sigmoid", "slexp", "sqexp", "edist",   "poly",   "ifthen",   "vcpower",   "cvpower", "ncpcm",
  };
  return names.count(name) != 0;
}

/// Words that only appear in indexed GAMS.
bool indexed_gams(const std::string& w)
{
  static const std::unordered_set<std::string> words {
      "set",  "sets",    "parameter", "parameters", "table", "alias", "scalar", "scalars", "loop",  "sum",
      "prod", "smax",    "smin",      "abort",      "display", "file", "put",   "repeat",  "while", "for",
  };
  return words.count(w) != 0;
}

/// One statement's lexemes, read front to back. Past the end it keeps
/// returning the statement's terminator (its `;`, or the end of the file), so
/// errors there carry that lexeme's line.
class Cursor
{
public:
  Cursor(const std::vector<Lexeme>& all, std::size_t first, std::size_t end)
      : all_(all)
      , pos_(first)
      , end_(end)
  {
  }

  const Lexeme& peek(std::size_t k = 0) const { return all_[std::min(pos_ + k, end_)]; }
  const Lexeme& take()
  {
    const Lexeme& t = peek();
    if (pos_ < end_) {
      ++pos_;
    }
    return t;
  }
  bool at(Lex k) const { return peek().kind == k; }
  bool at_word(std::string_view w) const { return at(Lex::Word) && peek().word == w; }
  bool done() const { return pos_ >= end_; }

  void expect(Lex k, const char* what)
  {
    if (!at(k)) {
      throw ParseError(peek().line, std::string("expected ") + what);
    }
    take();
  }

  void expect_end(const char* statement)
  {
    if (!done()) {
      throw ParseError(peek().line, std::string("expected ';' to end ") + statement);
    }
  }

private:
  const std::vector<Lexeme>& all_;
  std::size_t pos_;
  std::size_t end_;  // index of the terminator lexeme
};

/// Pratt parser over one statement's lexemes, building into the arena.
/// Binding powers: + and - 10, * and / 20, a sign 25 (so -x*y is (-x)*y and
/// -x**2 is -(x**2)), ** 30 and right-associative with a signed exponent.
class ExpressionParser
{
public:
  ExpressionParser(Model& m, Cursor& c)
      : m_(m)
      , c_(c)
  {
  }

  int parse(int min_power = 0)
  {
    int left = prefix();
    for (;;) {
      Lex const k = c_.peek().kind;
      int const power = infix_power(k);
      if (power <= min_power) {
        return left;
      }
      int const line = c_.take().line;
      if (k == Lex::Power) {
        int const exponent = parse(power - 1);
        left = m_.arena.call(Fn::Power, left, exponent, line);
        continue;
      }
      int const right = parse(power);
      Arena& a = m_.arena;
      left = k == Lex::Plus ? a.add(left, right)
          : k == Lex::Minus ? a.sub(left, right)
          : k == Lex::Times ? a.mul(left, right)
                            : a.div(left, right);
    }
  }

private:
  static constexpr int kSign = 25;

  static int infix_power(Lex k)
  {
    switch (k) {
      case Lex::Plus:
      case Lex::Minus: return 10;
      case Lex::Times:
      case Lex::Divide: return 20;
      case Lex::Power: return 30;
      default: return 0;
    }
  }

  int prefix()
  {
    const Lexeme& t = c_.peek();
    switch (t.kind) {
      case Lex::Minus: c_.take(); return m_.arena.neg(parse(kSign));
      case Lex::Plus: c_.take(); return parse(kSign);
      case Lex::Number: c_.take(); return m_.arena.literal(t.number, t.line);
      case Lex::Open: {
        c_.take();
        int const inner = parse();
        c_.expect(Lex::Close, "')'");
        return inner;
      }
      case Lex::Word: return word();
      default: throw ParseError(t.line, "expected a number, a name or '('");
    }
  }

  int word()
  {
    Lexeme const t = c_.take();
    if (t.word == "inf") {
      return m_.arena.literal(kInfinity, t.line);
    }
    if (t.word == "eps") {
      return m_.arena.literal(0.0, t.line);
    }
    if (t.word == "na" || t.word == "undf") {
      throw ParseError(t.line, "GAMS special value '" + t.word + "' has no numeric meaning",
                       ErrorKind::Unrepresentable);
    }
    if (c_.at(Lex::Open)) {
      return function(t);
    }
    auto const found = m_.index.find(t.word);
    if (found == m_.index.end()) {
      throw ParseError(t.line, "'" + t.word + "' is not a declared variable");
    }
    return m_.arena.symbol(found->second, t.line);
  }

  int function(const Lexeme& name)
  {
    auto const f = known_functions().find(name.word);
    if (f == known_functions().end()) {
      throw ParseError(name.line,
                       refused(name.word) ? "GAMS function '" + name.word + "' has no interval form here"
                                          : "'" + name.word + "' is not a function this reader knows",
                       ErrorKind::UnsupportedFunction);
    }
    c_.take();  // '('
    std::vector<int> args;
    args.push_back(parse());
    while (c_.at(Lex::Comma)) {
      c_.take();
      args.push_back(parse());
    }
    c_.expect(Lex::Close, "')' after the arguments");
    FunctionInfo const info = f->second;
    if (info.arity == 0) {
      int acc = args.front();
      for (std::size_t i = 1; i < args.size(); ++i) {
        acc = m_.arena.call(info.fn, acc, args[i], name.line);
      }
      return acc;
    }
    if (static_cast<int>(args.size()) != info.arity) {
      throw ParseError(name.line, "'" + name.word + "' takes " + std::to_string(info.arity) + " argument(s), got "
                                      + std::to_string(args.size()));
    }
    return m_.arena.call(info.fn, args[0], info.arity == 2 ? args[1] : -1, name.line);
  }

  Model& m_;
  Cursor& c_;
};

class StatementReader
{
public:
  explicit StatementReader(Model& m)
      : m_(m)
  {
    using Handler = void (StatementReader::*)(Cursor&);
    for (const char* w : {"free", "positive", "negative", "binary", "integer", "variable", "variables"}) {
      handlers_.emplace(w, static_cast<Handler>(&StatementReader::declare_variables));
    }
    for (const char* w : {"equation", "equations"}) {
      handlers_.emplace(w, static_cast<Handler>(&StatementReader::declare_equations));
    }
    for (const char* w : {"model", "models", "option", "options"}) {
      handlers_.emplace(w, static_cast<Handler>(&StatementReader::ignore));
    }
    handlers_.emplace("solve", static_cast<Handler>(&StatementReader::solve));
  }

  /// Every statement of the lexed file, in order.
  void read(const std::vector<Lexeme>& lexemes)
  {
    std::size_t first = 0;
    while (lexemes[first].kind != Lex::Eof) {
      std::size_t end = first;
      while (lexemes[end].kind != Lex::Semicolon && lexemes[end].kind != Lex::Eof) {
        ++end;
      }
      Cursor c(lexemes, first, end);
      if (!c.done()) {
        statement(c);
      }
      first = lexemes[end].kind == Lex::Eof ? end : end + 1;
    }
  }

private:
  void statement(Cursor& c)
  {
    const Lexeme& head = c.peek();
    if (head.kind != Lex::Word) {
      throw ParseError(head.line, "a statement must start with a name");
    }
    if (indexed_gams(head.word)) {
      throw ParseError(head.line,
                       "'" + head.word + "' belongs to indexed GAMS; convert the model to scalar form with GAMS "
                                         "Convert",
                       ErrorKind::NonScalar);
    }
    if (head.word == "sos1" || head.word == "sos2" || head.word == "semicont" || head.word == "semiint") {
      throw ParseError(head.line, "'" + head.word + "' variables have no representation here", ErrorKind::Discrete);
    }
    auto const h = handlers_.find(head.word);
    if (h != handlers_.end()) {
      (this->*(h->second))(c);
    } else if (c.peek(1).kind == Lex::Range) {
      define_equation(c);
    } else if (c.peek(1).kind == Lex::Dot) {
      assign_attribute(c);
    } else {
      throw ParseError(head.line, "no statement begins with '" + head.word + "'");
    }
  }

  void ignore(Cursor&) {}

  /// `[free|positive|negative|binary|integer] variable(s) a, b, ...`. A
  /// repeated declaration resets the bounds to its domain's and keeps the
  /// first position; integrality, once declared, stays.
  void declare_variables(Cursor& c)
  {
    Lexeme const head = c.take();
    if (c.at_word("variable") || c.at_word("variables")) {
      c.take();
    }
    double lo = -kInfinity;
    double up = kInfinity;
    std::optional<model::VarKind> discrete;
    if (head.word == "positive") {
      lo = 0.0;
    } else if (head.word == "negative") {
      up = 0.0;
    } else if (head.word == "binary") {
      lo = 0.0;
      up = 1.0;
      discrete = model::VarKind::Binary;
    } else if (head.word == "integer") {
      lo = 0.0;
      discrete = model::VarKind::Integer;
    }
    do {
      if (!c.at(Lex::Word)) {
        throw ParseError(c.peek().line, "expected a variable name");
      }
      Variable& v = m_.vars[static_cast<std::size_t>(variable(c.take().word, head.line))];
      v.lo = lo;
      v.up = up;
      if (discrete) {
        v.kind = *discrete;
        v.discrete_at = head.line;
      }
    } while (c.at(Lex::Comma) && (c.take(), true));
    c.expect_end("the variable declaration");
  }

  /// `equation(s) a, b, ...`: the names are not needed until a definition.
  void declare_equations(Cursor& c)
  {
    c.take();
    do {
      if (!c.at(Lex::Word)) {
        throw ParseError(c.peek().line, "expected an equation name");
      }
      c.take();
    } while (c.at(Lex::Comma) && (c.take(), true));
    c.expect_end("the equation declaration");
  }

  /// `name.. lhs =R= rhs`
  void define_equation(Cursor& c)
  {
    EquationDef e;
    Lexeme const name = c.take();
    e.name = name.word;
    e.line = name.line;
    c.take();  // ..
    ExpressionParser p(m_, c);
    e.lhs = p.parse();
    if (!c.at(Lex::Relation)) {
      throw ParseError(c.peek().line, "expected =E=, =L=, =G= or =N= between the two sides");
    }
    e.relation = c.take().relation;
    e.rhs = p.parse();
    c.expect_end("the equation");
    m_.equations.push_back(std::move(e));
  }

  /// `name.attribute = constant`. Only lo, up, fx and l of a declared variable
  /// are read; other attributes (marginals, model options) are passed over.
  void assign_attribute(Cursor& c)
  {
    Lexeme const name = c.take();
    c.take();  // .
    if (!c.at(Lex::Word)) {
      throw ParseError(name.line, "expected an attribute name after '.'");
    }
    std::string const attr = c.take().word;
    auto const found = m_.index.find(name.word);
    static const std::unordered_set<std::string> read {"lo", "up", "fx", "l"};
    if (found == m_.index.end() || read.count(attr) == 0) {
      return;
    }
    c.expect(Lex::Assign, "'=' after the attribute");
    ExpressionParser p(m_, c);
    std::optional<double> const value = m_.arena.fixed_value(p.parse());
    if (!value) {
      throw ParseError(name.line, "'" + name.word + "." + attr + "' must be given a constant");
    }
    c.expect_end("the assignment");
    Variable& v = m_.vars[static_cast<std::size_t>(found->second)];
    if (attr == "lo") {
      v.lo = *value;
    } else if (attr == "up") {
      v.up = *value;
    } else if (attr == "fx") {
      v.fixed = *value;
    } else {
      v.level = *value;
    }
  }

  /// `solve m using TYPE minimizing|maximizing objvar`: only the direction
  /// and the objective variable matter.
  void solve(Cursor& c)
  {
    int const line = c.take().line;
    static const std::unordered_map<std::string, bool> direction {
        {"minimizing", false}, {"minimising", false}, {"maximizing", true}, {"maximising", true}};
    bool found = false;
    while (!c.done()) {
      Lexeme const t = c.take();
      auto const d = t.kind == Lex::Word ? direction.find(t.word) : direction.end();
      if (d == direction.end()) {
        continue;
      }
      if (!c.at(Lex::Word)) {
        throw ParseError(c.peek().line, "expected the objective variable after '" + t.word + "'");
      }
      Lexeme const v = c.take();
      auto const idx = m_.index.find(v.word);
      if (idx == m_.index.end()) {
        throw ParseError(v.line, "the objective variable '" + v.word + "' is not declared");
      }
      m_.objective_var = idx->second;
      m_.maximise = d->second;
      found = true;
    }
    if (!found) {
      throw ParseError(line, "the Solve statement has no objective variable; a model with no objective cannot "
                             "be searched",
                       ErrorKind::Unrepresentable);
    }
    m_.solve_line = line;
  }

  int variable(const std::string& name, int line)
  {
    auto const [it, fresh] = m_.index.emplace(name, static_cast<int>(m_.vars.size()));
    if (fresh) {
      Variable v;
      v.name = name;
      v.declared_at = line;
      m_.vars.push_back(std::move(v));
    }
    return it->second;
  }

  Model& m_;
  std::unordered_map<std::string, void (StatementReader::*)(Cursor&)> handlers_;
};

// ===========================================================================
// 5. Solving for the objective variable.
// ===========================================================================

/// An equation rearranged as c * objvar + lhs_rest (relation) rhs_rest.
struct Collected
{
  double c;
  int lhs_rest;
  int rhs_rest;
};

/// Term i as coefficient * sym + rest with rest free of sym, for every term
/// up to `root`, in one forward sweep (children come first, so no recursion
/// on deep expressions). Pass one finds the coefficients: sym passes only
/// through +, -, negation, a product with a constant factor and a quotient by
/// a non-zero constant. Pass two builds the rests of the terms that mention
/// sym; a term that does not mention it is its own rest.
std::optional<std::pair<double, int>> split(Arena& a, int root, int sym)
{
  std::size_t const n = static_cast<std::size_t>(root) + 1;
  std::vector<double> coeff(n, 0.0);
  std::vector<char> linear(n, 1);
  for (std::size_t k = 0; k < n; ++k) {
    int const i = static_cast<int>(k);
    if (!a.mentions(i, sym)) {
      continue;
    }
    const Term& t = a.at(i);
    auto lin = [&](int j) { return linear[static_cast<std::size_t>(j)] != 0; };
    auto co = [&](int j) { return coeff[static_cast<std::size_t>(j)]; };
    bool ok = false;
    double c = 0.0;
    switch (t.kind) {
      case Kind::Symbol: ok = true; c = 1.0; break;
      case Kind::Neg: ok = lin(t.a); c = -co(t.a); break;
      case Kind::Add: ok = lin(t.a) && lin(t.b); c = co(t.a) + co(t.b); break;
      case Kind::Sub: ok = lin(t.a) && lin(t.b); c = co(t.a) - co(t.b); break;
      case Kind::Mul: {
        bool const left_has = a.mentions(t.a, sym);
        int const factor = left_has ? t.b : t.a;
        int const carrier = left_has ? t.a : t.b;
        auto const k_value = a.fixed_value(factor);
        ok = !a.mentions(factor, sym) && k_value.has_value() && lin(carrier);
        c = ok ? *k_value * co(carrier) : 0.0;
        break;
      }
      case Kind::Div: {
        auto const d = a.fixed_value(t.b);
        ok = !a.mentions(t.b, sym) && d.has_value() && !equal(*d, 0.0) && lin(t.a);
        c = ok ? co(t.a) / *d : 0.0;
        break;
      }
      default: break;
    }
    linear[k] = ok ? 1 : 0;
    coeff[k] = c;
  }
  if (!linear[n - 1]) {
    return std::nullopt;
  }
  std::vector<int> rest(n, -1);
  for (std::size_t k = 0; k < n; ++k) {
    int const i = static_cast<int>(k);
    if (!a.mentions(i, sym)) {
      rest[k] = i;
      continue;
    }
    if (!linear[k]) {
      continue;
    }
    Term const t = a.at(i);  // a copy: building below grows the arena
    auto r = [&](int j) { return rest[static_cast<std::size_t>(j)]; };
    switch (t.kind) {
      case Kind::Symbol: rest[k] = a.literal(0.0, 0); break;
      case Kind::Neg: rest[k] = a.neg(r(t.a)); break;
      case Kind::Add: rest[k] = a.add(r(t.a), r(t.b)); break;
      case Kind::Sub: rest[k] = a.sub(r(t.a), r(t.b)); break;
      case Kind::Mul:
        rest[k] = a.mentions(t.a, sym) ? a.mul(t.b, r(t.a)) : a.mul(t.a, r(t.b));
        break;
      case Kind::Div: rest[k] = a.div(r(t.a), t.b); break;
      default: break;
    }
  }
  return std::make_pair(coeff[n - 1], rest[n - 1]);
}

std::optional<Collected> collect(Model& m, const EquationDef& e, int sym)
{
  Arena& a = m.arena;
  if (!a.mentions(e.lhs, sym) && !a.mentions(e.rhs, sym)) {
    return std::nullopt;
  }
  auto const left = split(a, e.lhs, sym);
  auto const right = split(a, e.rhs, sym);
  if (!left || !right) {
    return std::nullopt;
  }
  double const c = left->first - right->first;
  if (std::fabs(c) < 1e-12) {
    return std::nullopt;  // objvar cancels
  }
  return Collected {c, left->second, right->second};
}

/// objvar = (rhs_rest - lhs_rest) / c
int solved_for(Arena& a, const Collected& q)
{
  return a.div(a.sub(q.rhs_rest, q.lhs_rest), a.literal(q.c, 0));
}

struct ObjectiveChoice
{
  std::optional<int> term;  // the objective expression, if objvar was solved for
  bool via_inequality = false;
};

/// The first =E= linear in objvar defines it. Failing that, an =L=/=G= that
/// bounds objvar against the direction of optimisation (from below when
/// minimising, from above when maximising), and is the only equation that
/// mentions objvar, holds with equality at every optimum and defines it too.
ObjectiveChoice choose_objective(Model& m, std::vector<Diagnostic>& warnings)
{
  int const sym = *m.objective_var;
  ObjectiveChoice out;
  int definitions = 0;
  for (EquationDef& e : m.equations) {
    if (e.relation != 'e') {
      continue;
    }
    auto const q = collect(m, e, sym);
    if (!q) {
      continue;
    }
    if (++definitions == 1) {
      out.term = solved_for(m.arena, *q);
      e.used_for_objective = true;
    }
  }
  if (definitions > 1) {
    warnings.push_back({0, "several equations define the objective variable; the first one is used"});
  }
  if (out.term) {
    return out;
  }
  for (std::size_t i = 0; i < m.equations.size(); ++i) {
    EquationDef& e = m.equations[i];
    if (e.relation != 'l' && e.relation != 'g') {
      continue;
    }
    auto const q = collect(m, e, sym);
    if (!q) {
      continue;
    }
    // c*objvar <= ... bounds objvar from above for c > 0 and from below for c < 0
    bool const bounds_from_below = (e.relation == 'l') == (q->c < 0.0);
    if (bounds_from_below == m.maximise) {
      continue;
    }
    bool mentioned_elsewhere = false;
    for (std::size_t j = 0; j < m.equations.size(); ++j) {
      const EquationDef& o = m.equations[j];
      if (j != i && o.relation != 'n' && (m.arena.mentions(o.lhs, sym) || m.arena.mentions(o.rhs, sym))) {
        mentioned_elsewhere = true;
        break;
      }
    }
    if (mentioned_elsewhere) {
      continue;
    }
    out.term = solved_for(m.arena, *q);
    out.via_inequality = true;
    e.used_for_objective = true;
    warnings.push_back({e.line, "'" + m.vars[static_cast<std::size_t>(sym)].name + "' is defined by inequality '"
                                    + e.name + "', which holds with equality at every optimum; substituted it"});
    return out;
  }
  return out;
}

// ===========================================================================
// 6. The DAG.
// ===========================================================================

class Emitter
{
public:
  Emitter(Model& m, const ParseOptions& opt, ParsedModel& out)
      : m_(m)
      , opt_(opt)
      , out_(out)
      , var_node_(m.vars.size(), kUnset)
  {
  }

  /// Variables in declaration order, except the substituted objective and
  /// (when folding) fixed ones. Unbounded sides get +-default_bound with a
  /// warning; discrete boxes snap inward to integers.
  void variables()
  {
    bool const with_levels = std::any_of(m_.vars.begin(), m_.vars.end(), [](const Variable& v) { return v.level; });
    for (std::size_t i = 0; i < m_.vars.size(); ++i) {
      const Variable& v = m_.vars[i];
      if (v.substituted || (v.fixed && opt_.fold_fixed_to_const)) {
        continue;
      }
      double lo = v.fixed.value_or(v.lo);
      double up = v.fixed.value_or(v.up);
      bool const open_lo = !std::isfinite(lo);
      bool const open_up = !std::isfinite(up);
      lo = open_lo ? -opt_.default_bound : lo;
      up = open_up ? opt_.default_bound : up;
      bool const discrete = v.kind != model::VarKind::Continuous;
      if (discrete) {
        lo = std::ceil(lo);
        up = std::floor(up);
      }
      if (up < lo) {
        throw ParseError(v.declared_at, "'" + v.name + "' has no value in its bounds [" + number_text(lo) + ", "
                                            + number_text(up) + "]");
      }
      std::string const note = discrete ? " (integer domain size " + number_text(up - lo + 1.0) + ")" : "";
      if (open_lo) {
        warn(v.declared_at, "'" + v.name + "' is unbounded below; using " + number_text(lo) + note);
      }
      if (open_up) {
        warn(v.declared_at, "'" + v.name + "' is unbounded above; using " + number_text(up) + note);
      }
      model::VarKind kind = v.kind;
      bool const zero_one = equal(lo, 0.0) && equal(up, 1.0);
      if (kind == model::VarKind::Binary && !zero_one) {
        kind = model::VarKind::Integer;
      }
      var_node_[i] = out_.problem.var(lo, up, kind).id();
      out_.var_names.push_back(v.name);
      if (with_levels) {
        double const start = v.level.value_or(std::clamp(0.0, lo, up));
        out_.initial_point.push_back(std::clamp(start, lo, up));
      }
    }
  }

  /// The objective; any mention of a substituted objective variable left in
  /// the constraints then reads its (un-negated) value.
  void objective(int term)
  {
    nodes_for({term});
    std::size_t const value = node_[static_cast<std::size_t>(term)];
    if (dag().nodes[value].op == model::Op::Const) {
      throw ParseError(0, "the objective has no variable in it", ErrorKind::Unrepresentable);
    }
    objective_value_ = value;
    var_node_[static_cast<std::size_t>(*m_.objective_var)] = value;
    std::size_t root = value;
    if (m_.maximise) {
      root = dag().emit(model::Op::Neg, {value});
      out_.sense = Sense::Maximise;
    }
    out_.problem.objective_root = root;
  }

  /// Every equation not used for the objective, as root (<= or ==) rhs:
  /// a >= b becomes b <= a; a constant right side is the rhs, otherwise the
  /// root is lhs - rhs against 0. =N= and variable-free equations are dropped
  /// with a warning.
  void constraints()
  {
    struct Pending
    {
      const EquationDef* e;
      model::Cmp cmp;
      int left;
      std::optional<int> right;
      double rhs;
    };
    std::vector<Pending> rows;
    std::vector<int> roots;
    for (const EquationDef& e : m_.equations) {
      if (e.used_for_objective) {
        continue;
      }
      if (e.relation == 'n') {
        warn(e.line, "'" + e.name + "' is an =N= equation, which relates nothing; dropped");
        continue;
      }
      bool const flip = e.relation == 'g';
      int const left = flip ? e.rhs : e.lhs;
      int const right = flip ? e.lhs : e.rhs;
      model::Cmp const cmp = e.relation == 'e' ? model::Cmp::EQ : model::Cmp::LE;
      auto const lv = m_.arena.fixed_value(left);
      auto const rv = m_.arena.fixed_value(right);
      if (lv && rv) {
        bool const holds = cmp == model::Cmp::EQ ? equal(*lv, *rv) : *lv <= *rv;
        warn(e.line, "'" + e.name + "' has no variables: "
                         + (holds ? std::string("trivially satisfied; dropped")
                                  : std::string("it does not hold, so the model as written is VIOLATED (infeasible)")));
        continue;
      }
      roots.push_back(left);
      if (rv) {
        rows.push_back({&e, cmp, left, std::nullopt, *rv});
      } else {
        roots.push_back(right);
        rows.push_back({&e, cmp, left, right, 0.0});
      }
    }
    nodes_for(roots);
    for (const Pending& p : rows) {
      std::size_t root = node_[static_cast<std::size_t>(p.left)];
      if (p.right) {
        root = binary(model::Op::Sub, root, node_[static_cast<std::size_t>(*p.right)]);
      }
      out_.problem.constraints.push_back({root, p.cmp, p.rhs});
      out_.constraint_names.push_back(p.e->name);
    }
  }

  /// Bounds the file put on a substituted objective variable bound the
  /// objective expression: up as f <= up, lo as -f <= -lo. After an
  /// inequality substitution the bound on the side that inequality already
  /// holds (lo when minimising, up when maximising) restricts nothing and is
  /// dropped, with a warning.
  void objective_bounds(bool via_inequality)
  {
    const Variable& v = m_.vars[static_cast<std::size_t>(*m_.objective_var)];
    if (!v.substituted) {
      return;
    }
    bool const keep_up = !(via_inequality && m_.maximise);
    bool const keep_lo = !(via_inequality && !m_.maximise);
    if (std::isfinite(v.up)) {
      if (keep_up) {
        out_.problem.constraints.push_back({objective_value_, model::Cmp::LE, v.up});
        out_.constraint_names.push_back(v.name + ".up");
      }
    }
    if (std::isfinite(v.lo) && keep_lo) {
      std::size_t const negated = dag().emit(model::Op::Neg, {objective_value_});
      out_.problem.constraints.push_back({negated, model::Cmp::LE, -v.lo});
      out_.constraint_names.push_back(v.name + ".lo");
    }
    for (auto [finite, kept, attr] : {std::tuple {std::isfinite(v.lo), keep_lo, ".lo"},
                                      std::tuple {std::isfinite(v.up), keep_up, ".up"}}) {
      if (finite && !kept) {
        warn(v.declared_at, "'" + v.name + attr + "' bounds only the substituted objective variable from the side "
                                                 "the defining inequality already holds; dropped");
      }
    }
  }

private:
  static constexpr std::size_t kUnset = std::numeric_limits<std::size_t>::max();

  model::ExprDAG<double>& dag() { return out_.problem.graph; }
  void warn(int line, std::string message) { out_.warnings.push_back({line, std::move(message)}); }

  std::size_t constant(double v)
  {
    model::DAGNodePayload<double> p;
    p.constant = v;
    return dag().emit(model::Op::Const, {}, p);
  }

  std::size_t unary(model::Op op, std::size_t x) { return dag().emit(op, {x}); }

  /// A product of a node with itself is its square: Sqr has the tight
  /// enclosure, a product of two independent factors does not.
  std::size_t binary(model::Op op, std::size_t x, std::size_t y)
  {
    return (op == model::Op::Mul && x == y) ? unary(model::Op::Sqr, x) : dag().emit(op, {x, y});
  }

  /// DAG nodes for every term the roots need. A constant subtree becomes one
  /// Const and is not entered; a constant exponent lives in its power's node.
  /// Marking runs down the arena and emission up it, both as loops (single
  /// MINLPLib expressions have hundreds of thousands of terms). Terms already
  /// emitted for an earlier call are reused.
  void nodes_for(const std::vector<int>& roots)
  {
    if (roots.empty()) {
      return;
    }
    int const top = *std::max_element(roots.begin(), roots.end());
    std::vector<char> needed(static_cast<std::size_t>(top) + 1, 0);
    for (int r : roots) {
      needed[static_cast<std::size_t>(r)] = 1;
    }
    for (int i = top; i >= 0; --i) {
      if (!needed[static_cast<std::size_t>(i)] || m_.arena.fixed_value(i)) {
        continue;
      }
      const Term& t = m_.arena.at(i);
      bool const exponent_folded = t.kind == Kind::Call && t.fn == Fn::Power && m_.arena.fixed_value(t.b);
      if (t.a >= 0) {
        needed[static_cast<std::size_t>(t.a)] = 1;
      }
      if (t.b >= 0 && !exponent_folded) {
        needed[static_cast<std::size_t>(t.b)] = 1;
      }
    }
    node_.resize(static_cast<std::size_t>(m_.arena.count()), kUnset);
    for (int i = 0; i <= top; ++i) {
      std::size_t& slot = node_[static_cast<std::size_t>(i)];
      if (needed[static_cast<std::size_t>(i)] && slot == kUnset) {
        slot = node_of_term(i);
      }
    }
  }

  std::size_t node_of_term(int i)
  {
    if (auto v = m_.arena.fixed_value(i)) {
      return constant(*v);
    }
    const Term& t = m_.arena.at(i);
    auto in = [&](int j) { return node_[static_cast<std::size_t>(j)]; };
    switch (t.kind) {
      case Kind::Symbol: {
        std::size_t const id = var_node_[static_cast<std::size_t>(t.symbol)];
        if (id == kUnset) {
          throw ParseError(t.line, "'" + m_.vars[static_cast<std::size_t>(t.symbol)].name + "' has no value here",
                           ErrorKind::Unrepresentable);
        }
        return id;
      }
      case Kind::Add: return binary(model::Op::Add, in(t.a), in(t.b));
      case Kind::Sub: return binary(model::Op::Sub, in(t.a), in(t.b));
      case Kind::Mul: return binary(model::Op::Mul, in(t.a), in(t.b));
      case Kind::Div: return binary(model::Op::Div, in(t.a), in(t.b));
      case Kind::Neg: return unary(model::Op::Neg, in(t.a));
      case Kind::Call: return call(t);
      case Kind::Literal: return constant(t.value);
    }
    throw ParseError(t.line, "internal: a term of unknown kind");
  }

  std::size_t call(const Term& t)
  {
    std::size_t const x = node_[static_cast<std::size_t>(t.a)];
    static const std::unordered_map<int, model::Op> direct {
        {static_cast<int>(Fn::Sqr), model::Op::Sqr}, {static_cast<int>(Fn::Sqrt), model::Op::Sqrt},
        {static_cast<int>(Fn::Exp), model::Op::Exp}, {static_cast<int>(Fn::Log), model::Op::Log},
        {static_cast<int>(Fn::Abs), model::Op::Abs}, {static_cast<int>(Fn::Sin), model::Op::Sin},
        {static_cast<int>(Fn::Cos), model::Op::Cos}, {static_cast<int>(Fn::Tanh), model::Op::Tanh},
    };
    if (auto d = direct.find(static_cast<int>(t.fn)); d != direct.end()) {
      return unary(d->second, x);
    }
    switch (t.fn) {
      case Fn::Log10:
      case Fn::Log2: {
        // log_b x = log x / log b, as a product with the constant 1 / log b
        std::size_t const ln = unary(model::Op::Log, x);
        double const base = t.fn == Fn::Log10 ? 10.0 : 2.0;
        return binary(model::Op::Mul, ln, constant(1.0 / std::log(base)));
      }
      case Fn::Min: return binary(model::Op::Min, x, node_[static_cast<std::size_t>(t.b)]);
      case Fn::Max: return binary(model::Op::Max, x, node_[static_cast<std::size_t>(t.b)]);
      case Fn::Power: return power(t, x);
      default: break;
    }
    throw ParseError(t.line, "internal: a function with no lowering");
  }

  /// x ** y is GAMS's exp(y * log x). A constant exponent of 0.5 is Sqrt and
  /// an integer one (|e| <= 1e9) is PowN; anything else keeps exp and log.
  std::size_t power(const Term& t, std::size_t x)
  {
    auto const e = m_.arena.fixed_value(t.b);
    if (!e) {
      std::size_t const ln = unary(model::Op::Log, x);
      return unary(model::Op::Exp, binary(model::Op::Mul, ln, node_[static_cast<std::size_t>(t.b)]));
    }
    if (equal(*e, 0.5)) {
      return unary(model::Op::Sqrt, x);
    }
    bool const integral = !(std::floor(*e) < *e);
    if (integral && std::fabs(*e) <= 1e9) {
      model::DAGNodePayload<double> p;
      p.int_exp = static_cast<int>(*e);
      return dag().emit(model::Op::PowN, {x}, p);
    }
    std::size_t const ln = unary(model::Op::Log, x);
    std::size_t const scaled = binary(model::Op::Mul, ln, constant(*e));
    return unary(model::Op::Exp, scaled);
  }

  Model& m_;
  const ParseOptions& opt_;
  ParsedModel& out_;
  std::vector<std::size_t> var_node_;  // per variable
  std::vector<std::size_t> node_;      // per arena term
  std::size_t objective_value_ = 0;
};

}  // namespace

ParsedModel parse(std::string_view source, const ParseOptions& options)
{
  Model m;
  StatementReader(m).read(Lexer(source).run());
  if (!m.objective_var) {
    throw ParseError(0, "no Solve statement names an objective", ErrorKind::Unrepresentable);
  }
  if (options.reject_discrete) {
    for (const Variable& v : m.vars) {
      if (v.kind != model::VarKind::Continuous) {
        throw ParseError(v.discrete_at, "'" + v.name + "' is discrete and discrete variables were refused",
                         ErrorKind::Discrete);
      }
    }
  }

  ParsedModel out;
  // Fixed variables become literals before anything inspects values (the
  // objective variable keeps its symbol).
  if (options.fold_fixed_to_const) {
    std::vector<std::optional<double>> values(m.vars.size());
    bool any = false;
    for (std::size_t i = 0; i < m.vars.size(); ++i) {
      if (m.vars[i].fixed && static_cast<int>(i) != *m.objective_var) {
        values[i] = m.vars[i].fixed;
        any = true;
      }
    }
    if (any) {
      m.arena.substitute(values);
    }
  }

  ObjectiveChoice choice = choose_objective(m, out.warnings);
  Variable& objvar = m.vars[static_cast<std::size_t>(*m.objective_var)];
  out.objvar_kept = !choice.term;
  if (choice.term) {
    objvar.substituted = true;
  } else {
    out.warnings.push_back({0, "no equation could be solved for '" + objvar.name
                                   + "'; keeping it as a variable, which costs one search dimension"});
    choice.term = m.arena.symbol(*m.objective_var, m.solve_line);
  }

  Emitter emit(m, options, out);
  emit.variables();
  emit.objective(*choice.term);
  emit.constraints();
  emit.objective_bounds(choice.via_inequality);
  try {
    out.problem.validate();
  } catch (const cmi::error& err) {
    throw ParseError(0, err.what(), ErrorKind::Unrepresentable);
  }
  return out;
}

ParsedModel parse_file(const std::filesystem::path& path, const ParseOptions& options)
{
  std::ifstream in(path, std::ios::binary);
  if (!in) {
    throw ParseError(0, "cannot read '" + path.string() + "'", ErrorKind::Io);
  }
  std::ostringstream text;
  text << in.rdbuf();
  return parse(text.str(), options);
}

std::string result_line(Sense sense, double lower, double upper)
{
  bool const flip = sense == Sense::Maximise;
  auto field = [&](bool present, double v) {
    if (!present) {
      return std::string("none");
    }
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.17g", flip ? -v : v);
    return std::string(buf);
  };
  return std::string("RESULT\tsense=") + (flip ? "max" : "min") + "\tprimal="
      + field(upper < std::numeric_limits<double>::max(), upper) + "\tdual="
      + field(lower > std::numeric_limits<double>::lowest(), lower);
}

}  // namespace cmi::gams
