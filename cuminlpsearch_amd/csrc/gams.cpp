// GAMS scalar-format frontend: scanner, parser into a term pool, objective
// elimination, lowering to Problem. Public interface and the rules it follows
// are described in cmi/gams.hpp.

#include "cmi/gams.hpp"

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
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

constexpr double kInf = std::numeric_limits<double>::infinity();

/// Equality as the peephole rules test it: !(a < b) && !(b < a), which also
/// holds when either side is NaN (reference ast.hpp feq).
constexpr bool same_value(double a, double b) { return !(a < b) && !(b < a); }

std::string folded(std::string_view s)
{
  std::string out(s);
  for (char& c : out) {
    c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
  }
  return out;
}

bool begins_with_folded(std::string_view line, std::string_view prefix)
{
  if (line.size() < prefix.size()) {
    return false;
  }
  return folded(line.substr(0, prefix.size())) == prefix;
}

// ---------------------------------------------------------------------------
// Tokens. GAMS scalar format never continues a token across a line break, so
// the scanner works a line at a time; two rules look at column 1 only (`*`
// starts a comment line, `$` a dollar-control line, `$ontext` .. `$offtext` a
// comment block), reference frontend.cpp:84-144.
// ---------------------------------------------------------------------------

enum class Tk : std::uint8_t
{
  End,
  Name,
  Number,
  Plus,
  Minus,
  Star,
  Slash,
  StarStar,
  LParen,
  RParen,
  Comma,
  Semi,
  Assign,
  Relation,
  DotDot,
  Dot
};

enum class Rel : std::uint8_t
{
  E,
  L,
  G,
  N
};

struct Token
{
  Tk kind = Tk::End;
  Rel rel = Rel::E;
  int line = 1;
  double number = 0.0;
  std::string name;  // lower case
};

bool name_start(char c) { return std::isalpha(static_cast<unsigned char>(c)) != 0 || c == '_'; }
bool name_char(char c) { return std::isalnum(static_cast<unsigned char>(c)) != 0 || c == '_'; }
bool digit(char c) { return std::isdigit(static_cast<unsigned char>(c)) != 0; }

void scan_line(const std::string& text, int line, std::vector<Token>& out)
{
  std::size_t i = 0;
  std::size_t const n = text.size();
  while (i < n) {
    char const c = text[i];
    if (std::isspace(static_cast<unsigned char>(c)) != 0) {
      ++i;
      continue;
    }
    Token t;
    t.line = line;
    if (name_start(c)) {
      std::size_t j = i;
      while (j < n && name_char(text[j])) {
        ++j;
      }
      t.kind = Tk::Name;
      t.name = folded(std::string_view(text).substr(i, j - i));
      i = j;
      out.push_back(std::move(t));
      continue;
    }
    if (digit(c) || (c == '.' && i + 1 < n && digit(text[i + 1]))) {
      char const* begin = text.c_str() + i;
      char* end = nullptr;
      t.number = std::strtod(begin, &end);
      if (end == begin) {
        throw ParseError(line, "malformed number");
      }
      i += static_cast<std::size_t>(end - begin);
      t.kind = Tk::Number;
      out.push_back(std::move(t));
      continue;
    }
    if (c == '%') {
      // %NLP% and other compile-time variables: one opaque name.
      std::size_t j = i + 1;
      while (j < n && text[j] != '%') {
        ++j;
      }
      if (j < n) {
        ++j;
      }
      t.kind = Tk::Name;
      t.name = folded(std::string_view(text).substr(i, j - i));
      i = j;
      out.push_back(std::move(t));
      continue;
    }
    std::size_t width = 1;
    switch (c) {
      case '+': t.kind = Tk::Plus; break;
      case '-': t.kind = Tk::Minus; break;
      case '/': t.kind = Tk::Slash; break;
      case '(': t.kind = Tk::LParen; break;
      case ')': t.kind = Tk::RParen; break;
      case ',': t.kind = Tk::Comma; break;
      case ';': t.kind = Tk::Semi; break;
      case '*':
        if (i + 1 < n && text[i + 1] == '*') {
          t.kind = Tk::StarStar;
          width = 2;
        } else {
          t.kind = Tk::Star;
        }
        break;
      case '.':
        if (i + 1 < n && text[i + 1] == '.') {
          t.kind = Tk::DotDot;
          width = 2;
        } else {
          t.kind = Tk::Dot;
        }
        break;
      case '=':
        if (i + 2 < n && std::isalpha(static_cast<unsigned char>(text[i + 1])) != 0 && text[i + 2] == '=') {
          t.kind = Tk::Relation;
          switch (std::tolower(static_cast<unsigned char>(text[i + 1]))) {
            case 'e': t.rel = Rel::E; break;
            case 'l': t.rel = Rel::L; break;
            case 'g': t.rel = Rel::G; break;
            case 'n': t.rel = Rel::N; break;
            default: throw ParseError(line, "unknown relational operator");
          }
          width = 3;
        } else {
          t.kind = Tk::Assign;
        }
        break;
      case '\'':
      case '"':
        throw ParseError(line,
                         "quoted set element: the file has indexed identifiers and is not in GAMS scalar "
                         "format; regenerate it with GAMS Convert",
                         ErrorKind::NonScalar);
      default: throw ParseError(line, std::string("unexpected character '") + c + "'");
    }
    i += width;
    out.push_back(std::move(t));
  }
}

std::vector<Token> scan(std::string_view src)
{
  std::vector<Token> out;
  int line = 0;
  bool in_block = false;
  std::size_t pos = 0;
  while (pos <= src.size()) {
    std::size_t end = src.find('\n', pos);
    if (end == std::string_view::npos) {
      end = src.size();
    }
    ++line;
    std::string_view const text = src.substr(pos, end - pos);
    pos = end + 1;
    if (in_block) {
      in_block = !begins_with_folded(text, "$offtext");
      continue;
    }
    if (!text.empty() && text[0] == '*') {
      continue;
    }
    if (!text.empty() && text[0] == '$') {
      in_block = begins_with_folded(text, "$ontext");
      continue;
    }
    scan_line(std::string(text), line, out);
  }
  Token end;
  end.line = line;
  out.push_back(end);
  return out;
}

// ---------------------------------------------------------------------------
// Term pool: the parsed expressions, children always before parents. Building
// applies only identities that are also exact for intervals (literal
// arithmetic, x+0, 0+x, x-0, 0-x = -x, x-(-y) = x+y, x*1, x*(-1), x/1,
// x/(-1), -(-x)); never 0*x or 0/x (reference ast.hpp:73-205).
// ---------------------------------------------------------------------------

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

enum class TermKind : std::uint8_t
{
  Literal,
  Variable,
  Add,
  Sub,
  Mul,
  Div,
  Neg,
  Call
};

struct Term
{
  TermKind kind = TermKind::Literal;
  Fn fn = Fn::Sqr;
  std::uint8_t n_args = 0;
  std::array<int, 2> arg {-1, -1};
  double value = 0.0;  // Literal
  int sym = -1;        // Variable
  int line = 0;
};

/// Value of a variable-free function application, or nothing where the
/// function is undefined there (reference ast.hpp:346-386).
std::optional<double> apply(Fn f, double a, double b)
{
  switch (f) {
    case Fn::Sqr: return a * a;
    case Fn::Sqrt: return a < 0.0 ? std::nullopt : std::optional<double>(std::sqrt(a));
    case Fn::Exp: return std::exp(a);
    case Fn::Log: return a <= 0.0 ? std::nullopt : std::optional<double>(std::log(a));
    case Fn::Log10: return a <= 0.0 ? std::nullopt : std::optional<double>(std::log10(a));
    case Fn::Log2: return a <= 0.0 ? std::nullopt : std::optional<double>(std::log2(a));
    case Fn::Abs: return std::fabs(a);
    case Fn::Sin: return std::sin(a);
    case Fn::Cos: return std::cos(a);
    case Fn::Tanh: return std::tanh(a);
    case Fn::Min: return std::fmin(a, b);
    case Fn::Max: return std::fmax(a, b);
    case Fn::Power:
      if (std::floor(b) < b && a <= 0.0) {
        return std::nullopt;
      }
      return std::pow(a, b);
  }
  return std::nullopt;
}

class Terms
{
public:
  const Term& operator[](int i) const { return t_[static_cast<std::size_t>(i)]; }
  int size() const { return static_cast<int>(t_.size()); }

  int literal(double v, int line = 0)
  {
    Term t;
    t.kind = TermKind::Literal;
    t.value = v;
    t.line = line;
    return push(t);
  }

  int variable(int sym, int line)
  {
    Term t;
    t.kind = TermKind::Variable;
    t.sym = sym;
    t.line = line;
    return push(t);
  }

  int call(Fn f, int a, int b, int line)
  {
    Term t;
    t.kind = TermKind::Call;
    t.fn = f;
    t.arg = {a, b};
    t.n_args = b < 0 ? 1 : 2;
    t.line = line;
    return push(t);
  }

  int negate(int a)
  {
    if (auto v = literal_value(a)) {
      return literal(-*v, line_of(a));
    }
    if ((*this)[a].kind == TermKind::Neg) {
      return (*this)[a].arg[0];
    }
    return node(TermKind::Neg, a, -1);
  }

  int plus(int a, int b)
  {
    auto const va = literal_value(a);
    auto const vb = literal_value(b);
    if (va && vb) {
      return literal(*va + *vb, line_of(a));
    }
    if (vb && same_value(*vb, 0.0)) {
      return a;
    }
    if (va && same_value(*va, 0.0)) {
      return b;
    }
    return node(TermKind::Add, a, b);
  }

  int minus(int a, int b)
  {
    auto const va = literal_value(a);
    auto const vb = literal_value(b);
    if (va && vb) {
      return literal(*va - *vb, line_of(a));
    }
    if (vb && same_value(*vb, 0.0)) {
      return a;
    }
    if (va && same_value(*va, 0.0)) {
      return negate(b);
    }
    if ((*this)[b].kind == TermKind::Neg) {
      return plus(a, (*this)[b].arg[0]);
    }
    return node(TermKind::Sub, a, b);
  }

  int times(int a, int b)
  {
    auto const va = literal_value(a);
    auto const vb = literal_value(b);
    if (va && vb) {
      return literal(*va * *vb, line_of(a));
    }
    if (vb && same_value(*vb, 1.0)) {
      return a;
    }
    if (va && same_value(*va, 1.0)) {
      return b;
    }
    if (vb && same_value(*vb, -1.0)) {
      return negate(a);
    }
    if (va && same_value(*va, -1.0)) {
      return negate(b);
    }
    return node(TermKind::Mul, a, b);
  }

  int over(int a, int b)
  {
    auto const va = literal_value(a);
    auto const vb = literal_value(b);
    if (va && vb && !same_value(*vb, 0.0)) {
      return literal(*va / *vb, line_of(a));
    }
    if (vb && same_value(*vb, 1.0)) {
      return a;
    }
    if (vb && same_value(*vb, -1.0)) {
      return negate(a);
    }
    return node(TermKind::Div, a, b);
  }

  std::optional<double> literal_value(int i) const
  {
    const Term& t = (*this)[i];
    return t.kind == TermKind::Literal ? std::optional<double>(t.value) : std::nullopt;
  }

  /// Value of a variable-free subtree (kept up to date as terms are added).
  std::optional<double> constant(int i) const { return value_[static_cast<std::size_t>(i)]; }

  /// Whether the subtree at `i` mentions symbol `sym`.
  bool mentions(int i, int sym) const
  {
    if (mention_sym_ != sym) {
      mention_sym_ = sym;
      mention_.clear();
    }
    while (mention_.size() < t_.size()) {
      const Term& t = t_[mention_.size()];
      bool m = t.kind == TermKind::Variable && t.sym == sym;
      for (int k = 0; k < t.n_args && !m; ++k) {
        m = mention_[static_cast<std::size_t>(t.arg[k])] != 0;
      }
      mention_.push_back(m ? 1 : 0);
    }
    return mention_[static_cast<std::size_t>(i)] != 0;
  }

  /// Turn every Variable term of a symbol with a value in `by` into that
  /// literal (one pass over the pool), then recompute the constant values.
  void replace_symbols(const std::vector<std::optional<double>>& by)
  {
    for (Term& t : t_) {
      if (t.kind == TermKind::Variable && static_cast<std::size_t>(t.sym) < by.size()
          && by[static_cast<std::size_t>(t.sym)])
      {
        t.value = *by[static_cast<std::size_t>(t.sym)];
        t.kind = TermKind::Literal;
        t.sym = -1;
      }
    }
    value_.clear();
    for (const Term& t : t_) {
      value_.push_back(evaluate(t));
    }
    mention_.clear();
    mention_sym_ = -1;
  }

private:
  int line_of(int i) const { return (*this)[i].line; }

  int node(TermKind k, int a, int b)
  {
    Term t;
    t.kind = k;
    t.arg = {a, b};
    t.n_args = b < 0 ? 1 : 2;
    t.line = line_of(a);
    return push(t);
  }

  int push(const Term& t)
  {
    t_.push_back(t);
    value_.push_back(evaluate(t));
    return static_cast<int>(t_.size()) - 1;
  }

  std::optional<double> evaluate(const Term& t) const
  {
    if (t.kind == TermKind::Literal) {
      return t.value;
    }
    if (t.kind == TermKind::Variable) {
      return std::nullopt;
    }
    std::array<double, 2> v {0.0, 0.0};
    for (int k = 0; k < t.n_args; ++k) {
      auto const& c = value_[static_cast<std::size_t>(t.arg[k])];
      if (!c) {
        return std::nullopt;
      }
      v[static_cast<std::size_t>(k)] = *c;
    }
    switch (t.kind) {
      case TermKind::Add: return v[0] + v[1];
      case TermKind::Sub: return v[0] - v[1];
      case TermKind::Mul: return v[0] * v[1];
      case TermKind::Div: return same_value(v[1], 0.0) ? std::nullopt : std::optional<double>(v[0] / v[1]);
      case TermKind::Neg: return -v[0];
      case TermKind::Call: return apply(t.fn, v[0], v[1]);
      default: return std::nullopt;
    }
  }

  std::vector<Term> t_;
  std::vector<std::optional<double>> value_;
  mutable std::vector<char> mention_;
  mutable int mention_sym_ = -1;
};

// ---------------------------------------------------------------------------
// The model as written (reference ast.hpp:405-450).
// ---------------------------------------------------------------------------

struct Symbol
{
  std::string name;
  int decl_line = 0;
  double lo = -kInf;
  double up = kInf;
  bool fixed = false;
  double fx = 0.0;
  bool has_level = false;
  double level = 0.0;
  model::VarKind kind = model::VarKind::Continuous;
  int integral_line = 0;  // where kind last became Integer/Binary
  bool eliminated = false;
};

struct Equation
{
  std::string name;
  int line = 0;
  Rel rel = Rel::E;
  int lhs = -1;
  int rhs = -1;
  bool consumed = false;
};

struct Written
{
  Terms terms;
  std::vector<Symbol> symbols;
  std::unordered_map<std::string, int> by_name;
  std::vector<Equation> equations;
  int objvar = -1;
  bool maximise = false;
  bool have_solve = false;
};

// ---------------------------------------------------------------------------
// Statement and expression grammar (reference frontend.cpp:335-812).
// ---------------------------------------------------------------------------

const std::unordered_map<std::string, Fn>& functions()
{
  static const std::unordered_map<std::string, Fn> table = {
      {"sqr", Fn::Sqr},   {"sqrt", Fn::Sqrt}, {"exp", Fn::Exp},     {"log", Fn::Log},      {"log10", Fn::Log10},
      {"log2", Fn::Log2}, {"abs", Fn::Abs},   {"sin", Fn::Sin},     {"cos", Fn::Cos},      {"tanh", Fn::Tanh},
      {"min", Fn::Min},   {"max", Fn::Max},   {"power", Fn::Power}, {"rpower", Fn::Power},
  };
  return table;
}

/// GAMS functions that are refused rather than rewritten in a looser form.
bool refused_function(const std::string& name)
{
  static const std::unordered_set<std::string> table = {
      "tan",    "arctan",  "arcsin", "arccos",   "errorf", "sign",    "signpower", "ceil",    "floor", "round",
      "trunc",  "mod",     "div",    "centropy", "sinh",   "cosh",    "gamma",     "loggamma", "beta", "entropy",
      "sigmoid", "slexp",  "sqexp",  "edist",    "poly",   "ifthen",  "vcpower",   "cvpower", "ncpcm",
  };
  return table.count(name) != 0;
}

/// Statement keywords of full (indexed) GAMS.
bool indexed_keyword(const std::string& w)
{
  static const std::unordered_set<std::string> table = {
      "set",   "sets",  "parameter", "parameters", "table", "alias",  "scalar", "scalars", "loop", "sum",
      "prod",  "smax",  "smin",      "abort",      "display", "file", "put",    "repeat",  "while", "for",
  };
  return table.count(w) != 0;
}

class Reader
{
public:
  Reader(std::vector<Token> tokens, Written& w) : tok_(std::move(tokens)), w_(w) {}

  void read_all()
  {
    while (at().kind != Tk::End) {
      statement();
    }
  }

private:
  const Token& at(std::size_t ahead = 0) const { return tok_[std::min(pos_ + ahead, tok_.size() - 1)]; }
  void step()
  {
    if (pos_ + 1 < tok_.size()) {
      ++pos_;
    }
  }
  bool word(std::string_view s) const { return at().kind == Tk::Name && at().name == s; }
  void require(Tk k, const char* what)
  {
    if (at().kind != k) {
      throw ParseError(at().line, std::string("expected ") + what);
    }
    step();
  }
  void skip_statement()
  {
    while (at().kind != Tk::Semi && at().kind != Tk::End) {
      step();
    }
    if (at().kind == Tk::Semi) {
      step();
    }
  }

  void statement()
  {
    if (at().kind == Tk::Semi) {
      step();
      return;
    }
    if (at().kind != Tk::Name) {
      throw ParseError(at().line, "expected the start of a statement");
    }
    const std::string w = at().name;
    if (indexed_keyword(w)) {
      throw ParseError(at().line,
                       "'" + w + "' means the file is not in GAMS scalar format; regenerate it with GAMS Convert",
                       ErrorKind::NonScalar);
    }
    if (w == "sos1" || w == "sos2" || w == "semicont" || w == "semiint") {
      throw ParseError(at().line, "'" + w + "' variables are not supported", ErrorKind::Discrete);
    }
    if (w == "free" || w == "positive" || w == "negative" || w == "binary" || w == "integer") {
      variables(w);
    } else if (w == "variables" || w == "variable") {
      variables("free");
    } else if (w == "equations" || w == "equation") {
      equation_names();
    } else if (w == "model" || w == "models" || w == "option" || w == "options") {
      skip_statement();
    } else if (w == "solve") {
      solve();
    } else if (at(1).kind == Tk::DotDot) {
      equation();
    } else if (at(1).kind == Tk::Dot) {
      attribute();
    } else {
      throw ParseError(at().line, "unrecognised statement starting at '" + w + "'");
    }
  }

  /// `[domain] Variable(s) a, b, ...;` -- a repeat declaration restates the
  /// bounds of an existing symbol (its index stays that of the first
  /// mention); integrality, once declared, is never cleared.
  void variables(const std::string& domain)
  {
    int const line = at().line;
    step();
    if (word("variables") || word("variable")) {
      step();
    }
    double lo = -kInf;
    double up = kInf;
    model::VarKind kind = model::VarKind::Continuous;
    if (domain == "positive") {
      lo = 0.0;
    } else if (domain == "negative") {
      up = 0.0;
    } else if (domain == "binary") {
      lo = 0.0;
      up = 1.0;
      kind = model::VarKind::Binary;
    } else if (domain == "integer") {
      lo = 0.0;
      kind = model::VarKind::Integer;
    }
    for (;;) {
      if (at().kind != Tk::Name) {
        throw ParseError(at().line, "variable name");
      }
      Symbol& s = w_.symbols[static_cast<std::size_t>(symbol(at().name, line))];
      s.lo = lo;
      s.up = up;
      if (kind != model::VarKind::Continuous) {
        s.kind = kind;
        s.integral_line = line;
      }
      step();
      if (at().kind != Tk::Comma) {
        break;
      }
      step();
    }
    require(Tk::Semi, "';' after a variable declaration");
  }

  void equation_names()
  {
    step();
    for (;;) {
      if (at().kind != Tk::Name) {
        throw ParseError(at().line, "equation name");
      }
      step();
      if (at().kind != Tk::Comma) {
        break;
      }
      step();
    }
    require(Tk::Semi, "';' after an equation declaration");
  }

  void equation()
  {
    Equation eq;
    eq.name = at().name;
    eq.line = at().line;
    step();
    require(Tk::DotDot, "'..' after an equation name");
    eq.lhs = sum();
    if (at().kind != Tk::Relation) {
      throw ParseError(at().line, "expected =E=, =L=, =G= or =N=");
    }
    eq.rel = at().rel;
    step();
    eq.rhs = sum();
    require(Tk::Semi, "';' after an equation definition");
    w_.equations.push_back(std::move(eq));
  }

  /// `name.attr = constant;` for a variable's lo/up/fx/l; anything else
  /// (model options, equation marginals, unknown attributes) is skipped.
  void attribute()
  {
    int const line = at().line;
    const std::string name = at().name;
    step();
    require(Tk::Dot, "'.'");
    if (at().kind != Tk::Name) {
      throw ParseError(line, "attribute name");
    }
    const std::string attr = at().name;
    step();
    auto const it = w_.by_name.find(name);
    if (it == w_.by_name.end() || (attr != "lo" && attr != "up" && attr != "fx" && attr != "l")) {
      skip_statement();
      return;
    }
    require(Tk::Assign, "'=' in an attribute assignment");
    auto const value = w_.terms.constant(sum());
    if (!value) {
      throw ParseError(line, "'" + name + "." + attr + "' must be a constant");
    }
    require(Tk::Semi, "';' after an attribute assignment");
    Symbol& s = w_.symbols[static_cast<std::size_t>(it->second)];
    if (attr == "lo") {
      s.lo = *value;
    } else if (attr == "up") {
      s.up = *value;
    } else if (attr == "fx") {
      s.fixed = true;
      s.fx = *value;
    } else {
      s.has_level = true;
      s.level = *value;
    }
  }

  void solve()
  {
    int const line = at().line;
    while (at().kind != Tk::Semi && at().kind != Tk::End) {
      if (word("minimizing") || word("minimising") || word("maximizing") || word("maximising")) {
        bool const maximise = at().name[1] == 'a';
        step();
        if (at().kind != Tk::Name) {
          throw ParseError(at().line, "objective variable name after minimizing/maximizing");
        }
        auto const it = w_.by_name.find(at().name);
        if (it == w_.by_name.end()) {
          throw ParseError(at().line, "undeclared objective variable '" + at().name + "'");
        }
        w_.objvar = it->second;
        w_.maximise = maximise;
        w_.have_solve = true;
      }
      step();
    }
    if (at().kind == Tk::Semi) {
      step();
    }
    if (!w_.have_solve) {
      throw ParseError(line, "Solve statement names no objective (a feasibility-only model has no form in Problem)",
                       ErrorKind::Unrepresentable);
    }
  }

  // sum := product (('+'|'-') product)*
  int sum()
  {
    int acc = product();
    while (at().kind == Tk::Plus || at().kind == Tk::Minus) {
      bool const add = at().kind == Tk::Plus;
      step();
      int const rhs = product();
      acc = add ? w_.terms.plus(acc, rhs) : w_.terms.minus(acc, rhs);
    }
    return acc;
  }

  // product := signed (('*'|'/') signed)*
  int product()
  {
    int acc = signed_factor();
    while (at().kind == Tk::Star || at().kind == Tk::Slash) {
      bool const mul = at().kind == Tk::Star;
      step();
      int const rhs = signed_factor();
      acc = mul ? w_.terms.times(acc, rhs) : w_.terms.over(acc, rhs);
    }
    return acc;
  }

  // signed := ('-'|'+') signed | power; a sign binds looser than '**', so
  // -x**2 is -(x**2).
  int signed_factor()
  {
    if (at().kind == Tk::Minus) {
      step();
      return w_.terms.negate(signed_factor());
    }
    if (at().kind == Tk::Plus) {
      step();
      return signed_factor();
    }
    return power();
  }

  // power := primary ['**' signed]  (right-associative, signed exponent)
  int power()
  {
    int const base = primary();
    if (at().kind != Tk::StarStar) {
      return base;
    }
    int const line = at().line;
    step();
    int const exponent = signed_factor();
    return w_.terms.call(Fn::Power, base, exponent, line);
  }

  int primary()
  {
    int const line = at().line;
    if (at().kind == Tk::Number) {
      double const v = at().number;
      step();
      return w_.terms.literal(v, line);
    }
    if (at().kind == Tk::LParen) {
      step();
      int const inner = sum();
      require(Tk::RParen, "')'");
      return inner;
    }
    if (at().kind != Tk::Name) {
      throw ParseError(line, "expected a number, identifier or '('");
    }
    const std::string name = at().name;
    if (name == "inf") {
      step();
      return w_.terms.literal(kInf, line);
    }
    if (name == "eps") {
      step();
      return w_.terms.literal(0.0, line);
    }
    if (name == "na" || name == "undf") {
      throw ParseError(line, "the special value '" + name + "' cannot be represented", ErrorKind::Unrepresentable);
    }
    if (at(1).kind == Tk::LParen) {
      auto const f = functions().find(name);
      if (f == functions().end()) {
        throw ParseError(line,
                         refused_function(name) ? "function '" + name + "' is not supported yet"
                                                : "unknown function '" + name + "'",
                         ErrorKind::UnsupportedFunction);
      }
      step();
      step();
      std::vector<int> args {sum()};
      while (at().kind == Tk::Comma) {
        step();
        args.push_back(sum());
      }
      require(Tk::RParen, "')' closing a function call");
      return call(f->second, args, line);
    }
    auto const it = w_.by_name.find(name);
    if (it == w_.by_name.end()) {
      throw ParseError(line, "undeclared identifier '" + name + "'");
    }
    step();
    return w_.terms.variable(it->second, line);
  }

  /// min/max of several arguments fold left into binary applications.
  int call(Fn f, const std::vector<int>& args, int line)
  {
    if (f == Fn::Min || f == Fn::Max) {
      int acc = args[0];
      for (std::size_t i = 1; i < args.size(); ++i) {
        acc = w_.terms.call(f, acc, args[i], line);
      }
      return acc;
    }
    std::size_t const arity = f == Fn::Power ? 2 : 1;
    if (args.size() != arity) {
      throw ParseError(line, "wrong number of arguments to this function");
    }
    return w_.terms.call(f, args[0], arity == 2 ? args[1] : -1, line);
  }

  int symbol(const std::string& name, int line)
  {
    auto const it = w_.by_name.find(name);
    if (it != w_.by_name.end()) {
      return it->second;
    }
    Symbol s;
    s.name = name;
    s.decl_line = line;
    w_.symbols.push_back(std::move(s));
    int const idx = static_cast<int>(w_.symbols.size()) - 1;
    w_.by_name.emplace(name, idx);
    return idx;
  }

  std::vector<Token> tok_;
  std::size_t pos_ = 0;
  Written& w_;
};

// ---------------------------------------------------------------------------
// Objective elimination (reference frontend.cpp:818-1082).
// ---------------------------------------------------------------------------

/// term == coeff * sym + rest, rest free of sym.
struct Linear
{
  double coeff;
  int rest;
};

/// Write the subtree at `root` as coeff * sym + rest, or nothing if it is not
/// linear in sym with constant coefficients. One forward sweep over the pool
/// (children precede parents), so expressions of any depth are handled
/// without recursion. Only +, -, negation, multiplication by a constant and
/// division by a non-zero constant pass sym through.
std::optional<Linear> linear_in(Terms& terms, int root, int sym)
{
  std::vector<std::optional<Linear>> at(static_cast<std::size_t>(root) + 1);
  for (int i = 0; i <= root; ++i) {
    std::optional<Linear>& out = at[static_cast<std::size_t>(i)];
    if (!terms.mentions(i, sym)) {
      out = Linear {0.0, i};
      continue;
    }
    Term const t = terms[i];  // a copy: the builders below grow the pool
    auto part = [&](int k) { return at[static_cast<std::size_t>(t.arg[static_cast<std::size_t>(k)])]; };
    switch (t.kind) {
      case TermKind::Variable: out = Linear {1.0, terms.literal(0.0)}; break;
      case TermKind::Neg:
        if (auto a = part(0)) {
          out = Linear {-a->coeff, terms.negate(a->rest)};
        }
        break;
      case TermKind::Add:
      case TermKind::Sub: {
        auto a = part(0);
        auto b = part(1);
        if (a && b) {
          bool const add = t.kind == TermKind::Add;
          out = Linear {add ? a->coeff + b->coeff : a->coeff - b->coeff,
                        add ? terms.plus(a->rest, b->rest) : terms.minus(a->rest, b->rest)};
        }
        break;
      }
      case TermKind::Mul: {
        int const k = terms.mentions(t.arg[0], sym) ? 1 : 0;  // the side that must be constant
        if (terms.mentions(t.arg[k], sym)) {
          break;
        }
        auto const c = terms.constant(t.arg[k]);
        auto a = part(1 - k);
        if (c && a) {
          out = Linear {*c * a->coeff, terms.times(t.arg[k], a->rest)};
        }
        break;
      }
      case TermKind::Div: {
        if (terms.mentions(t.arg[1], sym)) {
          break;
        }
        auto const c = terms.constant(t.arg[1]);
        auto a = part(0);
        if (c && !same_value(*c, 0.0) && a) {
          out = Linear {a->coeff / *c, terms.over(a->rest, t.arg[1])};
        }
        break;
      }
      default: break;
    }
  }
  return at[static_cast<std::size_t>(root)];
}

struct Elimination
{
  std::optional<int> objective;
  bool from_inequality = false;
};

/// An equation gathered as coeff * objvar + lhs_rest = rhs_rest.
struct Gathered
{
  double coeff;
  int lhs_rest;
  int rhs_rest;
};

/// The equation with objvar gathered on the left, if both sides are linear in
/// objvar and it does not cancel.
std::optional<Gathered> gather(Written& w, const Equation& eq)
{
  if (!w.terms.mentions(eq.lhs, w.objvar) && !w.terms.mentions(eq.rhs, w.objvar)) {
    return std::nullopt;
  }
  auto const l = linear_in(w.terms, eq.lhs, w.objvar);
  auto const r = linear_in(w.terms, eq.rhs, w.objvar);
  if (!l || !r) {
    return std::nullopt;
  }
  double const coeff = l->coeff - r->coeff;
  if (std::fabs(coeff) < 1e-12) {
    return std::nullopt;
  }
  return Gathered {coeff, l->rest, r->rest};
}

/// objvar = (rhs_rest - lhs_rest) / coeff.
int solved(Written& w, const Gathered& g)
{
  int const rest = w.terms.minus(g.rhs_rest, g.lhs_rest);
  return w.terms.over(rest, w.terms.literal(g.coeff));
}

/// Equalities first: any =E= linear in objvar defines it (the first one is
/// used). Otherwise an =L=/=G= that bounds objvar on the side the sense
/// pushes against -- from below when minimising, above when maximising --
/// and is the only equation mentioning objvar is tight at every optimum and
/// is used the same way.
Elimination eliminate(Written& w, std::vector<Diagnostic>& warnings)
{
  std::optional<int> objective;
  int defining = 0;
  for (Equation& eq : w.equations) {
    if (eq.rel != Rel::E) {
      continue;
    }
    auto const g = gather(w, eq);
    if (!g) {
      continue;
    }
    ++defining;
    if (!objective) {
      objective = solved(w, *g);
      eq.consumed = true;
    }
  }
  if (defining > 1) {
    warnings.push_back({0, "several equations define the objective variable; used the first"});
  }
  if (objective) {
    return {objective, false};
  }

  for (std::size_t i = 0; i < w.equations.size(); ++i) {
    Equation& eq = w.equations[i];
    if (eq.rel != Rel::L && eq.rel != Rel::G) {
      continue;
    }
    auto const g = gather(w, eq);
    if (!g) {
      continue;
    }
    // lhs <= rhs with objvar gathered on the left reads coeff*objvar <= ...;
    // a negative coefficient turns it into a lower bound.
    bool const from_below = (eq.rel == Rel::L) == (g->coeff < 0.0);
    if (from_below == w.maximise) {
      continue;
    }
    bool elsewhere = false;
    for (std::size_t j = 0; j < w.equations.size() && !elsewhere; ++j) {
      const Equation& other = w.equations[j];
      elsewhere = j != i && other.rel != Rel::N
          && (w.terms.mentions(other.lhs, w.objvar) || w.terms.mentions(other.rhs, w.objvar));
    }
    if (elsewhere) {
      continue;
    }
    Elimination out {solved(w, *g), true};
    eq.consumed = true;
    warnings.push_back({eq.line, "'" + w.symbols[static_cast<std::size_t>(w.objvar)].name
                                     + "' is defined by the inequality '" + eq.name
                                     + "', which is tight at every optimum; substituted it away"});
    return out;
  }
  return {};
}

// ---------------------------------------------------------------------------
// Lowering to Problem (reference frontend.cpp:1088-1590).
// ---------------------------------------------------------------------------

class Lowering
{
public:
  Lowering(Written& w, const ParseOptions& opt, ParsedModel& out)
      : w_(w), opt_(opt), out_(out), var_node_(w.symbols.size(), kNone)
  {
  }

  void variables()
  {
    bool const any_level = std::any_of(w_.symbols.begin(), w_.symbols.end(), [](const Symbol& s) { return s.has_level; });
    for (std::size_t i = 0; i < w_.symbols.size(); ++i) {
      const Symbol& s = w_.symbols[i];
      if (s.eliminated || (s.fixed && opt_.fold_fixed_to_const)) {
        continue;
      }
      double lo = s.fixed ? s.fx : s.lo;
      double up = s.fixed ? s.fx : s.up;
      bool const discrete = s.kind != model::VarKind::Continuous;
      bool const lo_default = !std::isfinite(lo);
      bool const up_default = !std::isfinite(up);
      if (lo_default) {
        lo = -opt_.default_bound;
      }
      if (up_default) {
        up = opt_.default_bound;
      }
      if (discrete) {  // snap inward to the integers
        lo = std::ceil(lo);
        up = std::floor(up);
      }
      if (lo > up) {
        throw ParseError(s.decl_line, "'" + s.name + "' has an empty box");
      }
      std::string const domain =
          discrete ? " (integer domain size " + std::to_string(up - lo + 1.0) + ")" : std::string();
      if (lo_default) {
        warn(s.decl_line, "'" + s.name + "' has no lower bound; using " + std::to_string(lo) + domain);
      }
      if (up_default) {
        warn(s.decl_line, "'" + s.name + "' has no upper bound; using " + std::to_string(up) + domain);
      }
      model::VarKind kind = s.kind;
      if (kind == model::VarKind::Binary && (lo < 0.0 || lo > 0.0 || up < 1.0 || up > 1.0)) {
        kind = model::VarKind::Integer;  // a binary narrowed off [0, 1]
      }
      var_node_[i] = out_.problem.var(lo, up, kind).id();
      out_.var_names.push_back(s.name);
      if (any_level) {
        double const start = s.has_level ? s.level : std::clamp(0.0, lo, up);
        out_.initial_point.push_back(std::clamp(start, lo, up));
      }
    }
  }

  /// The objective first: a remaining mention of an eliminated objective
  /// variable then resolves to the objective's (un-negated) root.
  void objective(int term)
  {
    lower({term});
    std::size_t root = node_of(term);
    if (graph().nodes[root].op == model::Op::Const) {
      throw ParseError(0, "the objective is constant", ErrorKind::Unrepresentable);
    }
    objective_root_ = root;
    if (w_.maximise) {
      root = graph().emit(model::Op::Neg, {root});
      out_.sense = Sense::Maximise;
    }
    var_node_[static_cast<std::size_t>(w_.objvar)] = objective_root_;
    out_.problem.objective_root = root;
  }

  void constraints()
  {
    struct Row
    {
      const Equation* eq;
      model::Cmp cmp;
      int lhs;
      int rhs;  // -1: folded into `bound`
      double bound;
    };
    std::vector<Row> rows;
    std::vector<int> roots;
    for (const Equation& eq : w_.equations) {
      if (eq.consumed) {
        continue;
      }
      if (eq.rel == Rel::N) {
        warn(eq.line, "'" + eq.name + "' is a =N= equation and asserts nothing; dropped");
        continue;
      }
      int lhs = eq.lhs;
      int rhs = eq.rhs;
      if (eq.rel == Rel::G) {
        std::swap(lhs, rhs);  // a >= b is b <= a
      }
      model::Cmp const cmp = eq.rel == Rel::E ? model::Cmp::EQ : model::Cmp::LE;
      auto const lc = w_.terms.constant(lhs);
      auto const rc = w_.terms.constant(rhs);
      if (lc && rc) {
        bool const holds = cmp == model::Cmp::EQ ? same_value(*lc, *rc) : *lc <= *rc;
        warn(eq.line, "'" + eq.name + "' has no variables and is "
                          + (holds ? "trivially satisfied; dropped" : "VIOLATED: the model as written is infeasible"));
        continue;
      }
      roots.push_back(lhs);
      if (rc) {
        rows.push_back({&eq, cmp, lhs, -1, *rc});
      } else {
        roots.push_back(rhs);
        rows.push_back({&eq, cmp, lhs, rhs, 0.0});
      }
    }
    lower(roots);
    for (const Row& r : rows) {
      std::size_t root = node_of(r.lhs);
      if (r.rhs >= 0) {
        root = binary(model::Op::Sub, root, node_of(r.rhs));
      }
      out_.problem.constraints.push_back({root, r.cmp, r.bound});
      out_.constraint_names.push_back(r.eq->name);
    }
  }

  /// Bounds stated on an eliminated objective variable constrain the
  /// objective expression -- except, after an inequality elimination, the one
  /// on the side the inequality already bounds (objvar.lo when minimising,
  /// objvar.up when maximising), which restricts nothing and is dropped with
  /// a warning (reference frontend.cpp:1253-1317).
  void objective_bounds(bool from_inequality)
  {
    const Symbol& s = w_.symbols[static_cast<std::size_t>(w_.objvar)];
    if (!s.eliminated) {
      return;
    }
    bool const drop_lo = from_inequality && !w_.maximise;
    bool const drop_up = from_inequality && w_.maximise;
    if (std::isfinite(s.up) && !drop_up) {
      out_.problem.constraints.push_back({objective_root_, model::Cmp::LE, s.up});
      out_.constraint_names.push_back(s.name + ".up");
    }
    if (std::isfinite(s.lo) && !drop_lo) {
      std::size_t const neg = graph().emit(model::Op::Neg, {objective_root_});
      out_.problem.constraints.push_back({neg, model::Cmp::LE, -s.lo});
      out_.constraint_names.push_back(s.name + ".lo");
    }
    if (drop_lo && std::isfinite(s.lo)) {
      warn(s.decl_line, "'" + s.name + ".lo' constrains only the objective variable, which was substituted away; dropped");
    }
    if (drop_up && std::isfinite(s.up)) {
      warn(s.decl_line, "'" + s.name + ".up' constrains only the objective variable, which was substituted away; dropped");
    }
  }

private:
  static constexpr std::size_t kNone = ~std::size_t(0);

  model::ExprDAG<double>& graph() { return out_.problem.graph; }
  void warn(int line, std::string msg) { out_.warnings.push_back({line, std::move(msg)}); }
  std::size_t node_of(int term) const { return node_[static_cast<std::size_t>(term)]; }

  std::size_t constant(double v)
  {
    model::DAGNodePayload<double> p;
    p.constant = v;
    return graph().emit(model::Op::Const, {}, p);
  }

  /// x * x of one node is Sqr (tighter than a product of independent factors).
  std::size_t binary(model::Op op, std::size_t a, std::size_t b)
  {
    if (op == model::Op::Mul && a == b) {
      return graph().emit(model::Op::Sqr, {a});
    }
    return graph().emit(op, {a, b});
  }

  /// Give a DAG node to every term reachable from `roots`: mark reachable
  /// terms walking down the pool (a variable-free subtree is one Const, and a
  /// constant exponent lives in its power's node, so neither is descended
  /// into), then emit the marked ones walking up. Both are plain loops:
  /// MINLPLib has single expressions with hundreds of thousands of terms.
  void lower(const std::vector<int>& roots)
  {
    if (roots.empty()) {
      return;
    }
    int const top = *std::max_element(roots.begin(), roots.end());
    std::vector<char> live(static_cast<std::size_t>(top) + 1, 0);
    for (int r : roots) {
      live[static_cast<std::size_t>(r)] = 1;
    }
    for (int i = top; i >= 0; --i) {
      if (!live[static_cast<std::size_t>(i)] || w_.terms.constant(i)) {
        continue;
      }
      const Term& t = w_.terms[i];
      int n = t.n_args;
      if (t.kind == TermKind::Call && t.fn == Fn::Power && w_.terms.constant(t.arg[1])) {
        n = 1;
      }
      for (int k = 0; k < n; ++k) {
        live[static_cast<std::size_t>(t.arg[static_cast<std::size_t>(k)])] = 1;
      }
    }
    node_.resize(static_cast<std::size_t>(w_.terms.size()), kNone);
    for (int i = 0; i <= top; ++i) {
      if (live[static_cast<std::size_t>(i)] && node_[static_cast<std::size_t>(i)] == kNone) {
        node_[static_cast<std::size_t>(i)] = emit(i);
      }
    }
  }

  std::size_t emit(int i)
  {
    if (auto v = w_.terms.constant(i)) {
      return constant(*v);
    }
    const Term& t = w_.terms[i];
    auto in = [&](int k) { return node_of(t.arg[static_cast<std::size_t>(k)]); };
    switch (t.kind) {
      case TermKind::Literal: return constant(t.value);
      case TermKind::Variable: {
        std::size_t const id = var_node_[static_cast<std::size_t>(t.sym)];
        if (id == kNone) {
          throw ParseError(t.line, "'" + w_.symbols[static_cast<std::size_t>(t.sym)].name + "' is used but has no value",
                           ErrorKind::Unrepresentable);
        }
        return id;
      }
      case TermKind::Add: return binary(model::Op::Add, in(0), in(1));
      case TermKind::Sub: return binary(model::Op::Sub, in(0), in(1));
      case TermKind::Mul: return binary(model::Op::Mul, in(0), in(1));
      case TermKind::Div: return binary(model::Op::Div, in(0), in(1));
      case TermKind::Neg: return graph().emit(model::Op::Neg, {in(0)});
      case TermKind::Call: return emit_call(t);
    }
    throw ParseError(t.line, "unhandled expression kind");
  }

  std::size_t emit_call(const Term& t)
  {
    std::size_t const a = node_of(t.arg[0]);
    auto unary = [&](model::Op op) { return graph().emit(op, {a}); };
    switch (t.fn) {
      case Fn::Sqr: return unary(model::Op::Sqr);
      case Fn::Sqrt: return unary(model::Op::Sqrt);
      case Fn::Exp: return unary(model::Op::Exp);
      case Fn::Log: return unary(model::Op::Log);
      case Fn::Abs: return unary(model::Op::Abs);
      case Fn::Sin: return unary(model::Op::Sin);
      case Fn::Cos: return unary(model::Op::Cos);
      case Fn::Tanh: return unary(model::Op::Tanh);
      case Fn::Log10:
      case Fn::Log2: {
        std::size_t const l = unary(model::Op::Log);
        double const scale = 1.0 / (t.fn == Fn::Log10 ? std::log(10.0) : std::log(2.0));
        return binary(model::Op::Mul, l, constant(scale));
      }
      case Fn::Min: return binary(model::Op::Min, a, node_of(t.arg[1]));
      case Fn::Max: return binary(model::Op::Max, a, node_of(t.arg[1]));
      case Fn::Power: {
        // GAMS defines x**y as exp(y*log(x)); a constant 0.5 is sqrt, a
        // constant integer (|e| <= 1e9) is PowN.
        auto const e = w_.terms.constant(t.arg[1]);
        if (!e) {
          std::size_t const scaled = binary(model::Op::Mul, unary(model::Op::Log), node_of(t.arg[1]));
          return graph().emit(model::Op::Exp, {scaled});
        }
        if (same_value(*e, 0.5)) {
          return unary(model::Op::Sqrt);
        }
        if (!(std::floor(*e) < *e) && std::fabs(*e) <= 1e9) {
          model::DAGNodePayload<double> p;
          p.int_exp = static_cast<int>(*e);
          return graph().emit(model::Op::PowN, {a}, p);
        }
        std::size_t const l = unary(model::Op::Log);
        std::size_t const scaled = binary(model::Op::Mul, l, constant(*e));
        return graph().emit(model::Op::Exp, {scaled});
      }
    }
    throw ParseError(t.line, "unhandled function");
  }

  Written& w_;
  const ParseOptions& opt_;
  ParsedModel& out_;
  std::vector<std::size_t> var_node_;  // per symbol
  std::vector<std::size_t> node_;      // per term
  std::size_t objective_root_ = 0;
};

}  // namespace

ParsedModel parse(std::string_view source, const ParseOptions& options)
{
  ParsedModel out;
  Written w;
  Reader(scan(source), w).read_all();
  if (!w.have_solve) {
    throw ParseError(0, "the file has no Solve statement", ErrorKind::Unrepresentable);
  }
  if (options.reject_discrete) {
    for (const Symbol& s : w.symbols) {
      if (s.kind != model::VarKind::Continuous) {
        throw ParseError(s.integral_line, "'" + s.name + "' is discrete; reject_discrete is set", ErrorKind::Discrete);
      }
    }
  }
  // Fixed variables become literals before anything inspects the terms, so
  // constant values see through them. The objective variable is exempt.
  if (options.fold_fixed_to_const) {
    std::vector<std::optional<double>> by(w.symbols.size());
    bool any = false;
    for (std::size_t i = 0; i < w.symbols.size(); ++i) {
      if (w.symbols[i].fixed && static_cast<int>(i) != w.objvar) {
        by[i] = w.symbols[i].fx;
        any = true;
      }
    }
    if (any) {
      w.terms.replace_symbols(by);
    }
  }

  Elimination e = eliminate(w, out.warnings);
  Symbol& objvar = w.symbols[static_cast<std::size_t>(w.objvar)];
  out.objvar_kept = !e.objective.has_value();
  if (e.objective) {
    objvar.eliminated = true;
  } else {
    out.warnings.push_back({0, "could not solve any equation for '" + objvar.name
                                   + "'; keeping it as a variable, which costs one search dimension"});
    e.objective = w.terms.variable(w.objvar, 0);
  }

  Lowering lowering(w, options, out);
  lowering.variables();
  lowering.objective(*e.objective);
  lowering.constraints();
  lowering.objective_bounds(e.from_inequality);
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
    throw ParseError(0, "cannot open '" + path.string() + "'", ErrorKind::Io);
  }
  std::ostringstream text;
  text << in.rdbuf();
  return parse(text.str(), options);
}

std::string result_line(Sense sense, double lower, double upper)
{
  bool const max = sense == Sense::Maximise;
  auto number = [](double v) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.17g", v);
    return std::string(buf);
  };
  std::string line = "RESULT\tsense=";
  line += max ? "max" : "min";
  line += "\tprimal=";
  line += upper < std::numeric_limits<double>::max() ? number(max ? -upper : upper) : "none";
  line += "\tdual=";
  line += lower > std::numeric_limits<double>::lowest() ? number(max ? -lower : lower) : "none";
  return line;
}

}  // namespace cmi::gams
