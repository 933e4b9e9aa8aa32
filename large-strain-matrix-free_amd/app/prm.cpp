// prm.cpp -- parameter-file reader and expression evaluator (host only).
#include "prm.hpp"

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace lsmf_app {

std::string trim(const std::string &s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace((unsigned char)s[a])) ++a;
  while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
  return s.substr(a, b - a);
}

std::vector<std::string> split_list(const std::string &value, char sep) {
  std::vector<std::string> out;
  const std::string v = trim(value);
  if (v.empty()) return out;
  std::string cur;
  for (char c : v) {
    if (c == sep) {
      out.push_back(trim(cur));
      cur.clear();
    } else {
      cur += c;
    }
  }
  out.push_back(trim(cur));
  return out;
}

std::map<int, std::vector<std::string>> split_id_map(const std::string &value) {
  std::map<int, std::vector<std::string>> out;
  for (const std::string &item : split_list(value, ';')) {
    if (item.empty()) continue;
    const size_t colon = item.find(':');
    if (colon == std::string::npos) throw PrmError("expected <id>:<values> in '" + item + "'");
    const std::string key = trim(item.substr(0, colon));
    char *end = nullptr;
    const long id = std::strtol(key.c_str(), &end, 10);
    if (key.empty() || *end || id < 0 || id > 255) throw PrmError("bad boundary id '" + key + "'");
    if (out.count((int)id)) throw PrmError("boundary id " + key + " given twice");
    out[(int)id] = split_list(item.substr(colon + 1), ',');
  }
  return out;
}

// ---- patterns --------------------------------------------------------------
namespace {

bool parse_ll(const std::string &s, long long &v) {
  const std::string t = trim(s);
  if (t.empty()) return false;
  char *end = nullptr;
  errno = 0;
  v = std::strtoll(t.c_str(), &end, 10);
  return errno == 0 && *end == '\0';
}

bool parse_double(const std::string &s, double &v) {
  const std::string t = trim(s);
  if (t.empty()) return false;
  char *end = nullptr;
  errno = 0;
  v = std::strtod(t.c_str(), &end);
  return errno == 0 && *end == '\0' && std::isfinite(v);
}

}  // namespace

Pattern Pattern::Integer(long long lo, long long hi) {
  Pattern p;
  p.kind = INTEGER;
  p.lo = (double)lo;
  p.hi = (double)hi;
  p.text = "[Integer range " + std::to_string(lo) + "..." + (hi >= (1ll << 62) ? "MAX_INT" : std::to_string(hi)) +
           " (inclusive)]";
  return p;
}

Pattern Pattern::Double(double lo, double hi) {
  Pattern p;
  p.kind = DOUBLE;
  p.lo = lo;
  p.hi = hi;
  std::ostringstream os;
  os << "[Double " << (lo <= -1e300 ? std::string("-MAX_DOUBLE") : std::to_string(lo)) << "..."
     << (hi >= 1e300 ? std::string("MAX_DOUBLE") : std::to_string(hi)) << " (inclusive)]";
  p.text = os.str();
  return p;
}

Pattern Pattern::Bool() {
  Pattern p = Selection("true|false");
  p.kind = BOOL;
  p.text = "[Bool]";
  return p;
}

Pattern Pattern::Selection(const std::string &bar) {
  Pattern p;
  p.kind = SELECTION;
  p.options = split_list(bar, '|');
  p.text = "[Selection " + bar + " ]";
  return p;
}

Pattern Pattern::Anything() {
  Pattern p;
  p.kind = ANYTHING;
  p.text = "[Anything]";
  return p;
}

Pattern Pattern::Custom(const std::string &text, std::function<bool(const std::string &)> check) {
  Pattern p;
  p.kind = CUSTOM;
  p.text = text;
  p.check = std::move(check);
  return p;
}

bool Pattern::match(const std::string &v) const {
  switch (kind) {
    case INTEGER: {
      long long x;
      return parse_ll(v, x) && x >= lo && x <= hi;
    }
    case DOUBLE: {
      double x;
      return parse_double(v, x) && x >= lo && x <= hi;
    }
    case BOOL:
    case SELECTION:
      for (const std::string &o : options)
        if (trim(v) == o) return true;
      return false;
    case CUSTOM:
      try {
        return check(v);
      } catch (const std::exception &) {
        return false;
      }
    default:
      return true;
  }
}

// ---- parameter handler -----------------------------------------------------
void ParameterHandler::declare(const std::string &section, const std::string &name, const std::string &default_value,
                               const Pattern &pattern, std::function<void(const std::string &)> action) {
  const std::string key = section + "/" + name;
  if (!pattern.match(default_value))
    throw PrmError("default value '" + default_value + "' of " + key + " does not match " + pattern.text);
  Entry e;
  e.value = e.default_value = default_value;
  e.pattern = pattern;
  e.action = std::move(action);
  if (!entries_.count(key)) order_.push_back(key);
  entries_[key] = std::move(e);
  // every prefix of the section path is a declared section
  std::string path;
  for (const std::string &part : split_list(section, '/')) {
    path = path.empty() ? part : path + "/" + part;
    sections_[path] = 1;
  }
}

void ParameterHandler::parse_file(const std::string &file, bool skip_undefined) {
  std::ifstream in(file);
  if (!in) throw PrmError("cannot open parameter file '" + file + "'");
  std::vector<std::string> lines;
  std::string line;
  while (std::getline(in, line)) lines.push_back(line);
  std::vector<std::string> path;
  int skipping = 0;
  parse_lines(lines, file, skip_undefined, 0, path, skipping);
  if (!path.empty() || skipping)
    throw PrmError("file <" + file + "> ends inside a subsection (unequal numbers of 'subsection' and 'end')");
}

void ParameterHandler::parse_string(const std::string &text, const std::string &origin, bool skip_undefined) {
  std::vector<std::string> lines;
  std::istringstream in(text);
  std::string line;
  while (std::getline(in, line)) lines.push_back(line);
  std::vector<std::string> path;
  int skipping = 0;
  parse_lines(lines, origin, skip_undefined, 0, path, skipping);
  if (!path.empty() || skipping)
    throw PrmError("input <" + origin + "> ends inside a subsection (unequal numbers of 'subsection' and 'end')");
}

namespace {

bool starts_with_keyword(const std::string &line, const char *kw, std::string &rest) {
  const size_t n = std::char_traits<char>::length(kw);
  if (line.size() < n) return false;
  for (size_t i = 0; i < n; ++i)
    if (std::tolower((unsigned char)line[i]) != kw[i]) return false;
  if (line.size() > n && !std::isspace((unsigned char)line[n])) return false;
  rest = trim(line.substr(n));
  return true;
}

}  // namespace

// path: the current subsection stack; skipping: depth inside an undeclared
// subsection being skipped.  Both carry through include (an included file is
// read in the subsection where the include line stands and must close the
// subsections it opens).
void ParameterHandler::parse_lines(const std::vector<std::string> &raw, const std::string &origin,
                                   bool skip_undefined, int depth, std::vector<std::string> &path, int &skipping) {
  if (depth > 16) throw PrmError("include nesting deeper than 16 at " + origin);
  const size_t path0 = path.size();
  const int skip0 = skipping;
  std::string joined;
  int joined_start = 0;
  for (size_t ln = 0; ln < raw.size(); ++ln) {
    std::string line = raw[ln];
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line.erase(hash);
    std::string t = trim(line);
    if (joined.empty()) joined_start = (int)ln + 1;
    if (!t.empty() && t.back() == '\\') {
      t.pop_back();
      joined += t + " ";
      continue;
    }
    t = trim(joined + t);
    joined.clear();
    if (t.empty()) continue;
    const std::string where = "Line <" + std::to_string(joined_start) + "> of file <" + origin + ">: ";
    std::string rest;
    if (starts_with_keyword(t, "subsection", rest)) {
      if (skipping) {
        ++skipping;
        continue;
      }
      std::string p;
      for (const std::string &s : path) p += s + "/";
      p += rest;
      if (!sections_.count(p)) {
        if (!skip_undefined) throw PrmError(where + "There is no subsection <" + rest + "> declared here");
        skipping = 1;
        continue;
      }
      path.push_back(rest);
    } else if (starts_with_keyword(t, "end", rest) && rest.empty()) {
      if (skipping) {
        --skipping;
        continue;
      }
      if (path.size() <= path0) throw PrmError(where + "There is no subsection to leave here");
      path.pop_back();
    } else if (starts_with_keyword(t, "set", rest)) {
      if (skipping) continue;
      const size_t eq = rest.find('=');
      if (eq == std::string::npos) throw PrmError(where + "expected 'set <entry> = <value>'");
      const std::string name = trim(rest.substr(0, eq));
      const std::string value = trim(rest.substr(eq + 1));
      std::string key;
      for (const std::string &s : path) key += s + "/";
      if (path.empty()) key = "/";
      key += name;
      auto it = entries_.find(key);
      if (it == entries_.end()) {
        if (skip_undefined) continue;
        throw PrmError(where + "No entry with name <" + name + "> was declared in the current subsection");
      }
      if (!it->second.pattern.match(value))
        throw PrmError(where + "The entry value <" + value + "> for the entry named <" + name +
                       "> does not match the given pattern " + it->second.pattern.text);
      it->second.value = value;
      it->second.set = true;
    } else if (starts_with_keyword(t, "include", rest)) {
      if (skipping) continue;
      std::ifstream in(rest);
      if (!in) throw PrmError(where + "Could not open file <" + rest + ">");
      std::vector<std::string> lines;
      std::string l;
      while (std::getline(in, l)) lines.push_back(l);
      parse_lines(lines, rest, skip_undefined, depth + 1, path, skipping);
    } else {
      throw PrmError(where + "This line matched nothing known ('set' or 'subsection' missing!?): <" + t + ">");
    }
  }
  if (!joined.empty()) throw PrmError("file <" + origin + "> ends with a continued line");
  if (path.size() != path0 || skipping != skip0)
    throw PrmError("file <" + origin + "> has unequal numbers of 'subsection' and 'end' statements");
}

void ParameterHandler::finalize() {
  for (const std::string &k : order_) {
    Entry &e = entries_[k];
    if (e.action) e.action(e.value);
  }
}

std::string ParameterHandler::get(const std::string &section, const std::string &name) const {
  auto it = entries_.find(section + "/" + name);
  if (it == entries_.end()) throw PrmError("undeclared entry " + section + "/" + name);
  return it->second.value;
}

bool ParameterHandler::was_set(const std::string &section, const std::string &name) const {
  auto it = entries_.find(section + "/" + name);
  return it != entries_.end() && it->second.set;
}

std::vector<std::string> ParameterHandler::entries() const { return order_; }

// ---- expressions -----------------------------------------------------------
struct Expression::Node {
  enum Op { NUM, VAR, NEG, ADD, SUB, MUL, DIV, POW, LT, GT, LE, GE, EQ, NE, AND, OR, NOT, TERN, FUN1, FUN2 } op;
  double value = 0;
  int var = 0;
  double (*f1)(double) = nullptr;
  double (*f2)(double, double) = nullptr;
  int depth = 1;  // height of the subtree (bounded: evaluation and destruction recurse)
  std::shared_ptr<Node> a, b, c;
};

namespace {

using NodeP = std::shared_ptr<Expression::Node>;

double f_sign(double x) { return (x > 0) - (x < 0); }
double f_rint(double x) { return std::nearbyint(x); }
double f_cot(double x) { return 1.0 / std::tan(x); }
double f_sec(double x) { return 1.0 / std::cos(x); }
double f_csc(double x) { return 1.0 / std::sin(x); }
double f_int(double x) { return (double)(long long)x; }
double f_min(double a, double b) { return a < b ? a : b; }
double f_max(double a, double b) { return a > b ? a : b; }
double f_pow(double a, double b) { return std::pow(a, b); }
double f_atan2(double a, double b) { return std::atan2(a, b); }
double f_fmod(double a, double b) { return std::fmod(a, b); }

struct Fun1 {
  const char *name;
  double (*f)(double);
};
const Fun1 kFun1[] = {{"sin", [](double x) { return std::sin(x); }},   {"cos", [](double x) { return std::cos(x); }},
                      {"tan", [](double x) { return std::tan(x); }},   {"asin", [](double x) { return std::asin(x); }},
                      {"acos", [](double x) { return std::acos(x); }}, {"atan", [](double x) { return std::atan(x); }},
                      {"sinh", [](double x) { return std::sinh(x); }}, {"cosh", [](double x) { return std::cosh(x); }},
                      {"tanh", [](double x) { return std::tanh(x); }}, {"asinh", [](double x) { return std::asinh(x); }},
                      {"acosh", [](double x) { return std::acosh(x); }},
                      {"atanh", [](double x) { return std::atanh(x); }},
                      {"log2", [](double x) { return std::log2(x); }}, {"log10", [](double x) { return std::log10(x); }},
                      {"log", [](double x) { return std::log(x); }},   {"ln", [](double x) { return std::log(x); }},
                      {"exp", [](double x) { return std::exp(x); }},   {"sqrt", [](double x) { return std::sqrt(x); }},
                      {"abs", [](double x) { return std::fabs(x); }},  {"floor", [](double x) { return std::floor(x); }},
                      {"ceil", [](double x) { return std::ceil(x); }}, {"sign", f_sign},
                      {"rint", f_rint},                                {"cot", f_cot},
                      {"sec", f_sec},                                  {"csc", f_csc},
                      {"int", f_int},                                  {"erf", [](double x) { return std::erf(x); }},
                      {"erfc", [](double x) { return std::erfc(x); }}};
struct Fun2 {
  const char *name;
  double (*f)(double, double);
};
const Fun2 kFun2[] = {{"min", f_min}, {"max", f_max}, {"pow", f_pow}, {"atan2", f_atan2}, {"fmod", f_fmod}};

class Parser {
 public:
  Parser(const std::string &s, const std::vector<std::string> &vars) : s_(s), vars_(vars) {}
  NodeP parse() {
    NodeP n = ternary();
    skip();
    if (i_ != s_.size()) fail("unexpected '" + s_.substr(i_) + "'");
    return n;
  }

 private:
  // nesting limits: parentheses, unary signs and ternaries recurse in the
  // parser, and the tree's height bounds the recursion of eval / destruction
  static constexpr int kMaxNest = 512;
  static constexpr int kMaxDepth = 1024;
  const std::string &s_;
  const std::vector<std::string> &vars_;
  size_t i_ = 0;
  int nest_ = 0;
  struct Nest {
    Parser &p;
    explicit Nest(Parser &q) : p(q) {
      if (++p.nest_ > kMaxNest) p.fail("nested more than " + std::to_string(kMaxNest) + " levels deep");
    }
    ~Nest() { --p.nest_; }
  };

  [[noreturn]] void fail(const std::string &m) const {
    const std::string shown = s_.size() > 120 ? s_.substr(0, 120) + "..." : s_;
    throw PrmError("cannot parse expression '" + shown + "': " + m);
  }
  void skip() {
    while (i_ < s_.size() && std::isspace((unsigned char)s_[i_])) ++i_;
  }
  bool eat(const char *tok) {
    skip();
    const size_t n = std::char_traits<char>::length(tok);
    if (s_.compare(i_, n, tok) == 0) {
      i_ += n;
      return true;
    }
    return false;
  }
  NodeP mk(Expression::Node::Op op, NodeP a = nullptr, NodeP b = nullptr, NodeP c = nullptr) {
    auto n = std::make_shared<Expression::Node>();
    for (const NodeP *ch : {&a, &b, &c})
      if (*ch) n->depth = std::max(n->depth, (*ch)->depth + 1);
    if (n->depth > kMaxDepth) fail("more than " + std::to_string(kMaxDepth) + " operations deep");
    n->op = op;
    n->a = std::move(a);
    n->b = std::move(b);
    n->c = std::move(c);
    return n;
  }
  NodeP ternary() {
    Nest guard(*this);
    NodeP c = logic_or();
    if (eat("?")) {
      NodeP a = ternary();
      if (!eat(":")) fail("expected ':'");
      NodeP b = ternary();
      return mk(Expression::Node::TERN, a, b, c);
    }
    return c;
  }
  NodeP logic_or() {
    NodeP a = logic_and();
    while (eat("||")) a = mk(Expression::Node::OR, a, logic_and());
    return a;
  }
  NodeP logic_and() {
    NodeP a = compare();
    while (eat("&&")) a = mk(Expression::Node::AND, a, compare());
    return a;
  }
  NodeP compare() {
    NodeP a = additive();
    for (;;) {
      if (eat("<=")) a = mk(Expression::Node::LE, a, additive());
      else if (eat(">=")) a = mk(Expression::Node::GE, a, additive());
      else if (eat("==")) a = mk(Expression::Node::EQ, a, additive());
      else if (eat("!=")) a = mk(Expression::Node::NE, a, additive());
      else if (eat("<")) a = mk(Expression::Node::LT, a, additive());
      else if (eat(">")) a = mk(Expression::Node::GT, a, additive());
      else return a;
    }
  }
  NodeP additive() {
    NodeP a = multiplicative();
    for (;;) {
      if (eat("+")) a = mk(Expression::Node::ADD, a, multiplicative());
      else if (eat("-")) a = mk(Expression::Node::SUB, a, multiplicative());
      else return a;
    }
  }
  NodeP multiplicative() {
    NodeP a = unary();
    for (;;) {
      if (eat("*")) a = mk(Expression::Node::MUL, a, unary());
      else if (eat("/")) a = mk(Expression::Node::DIV, a, unary());
      else return a;
    }
  }
  // unary signs bind looser than ^ (-2^2 = -4)
  NodeP unary() {
    Nest guard(*this);
    if (eat("-")) return mk(Expression::Node::NEG, unary());
    if (eat("+")) return unary();
    if (eat("!")) return mk(Expression::Node::NOT, unary());
    return power();
  }
  NodeP power() {
    NodeP a = primary();
    if (eat("^")) return mk(Expression::Node::POW, a, unary());  // right associative
    return a;
  }
  NodeP primary() {
    skip();
    if (i_ >= s_.size()) fail("unexpected end");
    const char c = s_[i_];
    if (std::isdigit((unsigned char)c) || c == '.') {
      const char *b = s_.c_str() + i_;
      char *e = nullptr;
      const double v = std::strtod(b, &e);
      if (e == b) fail("bad number");
      i_ += (size_t)(e - b);
      auto n = mk(Expression::Node::NUM);
      n->value = v;
      return n;
    }
    if (eat("(")) {
      NodeP n = ternary();
      if (!eat(")")) fail("expected ')'");
      return n;
    }
    if (std::isalpha((unsigned char)c) || c == '_') {
      size_t j = i_;
      while (j < s_.size() && (std::isalnum((unsigned char)s_[j]) || s_[j] == '_')) ++j;
      const std::string id = s_.substr(i_, j - i_);
      i_ = j;
      skip();
      if (i_ < s_.size() && s_[i_] == '(') {
        ++i_;
        NodeP a = ternary();
        if (eat(",")) {
          NodeP b = ternary();
          if (!eat(")")) fail("expected ')'");
          for (const Fun2 &f : kFun2)
            if (id == f.name) {
              auto n = mk(Expression::Node::FUN2, a, b);
              n->f2 = f.f;
              return n;
            }
          fail("unknown function " + id + "(a, b)");
        }
        if (!eat(")")) fail("expected ')'");
        for (const Fun1 &f : kFun1)
          if (id == f.name) {
            auto n = mk(Expression::Node::FUN1, a);
            n->f1 = f.f;
            return n;
          }
        fail("unknown function " + id + "(x)");
      }
      for (size_t k = 0; k < vars_.size(); ++k)
        if (id == vars_[k]) {
          auto n = mk(Expression::Node::VAR);
          n->var = (int)k;
          return n;
        }
      auto n = mk(Expression::Node::NUM);
      if (id == "pi" || id == "Pi" || id == "_pi") n->value = M_PI;
      else if (id == "_e") n->value = M_E;
      else fail("unknown variable or constant '" + id + "'");
      return n;
    }
    fail(std::string("unexpected '") + c + "'");
  }
};

double eval_node(const Expression::Node &n, const double *v) {
  using N = Expression::Node;
  switch (n.op) {
    case N::NUM: return n.value;
    case N::VAR: return v[n.var];
    case N::NEG: return -eval_node(*n.a, v);
    case N::NOT: return eval_node(*n.a, v) == 0.0 ? 1.0 : 0.0;
    case N::ADD: return eval_node(*n.a, v) + eval_node(*n.b, v);
    case N::SUB: return eval_node(*n.a, v) - eval_node(*n.b, v);
    case N::MUL: return eval_node(*n.a, v) * eval_node(*n.b, v);
    case N::DIV: return eval_node(*n.a, v) / eval_node(*n.b, v);
    case N::POW: return std::pow(eval_node(*n.a, v), eval_node(*n.b, v));
    case N::LT: return eval_node(*n.a, v) < eval_node(*n.b, v);
    case N::GT: return eval_node(*n.a, v) > eval_node(*n.b, v);
    case N::LE: return eval_node(*n.a, v) <= eval_node(*n.b, v);
    case N::GE: return eval_node(*n.a, v) >= eval_node(*n.b, v);
    case N::EQ: return eval_node(*n.a, v) == eval_node(*n.b, v);
    case N::NE: return eval_node(*n.a, v) != eval_node(*n.b, v);
    case N::AND: return (eval_node(*n.a, v) != 0.0) && (eval_node(*n.b, v) != 0.0);
    case N::OR: return (eval_node(*n.a, v) != 0.0) || (eval_node(*n.b, v) != 0.0);
    case N::TERN: return eval_node(*n.c, v) != 0.0 ? eval_node(*n.a, v) : eval_node(*n.b, v);
    case N::FUN1: return n.f1(eval_node(*n.a, v));
    case N::FUN2: return n.f2(eval_node(*n.a, v), eval_node(*n.b, v));
  }
  return 0.0;
}

}  // namespace

Expression::Expression(const std::string &text, const std::vector<std::string> &variables) : text_(text) {
  if (trim(text).empty()) throw PrmError("empty expression");
  root_ = Parser(text_, variables).parse();
}

double Expression::eval(const double *values) const { return eval_node(*root_, values); }

}  // namespace lsmf_app
