// prm.hpp -- parameter files and parsed expressions for the lsmf_run driver.
//
// ParameterHandler: the input syntax of deal.II's ParameterHandler as the
// reference's driver reads it (include/mf_elasticity.h:156-607,
// src/main.cc:71-80):
//   subsection <name> ... end       nested sections
//   set <entry> = <value>           value checked against the entry's pattern
//   include <file>                  the file's lines are read in place
//   # comment, trailing backslash joins the next line
// Undeclared sections or entries are errors, unless the parse is asked to
// skip them (the first pass of src/main.cc:80 does).
//
// Expression: the FunctionParser expressions of the boundary-condition keys
// ("Dirichlet IDs and expressions" ...), a muparser-style grammar with the
// variables x, y[, z] and t, constants pi/Pi/_pi/_e, + - * / ^, comparisons,
// && || and ?:, and the usual one- and two-argument functions.
#pragma once

#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace lsmf_app {

struct PrmError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// value pattern of one entry (Patterns::Integer / Double / Bool / Selection /
// Anything, plus a free-form check for the list and map entries)
struct Pattern {
  enum Kind { INTEGER, DOUBLE, BOOL, SELECTION, ANYTHING, CUSTOM } kind = ANYTHING;
  double lo = -1e300, hi = 1e300;
  std::vector<std::string> options;  // SELECTION
  std::function<bool(const std::string &)> check;  // CUSTOM
  std::string text;  // description for error messages

  static Pattern Integer(long long lo = -(1ll << 62), long long hi = (1ll << 62));
  static Pattern Double(double lo = -1e300, double hi = 1e300);
  static Pattern Bool();
  static Pattern Selection(const std::string &bar_separated);
  static Pattern Anything();
  static Pattern Custom(const std::string &text, std::function<bool(const std::string &)> check);
  bool match(const std::string &v) const;
};

class ParameterHandler {
 public:
  // entry path: "Section/Sub/Entry"; the action runs on every set (and once
  // with the default at finalize(), unless the entry was set)
  void declare(const std::string &section, const std::string &name, const std::string &default_value,
               const Pattern &pattern, std::function<void(const std::string &)> action);
  void parse_file(const std::string &file, bool skip_undefined = false);
  void parse_string(const std::string &text, const std::string &origin = "<string>", bool skip_undefined = false);
  // run every action: defaults first for entries never set, then set values
  void finalize();
  std::string get(const std::string &section, const std::string &name) const;
  bool was_set(const std::string &section, const std::string &name) const;
  std::vector<std::string> entries() const;  // "Section/Entry" in declaration order

 private:
  struct Entry {
    std::string value, default_value;
    Pattern pattern;
    std::function<void(const std::string &)> action;
    bool set = false;
  };
  std::map<std::string, Entry> entries_;
  std::vector<std::string> order_;
  std::map<std::string, int> sections_;  // declared section paths
  void parse_lines(const std::vector<std::string> &lines, const std::string &origin, bool skip_undefined,
                   int depth, std::vector<std::string> &path, int &skipping);
};

// ---- expressions -----------------------------------------------------------
class Expression {
 public:
  // variables: names bound, in order, to the values passed to eval()
  Expression(const std::string &text, const std::vector<std::string> &variables);
  double eval(const double *values) const;
  const std::string &text() const { return text_; }
  struct Node;  // parse tree (defined in prm.cpp)

 private:
  std::string text_;
  std::shared_ptr<Node> root_;
};

// "<id>:<f0>,<f1>[,<f2>];<id>:..." -> id -> component expressions
std::map<int, std::vector<std::string>> split_id_map(const std::string &value);
// comma-separated list, each item trimmed
std::vector<std::string> split_list(const std::string &value, char sep);
std::string trim(const std::string &s);

}  // namespace lsmf_app
