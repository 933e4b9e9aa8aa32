// prm_check.cpp -- host-only checker of the lsmf_run parameter-file front end
// (ParameterHandler syntax, the reference's entries, FunctionParser-style
// boundary expressions), built with AddressSanitizer and UBSan by the
// package Makefile's `sanitize` target.  It runs the same parsing code as
// lsmf_run without the library or a GPU:
//   prm_check file.prm        read the file as lsmf_run does (both passes of
//                             src/main.cc:71-80) and evaluate every boundary
//                             expression at a few points and times
//   prm_check --expr TEXT     parse TEXT with the variables x, y, z, t and
//                             evaluate it at (0.5, 0.25, 0.125, 1)
// Exit status 0 = accepted, 1 = rejected with a message on stderr (the
// "Exception on processing" path of lsmf_run); a memory or undefined-
// behaviour error aborts with the sanitizer's report instead.
#include <cstdio>
#include <exception>
#include <iostream>
#include <string>
#include <vector>

#include "parameters.hpp"
#include "prm.hpp"

namespace {

int check_expression(const std::string &text) {
  const lsmf_app::Expression e(text, {"x", "y", "z", "t"});
  const double v[4] = {0.5, 0.25, 0.125, 1.0};
  std::printf("%.17g\n", e.eval(v));
  return 0;
}

int check_file(const std::string &file) {
  int dim = 0, degree = 0, quad = 0;
  lsmf_app::read_dimension_and_degree(file, dim, degree, quad);
  const lsmf_app::Parameters P = lsmf_app::read_parameters(file, dim);
  std::vector<std::string> vars = {"x", "y"};
  if (dim == 3) vars.push_back("z");
  vars.push_back("t");
  int n_expr = 0;
  for (const auto *bc : {&P.dirichlet, &P.neumann})
    for (const auto &kv : *bc)
      for (const std::string &text : kv.second) {
        const lsmf_app::Expression e(text, vars);
        for (double t : {0.0, 0.5, 1.0}) {
          const double v[4] = {0.25, 0.75, dim == 3 ? 0.5 : t, t};
          (void)e.eval(v);
        }
        ++n_expr;
      }
  std::printf("dim %d degree %d quadrature %d: %d boundary expressions\n", dim, degree, quad, n_expr);
  std::fputs(lsmf_app::describe(file, dim).c_str(), stdout);
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc == 3 && std::string(argv[1]) == "--expr") {
    try {
      return check_expression(argv[2]);
    } catch (const std::exception &e) {
      std::cerr << e.what() << std::endl;
      return 1;
    }
  }
  if (argc != 2) {
    std::cerr << "usage: prm_check file.prm | prm_check --expr TEXT" << std::endl;
    return 2;
  }
  try {
    return check_file(argv[1]);
  } catch (const std::exception &e) {
    std::cerr << "Exception on processing: " << std::endl << e.what() << std::endl << "Aborting!" << std::endl;
    return 1;
  }
}
