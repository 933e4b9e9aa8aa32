// parameters.cpp -- declaration of the driver's entries (host only).
#include "parameters.hpp"

#include <cstdlib>
#include <sstream>

namespace lsmf_app {

namespace {

double to_d(const std::string &s) { return std::strtod(trim(s).c_str(), nullptr); }
int to_i(const std::string &s) { return (int)std::strtol(trim(s).c_str(), nullptr, 10); }
bool to_b(const std::string &s) { return trim(s) == "true"; }

// "x0, y0[, z0]; x1, y1 ..." (a list of Point<dim>)
std::vector<std::vector<double>> parse_points(const std::string &v, int dim) {
  std::vector<std::vector<double>> out;
  for (const std::string &item : split_list(v, ';')) {
    if (item.empty()) continue;
    const std::vector<std::string> c = split_list(item, ',');
    if ((int)c.size() != dim) throw PrmError("a point needs " + std::to_string(dim) + " coordinates: '" + item + "'");
    std::vector<double> p;
    for (const std::string &s : c) {
      char *end = nullptr;
      const double x = std::strtod(s.c_str(), &end);
      if (s.empty() || *end) throw PrmError("bad coordinate '" + s + "'");
      p.push_back(x);
    }
    out.push_back(p);
  }
  return out;
}

std::vector<std::string> variables(int dim) {
  return dim == 2 ? std::vector<std::string>{"x", "y", "t"} : std::vector<std::string>{"x", "y", "z", "t"};
}

// id -> dim expressions, each parseable
bool check_function_map(const std::string &v, int dim) {
  for (const auto &kv : split_id_map(v)) {
    if ((int)kv.second.size() != dim) return false;
    for (const std::string &e : kv.second) Expression(e, variables(dim));
  }
  return true;
}

bool check_mask_map(const std::string &v, int dim) {
  for (const auto &kv : split_id_map(v)) {
    if ((int)kv.second.size() != dim) return false;
    for (const std::string &e : kv.second)
      if (e != "true" && e != "false") return false;
  }
  return true;
}

void declare_all(ParameterHandler &prm, Parameters &P, int dim) {
  using PT = Pattern;
  const std::string fe = "Finite element system", geo = "Geometry", mat = "Material properties",
                    lin = "Linear solver", nl = "Nonlinear solver", tm = "Time", misc = "Misc",
                    bc = "Boundary conditions";
  prm.declare(fe, "Polynomial degree", "2", PT::Integer(1), [&](const std::string &v) { P.poly_degree = to_i(v); });
  prm.declare(fe, "Quadrature order", "3", PT::Integer(1), [&](const std::string &v) { P.quad_order = to_i(v); });

  prm.declare(geo, "Elements per edge", "32", PT::Integer(0),
              [&](const std::string &v) { P.elements_per_edge = to_i(v); });
  prm.declare(geo, "Global refinement", "0", PT::Integer(0),
              [&](const std::string &v) { P.n_global_refinement = to_i(v); });
  prm.declare(geo, "Grid scale", "1e-3", PT::Double(0.0), [&](const std::string &v) { P.scale = to_d(v); });
  prm.declare(geo, "Extrusion height", "1", PT::Double(0.0),
              [&](const std::string &v) { P.extrusion_height = to_d(v); });
  prm.declare(geo, "Extrusion slices", "5", PT::Integer(0),
              [&](const std::string &v) { P.extrusion_slices = to_i(v); });
  prm.declare(geo, "Dimension", "2", PT::Integer(2, 3), [&](const std::string &v) { P.dim = to_i(v); });
  prm.declare(geo, "Type", "Cook", PT::Selection("Cook|Holes|Box"), [&](const std::string &v) { P.type = trim(v); });

  prm.declare(mat, "Poisson's ratio", "0.3", PT::Double(-1.0, 0.5), [&](const std::string &v) { P.nu = to_d(v); });
  prm.declare(mat, "Shear modulus", "0.4225e6", PT::Double(0.), [&](const std::string &v) { P.mu = to_d(v); });
  prm.declare(mat, "Formulation", "0", PT::Integer(0, 1),
              [&](const std::string &v) { P.material_formulation = to_i(v); });
  prm.declare(mat, "Young's modulus", "0", PT::Double(0.), [&](const std::string &v) { P.youngs_modulus = to_d(v); });
  prm.declare(mat, "Material field", "uniform", PT::Selection("uniform|inclusions|lognormal"),
              [&](const std::string &v) { P.material_field = trim(v); });
  prm.declare(mat, "Inclusion count", "0", PT::Integer(0),
              [&](const std::string &v) { P.inclusion_count = to_i(v); });
  prm.declare(mat, "Inclusion radius", "0.05", PT::Double(0.),
              [&](const std::string &v) { P.inclusion_radius = to_d(v); });
  prm.declare(mat, "Inclusion factor", "10", PT::Double(0.),
              [&](const std::string &v) { P.inclusion_factor = to_d(v); });
  prm.declare(mat, "Log-normal sigma", "0.5", PT::Double(0.),
              [&](const std::string &v) { P.lognormal_sigma = to_d(v); });
  prm.declare(mat, "Random seed", "20261202", PT::Integer(0),
              [&](const std::string &v) { P.random_seed = std::strtoll(trim(v).c_str(), nullptr, 10); });

  prm.declare(lin, "Solver type", "CG", PT::Selection("CG|Direct|MF_CG|MF_AD_CG"),
              [&](const std::string &v) { P.type_lin = trim(v); });
  prm.declare(lin, "Residual", "1e-6", PT::Double(0.0), [&](const std::string &v) { P.tol_lin = to_d(v); });
  prm.declare(lin, "Max iteration multiplier", "1", PT::Integer(1),
              [&](const std::string &v) { P.max_iterations_lin = to_i(v); });
  prm.declare(lin, "Preconditioner type", "jacobi", PT::Selection("jacobi|ssor|amg|gmg|gmg-gc|none"),
              [&](const std::string &v) { P.preconditioner_type = trim(v); });
  prm.declare(lin, "Preconditioner relaxation", "0.65", PT::Double(0.0),
              [&](const std::string &v) { P.preconditioner_relaxation = to_d(v); });
  prm.declare(lin, "Preconditioner AMG aggregation threshold", "1e-4", PT::Double(0.0),
              [&](const std::string &v) { P.preconditioner_aggregation_threshold = to_d(v); });
  prm.declare(lin, "Condition number CG iterations", "20", PT::Integer(1),
              [&](const std::string &v) { P.cond_number_cg_iterations = to_i(v); });
  prm.declare(lin, "MF caching", "scalar", PT::Selection("scalar|scalar_referential|tensor2|tensor4|tensor4_ns"),
              [&](const std::string &v) { P.mf_caching = trim(v); });
  prm.declare(lin, "MF Chebyshev number CG iterations", "30", PT::Integer(2),
              [&](const std::string &v) { P.mf_chebyshev_n_cg_iterations = to_i(v); });
  prm.declare(lin, "MF Chebyshev coarse", "true", PT::Bool(),
              [&](const std::string &v) { P.mf_coarse_chebyshev = to_b(v); });
  prm.declare(lin, "MF Chebyshev coarse accurate eigenvalues", "true", PT::Bool(),
              [&](const std::string &v) { P.mf_coarse_chebyshev_accurate_eigenval = to_b(v); });
  prm.declare(lin, "MG level precision", "float", PT::Selection("float|double"),
              [&](const std::string &v) { P.mg_level_precision = trim(v); });

  prm.declare(nl, "Max iterations Newton-Raphson", "10", PT::Integer(0),
              [&](const std::string &v) { P.max_iterations_NR = to_i(v); });
  prm.declare(nl, "Tolerance force", "1e-9", PT::Double(0.0), [&](const std::string &v) { P.tol_f = to_d(v); });
  prm.declare(nl, "Absolute tolerance force", "0", PT::Double(0.0),
              [&](const std::string &v) { P.tol_f_abs = to_d(v); });
  prm.declare(nl, "Absolute tolerance displacement", "0", PT::Double(0.0),
              [&](const std::string &v) { P.tol_u_abs = to_d(v); });
  prm.declare(nl, "Tolerance displacement", "1e-6", PT::Double(0.0), [&](const std::string &v) { P.tol_u = to_d(v); });

  prm.declare(tm, "End time", "1", PT::Double(), [&](const std::string &v) { P.end_time = to_d(v); });
  prm.declare(tm, "Time step size", "0.1", PT::Double(0.), [&](const std::string &v) { P.delta_t = to_d(v); });

  prm.declare(misc, "Output folder", "", PT::Anything(), [&](const std::string &v) {
    P.output_folder = trim(v);
    if (!P.output_folder.empty() && P.output_folder.back() != '/') P.output_folder += "/";
  });
  prm.declare(misc, "Output solution", "true", PT::Bool(), [&](const std::string &v) { P.output_solution = to_b(v); });
  prm.declare(misc, "Always assemble tangent", "true", PT::Bool(),
              [&](const std::string &v) { P.always_assemble_tangent = to_b(v); });
  prm.declare(misc, "Output absolute norms", "false", PT::Bool(),
              [&](const std::string &v) { P.output_abs_norms = to_b(v); });
  prm.declare(misc, "Output points", "",
              PT::Custom("[List of Point<" + std::to_string(dim) + ">]",
                         [dim](const std::string &v) { return parse_points(v, dim), true; }),
              [&, dim](const std::string &v) { P.output_points = parse_points(v, dim); });
  prm.declare(misc, "Benchmark vmult", "0", PT::Integer(0),
              [&](const std::string &v) { P.benchmark_vmult = to_i(v); });

  prm.declare(bc, "Dirichlet IDs and expressions", "",
              PT::Custom("[Map of boundary id : FunctionParser<" + std::to_string(dim) + ">]",
                         [dim](const std::string &v) { return check_function_map(v, dim); }),
              [&](const std::string &v) { P.dirichlet = split_id_map(v); });
  prm.declare(bc, "Dirichlet IDs and component mask", "",
              PT::Custom("[Map of boundary id : ComponentMask]",
                         [dim](const std::string &v) { return check_mask_map(v, dim); }),
              [&](const std::string &v) {
                P.dirichlet_mask.clear();
                for (const auto &kv : split_id_map(v))
                  for (const std::string &b : kv.second) P.dirichlet_mask[kv.first].push_back(b == "true");
              });
  prm.declare(bc, "Neumann IDs and expressions", "",
              PT::Custom("[Map of boundary id : FunctionParser<" + std::to_string(dim) + ">]",
                         [dim](const std::string &v) { return check_function_map(v, dim); }),
              [&](const std::string &v) { P.neumann = split_id_map(v); });
}

}  // namespace

void read_dimension_and_degree(const std::string &file, int &dim, int &degree, int &quad) {
  ParameterHandler prm;
  Parameters P;
  prm.declare("Geometry", "Dimension", "2", Pattern::Integer(2, 3), [&](const std::string &v) { P.dim = to_i(v); });
  prm.declare("Finite element system", "Polynomial degree", "2", Pattern::Integer(1),
              [&](const std::string &v) { P.poly_degree = to_i(v); });
  prm.declare("Finite element system", "Quadrature order", "3", Pattern::Integer(1),
              [&](const std::string &v) { P.quad_order = to_i(v); });
  prm.parse_file(file, true);
  prm.finalize();
  dim = P.dim;
  degree = P.poly_degree;
  quad = P.quad_order;
}

Parameters read_parameters(const std::string &file, int dim) {
  ParameterHandler prm;
  Parameters P;
  declare_all(prm, P, dim);
  prm.parse_file(file, false);
  prm.finalize();
  if (P.dim != dim) throw PrmError("Dimension " + std::to_string(P.dim) + " != " + std::to_string(dim));
  return P;
}

std::string describe(const std::string &file, int dim) {
  ParameterHandler prm;
  Parameters P;
  declare_all(prm, P, dim);
  prm.parse_file(file, false);
  prm.finalize();
  std::ostringstream os;
  for (const std::string &k : prm.entries()) {
    const size_t slash = k.rfind('/');
    os << k << " = " << prm.get(k.substr(0, slash), k.substr(slash + 1)) << "\n";
  }
  return os.str();
}

}  // namespace lsmf_app
