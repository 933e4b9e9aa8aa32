// parameters.hpp -- the run-time parameters of the lsmf_run driver.
//
// Same sections, entry names, defaults and patterns as the reference's
// Parameters::AllParameters (include/mf_elasticity.h:156-607), so its .prm
// files read unchanged.  Entries marked "new" have no reference counterpart:
// the unit-cube geometry and the per-cell material fields of the benchmark
// workloads, and the multigrid level precision.
#pragma once

#include <array>
#include <map>
#include <string>
#include <vector>

#include "prm.hpp"

namespace lsmf_app {

struct Parameters {
  // Finite element system (:247-272)
  int poly_degree = 2;
  int quad_order = 3;
  // Geometry (:276-332); "Box" is new: the unit square / cube [0,1]^dim
  int elements_per_edge = 32;
  double scale = 1e-3;
  int dim = 2;
  int n_global_refinement = 0;
  int extrusion_slices = 5;
  double extrusion_height = 1;
  std::string type = "Cook";
  // Material properties (:336-377)
  double nu = 0.3;
  double mu = 0.4225e6;
  int material_formulation = 0;
  // new: per-cell fields ("uniform" is the reference's single material)
  double youngs_modulus = 0;  // > 0: mu = E / (2 (1 + nu)) replaces "Shear modulus"
  std::string material_field = "uniform";
  int inclusion_count = 0;
  double inclusion_radius = 0.05;
  double inclusion_factor = 10.0;
  double lognormal_sigma = 0.5;
  long long random_seed = 20261202;
  // Linear solver (:383-468)
  std::string type_lin = "CG";
  double tol_lin = 1e-6;
  int max_iterations_lin = 1;
  std::string preconditioner_type = "jacobi";
  double preconditioner_relaxation = 0.65;
  double preconditioner_aggregation_threshold = 1e-4;
  int cond_number_cg_iterations = 20;
  std::string mf_caching = "scalar";
  bool mf_coarse_chebyshev = true;
  bool mf_coarse_chebyshev_accurate_eigenval = true;
  int mf_chebyshev_n_cg_iterations = 30;
  std::string mg_level_precision = "float";  // new: the reference's levels are float
  // Nonlinear solver (:474-520)
  int max_iterations_NR = 10;
  double tol_f = 1e-9;
  double tol_f_abs = 0;
  double tol_u = 1e-6;
  double tol_u_abs = 0;
  // Time (:524-548)
  double delta_t = 0.1;
  double end_time = 1.0;
  // Misc (:158-211)
  std::string output_folder;
  bool output_solution = true;
  bool always_assemble_tangent = true;
  bool output_abs_norms = false;
  std::vector<std::vector<double>> output_points;
  // Boundary conditions (:213-242): id -> component expressions / mask
  std::map<int, std::vector<std::string>> dirichlet;
  std::map<int, std::vector<bool>> dirichlet_mask;
  std::map<int, std::vector<std::string>> neumann;
  // new: repetitions of the vmult timing loop per load step (the reference
  // times 10 per Newton step in solve_nonlinear_timestep, :2340-2450)
  int benchmark_vmult = 0;
};

// First pass of src/main.cc:71-80: only "Geometry/Dimension" and the FE
// degree are needed to pick the instantiation; undeclared entries skipped.
void read_dimension_and_degree(const std::string &file, int &dim, int &degree, int &quad);

// Full pass (AllParameters ctor, :583-604): undeclared entries are errors.
Parameters read_parameters(const std::string &file, int dim);

// every entry with its value after the parse, "Section/Entry = value" lines
std::string describe(const std::string &file, int dim);

}  // namespace lsmf_app
