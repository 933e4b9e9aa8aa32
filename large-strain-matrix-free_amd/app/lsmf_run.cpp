// lsmf_run -- run a large-strain Neo-Hookean problem from a parameter file.
//
//   lsmf_run [--dry-run] [--describe-setup] [--print-parameters]
//            [--dump-material FILE] file.prm
//
// The reference's executable (src/main.cc, Solid::run in
// include/mf_elasticity.h:1187-1222) on the library's GPU path: the same
// parameter file, the same blessed "output" text in the output folder
// (Grid / Triangulation / Timestep / Converged / Solution lines), a
// timings.txt report, and the same exception report and exit status 1 on
// failure.  --dry-run stops after the mesh and material set-up and prints
// the header lines only (no GPU is touched); --describe-setup adds, as one
// JSON line, the solver set-up derived from the file (levels, boundary
// conditions at the first load step); --print-parameters lists every entry
// with its value.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "lsmf.h"
#include "lsmf_fe.hpp"
#include "parameters.hpp"
#include "prm.hpp"

using namespace lsmf_app;

namespace {

using Clock = std::chrono::steady_clock;
double seconds_since(Clock::time_point t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }

void check(lsmf_status rc, const char *what) {
  if (rc != LSMF_OK) throw std::runtime_error(std::string(what) + ": " + lsmf_last_error());
}

// ---- seeded material fields (SplitMix64, Steele, Lea and Flood 2014) -------
struct SplitMix64 {
  uint64_t state;
  explicit SplitMix64(uint64_t s) : state(s) {}
  uint64_t next() {
    state += 0x9E3779B97F4A7C15ull;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  // Box-Muller on pairs: the n outputs are r cos(2 pi u2) for every pair,
  // followed by r sin(2 pi u2), truncated to n
  std::vector<double> normal(size_t n) {
    const size_t m = (n + 1) / 2;
    std::vector<double> r(m), a(m), out(2 * m);
    for (size_t k = 0; k < m; ++k) {
      const double u1 = uniform(), u2 = uniform();
      r[k] = std::sqrt(-2.0 * std::log(1.0 - u1));
      a[k] = 2.0 * M_PI * u2;
    }
    for (size_t k = 0; k < m; ++k) {
      out[k] = r[k] * std::cos(a[k]);
      out[m + k] = r[k] * std::sin(a[k]);
    }
    out.resize(n);
    return out;
  }
};

struct Problem {
  Parameters P;
  int dim = 2, degree = 1, n_q = 2;
  int n_cells_coarse[3] = {1, 1, 1}, n_levels = 1, n_fine[3] = {1, 1, 1};
  double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
  int geometry = LSMF_GEOMETRY_BOX;
  lsmf_mesh *mesh = nullptr;  // finest level, host
  std::vector<double> mu, lambda;
  lsmf_solver_options opt{};
  std::vector<std::vector<int>> id_sides;  // per boundary id (index) -> sides
  ~Problem() { lsmf_mesh_destroy(mesh); }
};

// boundary ids of the reference's make_grid (include/mf_elasticity.h:1339-1398):
// Cook: 1 = -X face, 11 = +X face, 2 = faces whose centre has |x| = 0.5
// (this is what the reference tests; with elements per edge != 48 no face
// qualifies), 0 = every other boundary face.  Box (new): id = face number
// 2 d + (0: low, 1: high), as deal.II's colorized hyper_rectangle.
std::vector<int> sides_of_id(const Problem &pr, int id) {
  std::vector<int> s;
  const int dim = pr.dim;
  if (pr.geometry == LSMF_GEOMETRY_COOK) {
    const double centre_x = 0.5 * 48.0 / pr.P.elements_per_edge;  // first column, unscaled
    const bool id2_hits = std::fabs(centre_x - 0.5) < 1e-6;
    if (id2_hits && (id == 0 || id == 2))
      throw std::runtime_error("Cook boundary id 2 covers part of a side for 48 elements per edge: not supported");
    if (id == 1) s.push_back(0);
    if (id == 11) s.push_back(1);
    if (id == 0)
      for (int f = 2; f < 2 * dim; ++f) s.push_back(f);
  } else {
    if (id >= 0 && id < 2 * dim) s.push_back(id);
  }
  return s;
}

void setup(Problem &pr, int dim, int degree, int n_q) {
  Parameters &P = pr.P;
  pr.dim = dim;
  pr.degree = degree;
  pr.n_q = n_q;
  if (P.type == "Holes")
    throw std::runtime_error("Geometry type Holes (curved manifold meshes) is not supported by this build");
  if (P.type_lin == "CG")
    std::cerr << "Note: Solver type CG (assembled matrix) runs on the matrix-free operator here" << std::endl;
  else if (P.type_lin == "MF_AD_CG")  // include/mf_ad_nh_operator.h: the same tangent by automatic differentiation
    std::cerr << "Note: Solver type MF_AD_CG runs on the matrix-free operator here (the same tangent)" << std::endl;
  else if (P.type_lin != "MF_CG")
    throw std::runtime_error("Solver type " + P.type_lin + " is not supported: use MF_CG");
  // gmg-gc (global coarsening, include/mf_elasticity.h:1597-1622): on the
  // uniformly refined meshes here its levels are those of gmg
  if (P.preconditioner_type == "gmg-gc") {
    std::cerr << "Note: Preconditioner type gmg-gc runs as gmg (uniform refinement: the same levels)" << std::endl;
    P.preconditioner_type = "gmg";
  }
  if (P.preconditioner_type != "gmg" && P.preconditioner_type != "jacobi" && P.preconditioner_type != "none")
    throw std::runtime_error("Preconditioner type " + P.preconditioner_type +
                             " is not supported: use gmg, gmg-gc, jacobi or none");
  int caching;
  if (P.mf_caching == "scalar") caching = LSMF_CACHING_SCALAR;
  else if (P.mf_caching == "tensor2") caching = LSMF_CACHING_TENSOR2;
  else if (P.mf_caching == "tensor4") caching = LSMF_CACHING_TENSOR4;
  else if (P.mf_caching == "tensor4_ns") caching = LSMF_CACHING_TENSOR4_NS;
  else caching = LSMF_CACHING_SCALAR_REFERENTIAL;
  // formulation 0 ignores the caching choice in the reference (its operator
  // caches J^{-1/d} only, include/mf_nh_operator.h:467-487); here it runs
  // with the scalar cache, or with the tensor4 cache when that is asked for
  if (P.material_formulation == 0 && caching != LSMF_CACHING_SCALAR && caching != LSMF_CACHING_TENSOR4) {
    std::cerr << "Note: Formulation 0 runs with scalar caching (MF caching = " << P.mf_caching << " asked)"
              << std::endl;
    caching = LSMF_CACHING_SCALAR;
  }

  // geometry of the finest level and of the coarse mesh
  const bool gmg = P.preconditioner_type == "gmg";
  const int E = P.elements_per_edge, R = P.n_global_refinement;
  if (E < 1) throw std::runtime_error("Elements per edge must be positive");
  int base[3] = {E, E, E};
  if (P.type == "Cook") {
    pr.geometry = LSMF_GEOMETRY_COOK;
    pr.lo[0] = 0.0, pr.lo[1] = 0.0, pr.lo[2] = -0.5;
    pr.hi[0] = 48.0, pr.hi[1] = 44.0, pr.hi[2] = 0.5;
    if (dim == 3) base[2] = 1;  // one element through the thickness
  } else {
    pr.geometry = LSMF_GEOMETRY_BOX;
  }
  for (int d = 0; d < 3; ++d) {
    pr.n_fine[d] = d < dim ? base[d] << R : 1;
    pr.n_cells_coarse[d] = d < dim ? (gmg ? base[d] : pr.n_fine[d]) : 1;
  }
  pr.n_levels = gmg ? R + 1 : 1;
  check(lsmf_mesh_create(dim, degree, pr.n_fine, pr.lo, pr.hi, pr.geometry, P.scale, &pr.mesh), "mesh");

  // materials (include/material.h:82-94: lambda = 2 mu nu / (1 - 2 nu))
  const double nu = P.nu;
  const double mu0 = P.youngs_modulus > 0 ? P.youngs_modulus / (2.0 * (1.0 + nu)) : P.mu;
  const double lam0 = 2.0 * mu0 * nu / (1.0 - 2.0 * nu);
  const size_t nc = pr.mesh->n_cells;
  pr.mu.assign(nc, mu0);
  pr.lambda.assign(nc, lam0);
  if (P.material_field == "inclusions") {
    SplitMix64 rng((uint64_t)P.random_seed);
    std::vector<double> c((size_t)P.inclusion_count * dim);
    for (double &v : c) v = rng.uniform();
    const int npc = (int)std::pow(degree + 1, dim);
    const double r2 = P.inclusion_radius * P.inclusion_radius;
    for (size_t cell = 0; cell < nc; ++cell) {
      double x[3] = {0, 0, 0};
      for (int a = 0; a < npc; ++a)
        for (int d = 0; d < dim; ++d) x[d] += pr.mesh->coords[(size_t)pr.mesh->cell_nodes[cell * npc + a] * dim + d];
      for (int d = 0; d < dim; ++d) x[d] /= npc;
      for (int k = 0; k < P.inclusion_count; ++k) {
        double s = 0;
        for (int d = 0; d < dim; ++d) {
          const double dd = x[d] - c[(size_t)k * dim + d];
          s += dd * dd;
        }
        if (s <= r2) {
          pr.mu[cell] = P.inclusion_factor * mu0;
          pr.lambda[cell] = P.inclusion_factor * lam0;
          break;
        }
      }
    }
  } else if (P.material_field == "lognormal") {
    SplitMix64 rng((uint64_t)P.random_seed);
    const std::vector<double> z = rng.normal(2 * nc);
    for (size_t cell = 0; cell < nc; ++cell) {
      pr.mu[cell] = mu0 * std::exp(P.lognormal_sigma * z[cell]);
      pr.lambda[cell] = lam0 * std::exp(P.lognormal_sigma * z[nc + cell]);
    }
  }

  // solver options
  lsmf_solver_options &o = pr.opt;
  check(lsmf_solver_options_default(&o), "options");
  o.dim = dim;
  o.degree = degree;
  o.n_q_1d = n_q;
  for (int d = 0; d < 3; ++d) {
    o.n_cells_coarse[d] = pr.n_cells_coarse[d];
    o.lo[d] = pr.lo[d];
    o.hi[d] = pr.hi[d];
  }
  o.n_levels = pr.n_levels;
  o.geometry = pr.geometry;
  o.scale = P.scale;
  o.caching = caching;
  o.mg_caching = caching;
  o.mg_precision = P.mg_level_precision == "double" ? LSMF_FP64 : LSMF_FP32;
  o.eig_cg_iterations = P.mf_chebyshev_n_cg_iterations;
  o.coarse_eig_iterations = P.mf_coarse_chebyshev_accurate_eigenval ? 0 : P.mf_chebyshev_n_cg_iterations;
  // "MF Chebyshev coarse": the Chebyshev iteration as coarse solver, or CG
  // preconditioned by the coarse level's smoother (include/mf_elasticity.h:2115-2175)
  o.coarse_solver = P.mf_coarse_chebyshev ? LSMF_COARSE_CHEBYSHEV : LSMF_COARSE_CG;
  o.tol_lin = P.tol_lin;
  const size_t n_dofs = (size_t)pr.mesh->n_nodes * dim;
  o.max_lin_iterations = (int)std::min<size_t>(n_dofs * (size_t)P.max_iterations_lin, 2000000000);
  o.max_newton = P.max_iterations_NR;
  o.tol_f = P.tol_f;
  o.tol_u = P.tol_u;
  o.tol_f_abs = P.tol_f_abs;
  o.tol_u_abs = P.tol_u_abs;
  o.dt = P.delta_t;
  o.t_end = P.end_time;
  o.preconditioner = gmg ? LSMF_PRECONDITIONER_GMG
                         : (P.preconditioner_type == "jacobi" ? LSMF_PRECONDITIONER_JACOBI : LSMF_PRECONDITIONER_NONE);
  o.preconditioner_relaxation = P.preconditioner_relaxation;
  o.formulation = P.material_formulation;
  o.cond_number_cg_iterations = P.cond_number_cg_iterations;
  for (const auto &kv : P.dirichlet) {
    auto m = P.dirichlet_mask.find(kv.first);
    if (m == P.dirichlet_mask.end())
      throw std::runtime_error("Could not find component mask for ID " + std::to_string(kv.first));
    for (int s : sides_of_id(pr, kv.first))
      for (int c = 0; c < dim; ++c)
        if (m->second[c]) o.dirichlet_components[s] |= 1u << c;
  }
  for (const auto &kv : P.neumann)
    for (int s : sides_of_id(pr, kv.first)) o.neumann_sides |= 1u << s;
}

// reference volume: sum over cells of the Q1 map's |det J| (GridTools::volume)
double reference_volume(const Problem &pr) {
  const lsmf_mesh *m = pr.mesh;
  const int dim = m->dim, p = m->degree, N = p + 1;
  const int npc = dim == 3 ? N * N * N : N * N;
  std::vector<double> qx, qw;
  lsmf::gauss_rule(2, qx, qw);
  double vol = 0;
  for (int cell = 0; cell < m->n_cells; ++cell) {
    const int32_t *cn = m->cell_nodes + (size_t)cell * npc;
    double X[8][3] = {};
    for (int v = 0; v < (1 << dim); ++v) {
      int a = 0, st = 1;
      for (int e = 0; e < dim; ++e) {
        a += ((v >> e) & 1) * p * st;
        st *= N;
      }
      for (int i = 0; i < dim; ++i) X[v][i] = m->coords[(size_t)cn[a] * dim + i];
    }
    const int nq = dim == 3 ? 8 : 4;
    for (int q = 0; q < nq; ++q) {
      double xi[3] = {qx[q & 1], qx[(q >> 1) & 1], qx[(q >> 2) & 1]};
      double w = qw[q & 1] * qw[(q >> 1) & 1] * (dim == 3 ? qw[(q >> 2) & 1] : 1.0);
      double J[3][3] = {};
      for (int v = 0; v < (1 << dim); ++v)
        for (int d = 0; d < dim; ++d) {
          double g = 1.0;
          for (int e = 0; e < dim; ++e) {
            const int ve = (v >> e) & 1;
            g *= (e == d) ? (ve ? 1.0 : -1.0) : (ve ? xi[e] : 1.0 - xi[e]);
          }
          for (int i = 0; i < dim; ++i) J[i][d] += X[v][i] * g;
        }
      const double det = dim == 2 ? J[0][0] * J[1][1] - J[0][1] * J[1][0]
                                  : J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                                        J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                                        J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
      vol += std::fabs(det) * w;
    }
  }
  return vol;
}

// the FE displacement at a point of the reference configuration: the cell
// whose Q1 map contains it (Newton inversion), then the Q_p basis at the
// local coordinates (GridTools::find_active_cell_around_point + FEValues,
// include/mf_elasticity.h:2623-2690)
bool point_value(const Problem &pr, const std::vector<double> &u, const std::vector<double> &pt, double *out) {
  const lsmf_mesh *m = pr.mesh;
  const int dim = m->dim, p = m->degree, N = p + 1;
  const int npc = dim == 3 ? N * N * N : N * N;
  const std::vector<double> nodes = lsmf::lobatto_points(p);
  for (int cell = 0; cell < m->n_cells; ++cell) {
    const int32_t *cn = m->cell_nodes + (size_t)cell * npc;
    double X[8][3] = {};
    double bmin[3] = {1e300, 1e300, 1e300}, bmax[3] = {-1e300, -1e300, -1e300};
    for (int v = 0; v < (1 << dim); ++v) {
      int a = 0, st = 1;
      for (int e = 0; e < dim; ++e) {
        a += ((v >> e) & 1) * p * st;
        st *= N;
      }
      for (int i = 0; i < dim; ++i) {
        X[v][i] = m->coords[(size_t)cn[a] * dim + i];
        bmin[i] = std::min(bmin[i], X[v][i]);
        bmax[i] = std::max(bmax[i], X[v][i]);
      }
    }
    bool inside_box = true;
    for (int i = 0; i < dim; ++i) {
      const double tol = 1e-10 * (bmax[i] - bmin[i] + 1e-300);
      if (pt[i] < bmin[i] - tol || pt[i] > bmax[i] + tol) inside_box = false;
    }
    if (!inside_box) continue;
    double xi[3] = {0.5, 0.5, 0.5};
    for (int it = 0; it < 50; ++it) {
      double x[3] = {}, J[3][3] = {};
      for (int v = 0; v < (1 << dim); ++v) {
        double phi = 1.0;
        for (int e = 0; e < dim; ++e) phi *= ((v >> e) & 1) ? xi[e] : 1.0 - xi[e];
        for (int i = 0; i < dim; ++i) x[i] += phi * X[v][i];
        for (int d = 0; d < dim; ++d) {
          double g = 1.0;
          for (int e = 0; e < dim; ++e) {
            const int ve = (v >> e) & 1;
            g *= (e == d) ? (ve ? 1.0 : -1.0) : (ve ? xi[e] : 1.0 - xi[e]);
          }
          for (int i = 0; i < dim; ++i) J[i][d] += X[v][i] * g;
        }
      }
      double r[3], dx[3];
      for (int i = 0; i < dim; ++i) r[i] = pt[i] - x[i];
      if (dim == 2) {
        const double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        dx[0] = (J[1][1] * r[0] - J[0][1] * r[1]) / det;
        dx[1] = (-J[1][0] * r[0] + J[0][0] * r[1]) / det;
      } else {
        double inv[3][3];
        const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                           J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) {
            const int i1 = (j + 1) % 3, i2 = (j + 2) % 3, j1 = (i + 1) % 3, j2 = (i + 2) % 3;
            inv[i][j] = (J[i1][j1] * J[i2][j2] - J[i1][j2] * J[i2][j1]) / det;
          }
        for (int i = 0; i < 3; ++i) dx[i] = inv[i][0] * r[0] + inv[i][1] * r[1] + inv[i][2] * r[2];
      }
      double step = 0;
      for (int i = 0; i < dim; ++i) {
        xi[i] += dx[i];
        step = std::max(step, std::fabs(dx[i]));
      }
      if (step < 1e-15) break;
    }
    bool inside = true;
    for (int i = 0; i < dim; ++i)
      if (xi[i] < -1e-10 || xi[i] > 1.0 + 1e-10) inside = false;
    if (!inside) continue;
    for (int c = 0; c < dim; ++c) out[c] = 0.0;
    for (int a = 0; a < npc; ++a) {
      int ai = a;
      double phi = 1.0;
      for (int d = 0; d < dim; ++d) {
        phi *= lsmf::lagrange_value(nodes, ai % N, xi[d]);
        ai /= N;
      }
      for (int c = 0; c < dim; ++c) out[c] += phi * u[(size_t)cn[a] * dim + c];
    }
    return true;
  }
  return false;
}

std::vector<std::string> expr_variables(int dim) {
  return dim == 2 ? std::vector<std::string>{"x", "y", "t"} : std::vector<std::string>{"x", "y", "z", "t"};
}

// Dirichlet increments at time t: every id in ascending order, its function
// at the support points of its sides; a DoF keeps the first value it gets
// (interpolate_boundary_values leaves constrained DoFs alone,
// include/mf_elasticity.h:2963-2990)
std::vector<double> dirichlet_increment(const Problem &pr, double t) {
  const lsmf_mesh *m = pr.mesh;
  const int dim = m->dim;
  const size_t n = (size_t)m->n_nodes * dim;
  std::vector<double> inc(n, 0.0);
  std::vector<char> done(n, 0);
  const auto vars = expr_variables(dim);
  for (const auto &kv : pr.P.dirichlet) {
    const std::vector<bool> &mask = pr.P.dirichlet_mask.at(kv.first);
    std::vector<Expression> f;
    for (const std::string &e : kv.second) f.emplace_back(e, vars);
    for (int s : sides_of_id(pr, kv.first)) {
      const int cnt = lsmf_mesh_boundary_nodes(m, s, nullptr);
      std::vector<int32_t> nodes(cnt);
      lsmf_mesh_boundary_nodes(m, s, nodes.data());
      for (int nd : nodes) {
        double v[4];
        for (int d = 0; d < dim; ++d) v[d] = m->coords[(size_t)nd * dim + d];
        v[dim] = t;
        for (int c = 0; c < dim; ++c) {
          const size_t dof = (size_t)nd * dim + c;
          if (!mask[c] || done[dof]) continue;
          inc[dof] = f[c].eval(v);
          done[dof] = 1;
        }
      }
    }
  }
  return inc;
}

// traction per side at time t; expressions must not vary along the side
void traction_at(const Problem &pr, double t, double (&tr)[6][3]) {
  const lsmf_mesh *m = pr.mesh;
  const int dim = m->dim;
  std::memset(tr, 0, sizeof(tr));
  const auto vars = expr_variables(dim);
  for (const auto &kv : pr.P.neumann) {
    std::vector<Expression> f;
    for (const std::string &e : kv.second) f.emplace_back(e, vars);
    for (int s : sides_of_id(pr, kv.first)) {
      const int cnt = lsmf_mesh_boundary_nodes(m, s, nullptr);
      std::vector<int32_t> nodes(cnt);
      lsmf_mesh_boundary_nodes(m, s, nodes.data());
      for (int c = 0; c < dim; ++c) {
        double first = 0;
        for (int k = 0; k < cnt; ++k) {
          double v[4];
          for (int d = 0; d < dim; ++d) v[d] = m->coords[(size_t)nodes[k] * dim + d];
          v[dim] = t;
          const double val = f[c].eval(v);
          if (k == 0) first = val;
          else if (std::fabs(val - first) > 1e-12 * std::max(std::fabs(first), std::fabs(val)))
            throw std::runtime_error("Neumann expression '" + kv.second[c] +
                                     "' varies along the boundary: only tractions uniform on each side are "
                                     "supported");
        }
        tr[s][c] += first;
      }
    }
  }
}

std::string fmt_time(double s) {
  std::ostringstream os;
  os << std::setprecision(3) << s;
  return os.str();
}

struct Section {
  std::string name;
  long calls = 0;
  double time = 0;
};

// TimerOutput::summary-style table (wall times)
void print_timer_table(std::ostream &os, double total, const std::vector<Section> &secs) {
  os << "\n\n+" << std::string(46, '-') << "+" << std::string(12, '-') << "+" << std::string(12, '-') << "+\n";
  os << "| " << std::left << std::setw(44) << "Total wallclock time elapsed since start" << " |" << std::right
     << std::setw(10) << fmt_time(total) << "s |            |\n";
  os << "|" << std::string(46, ' ') << "|            |            |\n";
  os << "| " << std::left << std::setw(32) << "Section" << " | no. calls |  wall time | % of total |\n";
  os << "+" << std::string(34, '-') << "+" << std::string(11, '-') << "+" << std::string(12, '-') << "+"
     << std::string(12, '-') << "+\n";
  for (const Section &s : secs) {
    std::ostringstream pct;
    const double f = total > 0 ? 100.0 * s.time / total : 0.0;
    pct << std::setprecision(2) << (f < 0.01 ? 0.0 : f);
    os << "| " << std::left << std::setw(32) << s.name << " | " << std::right << std::setw(9) << s.calls << " | "
       << std::setw(9) << fmt_time(s.time) << "s | " << std::setw(9) << pct.str() << "% |\n";
  }
  os << "+" << std::string(34, '-') << "+" << std::string(11, '-') << "+" << std::string(12, '-') << "+"
     << std::string(12, '-') << "+\n\n";
}

// a stream that writes to stdout and, when open, a file
struct Tee {
  std::ostream *a = nullptr;
  std::ofstream *b = nullptr;
  template <class T>
  Tee &operator<<(const T &v) {
    if (a) *a << v;
    if (b && b->is_open()) *b << v;
    return *this;
  }
  Tee &operator<<(std::ostream &(*f)(std::ostream &)) {
    if (a) f(*a);
    if (b && b->is_open()) f(*b);
    return *this;
  }
};

template <class S>
void print_point(S &os, const std::vector<double> &p) {
  for (size_t i = 0; i < p.size(); ++i) os << (i ? " " : "") << p[i];
}

void describe_setup(const Problem &pr) {
  const lsmf_solver_options &o = pr.opt;
  const int dim = pr.dim;
  const double t1 = pr.P.delta_t;
  const std::vector<double> inc = dirichlet_increment(pr, t1);
  double tr[6][3];
  traction_at(pr, t1, tr);
  size_t n_nonzero = 0;
  double inc_sum = 0;
  for (double v : inc) {
    n_nonzero += v != 0.0;
    inc_sum += v;
  }
  std::ostringstream js;
  js << std::setprecision(17) << "{\"dim\": " << dim << ", \"degree\": " << o.degree << ", \"n_q_1d\": " << o.n_q_1d
     << ", \"geometry\": " << o.geometry << ", \"n_levels\": " << o.n_levels << ", \"n_cells_coarse\": ["
     << o.n_cells_coarse[0] << ", " << o.n_cells_coarse[1] << ", " << o.n_cells_coarse[2] << "], \"n_cells_fine\": ["
     << pr.n_fine[0] << ", " << pr.n_fine[1] << ", " << pr.n_fine[2] << "], \"caching\": " << o.caching
     << ", \"mg_precision\": " << o.mg_precision << ", \"formulation\": " << o.formulation
     << ", \"preconditioner\": " << o.preconditioner << ", \"coarse_solver\": " << o.coarse_solver
     << ", \"relaxation\": " << o.preconditioner_relaxation << ", \"tol_lin\": " << o.tol_lin
     << ", \"max_lin_iterations\": " << o.max_lin_iterations << ", \"max_newton\": " << o.max_newton
     << ", \"tol_f\": " << o.tol_f << ", \"tol_u\": " << o.tol_u << ", \"dt\": " << o.dt << ", \"t_end\": " << o.t_end
     << ", \"eig_cg_iterations\": " << o.eig_cg_iterations << ", \"coarse_eig_iterations\": "
     << o.coarse_eig_iterations << ", \"dirichlet_components\": [";
  for (int s = 0; s < 2 * dim; ++s) js << (s ? ", " : "") << o.dirichlet_components[s];
  js << "], \"neumann_sides\": " << o.neumann_sides << ", \"traction\": [";
  for (int s = 0; s < 2 * dim; ++s) {
    js << (s ? ", " : "") << "[";
    for (int c = 0; c < dim; ++c) js << (c ? ", " : "") << tr[s][c];
    js << "]";
  }
  js << "], \"increment_nonzero\": " << n_nonzero << ", \"increment_sum\": " << inc_sum
     << ", \"mu_sum\": " << [&] { double a = 0; for (double v : pr.mu) a += v; return a; }()
     << ", \"n_dofs\": " << (size_t)pr.mesh->n_nodes * dim << "}";
  std::cerr << js.str() << std::endl;
}

int run(const std::string &file, bool dry_run, bool describe_json, bool print_params, const std::string &dump_material) {
  int dim, degree, quad;
  read_dimension_and_degree(file, dim, degree, quad);
  if (print_params) {
    std::cout << describe(file, dim);
    return 0;
  }
  // instantiated (degree, n_q_points_1d) pairs: n_q = degree + 1, degree 1, 2, 4
  if (!((degree == 1 || degree == 2 || degree == 4) && quad == degree + 1))
    throw std::runtime_error("Matrix-free calculations with degree=" + std::to_string(degree) +
                             " and n_q_points_1d=" + std::to_string(quad) + " are not supported.");
  Problem pr;
  pr.P = read_parameters(file, dim);
  const auto t_start = Clock::now();
  setup(pr, dim, degree, quad);
  const Parameters &P = pr.P;
  if (!dump_material.empty()) {
    std::ofstream out(dump_material, std::ios::binary);
    out.write((const char *)pr.mu.data(), pr.mu.size() * sizeof(double));
    out.write((const char *)pr.lambda.data(), pr.lambda.size() * sizeof(double));
    if (!out) throw std::runtime_error("cannot write " + dump_material);
  }
  std::ofstream blessed, timings;
  if (!dry_run) {
    blessed.open(P.output_folder + "output");
    timings.open(P.output_folder + "timings.txt");
    if (!blessed || !timings) throw std::runtime_error("can't create: " + P.output_folder);
  }
  Tee out{&std::cout, &blessed};
  const size_t n_dofs = (size_t)pr.mesh->n_nodes * dim;
  out << "Grid:\n  Reference volume: " << reference_volume(pr) << std::endl;
  out << "Triangulation:"
      << "\n  Number of active cells: " << pr.mesh->n_cells << "\n  Number of degrees of freedom: " << n_dofs
      << std::endl;
  if (describe_json) describe_setup(pr);
  if (dry_run) return 0;

  // validate the expressions on the host before the GPU is touched
  (void)dirichlet_increment(pr, P.delta_t);
  double tr[6][3];
  traction_at(pr, P.delta_t, tr);

  lsmf_solver *S = nullptr;
  const auto t_setup = Clock::now();
  check(lsmf_solver_create(&pr.opt, pr.mu.data(), pr.lambda.data(), &S), "solver");
  std::vector<Section> secs{{"Setup system", 1, seconds_since(t_setup)},
                            {"Assemble linear system"},
                            {"Linear solver"},
                            {"Setup MF: cache() and diagonal()"},
                            {"Setup MF: GMG setup"}};
  std::vector<Section> bench{{"vmult (MF)"}, {"vmult (MF) Cell loop"}, {"vmult (MF) RW"}, {"vmult (MF) RWSF"}};
  std::vector<double> u(n_dofs);
  long total_cg = 0, total_solves = 0;
  double t = 0.0;
  int step = 0;
  double *d_src = nullptr, *d_dst = nullptr;
  try {
    for (;;) {
      const double t_next = t + P.delta_t;  // Time::increment
      if (!(t_next <= P.end_time)) break;
      ++step;
      traction_at(pr, t_next, tr);
      check(lsmf_solver_set_traction(S, tr), "traction");
      const std::vector<double> inc = dirichlet_increment(pr, t_next);
      check(lsmf_solver_set_dirichlet_increment(S, inc.data()), "dirichlet");
      out << "\nTimestep " << step << " @ " << t_next << "s" << std::endl;
      lsmf_step_stats st{};
      double t_got = 0;
      int done = 0;
      check(lsmf_solver_step(S, &st, &t_got, &done), "step");
      if (done) break;
      t = t_got;
      out << "Converged in " << st.newton_iterations << " Newton iterations" << std::endl;
      if (st.cond_number > 0) std::cout << "  (condition number estimate " << st.cond_number << ")" << std::endl;
      secs[1].calls += st.newton_iterations + 1;
      secs[1].time += st.time_residual_s;
      secs[2].calls += st.newton_iterations;
      secs[2].time += st.time_linear_s;
      secs[3].calls += st.newton_iterations + 1;
      secs[3].time += st.time_setup_cache_s;
      // Chebyshev eigenvalue estimates (deal.II runs them in the first
      // smoother application) and the coarse solve's set-up
      if (pr.opt.preconditioner == LSMF_PRECONDITIONER_GMG) {
        secs[4].calls += st.newton_iterations;
        secs[4].time += st.time_setup_s - st.time_setup_cache_s;
      }
      total_cg += st.linear_iterations;
      total_solves += st.newton_iterations;
      check(lsmf_solver_get_solution(S, u.data()), "solution");
      for (const std::vector<double> &pt : P.output_points) {
        double d[3] = {0, 0, 0};
        if (!point_value(pr, u, pt, d)) throw std::runtime_error("Found no cell with point inside!");
        out << "Solution @ ";
        print_point(out, pt);
        out << std::endl << "  displacement: ";
        print_point(out, std::vector<double>(d, d + dim));
        out << std::endl;
      }
      if (P.benchmark_vmult > 0) {
        lsmf_op *op = lsmf_solver_fine_operator(S);
        if (!d_src) {
          if (hipMalloc(&d_src, n_dofs * sizeof(double)) != hipSuccess ||
              hipMalloc(&d_dst, n_dofs * sizeof(double)) != hipSuccess)
            throw std::runtime_error("hipMalloc failed");
          if (hipMemcpy(d_src, u.data(), n_dofs * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
            throw std::runtime_error("hipMemcpy failed");
        }
        const unsigned masks[4] = {LSMF_MASK_DEFAULT, LSMF_MASK_CELLLOOP, LSMF_MASK_RW, LSMF_MASK_RWSF};
        for (int b = 0; b < 4; ++b) {
          if (hipDeviceSynchronize() != hipSuccess) throw std::runtime_error("hipDeviceSynchronize failed");
          const auto t0 = Clock::now();
          for (int k = 0; k < P.benchmark_vmult; ++k) check(lsmf_vmult(op, d_dst, d_src, 0, masks[b], nullptr), "vmult");
          if (hipDeviceSynchronize() != hipSuccess) throw std::runtime_error("hipDeviceSynchronize failed");
          bench[b].calls += P.benchmark_vmult;
          bench[b].time += seconds_since(t0);
        }
      }
    }
  } catch (...) {
    if (d_src) (void)hipFree(d_src);
    if (d_dst) (void)hipFree(d_dst);
    lsmf_solver_destroy(S);
    throw;
  }
  if (d_src) (void)hipFree(d_src);
  if (d_dst) (void)hipFree(d_dst);

  // timings.txt (include/mf_elasticity.h:1099-1142, 1226-1231, TimerOutput)
  int hip_version = 0;
  (void)hipRuntimeGetVersion(&hip_version);
  hipDeviceProp_t prop{};
  int dev = 0;
  (void)hipGetDevice(&dev);
  (void)hipGetDeviceProperties(&prop, dev);
  lsmf_op *fine = lsmf_solver_fine_operator(S);
  timings << "-----------------------------------------------------------------------------\n"
          << "--     . running in OPTIMIZED mode\n"
          << "--     . running on 1 GPU (" << prop.gcnArchName << ", " << prop.multiProcessorCount << " CUs)\n"
          << "--     . HIP runtime " << hip_version << "\n"
          << "-----------------------------------------------------------------------------\n\n";
  timings << "dim   = " << dim << "\np     = " << degree << "\nq     = " << quad << "\ncells = " << pr.mesh->n_cells
          << "\ndofs  = " << n_dofs << "\n\n";
  timings << "MF cache memory = " << (lsmf_memory_consumption(fine) >> 20) << " Mb\n\n";
  timings << "Average CG iter = " << (total_solves ? total_cg / total_solves : 0) << "\nTotal CG iter = " << total_cg
          << "\nTotal CG solve = " << total_solves << std::endl;
  if (P.benchmark_vmult > 0) secs.insert(secs.end(), bench.begin(), bench.end());
  std::sort(secs.begin(), secs.end(), [](const Section &a, const Section &b) { return a.name < b.name; });
  print_timer_table(timings, seconds_since(t_start), secs);
  lsmf_solver_destroy(S);
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  bool dry_run = false, print_params = false, describe = false;
  std::string dump_material, file = "parameters.prm";
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if (a == "--dry-run") dry_run = true;
    else if (a == "--print-parameters") print_params = true;
    else if (a == "--describe-setup") describe = true;
    else if (a == "--dump-material" && i + 1 < argc) dump_material = argv[++i];
    else if (a == "-h" || a == "--help") {
      std::cout << "usage: lsmf_run [--dry-run] [--describe-setup] [--print-parameters] [--dump-material FILE] "
                   "file.prm\n";
      return 0;
    } else file = a;
  }
  try {
    return run(file, dry_run, describe, print_params, dump_material);
  } catch (std::exception &exc) {
    std::cerr << std::endl
              << std::endl
              << "----------------------------------------------------" << std::endl;
    std::cerr << "Exception on processing: " << std::endl
              << exc.what() << std::endl
              << "Aborting!" << std::endl
              << "----------------------------------------------------" << std::endl;
    return 1;
  }
}
