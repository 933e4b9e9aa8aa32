// lsmf_oracle.cpp -- CPU FP64 restatement of the reference's matrix-free
// Neo-Hookean tangent operator, residual, diagonal and Newton/CG driver.
//
// TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load it to compare
// against the HIP product path.  Nothing in the product links, loads or calls
// it, and it is never the thing measured as the GPU result.
//
// What it restates (reference paths relative to the reference repository):
//   * material law, formulations 1 and 0 ..... include/material.h:78-163, 171-250,
//                                              260-295 (Psi_vol and derivatives)
//   * NeoHookOperator::cache() ............... include/mf_nh_operator.h:447-592
//   * NeoHookOperator::vmult_add() ........... include/mf_nh_operator.h:646-706
//   * local_apply_cell / do_operation_on_cell  include/mf_nh_operator.h:710-785,
//                                              922-1499 (formulation 1 branches,
//                                              formulation 0 at :1001-1181)
//   * compute_diagonal / local_diagonal_cell . include/mf_nh_operator.h:789-918,
//                                              1503-1557
//   * residual in assemble_system() .......... include/mf_elasticity.h:2758-2804,
//                                              2859-2898, 2917-2921
//   * make_constraints (increment at it 0) ... include/mf_elasticity.h:2932-3007
//   * hanging-node (general affine) constraints in read_dof_values /
//     distribute_local_to_global, as deal.II's AffineConstraints resolve
//     them (include/mf_nh_operator.h:745-749, 769-780)
//   * solve_linear_system (CG, Jacobi) ....... include/mf_elasticity.h:3012-3214
//   * check_convergence / Newton loop ........ include/mf_elasticity.h:2245-2559
//   * Time stepping (t accumulates dt) ....... include/mf_elasticity.h:614-653,
//                                              1187-1222
//
// The arithmetic of deal.II's FEEvaluation (sum factorization, MappingQEulerian
// metric, read_dof_values / distribute_local_to_global) is restated from its
// documented semantics: gradients in the current configuration use the
// inverse Jacobian of x(xi) = sum_a (X_a + u_a) phi_a(xi) (the Eulerian Q_p
// mapping); reference quantities use the Q1 vertex mapping; constrained DoFs
// read as zero and are skipped on write; constrained rows act as identity.
//
// Parity status of each function is listed in the header of tests/ files that
// pin it (known-answer tests of tests/mf_neo_hook_operator_form1.cc restated,
// blessed Cook outputs, closed forms, finite differences).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// ---------------------------------------------------------------------------
// 1D finite element tables: Gauss points, Gauss-Lobatto support points,
// Lagrange shape values and derivatives.  deal.II's FE_Q(p) places its
// support points at Gauss-Lobatto points (equispaced for p <= 2) and the
// operator uses QGauss(n_q_points_1d) (include/mf_elasticity.h:1752).
// ---------------------------------------------------------------------------
static void legendre(int n, double x, double &P, double &dP) {
  double p0 = 1.0, p1 = x;
  if (n == 0) { P = 1.0; dP = 0.0; return; }
  for (int k = 2; k <= n; ++k) {
    const double pk = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1;
    p1 = pk;
  }
  P = p1;
  dP = n * (x * p1 - p0) / (x * x - 1.0);
}

// Gauss-Legendre points/weights on [0,1], ascending.
static void gauss01(int n, double *x, double *w) {
  for (int i = 0; i < n; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; ++it) {
      double P, dP;
      legendre(n, z, P, dP);
      const double dz = P / dP;
      z -= dz;
      if (std::abs(dz) < 1e-16) break;
    }
    double P, dP;
    legendre(n, z, P, dP);
    const double wi = 2.0 / ((1.0 - z * z) * dP * dP);
    x[n - 1 - i] = 0.5 * (z + 1.0);
    w[n - 1 - i] = 0.5 * wi;
  }
}

// Gauss-Lobatto points on [0,1] for degree p (p+1 points), ascending.
static void gll01(int p, double *x) {
  x[0] = 0.0;
  x[p] = 1.0;
  for (int i = 1; i < p; ++i) {
    // roots of P'_p; start from Chebyshev-Gauss-Lobatto guess
    double z = -std::cos(M_PI * i / p);
    for (int it = 0; it < 100; ++it) {
      double P, dP;
      legendre(p, z, P, dP);
      // (1-z^2) P'' = 2 z P' - p(p+1) P
      const double d2P = (2.0 * z * dP - p * (p + 1.0) * P) / (1.0 - z * z);
      const double dz = dP / d2P;
      z -= dz;
      if (std::abs(dz) < 1e-16) break;
    }
    x[i] = 0.5 * (z + 1.0);
  }
}

static double lagrange(int n_nodes, const double *nodes, int i, double x) {
  double v = 1.0;
  for (int j = 0; j < n_nodes; ++j)
    if (j != i) v *= (x - nodes[j]) / (nodes[i] - nodes[j]);
  return v;
}

static double lagrange_d(int n_nodes, const double *nodes, int i, double x) {
  double s = 0.0;
  for (int k = 0; k < n_nodes; ++k) {
    if (k == i) continue;
    double v = 1.0 / (nodes[i] - nodes[k]);
    for (int j = 0; j < n_nodes; ++j)
      if (j != i && j != k) v *= (x - nodes[j]) / (nodes[i] - nodes[j]);
    s += v;
  }
  return s;
}

constexpr int ipow(int b, int e) { return e == 0 ? 1 : b * ipow(b, e - 1); }

// ---------------------------------------------------------------------------
// small dense DIM x DIM helpers (row-major)
// ---------------------------------------------------------------------------
template <int D> static double det(const double *A) {
  if constexpr (D == 2) return A[0] * A[3] - A[1] * A[2];
  else
    return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
           A[2] * (A[3] * A[7] - A[4] * A[6]);
}
template <int D> static void inv(const double *A, double *B) {
  const double d = det<D>(A);
  if constexpr (D == 2) {
    B[0] = A[3] / d; B[1] = -A[1] / d; B[2] = -A[2] / d; B[3] = A[0] / d;
  } else {
    B[0] = (A[4] * A[8] - A[5] * A[7]) / d;
    B[1] = (A[2] * A[7] - A[1] * A[8]) / d;
    B[2] = (A[1] * A[5] - A[2] * A[4]) / d;
    B[3] = (A[5] * A[6] - A[3] * A[8]) / d;
    B[4] = (A[0] * A[8] - A[2] * A[6]) / d;
    B[5] = (A[2] * A[3] - A[0] * A[5]) / d;
    B[6] = (A[3] * A[7] - A[4] * A[6]) / d;
    B[7] = (A[1] * A[6] - A[0] * A[7]) / d;
    B[8] = (A[0] * A[4] - A[1] * A[3]) / d;
  }
}
template <int D> static void matmul(const double *A, const double *B, double *C) {
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j) {
      double s = 0;
      for (int k = 0; k < D; ++k) s += A[i * D + k] * B[k * D + j];
      C[i * D + j] = s;
    }
}
// C = A * B^T
template <int D> static void matmul_bt(const double *A, const double *B, double *C) {
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j) {
      double s = 0;
      for (int k = 0; k < D; ++k) s += A[i * D + k] * B[j * D + k];
      C[i * D + j] = s;
    }
}

enum Caching { SCALAR = 0, TENSOR2 = 1, TENSOR4 = 2, TENSOR4_NS = 3, SCALAR_REF = 4 };

// ---------------------------------------------------------------------------
// Operator interface (runtime-polymorphic over dim / degree / n_q_1d)
// ---------------------------------------------------------------------------
struct OpBase {
  virtual ~OpBase() {}
  virtual int set_material(const double *mu, const double *lambda) = 0;
  virtual int set_linearization(const double *u) = 0;
  virtual int cache() = 0;
  virtual int vmult(double *dst, const double *src, int add, int plain) = 0;
  virtual int diagonal(double *diag) = 0;
  virtual int residual(double *rhs) = 0;
  virtual int neumann(const int *face_cells, const int *face_ids, int n_faces,
                      const double *traction, double *rhs) = 0;
  virtual int qp_state(int cell, const double *src, double *out) = 0;
  virtual double volume() = 0;
  virtual int n_colours() = 0;
  virtual int min_detF(double *out) = 0;
  virtual int set_node_constraints(int n, const int *nodes, const int *ptr, const int *masters,
                                   const double *weights) = 0;
  int dim = 0, p = 0, nq = 0, n_cells = 0, n_nodes = 0, caching = 0;
  // Material_Compressible_Neo_Hook_One_Field's formulation (include/material.h:82-94):
  // 1 = mu/2 (tr C - d - 2 ln J) + lambda ln^2 J; 0 = volumetric/isochoric split
  // with Psi_vol = kappa/4 (J^2 - 1 - 2 ln J), kappa = lambda + 2 mu / 3
  int formulation = 1;
  std::vector<unsigned char> cons; // per DoF
};

template <int DIM, int P, int NQ> struct Op : OpBase {
  static constexpr int ND = P + 1;
  static constexpr int NDOF1 = ipow(ND, DIM);   // nodes per cell
  static constexpr int NQP = ipow(NQ, DIM);     // qp per cell
  static constexpr int DD = DIM * DIM;
  static constexpr int D4 = DD * DD;

  double S[NQ][ND], Dm[NQ][ND];   // shape values / derivatives at Gauss points
  double Sv1[NQ][2], Dv1[NQ][2];  // Q1 (vertex) shape functions at Gauss points
  double Cg[NQ][NQ];              // collocation derivative on Gauss points
  double W[NQP];                  // tensor-product quadrature weights
  double gl_nodes[ND], q_pts[NQ], q_w[NQ];

  std::vector<int> cell_nodes;
  std::vector<double> coords;
  std::vector<double> mu, lam, u;
  std::vector<int> colour_of, colour_ptr, colour_cells;
  // per qp metric
  std::vector<double> jinv_ref, jxw_ref, jinv_cur, jxw_cur;
  // node constraints: node n's values = sum_j hc_w[j] values(hc_m[j]),
  // j in [hc_ptr[k], hc_ptr[k+1]) with k = hang_of[n] (-1: free)
  std::vector<int> hang_of, hc_ptr, hc_m;
  std::vector<double> hc_w;
  // caches (restating the reference's Table<2,...> members)
  std::vector<double> c_s, c_s2, c_t2, c_t4, c_t4ns, c_uqp;
  bool cached = false;

  Op(int n_cells_, int n_nodes_, const int *cn, const double *xyz, const unsigned char *cons_, int caching_) {
    dim = DIM; p = P; nq = NQ; n_cells = n_cells_; n_nodes = n_nodes_; caching = caching_;
    cell_nodes.assign(cn, cn + (size_t)n_cells * NDOF1);
    coords.assign(xyz, xyz + (size_t)n_nodes * DIM);
    cons.assign((size_t)n_nodes * DIM, 0);
    if (cons_) std::memcpy(cons.data(), cons_, cons.size());
    gll01(P, gl_nodes);
    gauss01(NQ, q_pts, q_w);
    const double vtx[2] = {0.0, 1.0};
    for (int q = 0; q < NQ; ++q) {
      for (int a = 0; a < ND; ++a) {
        S[q][a] = lagrange(ND, gl_nodes, a, q_pts[q]);
        Dm[q][a] = lagrange_d(ND, gl_nodes, a, q_pts[q]);
      }
      for (int a = 0; a < 2; ++a) {
        Sv1[q][a] = lagrange(2, vtx, a, q_pts[q]);
        Dv1[q][a] = lagrange_d(2, vtx, a, q_pts[q]);
      }
      for (int r = 0; r < NQ; ++r) Cg[q][r] = lagrange_d(NQ, q_pts, r, q_pts[q]);
    }
    for (int q = 0; q < NQP; ++q) {
      double w = 1.0;
      int r = q;
      for (int d = 0; d < DIM; ++d) { w *= q_w[r % NQ]; r /= NQ; }
      W[q] = w;
    }
    mu.assign(n_cells, 0.0);
    lam.assign(n_cells, 0.0);
    u.assign((size_t)n_nodes * DIM, 0.0);
    colour();
    reference_metric();
  }

  // nodes a cell writes: its own nodes and the masters of its constrained ones
  std::vector<int> write_set(int c) const {
    std::vector<int> w;
    for (int a = 0; a < NDOF1; ++a) {
      const int n = cell_nodes[(size_t)c * NDOF1 + a];
      w.push_back(n);
      if (!hang_of.empty() && hang_of[n] >= 0)
        for (int j = hc_ptr[hang_of[n]]; j < hc_ptr[hang_of[n] + 1]; ++j) w.push_back(hc_m[j]);
    }
    std::sort(w.begin(), w.end());
    w.erase(std::unique(w.begin(), w.end()), w.end());
    return w;
  }

  // Greedy first-fit colouring: no two cells of one colour write a common
  // node, so a colour's scatter-adds can run in parallel without conflicts.
  void colour() {
    std::vector<uint64_t> used(n_nodes, 0);
    colour_of.assign(n_cells, 0);
    int nc = 0;
    for (int c = 0; c < n_cells; ++c) {
      const std::vector<int> ws = write_set(c);
      uint64_t m = 0;
      for (int n : ws) m |= used[n];
      int col = 0;
      while (m & (1ull << col)) ++col;
      colour_of[c] = col;
      nc = std::max(nc, col + 1);
      for (int n : ws) used[n] |= (1ull << col);
    }
    colour_ptr.assign(nc + 1, 0);
    for (int c = 0; c < n_cells; ++c) colour_ptr[colour_of[c] + 1]++;
    for (int k = 0; k < nc; ++k) colour_ptr[k + 1] += colour_ptr[k];
    colour_cells.assign(n_cells, 0);
    std::vector<int> fill(colour_ptr.begin(), colour_ptr.end() - 1);
    for (int c = 0; c < n_cells; ++c) colour_cells[fill[colour_of[c]]++] = c;
  }
  int n_colours() override { return (int)colour_ptr.size() - 1; }

  static int corner_local(int v) {
    // Q1 vertex v (lexicographic over {0,1}^DIM) -> local Q_p node index
    int a = 0, stride = 1;
    for (int d = 0; d < DIM; ++d) {
      a += ((v >> d) & 1) * P * stride;
      stride *= ND;
    }
    return a;
  }

  // Reference configuration metric: StaticMappingQ1 of the cell vertices
  // (include/mf_elasticity.h:1828-1829).
  void reference_metric() {
    jinv_ref.assign((size_t)n_cells * NQP * DD, 0.0);
    jxw_ref.assign((size_t)n_cells * NQP, 0.0);
    for (int c = 0; c < n_cells; ++c) {
      double X[1 << DIM][DIM];
      for (int v = 0; v < (1 << DIM); ++v)
        for (int i = 0; i < DIM; ++i)
          X[v][i] = coords[(size_t)cell_nodes[(size_t)c * NDOF1 + corner_local(v)] * DIM + i];
      for (int q = 0; q < NQP; ++q) {
        int qi[3] = {0, 0, 0};
        int r = q;
        for (int d = 0; d < DIM; ++d) { qi[d] = r % NQ; r /= NQ; }
        double J[DD] = {};
        for (int v = 0; v < (1 << DIM); ++v)
          for (int d = 0; d < DIM; ++d) {
            double g = 1.0;
            for (int e = 0; e < DIM; ++e) {
              const int ve = (v >> e) & 1;
              g *= (e == d) ? Dv1[qi[e]][ve] : Sv1[qi[e]][ve];
            }
            for (int i = 0; i < DIM; ++i) J[i * DIM + d] += X[v][i] * g;
          }
        inv<DIM>(J, &jinv_ref[((size_t)c * NQP + q) * DD]);
        jxw_ref[(size_t)c * NQP + q] = det<DIM>(J) * W[q];
      }
    }
  }

  double volume() override {
    double s = 0;
    for (double v : jxw_ref) s += v;
    return s;
  }

  // ---- sum factorization: nodal values v[c][a] -> G[q][c][d] = d v_c / d xi_d
  void grad(const double *v, double *G) const {
    for (int c = 0; c < DIM; ++c) {
      const double *vc = v + c * NDOF1;
      if constexpr (DIM == 2) {
        double t0[ND][NQ], t1[ND][NQ];
        for (int ay = 0; ay < ND; ++ay)
          for (int qx = 0; qx < NQ; ++qx) {
            double s0 = 0, s1 = 0;
            for (int ax = 0; ax < ND; ++ax) {
              s0 += S[qx][ax] * vc[ay * ND + ax];
              s1 += Dm[qx][ax] * vc[ay * ND + ax];
            }
            t0[ay][qx] = s0; t1[ay][qx] = s1;
          }
        for (int qy = 0; qy < NQ; ++qy)
          for (int qx = 0; qx < NQ; ++qx) {
            double gx = 0, gy = 0;
            for (int ay = 0; ay < ND; ++ay) {
              gx += S[qy][ay] * t1[ay][qx];
              gy += Dm[qy][ay] * t0[ay][qx];
            }
            const int q = qy * NQ + qx;
            G[(q * DIM + c) * DIM + 0] = gx;
            G[(q * DIM + c) * DIM + 1] = gy;
          }
      } else {
        double t0[ND][ND][NQ], t1[ND][ND][NQ];
        for (int az = 0; az < ND; ++az)
          for (int ay = 0; ay < ND; ++ay)
            for (int qx = 0; qx < NQ; ++qx) {
              double s0 = 0, s1 = 0;
              for (int ax = 0; ax < ND; ++ax) {
                const double x = vc[(az * ND + ay) * ND + ax];
                s0 += S[qx][ax] * x;
                s1 += Dm[qx][ax] * x;
              }
              t0[az][ay][qx] = s0; t1[az][ay][qx] = s1;
            }
        double s00[ND][NQ][NQ], s10[ND][NQ][NQ], s01[ND][NQ][NQ];
        for (int az = 0; az < ND; ++az)
          for (int qy = 0; qy < NQ; ++qy)
            for (int qx = 0; qx < NQ; ++qx) {
              double a = 0, b = 0, d = 0;
              for (int ay = 0; ay < ND; ++ay) {
                a += S[qy][ay] * t0[az][ay][qx];
                b += S[qy][ay] * t1[az][ay][qx];
                d += Dm[qy][ay] * t0[az][ay][qx];
              }
              s00[az][qy][qx] = a; s10[az][qy][qx] = b; s01[az][qy][qx] = d;
            }
        for (int qz = 0; qz < NQ; ++qz)
          for (int qy = 0; qy < NQ; ++qy)
            for (int qx = 0; qx < NQ; ++qx) {
              double gx = 0, gy = 0, gz = 0;
              for (int az = 0; az < ND; ++az) {
                gx += S[qz][az] * s10[az][qy][qx];
                gy += S[qz][az] * s01[az][qy][qx];
                gz += Dm[qz][az] * s00[az][qy][qx];
              }
              const int q = (qz * NQ + qy) * NQ + qx;
              G[(q * DIM + c) * DIM + 0] = gx;
              G[(q * DIM + c) * DIM + 1] = gy;
              G[(q * DIM + c) * DIM + 2] = gz;
            }
      }
    }
  }

  // transpose: out[c][a] = sum_q sum_d (d phi_a / d xi_d)(q) Phi[q][c][d]
  void grad_T(const double *Phi, double *out) const {
    for (int c = 0; c < DIM; ++c) {
      double *oc = out + c * NDOF1;
      if constexpr (DIM == 2) {
        double p0[ND][NQ], p1[ND][NQ];
        for (int ay = 0; ay < ND; ++ay)
          for (int qx = 0; qx < NQ; ++qx) {
            double a = 0, b = 0;
            for (int qy = 0; qy < NQ; ++qy) {
              const int q = qy * NQ + qx;
              a += Dm[qy][ay] * Phi[(q * DIM + c) * DIM + 1];
              b += S[qy][ay] * Phi[(q * DIM + c) * DIM + 0];
            }
            p0[ay][qx] = a; p1[ay][qx] = b;
          }
        for (int ay = 0; ay < ND; ++ay)
          for (int ax = 0; ax < ND; ++ax) {
            double s = 0;
            for (int qx = 0; qx < NQ; ++qx) s += S[qx][ax] * p0[ay][qx] + Dm[qx][ax] * p1[ay][qx];
            oc[ay * ND + ax] = s;
          }
      } else {
        double r10[ND][NQ][NQ], r01[ND][NQ][NQ], r00[ND][NQ][NQ];
        for (int az = 0; az < ND; ++az)
          for (int qy = 0; qy < NQ; ++qy)
            for (int qx = 0; qx < NQ; ++qx) {
              double a = 0, b = 0, d = 0;
              for (int qz = 0; qz < NQ; ++qz) {
                const int q = (qz * NQ + qy) * NQ + qx;
                a += S[qz][az] * Phi[(q * DIM + c) * DIM + 0];
                b += S[qz][az] * Phi[(q * DIM + c) * DIM + 1];
                d += Dm[qz][az] * Phi[(q * DIM + c) * DIM + 2];
              }
              r10[az][qy][qx] = a; r01[az][qy][qx] = b; r00[az][qy][qx] = d;
            }
        double p0[ND][ND][NQ], p1[ND][ND][NQ];
        for (int az = 0; az < ND; ++az)
          for (int ay = 0; ay < ND; ++ay)
            for (int qx = 0; qx < NQ; ++qx) {
              double a = 0, b = 0;
              for (int qy = 0; qy < NQ; ++qy) {
                a += S[qy][ay] * r00[az][qy][qx] + Dm[qy][ay] * r01[az][qy][qx];
                b += S[qy][ay] * r10[az][qy][qx];
              }
              p0[az][ay][qx] = a; p1[az][ay][qx] = b;
            }
        for (int az = 0; az < ND; ++az)
          for (int ay = 0; ay < ND; ++ay)
            for (int ax = 0; ax < ND; ++ax) {
              double s = 0;
              for (int qx = 0; qx < NQ; ++qx) s += S[qx][ax] * p0[az][ay][qx] + Dm[qx][ax] * p1[az][ay][qx];
              oc[(az * ND + ay) * ND + ax] = s;
            }
      }
    }
  }

  // nodal values -> values at qp: out[q][c]
  void values(const double *v, double *out) const {
    for (int q = 0; q < NQP; ++q) {
      int qi[3] = {0, 0, 0};
      int r = q;
      for (int d = 0; d < DIM; ++d) { qi[d] = r % NQ; r /= NQ; }
      for (int c = 0; c < DIM; ++c) {
        double s = 0;
        for (int a = 0; a < NDOF1; ++a) {
          int ai = a;
          double f = 1.0;
          for (int d = 0; d < DIM; ++d) { f *= S[qi[d]][ai % ND]; ai /= ND; }
          s += f * v[c * NDOF1 + a];
        }
        out[q * DIM + c] = s;
      }
    }
  }

  // collocation gradient of values given at the Gauss points (only NQ == ND):
  // G[q][c][d] from vq[c][q] (used by scalar_referential,
  // include/mf_nh_operator.h:1276-1310)
  void colloc_grad(const double *vq, double *G) const {
    for (int c = 0; c < DIM; ++c)
      for (int q = 0; q < NQP; ++q) {
        int qi[3] = {0, 0, 0};
        int r = q;
        for (int d = 0; d < DIM; ++d) { qi[d] = r % NQ; r /= NQ; }
        for (int d = 0; d < DIM; ++d) {
          double s = 0;
          for (int k = 0; k < NQ; ++k) {
            int qq[3] = {qi[0], qi[1], qi[2]};
            qq[d] = k;
            int idx = 0, st = 1;
            for (int e = 0; e < DIM; ++e) { idx += qq[e] * st; st *= NQ; }
            s += Cg[qi[d]][k] * vq[c * NQP + idx];
          }
          G[(q * DIM + c) * DIM + d] = s;
        }
      }
  }

  bool hanging(int n) const { return !hang_of.empty() && hang_of[n] >= 0; }

  // read_dof_values: Dirichlet DoFs read as zero (apply_constraints),
  // constrained nodes as the weighted sum of their masters
  void gather(const double *vec, int cell, double *loc, bool apply_constraints) const {
    for (int a = 0; a < NDOF1; ++a) {
      const int n = cell_nodes[(size_t)cell * NDOF1 + a];
      for (int c = 0; c < DIM; ++c) {
        const size_t dof = (size_t)n * DIM + c;
        double v;
        if (hanging(n)) {
          // node constraints belong to the space: applied also without the
          // Dirichlet ones (the plain product of the Dirichlet lifting)
          v = 0.0;
          for (int j = hc_ptr[hang_of[n]]; j < hc_ptr[hang_of[n] + 1]; ++j) {
            const size_t md = (size_t)hc_m[j] * DIM + c;
            if (!(apply_constraints && cons[md])) v += hc_w[j] * vec[md];
          }
        } else {
          v = (apply_constraints && cons[dof]) ? 0.0 : vec[dof];
        }
        loc[c * NDOF1 + a] = v;
      }
    }
  }
  // distribute_local_to_global: Dirichlet rows skipped (apply_constraints),
  // constrained nodes' entries distributed to their masters (hanging)
  void scatter(double *vec, int cell, const double *loc, bool apply_constraints, bool hang = true) const {
    for (int a = 0; a < NDOF1; ++a) {
      const int n = cell_nodes[(size_t)cell * NDOF1 + a];
      for (int c = 0; c < DIM; ++c) {
        const size_t dof = (size_t)n * DIM + c;
        if (hang && hanging(n)) {
          for (int j = hc_ptr[hang_of[n]]; j < hc_ptr[hang_of[n] + 1]; ++j) {
            const size_t md = (size_t)hc_m[j] * DIM + c;
            if (apply_constraints && cons[md]) continue;
            vec[md] += hc_w[j] * loc[c * NDOF1 + a];
          }
          continue;
        }
        if (apply_constraints && cons[dof]) continue;
        vec[dof] += loc[c * NDOF1 + a];
      }
    }
  }

  int set_node_constraints(int n, const int *nodes, const int *ptr, const int *masters,
                           const double *weights) override {
    hang_of.assign(n_nodes, -1);
    hc_ptr.assign(ptr, ptr + n + 1);
    hc_m.assign(masters, masters + ptr[n]);
    hc_w.assign(weights, weights + ptr[n]);
    for (int k = 0; k < n; ++k) hang_of[nodes[k]] = k;
    for (int j = 0; j < ptr[n]; ++j)
      if (hang_of[masters[j]] >= 0) return -1;  // masters must be free nodes
    colour();
    cached = false;
    return 0;
  }

  int set_material(const double *mu_, const double *lambda_) override {
    mu.assign(mu_, mu_ + n_cells);
    lam.assign(lambda_, lambda_ + n_cells);
    cached = false;
    return 0;
  }
  int set_linearization(const double *u_) override {
    u.assign(u_, u_ + (size_t)n_nodes * DIM);
    cached = false;
    return 0;
  }

  // Kinematics at qp from the displacement: F = I + Grad u (referential).
  void kinematics(int cell, double *Fq /*NQP*DD*/) const {
    double ul[DIM * NDOF1], G[NQP * DD];
    gather(u.data(), cell, ul, false);  // read_dof_values_plain
    grad(ul, G);
    for (int q = 0; q < NQP; ++q) {
      const double *Ji = &jinv_ref[((size_t)cell * NQP + q) * DD];
      matmul<DIM>(&G[q * DD], Ji, &Fq[q * DD]);
      for (int d = 0; d < DIM; ++d) Fq[q * DD + d * DIM + d] += 1.0;
    }
  }

  int min_detF(double *out) override {
    double m = 1e300;
    for (int c = 0; c < n_cells; ++c) {
      double Fq[NQP * DD];
      kinematics(c, Fq);
      for (int q = 0; q < NQP; ++q) m = std::min(m, det<DIM>(&Fq[q * DD]));
    }
    *out = m;
    return 0;
  }

  // NeoHookOperator::cache(), formulation 1 (include/mf_nh_operator.h:488-583)
  // plus the Eulerian metric that MatrixFree computes from MappingQEulerian
  // (include/mf_elasticity.h:1821-1831, 1949-1951).
  int cache() override {
    const size_t nqp = (size_t)n_cells * NQP;
    jinv_cur.assign(nqp * DD, 0.0);
    jxw_cur.assign(nqp, 0.0);
    c_s.assign(nqp, 0.0);
    c_s2.clear(); c_t2.clear(); c_t4.clear(); c_t4ns.clear(); c_uqp.clear();
    if (caching == TENSOR2) { c_s2.assign(nqp, 0.0); c_t2.assign(nqp * DD, 0.0); }
    if (caching == TENSOR4) { c_t2.assign(nqp * DD, 0.0); c_t4.assign(nqp * D4, 0.0); }
    if (caching == TENSOR4_NS) c_t4ns.assign(nqp * D4, 0.0);
    if (caching == SCALAR_REF) c_uqp.assign(nqp * DIM, 0.0);
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
    for (int cell = 0; cell < n_cells; ++cell) {
      // Eulerian metric: x = X + u at the Q_p support points
      double xl[DIM * NDOF1], Gx[NQP * DD];
      for (int a = 0; a < NDOF1; ++a) {
        const int n = cell_nodes[(size_t)cell * NDOF1 + a];
        for (int c = 0; c < DIM; ++c)
          xl[c * NDOF1 + a] = coords[(size_t)n * DIM + c] + u[(size_t)n * DIM + c];
      }
      grad(xl, Gx);
      for (int q = 0; q < NQP; ++q) {
        const size_t cq = (size_t)cell * NQP + q;
        inv<DIM>(&Gx[q * DD], &jinv_cur[cq * DD]);
        jxw_cur[cq] = det<DIM>(&Gx[q * DD]) * W[q];
      }
      double Fq[NQP * DD];
      kinematics(cell, Fq);
      double uq[NQP * DIM];
      if (caching == SCALAR_REF) {
        double ul[DIM * NDOF1];
        gather(u.data(), cell, ul, false);
        values(ul, uq);
      }
      const double m = mu[cell], l = lam[cell];
      for (int q = 0; q < NQP; ++q) {
        const size_t cq = (size_t)cell * NQP + q;
        const double *F = &Fq[q * DD];
        const double J = det<DIM>(F);
        if (!(J > 0)) { ++bad; continue; }
        if (formulation == 0) {
          // formulation 0 caches J^{-1/dim} whatever the caching
          // (include/mf_nh_operator.h:467-487)
          c_s[cq] = std::pow(J, -1.0 / DIM);
          continue;
        }
        const double scalar = m - 2.0 * l * std::log(J);
        double tau[DD];
        matmul_bt<DIM>(F, F, tau);  // b = F F^T
        for (int k = 0; k < DD; ++k) tau[k] *= m;
        for (int d = 0; d < DIM; ++d) tau[d * DIM + d] -= scalar;
        if (caching == SCALAR) {
          c_s[cq] = scalar;
        } else if (caching == SCALAR_REF) {
          c_s[cq] = scalar;
          for (int c = 0; c < DIM; ++c) c_uqp[(size_t)cell * NQP * DIM + c * NQP + q] = uq[q * DIM + c];
        } else if (caching == TENSOR2) {
          c_s[cq] = scalar * 2.0 / J;
          c_s2[cq] = 2.0 * l / J;
          for (int k = 0; k < DD; ++k) c_t2[cq * DD + k] = tau[k] / J;
        } else if (caching == TENSOR4) {
          for (int k = 0; k < DD; ++k) c_t2[cq * DD + k] = tau[k] / J;
          const double a = scalar * 2.0 / J, b = l * 2.0 / J;
          double *C = &c_t4[cq * D4];
          for (int i = 0; i < DIM; ++i)
            for (int j = 0; j < DIM; ++j)
              for (int k = 0; k < DIM; ++k)
                for (int ll = 0; ll < DIM; ++ll) {
                  const double Sijkl = 0.5 * ((i == k && j == ll) + (i == ll && j == k));
                  const double IxI = (i == j && k == ll) ? 1.0 : 0.0;
                  C[((i * DIM + j) * DIM + k) * DIM + ll] = a * Sijkl + b * IxI;
                }
        } else if (caching == TENSOR4_NS) {
          double Fi[DD], Fit[DD];
          inv<DIM>(F, Fi);
          for (int i = 0; i < DIM; ++i)
            for (int j = 0; j < DIM; ++j) Fit[i * DIM + j] = Fi[j * DIM + i];
          double *A = &c_t4ns[cq * D4];
          for (int i = 0; i < DIM; ++i)
            for (int j = 0; j < DIM; ++j)
              for (int k = 0; k < DIM; ++k)
                for (int ll = 0; ll < DIM; ++ll)
                  A[((i * DIM + j) * DIM + k) * DIM + ll] =
                      2.0 * l * Fit[i * DIM + j] * Fit[k * DIM + ll] +
                      scalar * Fit[i * DIM + ll] * Fi[j * DIM + k] +
                      m * ((i == k && j == ll) ? 1.0 : 0.0);
        }
      }
    }
    cached = (bad == 0);
    return bad ? -4 : 0;
  }

  // do_operation_on_cell for formulation 1 (include/mf_nh_operator.h:1182-1455)
  // Kirchhoff stress and the action of J c on sym(g) for formulation 0
  // (include/material.h:116-141, 171-226; include/mf_nh_operator.h:1005-1180):
  // tau = (J dPsi/dJ - tr(tau_bar)/d) I + tau_bar, tau_bar = mu b_bar,
  // b_bar = J^{-2/d} b.  s_bar = J^{-1/d}.
  static void form0_tau(const double *F, double s_bar, double m, double l, double *tau, double *tau_iso,
                        double &dpsi_J, double &kappa, double &tr_tau_bar) {
    const double J = det<DIM>(F);
    kappa = l + 2.0 * m / 3.0;
    double b[DD];
    matmul_bt<DIM>(F, F, b);
    tr_tau_bar = 0;
    for (int k = 0; k < DD; ++k) tau[k] = m * s_bar * s_bar * b[k];  // tau_bar
    for (int d = 0; d < DIM; ++d) tr_tau_bar += tau[d * DIM + d];
    for (int k = 0; k < DD; ++k) tau_iso[k] = tau[k];
    for (int d = 0; d < DIM; ++d) tau_iso[d * DIM + d] -= tr_tau_bar / DIM;
    dpsi_J = 0.5 * kappa * (J - 1.0 / J) * J;
    for (int d = 0; d < DIM; ++d) tau[d * DIM + d] += dpsi_J - tr_tau_bar / DIM;
  }

  void cell_op_form0(int cell, const double *in, double *out) const {
    double G[NQP * DD], Phi[NQP * DD], Gu[NQP * DD], ul[DIM * NDOF1];
    grad(in, G);
    gather(u.data(), cell, ul, false);
    grad(ul, Gu);
    const double m = mu[cell], l = lam[cell];
    for (int q = 0; q < NQP; ++q) {
      const size_t cq = (size_t)cell * NQP + q;
      double F[DD], g[DD], s[DD], tau[DD], tau_iso[DD], jc[DD], t1[DD], Q[DD];
      matmul<DIM>(&Gu[q * DD], &jinv_ref[cq * DD], F);
      for (int d = 0; d < DIM; ++d) F[d * DIM + d] += 1.0;
      const double J = det<DIM>(F);
      double dpsi_J, kappa, tr_tau_bar;
      form0_tau(F, c_s[cq], m, l, tau, tau_iso, dpsi_J, kappa, tr_tau_bar);
      matmul<DIM>(&G[q * DD], &jinv_cur[cq * DD], g);
      for (int i = 0; i < DIM; ++i)
        for (int j = 0; j < DIM; ++j) s[i * DIM + j] = 0.5 * (g[i * DIM + j] + g[j * DIM + i]);
      double tr = 0, tiso_s = 0;
      for (int d = 0; d < DIM; ++d) tr += s[d * DIM + d];
      for (int k = 0; k < DD; ++k) tiso_s += tau_iso[k] * s[k];
      const double d2psi = 0.5 * kappa * (1.0 + 1.0 / (J * J));
      const double dpsi = 0.5 * kappa * (J - 1.0 / J);
      for (int i = 0; i < DIM; ++i)
        for (int j = 0; j < DIM; ++j) {
          const double dev = s[i * DIM + j] - (i == j ? tr / DIM : 0.0);
          jc[i * DIM + j] = -2.0 * dpsi_J * s[i * DIM + j] + (i == j ? J * (dpsi + J * d2psi) * tr : 0.0) +
                            (2.0 / DIM) * tr_tau_bar * dev - (2.0 / DIM) * tr * tau_iso[i * DIM + j] -
                            (i == j ? (2.0 / DIM) * tiso_s : 0.0);
        }
      matmul<DIM>(g, tau, t1);  // geometric part grad v tau
      for (int k = 0; k < DD; ++k) Q[k] = (jc[k] + t1[k]) / J;
      matmul_bt<DIM>(Q, &jinv_cur[cq * DD], t1);
      for (int k = 0; k < DD; ++k) Phi[q * DD + k] = t1[k] * jxw_cur[cq];
    }
    grad_T(Phi, out);
  }

  void cell_op(int cell, const double *in, double *out) const {
    if (formulation == 0) {
      cell_op_form0(cell, in, out);
      return;
    }
    double G[NQP * DD], Phi[NQP * DD];
    grad(in, G);
    double Gu[NQP * DD];
    if (caching == SCALAR) {
      double ul[DIM * NDOF1];
      gather(u.data(), cell, ul, false);
      grad(ul, Gu);
    } else if (caching == SCALAR_REF) {
      colloc_grad(&c_uqp[(size_t)cell * NQP * DIM], Gu);
    }
    const double m = mu[cell], l = lam[cell];
    for (int q = 0; q < NQP; ++q) {
      const size_t cq = (size_t)cell * NQP + q;
      const double *Gq = &G[q * DD];
      double *Ph = &Phi[q * DD];
      if (caching == TENSOR4_NS) {
        // referential gradient and dP/dF contraction <2,0,3,1>
        const double *Jr = &jinv_ref[cq * DD];
        double Gr[DD], Pp[DD], tmp[DD];
        matmul<DIM>(Gq, Jr, Gr);
        const double *A = &c_t4ns[cq * D4];
        for (int i = 0; i < DIM; ++i)
          for (int j = 0; j < DIM; ++j) {
            double s = 0;
            for (int k = 0; k < DIM; ++k)
              for (int ll = 0; ll < DIM; ++ll) s += A[((i * DIM + j) * DIM + k) * DIM + ll] * Gr[k * DIM + ll];
            Pp[i * DIM + j] = s;
          }
        matmul_bt<DIM>(Pp, Jr, tmp);  // submit_gradient: flux * J^{-T}
        for (int k = 0; k < DD; ++k) Ph[k] = tmp[k] * jxw_ref[cq];
        continue;
      }
      if (caching == SCALAR_REF) {
        const double *Jr = &jinv_ref[cq * DD];
        double F[DD], Fi[DD], Gr[DD], g[DD], b[DD], tau[DD], Qp[DD], t1[DD], t2[DD];
        matmul<DIM>(&Gu[q * DD], Jr, F);
        for (int d = 0; d < DIM; ++d) F[d * DIM + d] += 1.0;
        matmul<DIM>(Gq, Jr, Gr);
        inv<DIM>(F, Fi);
        matmul<DIM>(Gr, Fi, g);
        matmul_bt<DIM>(F, F, b);
        const double cs = c_s[cq];
        for (int k = 0; k < DD; ++k) tau[k] = m * b[k];
        for (int d = 0; d < DIM; ++d) tau[d * DIM + d] -= cs;
        double tr = 0;
        for (int d = 0; d < DIM; ++d) tr += g[d * DIM + d];
        matmul<DIM>(g, tau, t1);
        for (int i = 0; i < DIM; ++i)
          for (int j = 0; j < DIM; ++j)
            Qp[i * DIM + j] = cs * (g[i * DIM + j] + g[j * DIM + i]) + (i == j ? 2.0 * l * tr : 0.0) + t1[i * DIM + j];
        matmul_bt<DIM>(Qp, Fi, t2);   // queued * F^{-T}
        matmul_bt<DIM>(t2, Jr, t1);   // submit_gradient on reference metric
        for (int k = 0; k < DD; ++k) Ph[k] = t1[k] * jxw_ref[cq];
        continue;
      }
      // spatial variants: gradient through the Eulerian metric
      const double *Jc = &jinv_cur[cq * DD];
      double g[DD], Q[DD], t1[DD];
      matmul<DIM>(Gq, Jc, g);
      double s[DD];
      for (int i = 0; i < DIM; ++i)
        for (int j = 0; j < DIM; ++j) s[i * DIM + j] = 0.5 * (g[i * DIM + j] + g[j * DIM + i]);
      double trs = 0;
      for (int d = 0; d < DIM; ++d) trs += s[d * DIM + d];
      if (caching == SCALAR) {
        const double *Jr = &jinv_ref[cq * DD];
        double F[DD], b[DD], tau[DD];
        matmul<DIM>(&Gu[q * DD], Jr, F);
        for (int d = 0; d < DIM; ++d) F[d * DIM + d] += 1.0;
        matmul_bt<DIM>(F, F, b);
        const double cs = c_s[cq];
        for (int k = 0; k < DD; ++k) tau[k] = m * b[k];
        for (int d = 0; d < DIM; ++d) tau[d * DIM + d] -= cs;
        const double J = det<DIM>(F);
        matmul<DIM>(g, tau, t1);
        for (int i = 0; i < DIM; ++i)
          for (int j = 0; j < DIM; ++j)
            Q[i * DIM + j] = (2.0 * cs * s[i * DIM + j] + (i == j ? 2.0 * l * trs : 0.0) + t1[i * DIM + j]) * (1.0 / J);
      } else if (caching == TENSOR2) {
        const double c1 = c_s[cq], c2 = c_s2[cq];
        matmul<DIM>(g, &c_t2[cq * DD], t1);
        for (int i = 0; i < DIM; ++i)
          for (int j = 0; j < DIM; ++j)
            Q[i * DIM + j] = c1 * s[i * DIM + j] + (i == j ? c2 * trs : 0.0) + t1[i * DIM + j];
      } else {  // TENSOR4
        matmul<DIM>(g, &c_t2[cq * DD], t1);
        const double *C = &c_t4[cq * D4];
        for (int i = 0; i < DIM; ++i)
          for (int j = 0; j < DIM; ++j) {
            double cs = 0;
            for (int k = 0; k < DIM; ++k)
              for (int ll = 0; ll < DIM; ++ll) cs += C[((i * DIM + j) * DIM + k) * DIM + ll] * s[k * DIM + ll];
            Q[i * DIM + j] = t1[i * DIM + j] + cs;
          }
      }
      matmul_bt<DIM>(Q, Jc, t1);
      for (int k = 0; k < DD; ++k) Ph[k] = t1[k] * jxw_cur[cq];
    }
    grad_T(Phi, out);
  }

  // vmult / vmult_add (include/mf_nh_operator.h:612-706); plain = no
  // constraints (used for the Dirichlet lifting -K g of the residual).
  int vmult(double *dst, const double *src, int add, int plain) override {
    if (!cached) return -5;
    const size_t nd = (size_t)n_nodes * DIM;
    if (!add) std::fill(dst, dst + nd, 0.0);
    const bool ac = !plain;
    for (int col = 0; col + 1 < (int)colour_ptr.size(); ++col) {
#pragma omp parallel for schedule(static)
      for (int k = colour_ptr[col]; k < colour_ptr[col + 1]; ++k) {
        const int cell = colour_cells[k];
        double in[DIM * NDOF1], out[DIM * NDOF1];
        gather(src, cell, in, ac);
        cell_op(cell, in, out);
        scatter(dst, cell, out, ac);
      }
    }
    if (ac)
      for (size_t i = 0; i < nd; ++i)
        if (cons[i] || hanging((int)(i / DIM))) dst[i] += src[i];
    return 0;
  }

  // compute_diagonal by unit-vector probing (include/mf_nh_operator.h:862-918,
  // 1503-1557): constrained entries are one.
  int diagonal(double *diag) override {
    if (!cached) return -5;
    const size_t nd = (size_t)n_nodes * DIM;
    std::fill(diag, diag + nd, 0.0);
    for (int col = 0; col + 1 < (int)colour_ptr.size(); ++col) {
#pragma omp parallel for schedule(dynamic, 4)
      for (int k = colour_ptr[col]; k < colour_ptr[col + 1]; ++k) {
        const int cell = colour_cells[k];
        double in[DIM * NDOF1], out[DIM * NDOF1], dl[DIM * NDOF1];
        for (int i = 0; i < DIM * NDOF1; ++i) {
          std::fill(in, in + DIM * NDOF1, 0.0);
          in[i] = 1.0;
          cell_op(cell, in, out);
          dl[i] = out[i];
        }
        if (hang_of.empty()) {
          scatter(diag, cell, dl, true);
          continue;
        }
        // with node constraints probe the constrained space: for every DoF
        // (m, c) the cell writes, in = C_cell e_(m,c), diag += in . A_cell in
        for (int m : write_set(cell)) {
          if (hanging(m)) continue;
          for (int c = 0; c < DIM; ++c) {
            if (cons[(size_t)m * DIM + c]) continue;
            std::fill(in, in + DIM * NDOF1, 0.0);
            bool any = false;
            for (int a = 0; a < NDOF1; ++a) {
              const int n = cell_nodes[(size_t)cell * NDOF1 + a];
              if (n == m) in[c * NDOF1 + a] = 1.0, any = true;
              if (hanging(n))
                for (int j = hc_ptr[hang_of[n]]; j < hc_ptr[hang_of[n] + 1]; ++j)
                  if (hc_m[j] == m) in[c * NDOF1 + a] += hc_w[j], any = true;
            }
            if (!any) continue;
            cell_op(cell, in, out);
            double v = 0;
            for (int i = 0; i < DIM * NDOF1; ++i) v += in[i] * out[i];
            diag[(size_t)m * DIM + c] += v;
          }
        }
      }
    }
    for (size_t i = 0; i < nd; ++i)
      if (cons[i] || hanging((int)(i / DIM))) diag[i] = 1.0;
    return 0;
  }

  // Internal-force part of the residual, include/mf_elasticity.h:2769-2804:
  // rhs_i -= sum_q sym(Grad N_i F^-1) : tau JxW_ref  (no constraints applied)
  int residual(double *rhs) override {
    const size_t nd = (size_t)n_nodes * DIM;
    std::fill(rhs, rhs + nd, 0.0);
    for (int col = 0; col + 1 < (int)colour_ptr.size(); ++col) {
#pragma omp parallel for schedule(static)
      for (int k = colour_ptr[col]; k < colour_ptr[col + 1]; ++k) {
        const int cell = colour_cells[k];
        double Fq[NQP * DD], Phi[NQP * DD], out[DIM * NDOF1];
        kinematics(cell, Fq);
        const double m = mu[cell], l = lam[cell];
        for (int q = 0; q < NQP; ++q) {
          const size_t cq = (size_t)cell * NQP + q;
          const double *F = &Fq[q * DD];
          double Fi[DD], tau[DD], P1[DD], t1[DD];
          const double J = det<DIM>(F);
          inv<DIM>(F, Fi);
          if (formulation == 0) {
            double tau_iso[DD], dpsi_J, kappa, tr_tau_bar;
            form0_tau(F, std::pow(J, -1.0 / DIM), m, l, tau, tau_iso, dpsi_J, kappa, tr_tau_bar);
          } else {
            matmul_bt<DIM>(F, F, tau);
            for (int k2 = 0; k2 < DD; ++k2) tau[k2] *= m;
            const double scalar = m - 2.0 * l * std::log(J);
            for (int d = 0; d < DIM; ++d) tau[d * DIM + d] -= scalar;
          }
          matmul_bt<DIM>(tau, Fi, P1);                       // P = tau F^{-T}
          matmul_bt<DIM>(P1, &jinv_ref[cq * DD], t1);        // to unit-cell gradients
          for (int k2 = 0; k2 < DD; ++k2) Phi[q * DD + k2] = -t1[k2] * jxw_ref[cq];
        }
        grad_T(Phi, out);
        scatter(rhs, cell, out, false);
      }
    }
    return 0;
  }

  // Neumann traction in the reference configuration,
  // include/mf_elasticity.h:2859-2892: rhs_i += sum_fq N_i t JxW_face.
  // face_ids[k] in [0, 2*DIM): face 2d+s is xi_d = s.
  int neumann(const int *face_cells, const int *face_ids, int n_faces, const double *t, double *rhs) override {
    for (int k = 0; k < n_faces; ++k) {
      const int cell = face_cells[k], f = face_ids[k];
      const int fd = f / 2, side = f % 2;
      double X[1 << DIM][DIM];
      for (int v = 0; v < (1 << DIM); ++v)
        for (int i = 0; i < DIM; ++i)
          X[v][i] = coords[(size_t)cell_nodes[(size_t)cell * NDOF1 + corner_local(v)] * DIM + i];
      const int nfq = ipow(NQ, DIM - 1);
      for (int fq = 0; fq < nfq; ++fq) {
        // face point in unit coordinates
        double xi[3] = {0, 0, 0};
        double w = 1.0;
        int r = fq;
        for (int d = 0; d < DIM; ++d) {
          if (d == fd) { xi[d] = side; continue; }
          xi[d] = q_pts[r % NQ];
          w *= q_w[r % NQ];
          r /= NQ;
        }
        // Q1 Jacobian at xi
        double J[DD] = {};
        for (int v = 0; v < (1 << DIM); ++v)
          for (int d = 0; d < DIM; ++d) {
            double g = 1.0;
            for (int e = 0; e < DIM; ++e) {
              const int ve = (v >> e) & 1;
              const double s = ve ? xi[e] : 1.0 - xi[e];
              const double ds = ve ? 1.0 : -1.0;
              g *= (e == d) ? ds : s;
            }
            for (int i = 0; i < DIM; ++i) J[i * DIM + d] += X[v][i] * g;
          }
        // face measure: norm of cofactor column fd
        double Ji[DD];
        inv<DIM>(J, Ji);
        const double dJ = det<DIM>(J);
        double nrm = 0;
        for (int i = 0; i < DIM; ++i) nrm += Ji[fd * DIM + i] * Ji[fd * DIM + i];
        const double jxw = std::abs(dJ) * std::sqrt(nrm) * w;
        for (int a = 0; a < NDOF1; ++a) {
          int ai = a;
          double phi = 1.0;
          for (int d = 0; d < DIM; ++d) {
            phi *= lagrange(ND, gl_nodes, ai % ND, xi[d]);
            ai /= ND;
          }
          const int n = cell_nodes[(size_t)cell * NDOF1 + a];
          for (int c = 0; c < DIM; ++c) rhs[(size_t)n * DIM + c] += phi * t[c] * jxw;
        }
      }
    }
    return 0;
  }

  // Per-qp intermediate state for the single-cell known-answer tests
  // (tests/mf_neo_hook_operator_form1.cc:339-541): out[q] = {Grad u (DD),
  // grad v (DD), Jc action (DD, without 1/J), geometric action g tau (DD),
  // JxW_ref, JxW_cur, det F}.
  int qp_state(int cell, const double *src, double *out) override {
    double in[DIM * NDOF1], G[NQP * DD], Gu[NQP * DD], ul[DIM * NDOF1];
    gather(src, cell, in, false);
    grad(in, G);
    gather(u.data(), cell, ul, false);
    grad(ul, Gu);
    const double m = mu[cell], l = lam[cell];
    const int stride = 4 * DD + 3;
    for (int q = 0; q < NQP; ++q) {
      const size_t cq = (size_t)cell * NQP + q;
      double *o = out + q * stride;
      double Gr[DD], g[DD], F[DD], b[DD], tau[DD], t1[DD];
      matmul<DIM>(&Gu[q * DD], &jinv_ref[cq * DD], Gr);
      for (int k = 0; k < DD; ++k) o[k] = Gr[k];
      matmul<DIM>(&G[q * DD], &jinv_cur[cq * DD], g);
      for (int k = 0; k < DD; ++k) o[DD + k] = g[k];
      for (int k = 0; k < DD; ++k) F[k] = Gr[k];
      for (int d = 0; d < DIM; ++d) F[d * DIM + d] += 1.0;
      const double J = det<DIM>(F);
      const double scalar = m - 2.0 * l * std::log(J);
      double trs = 0;
      for (int d = 0; d < DIM; ++d) trs += g[d * DIM + d];
      for (int i = 0; i < DIM; ++i)
        for (int j = 0; j < DIM; ++j)
          o[2 * DD + i * DIM + j] = scalar * (g[i * DIM + j] + g[j * DIM + i]) + (i == j ? 2.0 * l * trs : 0.0);
      matmul_bt<DIM>(F, F, b);
      for (int k = 0; k < DD; ++k) tau[k] = m * b[k];
      for (int d = 0; d < DIM; ++d) tau[d * DIM + d] -= scalar;
      matmul<DIM>(g, tau, t1);
      for (int k = 0; k < DD; ++k) o[3 * DD + k] = t1[k];
      o[4 * DD + 0] = jxw_ref[cq];
      o[4 * DD + 1] = jxw_cur.empty() ? 0.0 : jxw_cur[cq];
      o[4 * DD + 2] = J;
    }
    return 0;
  }
};

static OpBase *make_op(int dim, int p, int nq, int n_cells, int n_nodes, const int *cn, const double *xyz,
                       const unsigned char *cons, int caching) {
#define ORC_CASE(D, PP, QQ) \
  if (dim == D && p == PP && nq == QQ) return new Op<D, PP, QQ>(n_cells, n_nodes, cn, xyz, cons, caching);
  ORC_CASE(2, 1, 2)
  ORC_CASE(2, 2, 3)
  ORC_CASE(2, 3, 4)
  ORC_CASE(2, 4, 5)
  ORC_CASE(3, 1, 2)
  ORC_CASE(3, 2, 3)
  ORC_CASE(3, 3, 4)
  ORC_CASE(3, 4, 5)
#undef ORC_CASE
  return nullptr;
}

// ---------------------------------------------------------------------------
// Linear algebra for the Newton driver
// ---------------------------------------------------------------------------
static double dot(const std::vector<double> &a, const std::vector<double> &b) {
  double s = 0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}

// Preconditioned CG as deal.II's SolverCG (include/mf_elasticity.h:3049-3061,
// 3146-3150): stops when ||r|| <= tol; Jacobi preconditioner D^{-1}
// (relaxation does not change CG iterates).  Returns iterations, or -1.
static int pcg_jacobi(OpBase *op, std::vector<double> &x, const std::vector<double> &b,
                      const std::vector<double> &inv_diag, double tol, int max_it, double *res_out) {
  const size_t n = b.size();
  std::vector<double> r(n), z(n), pdir(n), Ap(n);
  op->vmult(Ap.data(), x.data(), 0, 0);
  for (size_t i = 0; i < n; ++i) r[i] = b[i] - Ap[i];
  double rn = std::sqrt(dot(r, r));
  if (rn <= tol) { *res_out = rn; return 0; }
  for (size_t i = 0; i < n; ++i) z[i] = inv_diag[i] * r[i];
  pdir = z;
  double rz = dot(r, z);
  int it = 0;
  while (it < max_it) {
    ++it;
    op->vmult(Ap.data(), pdir.data(), 0, 0);
    const double alpha = rz / dot(pdir, Ap);
    for (size_t i = 0; i < n; ++i) { x[i] += alpha * pdir[i]; r[i] -= alpha * Ap[i]; }
    rn = std::sqrt(dot(r, r));
    if (rn <= tol) break;
    for (size_t i = 0; i < n; ++i) z[i] = inv_diag[i] * r[i];
    const double rz_new = dot(r, z);
    const double beta = rz_new / rz;
    rz = rz_new;
    for (size_t i = 0; i < n; ++i) pdir[i] = z[i] + beta * pdir[i];
  }
  *res_out = rn;
  return rn <= tol ? it : -1;
}

}  // namespace orc

// ---------------------------------------------------------------------------
// extern "C" surface for ctypes
// ---------------------------------------------------------------------------
extern "C" {

int orc_fe_tables(int p, int nq, double *nodes, double *qpts, double *qw, double *S, double *D) {
  orc::gll01(p, nodes);
  orc::gauss01(nq, qpts, qw);
  for (int q = 0; q < nq; ++q)
    for (int a = 0; a <= p; ++a) {
      S[q * (p + 1) + a] = orc::lagrange(p + 1, nodes, a, qpts[q]);
      D[q * (p + 1) + a] = orc::lagrange_d(p + 1, nodes, a, qpts[q]);
    }
  return 0;
}

void *orc_op_create(int dim, int p, int nq, int n_cells, int n_nodes, const int *cell_nodes, const double *coords,
                    const unsigned char *constrained, int caching) {
  return orc::make_op(dim, p, nq, n_cells, n_nodes, cell_nodes, coords, constrained, caching);
}
void orc_op_destroy(void *op) { delete static_cast<orc::OpBase *>(op); }
int orc_op_set_material(void *op, const double *mu, const double *lambda) {
  return static_cast<orc::OpBase *>(op)->set_material(mu, lambda);
}
int orc_op_set_linearization(void *op, const double *u) { return static_cast<orc::OpBase *>(op)->set_linearization(u); }
int orc_op_cache(void *op) { return static_cast<orc::OpBase *>(op)->cache(); }
int orc_op_set_node_constraints(void *op, int n, const int *nodes, const int *ptr, const int *masters,
                                const double *weights) {
  return static_cast<orc::OpBase *>(op)->set_node_constraints(n, nodes, ptr, masters, weights);
}
int orc_op_set_formulation(void *op, int formulation) {
  if (formulation != 0 && formulation != 1) return -1;
  static_cast<orc::OpBase *>(op)->formulation = formulation;
  return 0;
}
int orc_op_vmult(void *op, double *dst, const double *src, int add, int plain) {
  return static_cast<orc::OpBase *>(op)->vmult(dst, src, add, plain);
}
int orc_op_diagonal(void *op, double *diag) { return static_cast<orc::OpBase *>(op)->diagonal(diag); }
int orc_op_residual(void *op, double *rhs) { return static_cast<orc::OpBase *>(op)->residual(rhs); }
int orc_op_neumann(void *op, const int *face_cells, const int *face_ids, int n_faces, const double *t, double *rhs) {
  return static_cast<orc::OpBase *>(op)->neumann(face_cells, face_ids, n_faces, t, rhs);
}
int orc_op_qp_state(void *op, int cell, const double *src, double *out) {
  return static_cast<orc::OpBase *>(op)->qp_state(cell, src, out);
}
double orc_op_volume(void *op) { return static_cast<orc::OpBase *>(op)->volume(); }
int orc_op_n_colours(void *op) { return static_cast<orc::OpBase *>(op)->n_colours(); }
int orc_op_min_detF(void *op, double *out) { return static_cast<orc::OpBase *>(op)->min_detF(out); }
int orc_num_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// Newton driver restating Solid::run / solve_nonlinear_timestep with the
// MF_CG + Jacobi linear solver (include/mf_elasticity.h:1187-1222,
// 2289-2559, 2932-3007, 3012-3214).
//   dirichlet_inc[n_dofs]: value of the Dirichlet expression per constrained
//     DoF (applied as the increment of the Newton update at iteration 0).
//   neumann faces with traction = t * traction_per_t (constant in space).
// Outputs per load step: newton_its[step], point displacement history of the
// requested DoFs (n_out_dofs per step) and the final total displacement u.
int orc_newton(void *opv, double dt, double t_end, const double *dirichlet_inc, int n_faces, const int *face_cells,
               const int *face_ids, const double *traction_per_t, int max_newton, double tol_f, double tol_u,
               double tol_lin, int n_out_dofs, const int *out_dofs, int *newton_its, double *out_hist, double *u_out,
               int max_steps, int verbose) {
  orc::OpBase *op = static_cast<orc::OpBase *>(opv);
  const int D = op->dim;
  const size_t nd = (size_t)op->n_nodes * D;
  std::vector<double> u_n(nd, 0.0), delta(nd, 0.0), u_tot(nd), rhs(nd), du(nd), diag(nd), inv_diag(nd), g(nd), Ag(nd);
  double t = 0.0;
  int step = 0;
  t += dt;
  while (t <= t_end && step < max_steps) {
    std::fill(delta.begin(), delta.end(), 0.0);
    double res0 = 0, upd0 = 0;
    int it = 0;
    bool converged = false;
    for (; it < max_newton; ++it) {
      // make_constraints: inhomogeneous values only at it 0
      for (size_t i = 0; i < nd; ++i) g[i] = (it == 0 && op->cons[i]) ? dirichlet_inc[i] : 0.0;
      for (size_t i = 0; i < nd; ++i) u_tot[i] = u_n[i] + delta[i];
      op->set_linearization(u_tot.data());
      if (op->cache() != 0) return -4;
      // assemble_system: residual + Neumann, lifting -K g on free rows
      op->residual(rhs.data());
      if (n_faces > 0) {
        double tr[3] = {0, 0, 0};
        for (int c = 0; c < D; ++c) tr[c] = t * traction_per_t[c];
        op->neumann(face_cells, face_ids, n_faces, tr, rhs.data());
      }
      bool any_g = false;
      for (size_t i = 0; i < nd; ++i) any_g |= (g[i] != 0.0);
      if (any_g) {
        op->vmult(Ag.data(), g.data(), 0, 1);
        for (size_t i = 0; i < nd; ++i) rhs[i] -= Ag[i];
      }
      for (size_t i = 0; i < nd; ++i)
        if (op->cons[i]) rhs[i] = 0.0;
      const double res = std::sqrt(orc::dot(rhs, rhs));
      if (it == 0) res0 = res;
      // check_convergence (include/mf_elasticity.h:2245-2284): at least 3
      // iterations; residual normalised by its it-0 value, update norm of the
      // previous iteration normalised by the it-0 update norm.
      if (it > 2) {
        const double rn = res0 > 0 ? res / res0 : res;
        double s = 0;
        for (size_t i = 0; i < nd; ++i)
          if (!op->cons[i]) s += du[i] * du[i];
        const double un = upd0 > 0 ? std::sqrt(s) / upd0 : std::sqrt(s);
        if (res <= 0.0) converged = true;  // tol_f_abs = 0
        if (rn <= tol_f && un <= tol_u) converged = true;
      }
      if (converged) break;
      op->diagonal(diag.data());
      for (size_t i = 0; i < nd; ++i) inv_diag[i] = 1.0 / diag[i];
      std::fill(du.begin(), du.end(), 0.0);
      const double tol = tol_lin * res;
      double lin_res = 0;
      const int lin_it = orc::pcg_jacobi(op, du, rhs, inv_diag, tol, (int)nd, &lin_res);
      if (lin_it < 0) return -6;
      for (size_t i = 0; i < nd; ++i)
        if (op->cons[i]) du[i] = 0.0;
      double s = 0;
      for (size_t i = 0; i < nd; ++i) s += du[i] * du[i];
      if (it == 0) upd0 = std::sqrt(s);
      if (verbose)
        std::fprintf(stderr, "step %d it %d res %.6e lin_it %d |du| %.6e\n", step + 1, it, res, lin_it, std::sqrt(s));
      // constraints.distribute: increment on constrained DoFs
      for (size_t i = 0; i < nd; ++i)
        if (op->cons[i]) du[i] = g[i];
      for (size_t i = 0; i < nd; ++i) delta[i] += du[i];
      // keep du with constrained zeroed for the next convergence check
      for (size_t i = 0; i < nd; ++i)
        if (op->cons[i]) du[i] = 0.0;
    }
    if (!converged) return -7;
    newton_its[step] = it;
    for (size_t i = 0; i < nd; ++i) u_n[i] += delta[i];
    for (int k = 0; k < n_out_dofs; ++k) out_hist[(size_t)step * n_out_dofs + k] = u_n[out_dofs[k]];
    ++step;
    t += dt;
  }
  std::memcpy(u_out, u_n.data(), nd * sizeof(double));
  return step;
}

}  // extern "C"
