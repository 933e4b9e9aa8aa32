// cpu_baseline.cpp -- the reference's CPU MatrixFree algorithm, restated for
// timing on the host: the CPU baseline of bench.py.
//
// TEST INFRASTRUCTURE ONLY.  bench.py's cpu_baseline leg times it and
// tests/test_cpu_baseline.py checks it against the oracle (lsmf_oracle.cpp);
// nothing in the product links, loads or calls it.
//
// What the reference runs on the CPU (deal.II MatrixFree through
// NeoHookOperator, include/mf_nh_operator.h:612-785, 922-1499) and how it is
// restated here:
//   * cell batches: W cells per SIMD vector (deal.II's VectorizedArray: 8
//     doubles with AVX-512, 4 with AVX2), every operation below acts on all
//     W lanes at once; batches hold cells of one colour of the mesh's parity
//     colouring, so the batches of a colour share no node and are spread
//     over OpenMP threads (deal.II's cell loop over colour partitions)
//   * read_dof_values / distribute_local_to_global with homogeneous
//     Dirichlet DoFs read as zero and skipped on write, identity on
//     constrained rows (:703-705)
//   * sum factorization: nodal values (Gauss-Lobatto support points) ->
//     reference gradients at the Gauss points, 1D matrices per direction,
//     and its transpose for integration (:975-989, :1484-1495)
//   * the caches of `tensor4` (:1425-1440) and `tensor2` (:1401-1424),
//     stored per batch as [field][point][lane] with the Jacobian weight
//     folded in: K = J_cur^-1 (9), T = tau det(J_ref) w (Voigt 6) and either
//     M = (2c S + 2 lambda IxI) det(J_ref) w (Voigt 6x6, 21) or
//     c1 = 2c det(J_ref) w, c2 = 2 lambda det(J_ref) w; the point action is
//     Q = g T + M : sym(g) (tensor2: + c1 sym(g) + c2 tr(g) I) with
//     g = Grad_xi v K and flux Q K^T -- the reference's
//     [g tau + Jc : sym g] / J integrated with JxW_cur = J det(J_ref) w
//   * cache(): F = I + Grad_xi u J_ref^-1 (Q1 vertex map, StaticMappingQ1),
//     c = mu - 2 lambda ln J, tau = mu b - c I (formulation 1,
//     include/material.h:108-163; :447-562)
// Built twice, for x86-64-v4 (W = 8) and x86-64-v3 (W = 4); the loader
// picks the one the host supports.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace cpb {

#if defined(__AVX512F__)
constexpr int W = 8;
#else
constexpr int W = 4;
#endif
typedef double vd __attribute__((vector_size(8 * W)));

enum { TENSOR2 = 1, TENSOR4 = 2 };
constexpr int NV = 6;
constexpr int NF2 = 9 + 6 + 2, NF4 = 9 + 6 + 21;

inline int vidx(int i, int j) { return i == j ? i : (i + j == 1 ? 5 : (i + j == 2 ? 4 : 3)); }
inline int sym_idx(int v, int w) {
  if (v > w) std::swap(v, w);
  return v * NV - v * (v - 1) / 2 + (w - v);
}

struct Op {
  int p = 0, N = 0, NQP = 0, caching = TENSOR4, nf = 0;
  int n_cells = 0, n_nodes = 0, n_batches = 0;
  std::vector<int32_t> nodes;     // [batch][NQP][W] node id, -1 on idle lanes
  std::vector<int32_t> cells;     // [batch][W] mesh cell, -1 on idle lanes
  std::vector<uint8_t> cons;      // [n_nodes * 3]
  std::vector<int32_t> cons_list; // constrained DoFs
  std::vector<int> colour_ptr;    // batches of colour k: [colour_ptr[k], colour_ptr[k+1])
  std::vector<double> coords;     // [n_nodes][3]
  std::vector<double> S, D, qw;   // S[q][a], D[q][a] on [0,1], Gauss weights
  std::vector<double> mu, lambda; // per mesh cell
  std::vector<double> u;          // linearisation state
  std::vector<vd> cache;          // [batch][field][NQP]
  int threads = 1;
};

// ---- sum factorization on W lanes -----------------------------------------
// in: nodal values v[NQP] (x fastest) -> gradient g[3][NQP] (d/dxi_x, y, z)
template <int N> inline void evaluate(const Op &o, const vd *v, vd *g0, vd *g1, vd *g2, vd *t) {
  constexpr int NQ = N * N * N;
  const double *S = o.S.data(), *D = o.D.data();
  vd *ux = t, *dx = t + NQ, *uxy = t + 2 * NQ, *dxy = t + 3 * NQ, *uxdy = t + 4 * NQ;
  for (int k = 0; k < N; ++k)
    for (int j = 0; j < N; ++j) {
      const vd *line = v + (k * N + j) * N;
      for (int q = 0; q < N; ++q) {
        vd a = {}, b = {};
        for (int i = 0; i < N; ++i) {
          a += S[q * N + i] * line[i];
          b += D[q * N + i] * line[i];
        }
        ux[(k * N + j) * N + q] = a;
        dx[(k * N + j) * N + q] = b;
      }
    }
  for (int k = 0; k < N; ++k)
    for (int q = 0; q < N; ++q)
      for (int i = 0; i < N; ++i) {
        vd a = {}, b = {}, c = {};
        for (int j = 0; j < N; ++j) {
          const int src = (k * N + j) * N + i;
          a += S[q * N + j] * ux[src];
          b += S[q * N + j] * dx[src];
          c += D[q * N + j] * ux[src];
        }
        uxy[(k * N + q) * N + i] = a;
        dxy[(k * N + q) * N + i] = b;
        uxdy[(k * N + q) * N + i] = c;
      }
  for (int q = 0; q < N; ++q)
    for (int j = 0; j < N; ++j)
      for (int i = 0; i < N; ++i) {
        vd a = {}, b = {}, c = {};
        for (int k = 0; k < N; ++k) {
          const int src = (k * N + j) * N + i;
          a += S[q * N + k] * dxy[src];
          b += S[q * N + k] * uxdy[src];
          c += D[q * N + k] * uxy[src];
        }
        const int dst = (q * N + j) * N + i;
        g0[dst] = a;
        g1[dst] = b;
        g2[dst] = c;
      }
}

// transpose: fluxes f0, f1, f2 at the points (per xi direction) -> nodal
// out[a] = sum_q (dphi_a/dxi_x f0 + dphi_a/dxi_y f1 + dphi_a/dxi_z f2)
template <int N> inline void integrate(const Op &o, const vd *f0, const vd *f1, const vd *f2, vd *out, vd *t) {
  constexpr int NQ = N * N * N;
  const double *S = o.S.data(), *D = o.D.data();
  vd *a0 = t, *a1 = t + NQ, *a2 = t + 2 * NQ, *p1 = t + 3 * NQ, *p2 = t + 4 * NQ;
  for (int c = 0; c < N; ++c)  // z: a0 = Sz^T f0, a1 = Sz^T f1, a2 = Dz^T f2
    for (int j = 0; j < N; ++j)
      for (int i = 0; i < N; ++i) {
        vd x = {}, y = {}, z = {};
        for (int q = 0; q < N; ++q) {
          const int src = (q * N + j) * N + i;
          x += S[q * N + c] * f0[src];
          y += S[q * N + c] * f1[src];
          z += D[q * N + c] * f2[src];
        }
        const int dst = (c * N + j) * N + i;
        a0[dst] = x;
        a1[dst] = y;
        a2[dst] = z;
      }
  for (int c = 0; c < N; ++c)  // y: p1 = Sy^T a0, p2 = Dy^T a1 + Sy^T a2
    for (int b = 0; b < N; ++b)
      for (int i = 0; i < N; ++i) {
        vd x = {}, y = {};
        for (int q = 0; q < N; ++q) {
          const int src = (c * N + q) * N + i;
          x += S[q * N + b] * a0[src];
          y += D[q * N + b] * a1[src] + S[q * N + b] * a2[src];
        }
        p1[(c * N + b) * N + i] = x;
        p2[(c * N + b) * N + i] = y;
      }
  for (int c = 0; c < N; ++c)  // x: out = Dx^T p1 + Sx^T p2
    for (int b = 0; b < N; ++b)
      for (int a = 0; a < N; ++a) {
        vd x = {};
        for (int q = 0; q < N; ++q) {
          const int src = (c * N + b) * N + q;
          x += D[q * N + a] * p1[src] + S[q * N + a] * p2[src];
        }
        out[(c * N + b) * N + a] = x;
      }
}

inline double inv3(const double A[3][3], double B[3][3]) {
  const double c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
  const double c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
  const double c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
  const double d = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
  const double id = 1.0 / d;
  B[0][0] = c00 * id;
  B[1][0] = c01 * id;
  B[2][0] = c02 * id;
  B[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) * id;
  B[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) * id;
  B[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) * id;
  B[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) * id;
  B[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) * id;
  B[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) * id;
  return d;
}

// gather a batch's component values (constrained DoFs read zero)
inline void gather(const Op &o, int b, const double *src, vd *v /* [3][NQP] */) {
  const int32_t *nd = o.nodes.data() + (size_t)b * o.NQP * W;
  for (int q = 0; q < o.NQP; ++q)
    for (int l = 0; l < W; ++l) {
      const int n = nd[q * W + l];
      for (int c = 0; c < 3; ++c)
        v[c * o.NQP + q][l] = (n < 0 || o.cons[(size_t)n * 3 + c]) ? 0.0 : src[(size_t)n * 3 + c];
    }
}

template <int N> void cache_t(Op &o) {
  constexpr int NQ = N * N * N;
  const int nf = o.nf;
  o.cache.assign((size_t)o.n_batches * nf * NQ, vd{});
#pragma omp parallel num_threads(o.threads)
  {
    std::vector<vd> v(3 * NQ), g(9 * NQ), t(5 * NQ);
#pragma omp for schedule(static)
    for (int b = 0; b < o.n_batches; ++b) {
      const int32_t *nd = o.nodes.data() + (size_t)b * NQ * W;
      for (int q = 0; q < NQ; ++q)
        for (int l = 0; l < W; ++l) {
          const int n = nd[q * W + l];
          for (int c = 0; c < 3; ++c) v[c * NQ + q][l] = n < 0 ? 0.0 : o.u[(size_t)n * 3 + c];
        }
      for (int c = 0; c < 3; ++c)
        evaluate<N>(o, &v[c * NQ], &g[(3 * c + 0) * NQ], &g[(3 * c + 1) * NQ], &g[(3 * c + 2) * NQ], t.data());
      vd *out = o.cache.data() + (size_t)b * nf * NQ;
      for (int l = 0; l < W; ++l) {
        const int cell = o.cells[(size_t)b * W + l];
        if (cell < 0) continue;
        double X[8][3];  // Q1 corners, lexicographic
        for (int vtx = 0; vtx < 8; ++vtx) {
          const int ix = (vtx & 1) ? N - 1 : 0, iy = (vtx & 2) ? N - 1 : 0, iz = (vtx & 4) ? N - 1 : 0;
          const int n = nd[((iz * N + iy) * N + ix) * W + l];
          for (int d = 0; d < 3; ++d) X[vtx][d] = o.coords[(size_t)n * 3 + d];
        }
        const double m = o.mu[cell], lam = o.lambda[cell];
        for (int q = 0; q < NQ; ++q) {
          const int q0 = q % N, q1 = (q / N) % N, q2 = q / (N * N);
          // reference Jacobian of the Q1 vertex map at the Gauss point
          const double *gp = o.qw.data() + N;  // Gauss points stored after the weights
          const double x[3] = {gp[q0], gp[q1], gp[q2]};
          double Jr[3][3] = {};
          for (int vtx = 0; vtx < 8; ++vtx) {
            const double b0 = (vtx & 1) ? x[0] : 1.0 - x[0], b1 = (vtx & 2) ? x[1] : 1.0 - x[1],
                         b2 = (vtx & 4) ? x[2] : 1.0 - x[2];
            const double s0 = (vtx & 1) ? 1.0 : -1.0, s1 = (vtx & 2) ? 1.0 : -1.0, s2 = (vtx & 4) ? 1.0 : -1.0;
            const double dN[3] = {s0 * b1 * b2, b0 * s1 * b2, b0 * b1 * s2};
            for (int i = 0; i < 3; ++i)
              for (int d = 0; d < 3; ++d) Jr[i][d] += X[vtx][i] * dN[d];
          }
          double Jri[3][3];
          const double detJr = inv3(Jr, Jri);
          double F[3][3];
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
              double s = 0.0;
              for (int d = 0; d < 3; ++d) s += g[(3 * i + d) * NQ + q][l] * Jri[d][j];
              F[i][j] = s + (i == j ? 1.0 : 0.0);
            }
          double Fi[3][3];
          const double J = inv3(F, Fi);
          const double cs = m - 2.0 * lam * std::log(J);
          const double dw = detJr * o.qw[q0] * o.qw[q1] * o.qw[q2];
          double K[3][3];
          for (int d = 0; d < 3; ++d)
            for (int j = 0; j < 3; ++j) {
              double s = 0.0;
              for (int e = 0; e < 3; ++e) s += Jri[d][e] * Fi[e][j];
              K[d][j] = s;
            }
          auto put = [&](int f, double val) { out[(size_t)f * NQ + q][l] = val; };
          for (int d = 0; d < 3; ++d)
            for (int j = 0; j < 3; ++j) put(3 * d + j, K[d][j]);
          for (int i = 0; i < 3; ++i)
            for (int j = i; j < 3; ++j) {
              double s = 0.0;
              for (int k = 0; k < 3; ++k) s += F[i][k] * F[j][k];
              put(9 + vidx(i, j), (m * s - (i == j ? cs : 0.0)) * dw);
            }
          if (o.caching == TENSOR2) {
            put(15, 2.0 * cs * dw);
            put(16, 2.0 * lam * dw);
          } else {
            for (int v1 = 0; v1 < NV; ++v1)
              for (int v2 = v1; v2 < NV; ++v2) {
                double M = 0.0;
                if (v1 < 3 && v2 < 3) M = 2.0 * lam + (v1 == v2 ? 2.0 * cs : 0.0);
                else if (v1 == v2) M = cs;  // engineering shear strain
                put(15 + sym_idx(v1, v2), M * dw);
              }
          }
        }
      }
    }
  }
}

template <int N, int CACHING> void vmult_t(const Op &o, double *dst, const double *src) {
  constexpr int NQ = N * N * N;
  constexpr int nf = CACHING == TENSOR2 ? NF2 : NF4;
  std::memset(dst, 0, sizeof(double) * (size_t)o.n_nodes * 3);
  for (size_t k = 0; k + 1 < o.colour_ptr.size(); ++k) {
#pragma omp parallel num_threads(o.threads)
    {
      std::vector<vd> v(3 * NQ), g(9 * NQ), f(9 * NQ), t(5 * NQ), r(NQ);
#pragma omp for schedule(static)
      for (int b = o.colour_ptr[k]; b < o.colour_ptr[k + 1]; ++b) {
        gather(o, b, src, v.data());
        for (int c = 0; c < 3; ++c)
          evaluate<N>(o, &v[c * NQ], &g[(3 * c + 0) * NQ], &g[(3 * c + 1) * NQ], &g[(3 * c + 2) * NQ], t.data());
        const vd *cc = o.cache.data() + (size_t)b * nf * NQ;
        for (int q = 0; q < NQ; ++q) {
          vd K[3][3], G[3][3], gg[3][3], Q[3][3];
          for (int d = 0; d < 3; ++d)
            for (int j = 0; j < 3; ++j) K[d][j] = cc[(size_t)(3 * d + j) * NQ + q];
          for (int i = 0; i < 3; ++i)
            for (int d = 0; d < 3; ++d) G[i][d] = g[(3 * i + d) * NQ + q];
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) gg[i][j] = G[i][0] * K[0][j] + G[i][1] * K[1][j] + G[i][2] * K[2][j];
          vd T[3][3];
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) T[i][j] = cc[(size_t)(9 + vidx(i, j)) * NQ + q];
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) Q[i][j] = gg[i][0] * T[0][j] + gg[i][1] * T[1][j] + gg[i][2] * T[2][j];
          if constexpr (CACHING == TENSOR2) {
            const vd c1 = cc[(size_t)15 * NQ + q], c2 = cc[(size_t)16 * NQ + q];
            const vd tr = gg[0][0] + gg[1][1] + gg[2][2];
            for (int i = 0; i < 3; ++i)
              for (int j = 0; j < 3; ++j) Q[i][j] += 0.5 * c1 * (gg[i][j] + gg[j][i]) + (i == j ? c2 * tr : vd{});
          } else {
            const vd s[NV] = {gg[0][0], gg[1][1], gg[2][2], gg[1][2] + gg[2][1], gg[0][2] + gg[2][0],
                              gg[0][1] + gg[1][0]};
            vd sig[NV];
            for (int v1 = 0; v1 < NV; ++v1) {
              vd a = {};
              for (int v2 = 0; v2 < NV; ++v2) a += cc[(size_t)(15 + sym_idx(v1, v2)) * NQ + q] * s[v2];
              sig[v1] = a;
            }
            for (int i = 0; i < 3; ++i)
              for (int j = 0; j < 3; ++j) Q[i][j] += sig[vidx(i, j)];
          }
          for (int c = 0; c < 3; ++c)
            for (int d = 0; d < 3; ++d) f[(3 * c + d) * NQ + q] = Q[c][0] * K[d][0] + Q[c][1] * K[d][1] + Q[c][2] * K[d][2];
        }
        const int32_t *nd = o.nodes.data() + (size_t)b * NQ * W;
        for (int c = 0; c < 3; ++c) {
          integrate<N>(o, &f[(3 * c + 0) * NQ], &f[(3 * c + 1) * NQ], &f[(3 * c + 2) * NQ], r.data(), t.data());
          for (int q = 0; q < NQ; ++q)
            for (int l = 0; l < W; ++l) {
              const int n = nd[q * W + l];
              if (n >= 0 && !o.cons[(size_t)n * 3 + c]) dst[(size_t)n * 3 + c] += r[q][l];
            }
        }
      }
    }
  }
  for (int32_t d : o.cons_list) dst[d] = src[d];
}

void cache(Op &o) {
  switch (o.N) {
    case 2: cache_t<2>(o); break;
    case 3: cache_t<3>(o); break;
    case 4: cache_t<4>(o); break;
    default: cache_t<5>(o); break;
  }
}

template <int CACHING> void vmult_c(const Op &o, double *dst, const double *src) {
  switch (o.N) {
    case 2: vmult_t<2, CACHING>(o, dst, src); break;
    case 3: vmult_t<3, CACHING>(o, dst, src); break;
    case 4: vmult_t<4, CACHING>(o, dst, src); break;
    default: vmult_t<5, CACHING>(o, dst, src); break;
  }
}

void vmult(const Op &o, double *dst, const double *src) {
  if (o.caching == TENSOR2) vmult_c<TENSOR2>(o, dst, src);
  else vmult_c<TENSOR4>(o, dst, src);
}

}  // namespace cpb

extern "C" {

int cpb_lanes() { return cpb::W; }

// p: FE degree (n_q_1d = p + 1); tables: S[q][a], D[q][a] (N x N each),
// Gauss weights then points (2N); cell_colour: colouring with no node shared
// inside a colour
void *cpb_create(int p, int n_cells, int n_nodes, const int32_t *cell_nodes, const double *coords,
                 const uint8_t *cons, const int32_t *cell_colour, int caching, int threads, const double *S,
                 const double *D, const double *qw_qp) {
  using namespace cpb;
  if (p < 1 || p > 4 || (caching != TENSOR2 && caching != TENSOR4) || n_cells <= 0) return nullptr;
  Op *o = new Op;
  o->p = p;
  o->N = p + 1;
  o->NQP = o->N * o->N * o->N;
  o->caching = caching;
  o->nf = caching == TENSOR2 ? NF2 : NF4;
  o->n_cells = n_cells;
  o->n_nodes = n_nodes;
  o->threads = threads > 0 ? threads : 1;
  o->coords.assign(coords, coords + (size_t)n_nodes * 3);
  o->cons.assign(cons, cons + (size_t)n_nodes * 3);
  for (size_t d = 0; d < o->cons.size(); ++d)
    if (o->cons[d]) o->cons_list.push_back((int32_t)d);
  o->S.assign(S, S + o->N * o->N);
  o->D.assign(D, D + o->N * o->N);
  o->qw.assign(qw_qp, qw_qp + 2 * o->N);
  int n_col = 0;
  for (int c = 0; c < n_cells; ++c) n_col = std::max(n_col, cell_colour[c] + 1);
  std::vector<std::vector<int>> by(n_col);
  for (int c = 0; c < n_cells; ++c) by[cell_colour[c]].push_back(c);
  o->colour_ptr.push_back(0);
  for (auto &list : by) {
    for (size_t s = 0; s < list.size(); s += W) {
      for (int l = 0; l < W; ++l) o->cells.push_back(s + l < list.size() ? list[s + l] : -1);
      ++o->n_batches;
    }
    o->colour_ptr.push_back(o->n_batches);
  }
  const int NQ = o->NQP;
  o->nodes.assign((size_t)o->n_batches * NQ * W, -1);
  for (int b = 0; b < o->n_batches; ++b)
    for (int l = 0; l < W; ++l) {
      const int cell = o->cells[(size_t)b * W + l];
      if (cell < 0) continue;
      for (int q = 0; q < NQ; ++q) o->nodes[((size_t)b * NQ + q) * W + l] = cell_nodes[(size_t)cell * NQ + q];
    }
  return o;
}

void cpb_destroy(void *op) { delete static_cast<cpb::Op *>(op); }

void cpb_set_material(void *op, const double *mu, const double *lambda) {
  cpb::Op *o = static_cast<cpb::Op *>(op);
  o->mu.assign(mu, mu + o->n_cells);
  o->lambda.assign(lambda, lambda + o->n_cells);
}

void cpb_set_linearization(void *op, const double *u) {
  cpb::Op *o = static_cast<cpb::Op *>(op);
  o->u.assign(u, u + (size_t)o->n_nodes * 3);
}

void cpb_cache(void *op) { cpb::cache(*static_cast<cpb::Op *>(op)); }

void cpb_vmult(void *op, double *dst, const double *src) { cpb::vmult(*static_cast<cpb::Op *>(op), dst, src); }

}  // extern "C"
