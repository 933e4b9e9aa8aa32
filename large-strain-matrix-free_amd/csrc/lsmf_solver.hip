// lsmf_solver.hip -- GMG-preconditioned CG and the Newton driver (host C++
// over HIP kernels), C ABI part "lsmf_solver_*" of include/lsmf.h.
//
// Reference behaviour (include/mf_elasticity.h):
//   * Newton loop, convergence test (>= 3 iterations; relative residual and
//     update norms), Dirichlet increment at iteration 0 ......... :2245-2559
//   * make_constraints: inhomogeneous only at it 0, levels homogeneous :2932-3007
//   * residual incl. Neumann and the lifting of prescribed values .. :2708-2922
//   * setup_matrix_free: level operators linearised at the displacement
//     interpolated to each level, caches, diagonals, Chebyshev data,
//     coarse Chebyshev as a solver, V-cycle ........................ :1712-2241
//   * solve_linear_system: FP64 CG, tolerance relative to |rhs|,
//     constraints.distribute of the update ......................... :3012-3214
// deal.II algorithms restated: Multigrid::level_v_step (pre-smooth from zero,
// residual, restrict, coarse, prolongate-add, post-smooth), PreconditionChebyshev
// (vmult: degree-1 products from a zero start, step: degree products,
// interval [1.2 lambda_max / range, 1.2 lambda_max], degree for a target
// reduction when used as a solver), SolverCG's Lanczos eigenvalue estimate.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <type_traits>
#include <vector>

#include "lsmf.h"
#include "lsmf_error.hpp"
#include "lsmf_fe.hpp"
#include "lsmf_solver.hpp"
#include "lsmf_vec.hpp"

namespace lsmf {

namespace {
inline unsigned nblk(size_t n) { return (unsigned)((n + 255) / 256); }
using Clock = std::chrono::steady_clock;
inline double seconds_since(Clock::time_point t0) {
  return std::chrono::duration<double>(Clock::now() - t0).count();
}
}  // namespace

// ---------------------------------------------------------------- Reducer
Reducer::Reducer(hipStream_t s) : s_(s), part_(kRedBlocks * 2), res_(2) {}

double Reducer::read(int nv, double *out) {
  if (nv == 1)
    hipLaunchKernelGGL(k_reduce_final<1>, dim3(1), dim3(kRedThreads), 0, s_, part_.p, res_.p);
  else
    hipLaunchKernelGGL(k_reduce_final<2>, dim3(1), dim3(kRedThreads), 0, s_, part_.p, res_.p);
  hip_check(hipMemcpyAsync(out, res_.p, nv * sizeof(double), hipMemcpyDeviceToHost, s_), "reduce read");
  hip_check(hipStreamSynchronize(s_), "reduce sync");
  return out[0];
}

template <typename TA, typename TB> double Reducer::dot(const TA *a, const TB *b, size_t n) {
  hipLaunchKernelGGL((k_dot<TA, TB>), dim3(kRedBlocks), dim3(kRedThreads), 0, s_, a, b, n, part_.p);
  double r[2];
  return read(1, r);
}

template <typename T>
void Reducer::dot2(const T *a, const T *b, const T *c, const T *d, size_t n, double &r0, double &r1) {
  hipLaunchKernelGGL(k_dot2<T>, dim3(kRedBlocks), dim3(kRedThreads), 0, s_, a, b, c, d, n, part_.p);
  double r[2];
  read(2, r);
  r0 = r[0];
  r1 = r[1];
}

template <typename TA, typename TB> void Reducer::dot_to(const TA *a, const TB *b, size_t n, double *dst) {
  hipLaunchKernelGGL((k_dot<TA, TB>), dim3(kRedBlocks), dim3(kRedThreads), 0, s_, a, b, n, part_.p);
  hipLaunchKernelGGL(k_reduce_final<1>, dim3(1), dim3(kRedThreads), 0, s_, part_.p, dst);
}

template <typename T>
void Reducer::cg_update_dev(T *x, T *r, const T *p, const T *Ap, const double *st, size_t n, double *dst) {
  hipLaunchKernelGGL(k_cg_update_dev<T>, dim3(kRedBlocks), dim3(kRedThreads), 0, s_, x, r, p, Ap, st, n, part_.p);
  hipLaunchKernelGGL(k_reduce_final<1>, dim3(1), dim3(kRedThreads), 0, s_, part_.p, dst);
}

template <typename T> double Reducer::cg_update(T *x, T *r, const T *p, const T *Ap, double alpha, size_t n) {
  hipLaunchKernelGGL(k_cg_update<T>, dim3(kRedBlocks), dim3(kRedThreads), 0, s_, x, r, p, Ap, alpha, n, part_.p);
  double rr[2];
  return read(1, rr);
}

// ---------------------------------------------------------------- transfer
// banded 1D operator: out[o] = sum_k w[o*W + k] in[start[o] + k]
struct Band {
  std::vector<int32_t> start;
  std::vector<double> w;
  int W = 0, n_out = 0, n_in = 0;
};

static double node_pos(const std::vector<double> &gl, int p, int n_cells, int i) {
  const int k = std::min(i / p, n_cells - 1);
  return k + gl[i - k * p];
}

// Q_p function on n_c coarse cells evaluated at the support points of 2 n_c
// fine cells (MGTransferMatrixFree::prolongate, FE embedding)
static Band prolongation_band(int p, int n_c) {
  const std::vector<double> gl = lobatto_points(p);
  Band b;
  const int n_f = 2 * n_c;
  b.n_out = n_f * p + 1;
  b.n_in = n_c * p + 1;
  b.W = p + 1;
  b.start.resize(b.n_out);
  b.w.resize((size_t)b.n_out * b.W);
  for (int i = 0; i < b.n_out; ++i) {
    const double x = 0.5 * node_pos(gl, p, n_f, i);  // coarse cell units
    const int kc = std::min((int)std::floor(x), n_c - 1);
    const double xi = x - kc;
    b.start[i] = kc * p;
    for (int a = 0; a <= p; ++a) b.w[(size_t)i * b.W + a] = lagrange_value(gl, a, xi);
  }
  return b;
}

// fine function evaluated at the coarse support points (interpolate_to_mg)
static Band interpolation_band(int p, int n_c) {
  const std::vector<double> gl = lobatto_points(p);
  Band b;
  const int n_f = 2 * n_c;
  b.n_out = n_c * p + 1;
  b.n_in = n_f * p + 1;
  b.W = p + 1;
  b.start.resize(b.n_out);
  b.w.resize((size_t)b.n_out * b.W);
  for (int I = 0; I < b.n_out; ++I) {
    const double x = 2.0 * node_pos(gl, p, n_c, I);  // fine cell units
    const int kf = std::min((int)std::floor(x + 1e-14), n_f - 1);
    const double xi = x - kf;
    b.start[I] = kf * p;
    for (int a = 0; a <= p; ++a) b.w[(size_t)I * b.W + a] = lagrange_value(gl, a, xi);
  }
  return b;
}

static Band transpose_band(const Band &P) {
  Band t;
  t.n_out = P.n_in;
  t.n_in = P.n_out;
  std::vector<int> lo(t.n_out, 1 << 30), hi(t.n_out, -1);
  for (int i = 0; i < P.n_out; ++i)
    for (int k = 0; k < P.W; ++k) {
      const int I = P.start[i] + k;
      lo[I] = std::min(lo[I], i);
      hi[I] = std::max(hi[I], i);
    }
  for (int I = 0; I < t.n_out; ++I) t.W = std::max(t.W, hi[I] - lo[I] + 1);
  t.start.assign(t.n_out, 0);
  t.w.assign((size_t)t.n_out * t.W, 0.0);
  for (int I = 0; I < t.n_out; ++I) {
    t.start[I] = std::min(lo[I], t.n_in - t.W);
    t.start[I] = std::max(t.start[I], 0);
  }
  for (int i = 0; i < P.n_out; ++i)
    for (int k = 0; k < P.W; ++k) {
      const int I = P.start[i] + k;
      t.w[(size_t)I * t.W + (i - t.start[I])] += P.w[(size_t)i * P.W + k];
    }
  return t;
}

// a band on the device: rows trimmed of their zero end taps (start moved
// past leading zeros, len up to the last nonzero), weights padded to the
// widest row W, in FP64 and FP32 (the FP32 levels transfer in FP32)
struct DevBand {
  DeviceArray<int32_t> start, len;
  DeviceArray<double> w;
  DeviceArray<float> wf;
  int W = 0, n_out = 0, n_in = 0;
  void set(const Band &b) {
    n_out = b.n_out;
    n_in = b.n_in;
    std::vector<int32_t> st(n_out), ln(n_out);
    for (int o = 0; o < n_out; ++o) {
      const double *r = &b.w[(size_t)o * b.W];
      int lo = 0, hi = b.W;
      while (lo < hi && r[lo] == 0.0) ++lo;
      while (hi > lo && r[hi - 1] == 0.0) --hi;
      st[o] = b.start[o] + lo;
      ln[o] = hi - lo;
    }
    W = 1;
    for (int o = 0; o < n_out; ++o) W = std::max(W, ln[o]);
    std::vector<double> wd((size_t)n_out * W, 0.0);
    for (int o = 0; o < n_out; ++o) {
      const int lo = st[o] - b.start[o];
      for (int k = 0; k < ln[o]; ++k) wd[(size_t)o * W + k] = b.w[(size_t)o * b.W + lo + k];
      // padded taps read valid input (weight 0) when the row ends early
      if (st[o] + W > n_in) {
        const int shift = st[o] + W - n_in;
        for (int k = ln[o] - 1; k >= 0; --k) wd[(size_t)o * W + k + shift] = wd[(size_t)o * W + k];
        for (int k = 0; k < shift; ++k) wd[(size_t)o * W + k] = 0.0;
        st[o] -= shift;
        ln[o] += shift;
      }
    }
    std::vector<float> wfh(wd.begin(), wd.end());
    start.resize(n_out);
    start.upload(st.data(), st.size());
    len.resize(n_out);
    len.upload(ln.data(), ln.size());
    w.resize(wd.size());
    w.upload(wd.data(), wd.size());
    wf.resize(wfh.size());
    wf.upload(wfh.data(), wfh.size());
  }
  template <typename T> const T *weights() const {
    if constexpr (std::is_same<T, double>::value) return w.p;
    else return wf.p;
  }
};

template <typename T, int C>
static void band_launch(const DevBand &b, int d, const int *dims, const T *in, T *out, bool add, hipStream_t s) {
  int od[3] = {dims[0], dims[1], dims[2]};
  od[d] = b.n_out;
  const T *wt = b.weights<T>();
  const int a = add ? 1 : 0;
  constexpr int R = 4;  // outputs per lane along y / z
  const size_t line_bytes = (size_t)dims[0] * C * sizeof(T);
  if (d == 0 && line_bytes <= 65536 && (size_t)dims[1] * dims[2] <= (size_t)INT32_MAX) {
    hipLaunchKernelGGL((k_band_line<T, C>), dim3((unsigned)((size_t)dims[1] * dims[2])), dim3(256), line_bytes, s, in,
                       out, b.start.p, wt, b.W, dims[0], b.n_out, a);
  } else if (d == 0) {
    if (od[2] > 65535 || (size_t)od[0] * od[1] > (size_t)INT32_MAX)
      throw HipError(LSMF_ERR_UNSUPPORTED, "transfer: level grid too large");
    const dim3 grid((unsigned)(((size_t)od[0] * od[1] + 255) / 256), (unsigned)od[2]);
    hipLaunchKernelGGL((k_band<T, 0, C>), grid, dim3(256), 0, s, in, out, b.start.p, b.len.p, wt, b.W, dims[0], dims[1],
                       b.n_out, a);
  } else {
    const size_t other = (size_t)dims[0] * (d == 1 ? dims[2] : dims[1]);
    const int oblocks = (b.n_out + R - 1) / R;
    if (oblocks > 65535 || other > (size_t)INT32_MAX)
      throw HipError(LSMF_ERR_UNSUPPORTED, "transfer: level grid too large");
    const dim3 grid((unsigned)((other + 255) / 256), (unsigned)oblocks);
    if (d == 1)
      hipLaunchKernelGGL((k_band_rows<T, 1, C, R>), grid, dim3(256), 0, s, in, out, b.start.p, b.len.p, wt, b.W,
                         dims[0], dims[1], dims[2], b.n_out, a);
    else
      hipLaunchKernelGGL((k_band_rows<T, 2, C, R>), grid, dim3(256), 0, s, in, out, b.start.p, b.len.p, wt, b.W,
                         dims[0], dims[1], dims[2], b.n_out, a);
  }
}

// apply one band per direction (tensor product), in has node dims nin[0..dim)
template <typename T>
static void tensor_apply(const DevBand *const *bands, int dim, const int *nin, const T *in, T *out, T *tmpA, T *tmpB,
                         bool add, hipStream_t s) {
  int dims[3] = {nin[0], nin[1], dim == 3 ? nin[2] : 1};
  const T *cur = in;
  for (int d = 0; d < dim; ++d) {
    T *dst = (d == dim - 1) ? out : ((d % 2 == 0) ? tmpA : tmpB);
    const DevBand &b = *bands[d];
    const bool a = add && d == dim - 1;
    if (dim == 3) band_launch<T, 3>(b, d, dims, cur, dst, a, s);
    else band_launch<T, 2>(b, d, dims, cur, dst, a, s);
    dims[d] = b.n_out;
    cur = dst;
  }
}

// ---------------------------------------------------------------- levels
static double sturm_count_below(const std::vector<double> &a, const std::vector<double> &b, double x) {
  // number of eigenvalues of the symmetric tridiagonal (a diag, b offdiag) < x
  int cnt = 0;
  double q = a[0] - x;
  if (q < 0) ++cnt;
  for (size_t i = 1; i < a.size(); ++i) {
    const double qq = (q != 0.0) ? q : 1e-300;
    q = a[i] - x - b[i - 1] * b[i - 1] / qq;
    if (q < 0) ++cnt;
  }
  return cnt;
}

static void tridiag_extremes(const std::vector<double> &a, const std::vector<double> &b, double &lmin, double &lmax) {
  const size_t n = a.size();
  double lo = 1e300, hi = -1e300;
  for (size_t i = 0; i < n; ++i) {
    const double r = (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < n ? std::fabs(b[i]) : 0.0);
    lo = std::min(lo, a[i] - r);
    hi = std::max(hi, a[i] + r);
  }
  auto kth = [&](int k) {  // k-th smallest eigenvalue (0-based)
    double l = lo, h = hi;
    for (int it = 0; it < 200; ++it) {
      const double m = 0.5 * (l + h);
      if (sturm_count_below(a, b, m) > k) h = m;
      else l = m;
      if (h - l <= 1e-14 * std::max(std::fabs(l), std::fabs(h))) break;
    }
    return 0.5 * (l + h);
  };
  lmin = kth(0);
  lmax = kth((int)n - 1);
}

struct GMGBase {
  virtual ~GMGBase() {}
  virtual void setup(const double *u_fine, double *tmp64a, double *tmp64b) = 0;
  // wall time of the last setup's caches and diagonals (all levels)
  virtual double setup_cache_seconds() const = 0;
  virtual void precondition(double *z, const double *r) = 0;
  virtual void info(int level, double &lmax, double &lmin, int &degree) const = 0;
  virtual int n_levels() const = 0;
  virtual size_t level_dofs(int l) const = 0;
  virtual void transfer64(int level, int mode, double *dst, const double *src) = 0;
};

struct LevelGeom {
  lsmf_mesh *mesh = nullptr;
  int nn[3] = {1, 1, 1};
  size_t n_dofs = 0;
  std::vector<uint8_t> cons;
  std::vector<int32_t> cons_list_h;
  DeviceArray<int32_t> cons_list;
  std::vector<double> mu, lam;
  ~LevelGeom() {
    if (mesh) lsmf_mesh_destroy(mesh);
  }
};

template <typename T> struct Level {
  LevelGeom *g = nullptr;
  std::unique_ptr<NeoHookOperatorHIP<T>> own;
  NeoHookOperatorHIP<T> *op = nullptr;
  DeviceArray<T> sol, rhs, t, d, z;
  DeviceArray<T> cg_r, cg_p, cg_q;  // coarse CG (LSMF_COARSE_CG, level 0 only)
  DeviceArray<double> u64;
  DeviceArray<T> tmpA, tmpB;  // transfer temporaries (from level-1 to this level)
  DevBand P[3], PT[3], R[3];
  double lmax_est = 0, lmin_est = 0, theta = 0, delta = 0;
  int degree = 0;
};

template <typename T> class GMG : public GMGBase {
 public:
  GMG(const lsmf_solver_options &o, std::vector<std::unique_ptr<LevelGeom>> &geoms, NeoHookOperatorHIP<double> *fine,
      hipStream_t s)
      : opt_(o), s_(s), red_(s) {
    const int L = (int)geoms.size();
    lv_.resize(L);
    for (int l = 0; l < L; ++l) {
      Level<T> &v = lv_[l];
      v.g = geoms[l].get();
      const size_t n = v.g->n_dofs;
      if constexpr (std::is_same<T, double>::value) {
        if (l == L - 1 && fine && o.mg_caching == o.caching) v.op = fine;
      }
      if (!v.op) {
        v.own.reset(new NeoHookOperatorHIP<T>());
        lsmf_mesh_desc md{o.dim, o.degree, v.g->mesh->n_cells, v.g->mesh->n_nodes, v.g->mesh->cell_nodes,
                          v.g->mesh->coords, v.g->cons.data(), v.g->mesh->cell_colour};
        v.own->initialize(md, o.n_q_1d, o.mg_caching, s);
        v.own->set_formulation(o.formulation);
        v.own->set_material(v.g->mu.data(), v.g->lam.data());
        v.op = v.own.get();
      }
      v.sol.resize(n);
      v.rhs.resize(n);
      v.t.resize(n);
      v.d.resize(n);
      v.z.resize(n);
      if (l == 0 && o.coarse_solver == LSMF_COARSE_CG) {
        v.cg_r.resize(n);
        v.cg_p.resize(n);
        v.cg_q.resize(n);
      }
      if (l < L - 1) v.u64.resize(n);
      if (l > 0) {
        const LevelGeom &c = *geoms[l - 1];
        size_t tmax = 0;
        for (int d = 0; d < o.dim; ++d) {
          const int nc = c.mesh->n_cells_dir[d];
          Band P = prolongation_band(o.degree, nc);
          v.P[d].set(P);
          v.PT[d].set(transpose_band(P));
          v.R[d].set(interpolation_band(o.degree, nc));
        }
        // largest intermediate of a prolongation / restriction sequence
        int dims[3] = {c.nn[0], c.nn[1], o.dim == 3 ? c.nn[2] : 1};
        for (int d = 0; d < o.dim; ++d) {
          dims[d] = v.g->nn[d];
          tmax = std::max(tmax, (size_t)dims[0] * dims[1] * dims[2] * o.dim);
        }
        int dims2[3] = {v.g->nn[0], v.g->nn[1], o.dim == 3 ? v.g->nn[2] : 1};
        for (int d = 0; d < o.dim; ++d) {
          dims2[d] = c.nn[d];
          tmax = std::max(tmax, (size_t)dims2[0] * dims2[1] * dims2[2] * o.dim);
        }
        v.tmpA.resize(tmax);
        v.tmpB.resize(tmax);
        tmax64_ = std::max(tmax64_, tmax);
      }
    }
    if (o.coarse_solver == LSMF_COARSE_DIRECT) {
      const LevelGeom &g0 = *geoms[0];
      if (g0.n_dofs > 4096)
        throw HipError(LSMF_ERR_INVALID_ARGUMENT, "direct coarse solve needs a coarsest level with <= 4096 DoFs");
      if constexpr (std::is_same<T, double>::value) {
        coarse64_ = lv_[0].op;
      } else {
        coarse64_own_.reset(new NeoHookOperatorHIP<double>());
        lsmf_mesh_desc md{o.dim, o.degree, g0.mesh->n_cells, g0.mesh->n_nodes, g0.mesh->cell_nodes, g0.mesh->coords,
                          g0.cons.data(), g0.mesh->cell_colour};
        coarse64_own_->initialize(md, o.n_q_1d, o.mg_caching, s);
        coarse64_own_->set_formulation(o.formulation);
        coarse64_own_->set_material(g0.mu.data(), g0.lam.data());
        coarse64_ = coarse64_own_.get();
      }
      ainv_.resize(g0.n_dofs * g0.n_dofs);
    }
  }
  ~GMG() override { drop_graph(); }

  double setup_cache_seconds() const override { return cache_s_; }
  int n_levels() const override { return (int)lv_.size(); }
  size_t level_dofs(int l) const override { return lv_[l].g->n_dofs; }
  void info(int l, double &lmax, double &lmin, int &degree) const override {
    lmax = lv_[l].lmax_est;
    lmin = lv_[l].lmin_est;
    degree = lv_[l].degree;
  }
  size_t tmax64() const { return tmax64_; }

  void prolongate_add(int l, T *fine, const T *coarse) {
    Level<T> &v = lv_[l];
    const DevBand *b[3] = {&v.P[0], &v.P[1], &v.P[2]};
    tensor_apply<T>(b, opt_.dim, lv_[l - 1].g->nn, coarse, fine, v.tmpA.p, v.tmpB.p, true, s_);
  }
  void restrict_(int l, T *coarse, const T *fine) {
    Level<T> &v = lv_[l];
    const DevBand *b[3] = {&v.PT[0], &v.PT[1], &v.PT[2]};
    tensor_apply<T>(b, opt_.dim, v.g->nn, fine, coarse, v.tmpA.p, v.tmpB.p, false, s_);
    zero_constrained(l - 1, coarse);
  }
  void zero_constrained(int l, T *x) {
    const LevelGeom &g = *lv_[l].g;
    if (!g.cons_list_h.empty())
      hipLaunchKernelGGL(k_set_list<T>, dim3(nblk(g.cons_list_h.size())), dim3(256), 0, s_, x, g.cons_list.p,
                         (const double *)nullptr, 0.0, (int)g.cons_list_h.size());
  }

  void transfer64(int l, int mode, double *dst, const double *src) override {
    Level<T> &v = lv_[l];
    DeviceArray<double> a(tmax64_), b(tmax64_);
    if (mode == 0) {
      const DevBand *bb[3] = {&v.P[0], &v.P[1], &v.P[2]};
      tensor_apply<double>(bb, opt_.dim, lv_[l - 1].g->nn, src, dst, a.p, b.p, false, s_);
    } else {
      const DevBand *bb[3] = {mode == 1 ? &v.PT[0] : &v.R[0], mode == 1 ? &v.PT[1] : &v.R[1],
                              mode == 1 ? &v.PT[2] : &v.R[2]};
      tensor_apply<double>(bb, opt_.dim, v.g->nn, src, dst, a.p, b.p, false, s_);
    }
    hip_check(hipStreamSynchronize(s_), "transfer");
  }

  // linearise every level at the fine displacement: interpolate_to_mg
  // (include/mf_elasticity.h:1926-1942), level caches and diagonals
  // (:2055-2073), Chebyshev data (:2081-2113)
  void setup(const double *u_fine, double *ta, double *tb) override {
    drop_graph();
    const int L = (int)lv_.size();
    // caches and diagonals of every level first (timed on their own), then
    // the Chebyshev data and the coarse solve
    const auto t0 = Clock::now();
    const double *ucur = u_fine;
    for (int l = L - 1; l >= 0; --l) {
      Level<T> &v = lv_[l];
      if (l < L - 1) {
        Level<T> &f = lv_[l + 1];
        const DevBand *bb[3] = {&f.R[0], &f.R[1], &f.R[2]};
        tensor_apply<double>(bb, opt_.dim, f.g->nn, ucur, v.u64.p, ta, tb, false, s_);
        ucur = v.u64.p;
      }
      if (v.own) {
        v.op->set_linearization(ucur);
        v.op->cache();
      }
      v.op->compute_diagonal();
    }
    hip_check(hipStreamSynchronize(s_), "level caches");
    cache_s_ = seconds_since(t0);
    for (int l = L - 1; l >= 0; --l) {
      Level<T> &v = lv_[l];
      if (l == 0 && opt_.coarse_solver == LSMF_COARSE_DIRECT) {
        assemble_coarse_inverse(l < L - 1 ? lv_[0].u64.p : u_fine);
      } else if (l == 0 && opt_.coarse_solver == LSMF_COARSE_CHEBYSHEV) {
        const int its = opt_.coarse_eig_iterations > 0 ? opt_.coarse_eig_iterations : (int)v.g->n_dofs;
        estimate_eigenvalues(v, its);
        const double lmax = 1.2 * v.lmax_est;
        const double alpha = std::min(0.9 * lmax, v.lmin_est);
        v.theta = 0.5 * (lmax + alpha);
        v.delta = 0.5 * (lmax - alpha);
        const double range = lmax / alpha;
        const double sigma = (1.0 - std::sqrt(1.0 / range)) / (1.0 + std::sqrt(1.0 / range));
        const double eps = opt_.coarse_reduction;
        v.degree = 1 + (int)(std::log(1.0 / eps + std::sqrt(1.0 / eps / eps - 1.0)) / std::log(1.0 / sigma));
        v.degree = std::min(std::max(v.degree, 1), 2000);
      } else {
        estimate_eigenvalues(v, opt_.eig_cg_iterations);
        const double lmax = 1.2 * v.lmax_est;
        const double alpha = lmax / opt_.smoothing_range;
        v.theta = 0.5 * (lmax + alpha);
        v.delta = 0.5 * (lmax - alpha);
        v.degree = opt_.smoother_degree;
      }
    }
  }

  // z = M r, the V-cycle preconditioner between FP64 vectors of the finest
  // level (copy_to_mg / copy_from_mg convert FP64 <-> level precision)
  void precondition(double *z, const double *r) override {
    if (!opt_.use_graph || opt_.coarse_solver == LSMF_COARSE_CG) {
      precondition_eager(z, r);
      return;
    }
    // the V-cycle is a fixed sequence of launches between two setups:
    // capture it once and replay it (launch overhead of the small levels)
    if (!graph_exec_ || z != graph_z_ || r != graph_r_) {
      drop_graph();
      hipGraph_t g = nullptr;
      hip_check(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal), "begin capture");
      precondition_eager(z, r);
      hip_check(hipStreamEndCapture(s_, &g), "end capture");
      hip_check(hipGraphInstantiate(&graph_exec_, g, nullptr, nullptr, 0), "graph instantiate");
      (void)hipGraphDestroy(g);
      graph_z_ = z;
      graph_r_ = r;
    }
    hip_check(hipGraphLaunch(graph_exec_, s_), "graph launch");
  }

 private:
  void precondition_eager(double *z, const double *r) {
    const int L = (int)lv_.size() - 1;
    Level<T> &f = lv_[L];
    const size_t n = f.g->n_dofs;
    hipLaunchKernelGGL((k_convert<double, T>), dim3(nblk(n)), dim3(256), 0, s_, r, f.rhs.p, n);
    vcycle(L);
    hipLaunchKernelGGL((k_convert<T, double>), dim3(nblk(n)), dim3(256), 0, s_, f.sol.p, z, n);
  }

  void drop_graph() {
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    graph_exec_ = nullptr;
    graph_z_ = nullptr;
    graph_r_ = nullptr;
  }

  // FP64 coarse matrix by probing with the n unit vectors in one batched
  // product (column j = A e_j, stored as row j: A is symmetric),
  // symmetrised and inverted in place on the device (blocked Gauss-Jordan)
  void assemble_coarse_inverse(const double *u0) {
    Level<T> &v = lv_[0];
    const int n = (int)v.g->n_dofs;
    if (coarse64_own_) {
      coarse64_->set_linearization(u0);
      coarse64_->cache();
    }
    const size_t nn = (size_t)n * n;
    DeviceArray<double> e(nn), W((size_t)kGJB * kGJB), Cold((size_t)n * kGJB), Cnew((size_t)n * kGJB),
        R((size_t)kGJB * n);
    DeviceArray<int> bad(1);
    double *A = ainv_.p;
    hipLaunchKernelGGL(k_identity, dim3(nblk(nn)), dim3(256), 0, s_, e.p, n);
    coarse64_->vmult_batched(A, e.p, n);
    hipLaunchKernelGGL(k_sym_average, dim3(nblk(nn)), dim3(256), 0, s_, A, n);
    hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(1), 0, s_, bad.p, 0, (size_t)1);
    const dim3 tiles((unsigned)((n + 63) / 64), (unsigned)((n + 15) / 16));
    for (int K0 = 0; K0 < n; K0 += kGJB) {
      const int bk = std::min(kGJB, n - K0);
      hipLaunchKernelGGL(k_bgj_diag, dim3(1), dim3(256), 0, s_, (const double *)A, n, K0, bk, W.p, bad.p);
      hipLaunchKernelGGL(k_bgj_panels, dim3(nblk(2 * (size_t)n * kGJB)), dim3(256), 0, s_, (const double *)A, n, K0,
                         bk, (const double *)W.p, Cold.p, Cnew.p, R.p);
      hipLaunchKernelGGL(k_bgj_update, tiles, dim3(256), 0, s_, A, n, K0, bk, (const double *)W.p,
                         (const double *)Cold.p, (const double *)Cnew.p, (const double *)R.p);
    }
    int bad_h = 0;
    hip_check(hipMemcpyAsync(&bad_h, bad.p, sizeof(int), hipMemcpyDeviceToHost, s_), "coarse inverse");
    hip_check(hipStreamSynchronize(s_), "coarse inverse");
    if (bad_h) throw HipError(LSMF_ERR_NOT_CONVERGED, "coarse matrix is not positive definite");
    v.degree = 0;
    v.lmax_est = v.lmin_est = 0.0;
  }


  // Chebyshev smoother / coarse solver on level v: x from zero (vmult) or
  // from its current value (step)
  void chebyshev(Level<T> &v, T *x, const T *b, bool x_zero) {
    const size_t n = v.g->n_dofs;
    const T *invd = v.op->get_matrix_diagonal_inverse();
    if (x_zero) {
      launch_cheb_first<T>(s_, x, v.d.p, b, (const T *)nullptr, invd, 1.0 / v.theta, 1, n);
    } else {
      v.op->vmult(v.t.p, x);
      launch_cheb_first<T>(s_, x, v.d.p, b, (const T *)v.t.p, invd, 1.0 / v.theta, 0, n);
    }
    if (v.degree < 2 || std::fabs(v.delta) < 1e-40) return;
    double rhok = v.delta / v.theta;
    const double sigma = v.theta / v.delta;
    for (int k = 0; k < v.degree - 1; ++k) {
      v.op->vmult(v.t.p, x);
      const double rhokp = 1.0 / (2.0 * sigma - rhok);
      const double f1 = rhokp * rhok, f2 = 2.0 * rhokp / v.delta;
      rhok = rhokp;
      launch_cheb_update<T>(s_, x, v.d.p, b, (const T *)v.t.p, invd, f1, f2, n);
    }
  }

  // Multigrid::level_v_step
  void vcycle(int l) {
    Level<T> &v = lv_[l];
    const size_t n = v.g->n_dofs;
    if (l == 0) {
      if (opt_.coarse_solver == LSMF_COARSE_DIRECT)
        hipLaunchKernelGGL(k_dense_matvec<T>, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s_,
                           (const double *)ainv_.p, (const T *)v.rhs.p, v.sol.p, (int)n);
      else if (opt_.coarse_solver == LSMF_COARSE_CG)
        coarse_cg(v);
      else
        chebyshev(v, v.sol.p, v.rhs.p, true);
      return;
    }
    chebyshev(v, v.sol.p, v.rhs.p, true);  // pre-smoothing from zero
    v.op->vmult(v.t.p, v.sol.p);
    launch_residual<T>(s_, v.t.p, (const T *)v.rhs.p, n);
    Level<T> &c = lv_[l - 1];
    restrict_(l, c.rhs.p, v.t.p);
    vcycle(l - 1);
    prolongate_add(l, v.sol.p, c.sol.p);
    zero_constrained(l, v.sol.p);
    chebyshev(v, v.sol.p, v.rhs.p, false);  // post-smoothing
  }

  // MGCoarseGridIterativeSolver (include/mf_elasticity.h:2115-2127): x = 0,
  // then SolverCG preconditioned by the level's Chebyshev smoother (applied
  // from zero) until |r| <= max(1e-10, 1e-3 |b|), at most max(n, 100) steps
  void coarse_cg(Level<T> &v) {
    const size_t n = v.g->n_dofs;
    T *x = v.sol.p, *r = v.cg_r.p, *p = v.cg_p.p, *q = v.cg_q.p, *z = v.z.p;
    const T *b = v.rhs.p;
    hipLaunchKernelGGL(k_fill<T>, dim3(nblk(n)), dim3(256), 0, s_, x, T(0), n);
    hip_check(hipMemcpyAsync(r, b, n * sizeof(T), hipMemcpyDeviceToDevice, s_), "coarse cg");
    const double r0 = std::sqrt(red_.dot(r, r, n));
    const double tol = std::max(1e-10, 1e-3 * r0);
    if (r0 <= tol) return;
    const int max_it = std::max<int>((int)n, 100);
    chebyshev(v, z, r, true);
    hip_check(hipMemcpyAsync(p, z, n * sizeof(T), hipMemcpyDeviceToDevice, s_), "coarse cg");
    double rz = red_.dot(r, z, n);
    for (int it = 0; it < max_it; ++it) {
      v.op->vmult(q, p);
      const double pq = red_.dot(p, q, n);
      if (!(pq > 0)) throw HipError(LSMF_ERR_NOT_CONVERGED, "coarse CG: operator not positive definite");
      const double a = rz / pq;
      const double rr = red_.cg_update(x, r, p, q, a, n);
      if (std::sqrt(rr) <= tol) return;
      chebyshev(v, z, r, true);
      const double rz_new = red_.dot(r, z, n);
      const double beta = rz_new / rz;
      rz = rz_new;
      launch_xpby<T>(s_, p, (const T *)z, beta, n);
    }
    throw HipError(LSMF_ERR_NOT_CONVERGED, "coarse CG: no convergence");
  }

  // CG on D^-1 A with the Lanczos tridiagonal of its coefficients
  // (SolverCG's condition-number / eigenvalue estimate)
  void estimate_eigenvalues(Level<T> &v, int its) {
    const size_t n = v.g->n_dofs;
    T *x = v.sol.p, *r = v.rhs.p, *p = v.d.p, *Ap = v.t.p, *z = v.z.p;
    const T *invd = v.op->get_matrix_diagonal_inverse();
    if (eig_ab_.n < (size_t)2 * its) eig_ab_.resize((size_t)2 * its);
    double *st = eig_st_.p, *al = eig_ab_.p, *be = eig_ab_.p + its;
    hipLaunchKernelGGL(k_fill<T>, dim3(nblk(n)), dim3(256), 0, s_, x, T(0), n);
    hipLaunchKernelGGL(k_pseudo_random<T>, dim3(nblk(n)), dim3(256), 0, s_, r, n, (uint64_t)20261202ull);
    zero_constrained((int)(&v - lv_.data()), r);
    launch_jacobi<T>(s_, z, (const T *)r, invd, n);
    hip_check(hipMemcpyAsync(p, z, n * sizeof(T), hipMemcpyDeviceToDevice, s_), "eig copy");
    red_.dot_to(r, z, n, st + kEigRz);
    red_.dot_to(r, r, n, st + kEigR0);
    hipLaunchKernelGGL(k_eig_init, dim3(1), dim3(1), 0, s_, st);
    // the CG steps with their scalars on the device; the recording stops at
    // a non-positive p.Ap or once the residual is below 1e-10 r0, and the
    // remaining steps' updates are then no-ops
    for (int k = 0; k < its; ++k) {
      v.op->vmult(Ap, p);
      red_.dot_to(p, Ap, n, st + kEigPAp);
      hipLaunchKernelGGL(k_eig_alpha, dim3(1), dim3(1), 0, s_, st, al, k);
      red_.cg_update_dev(x, r, p, Ap, st, n, st + kEigRr);
      hipLaunchKernelGGL(k_eig_check, dim3(1), dim3(1), 0, s_, st);
      launch_jacobi<T>(s_, z, (const T *)r, invd, n);
      red_.dot_to(r, z, n, st + kEigRzNew);
      hipLaunchKernelGGL(k_eig_beta, dim3(1), dim3(1), 0, s_, st, be, k);
      hipLaunchKernelGGL(k_xpby_dev<T>, dim3(nblk(n)), dim3(256), 0, s_, p, (const T *)z, (const double *)st, n);
    }
    std::vector<double> hst(kEigSlots), hab((size_t)2 * its);
    hip_check(hipMemcpyAsync(hst.data(), st, kEigSlots * sizeof(double), hipMemcpyDeviceToHost, s_), "eig read");
    hip_check(hipMemcpyAsync(hab.data(), eig_ab_.p, (size_t)2 * its * sizeof(double), hipMemcpyDeviceToHost, s_),
              "eig read");
    hip_check(hipStreamSynchronize(s_), "eig sync");
    const int mm = (int)hst[kEigM];
    std::vector<double> alpha(hab.begin(), hab.begin() + mm), beta(hab.begin() + its, hab.begin() + its + mm);
    const size_t m = alpha.size();
    if (m == 0) throw HipError(LSMF_ERR_HIP, "eigenvalue estimate: no CG step");
    std::vector<double> diag(m), off(m > 1 ? m - 1 : 0);
    diag[0] = 1.0 / alpha[0];
    for (size_t k = 1; k < m; ++k) {
      diag[k] = 1.0 / alpha[k] + beta[k - 1] / alpha[k - 1];
      off[k - 1] = std::sqrt(beta[k - 1]) / alpha[k - 1];
    }
    if (m == 1) {
      v.lmax_est = v.lmin_est = diag[0];
    } else {
      tridiag_extremes(diag, off, v.lmin_est, v.lmax_est);
    }
  }

  lsmf_solver_options opt_;
  hipStream_t s_;
  Reducer red_;
  DeviceArray<double> eig_st_{kEigSlots}, eig_ab_;  // device-resident eigenvalue-estimate CG
  std::vector<Level<T>> lv_;
  size_t tmax64_ = 0;
  double cache_s_ = 0.0;
  NeoHookOperatorHIP<double> *coarse64_ = nullptr;
  std::unique_ptr<NeoHookOperatorHIP<double>> coarse64_own_;
  DeviceArray<double> ainv_;
  hipGraphExec_t graph_exec_ = nullptr;
  double *graph_z_ = nullptr;
  const double *graph_r_ = nullptr;
};

}  // namespace lsmf

using namespace lsmf;

struct lsmf_solver {
  lsmf_solver_options opt{};
  hipStream_t s = nullptr;
  std::vector<std::unique_ptr<LevelGeom>> geoms;  // coarsest .. finest
  NeoHookOperatorHIP<double> fine;
  std::unique_ptr<GMGBase> mg;
  std::unique_ptr<Reducer> red;
  size_t n = 0;
  DeviceArray<double> u_n, delta, u_tot, rhs, du, g, Ag, neumann, r, z, p, Ap, ta, tb;
  std::vector<int32_t> cons_list;
  std::vector<double> cons_inc;  // increment per constrained DoF
  DeviceArray<int32_t> d_cons_list;
  DeviceArray<double> d_cons_inc;
  bool has_increment = false, has_neumann = false;
  bool neumann_absolute = false;  // traction set for the step (not scaled by t)
  double t_prev = 0.0;  // time of the last completed load step
  int n_steps_done = 0;
  ~lsmf_solver() {
    mg.reset();
    fine.clear();
    if (s) (void)hipStreamDestroy(s);
  }
};

namespace {

void build_level_geom(const lsmf_solver_options &o, int level, LevelGeom &g) {
  int nc[3] = {1, 1, 1};
  for (int d = 0; d < o.dim; ++d) nc[d] = o.n_cells_coarse[d] << level;
  abi_check(lsmf_mesh_create(o.dim, o.degree, nc, o.lo, o.hi, o.geometry, o.scale, &g.mesh), "mesh");
  for (int d = 0; d < o.dim; ++d) g.nn[d] = nc[d] * o.degree + 1;
  g.n_dofs = (size_t)g.mesh->n_nodes * o.dim;
  g.cons.assign(g.n_dofs, 0);
  for (int s = 0; s < 2 * o.dim; ++s) {
    if (!o.dirichlet_components[s]) continue;
    const int cnt = lsmf_mesh_boundary_nodes(g.mesh, s, nullptr);
    std::vector<int32_t> nodes(cnt);
    lsmf_mesh_boundary_nodes(g.mesh, s, nodes.data());
    for (int nd : nodes)
      for (int c = 0; c < o.dim; ++c)
        if ((o.dirichlet_components[s] >> c) & 1u) g.cons[(size_t)nd * o.dim + c] = 1;
  }
  for (size_t i = 0; i < g.n_dofs; ++i)
    if (g.cons[i]) g.cons_list_h.push_back((int32_t)i);
  g.cons_list.resize(g.cons_list_h.size());
  if (!g.cons_list_h.empty()) g.cons_list.upload(g.cons_list_h.data(), g.cons_list_h.size());
}

// cell (i,j,k) of level l-1 averages its 2^dim children of level l
void coarsen_material(int dim, const int *nc_fine, const std::vector<double> &f, std::vector<double> &c) {
  const int n0 = nc_fine[0], n1 = nc_fine[1], n2 = dim == 3 ? nc_fine[2] : 1;
  const int m0 = n0 / 2, m1 = n1 / 2, m2 = dim == 3 ? n2 / 2 : 1;
  c.assign((size_t)m0 * m1 * m2, 0.0);
  const double w = 1.0 / (1 << dim);
  for (int k = 0; k < n2; ++k)
    for (int j = 0; j < n1; ++j)
      for (int i = 0; i < n0; ++i)
        c[(size_t)(i / 2) + (size_t)m0 * ((j / 2) + (size_t)m1 * (dim == 3 ? k / 2 : 0))] +=
            w * f[(size_t)i + (size_t)n0 * (j + (size_t)n1 * k)];
}

// Neumann load for t = 1 on the finest mesh (reference configuration)
std::vector<double> neumann_vector(const lsmf_solver_options &o, const double (&traction)[6][3], unsigned sides,
                                   const LevelGeom &g) {
  std::vector<double> f(g.n_dofs, 0.0);
  if (!sides) return f;
  Tables1D t;
  t.init(o.degree, o.n_q_1d);
  const int dim = o.dim, p = o.degree, N = p + 1;
  const int npc = dim == 3 ? N * N * N : N * N;
  for (int s = 0; s < 2 * dim; ++s) {
    if (!((sides >> s) & 1u)) continue;
    const int cnt = lsmf_mesh_boundary_faces(g.mesh, s, nullptr, nullptr);
    std::vector<int32_t> cells(cnt), faces(cnt);
    lsmf_mesh_boundary_faces(g.mesh, s, cells.data(), faces.data());
    const int fd = s / 2, side = s % 2;
    for (int k = 0; k < cnt; ++k) {
      const int32_t *cn = g.mesh->cell_nodes + (size_t)cells[k] * npc;
      double X[8][3] = {};
      for (int v = 0; v < (1 << dim); ++v) {
        int a = 0, st = 1;
        for (int e = 0; e < dim; ++e) {
          a += ((v >> e) & 1) * p * st;
          st *= N;
        }
        for (int i = 0; i < dim; ++i) X[v][i] = g.mesh->coords[(size_t)cn[a] * dim + i];
      }
      const int nfq = dim == 3 ? o.n_q_1d * o.n_q_1d : o.n_q_1d;
      for (int fq = 0; fq < nfq; ++fq) {
        double xi[3] = {0, 0, 0}, w = 1.0;
        int r = fq;
        for (int d = 0; d < dim; ++d) {
          if (d == fd) {
            xi[d] = side;
            continue;
          }
          xi[d] = t.qpts[r % o.n_q_1d];
          w *= t.qw[r % o.n_q_1d];
          r /= o.n_q_1d;
        }
        double J[3][3] = {};
        for (int v = 0; v < (1 << dim); ++v)
          for (int d = 0; d < dim; ++d) {
            double gv = 1.0;
            for (int e = 0; e < dim; ++e) {
              const int ve = (v >> e) & 1;
              gv *= (e == d) ? (ve ? 1.0 : -1.0) : (ve ? xi[e] : 1.0 - xi[e]);
            }
            for (int i = 0; i < dim; ++i) J[i][d] += X[v][i] * gv;
          }
        // face measure: |cofactor column of J for direction fd|
        double meas;
        if (dim == 2) {
          const int td = 1 - fd;
          meas = std::sqrt(J[0][td] * J[0][td] + J[1][td] * J[1][td]);
        } else {
          const int a1 = (fd + 1) % 3, a2 = (fd + 2) % 3;
          const double cx = J[1][a1] * J[2][a2] - J[2][a1] * J[1][a2];
          const double cy = J[2][a1] * J[0][a2] - J[0][a1] * J[2][a2];
          const double cz = J[0][a1] * J[1][a2] - J[1][a1] * J[0][a2];
          meas = std::sqrt(cx * cx + cy * cy + cz * cz);
        }
        const double jxw = meas * w;
        for (int a = 0; a < npc; ++a) {
          int ai = a;
          double phi = 1.0;
          for (int d = 0; d < dim; ++d) {
            phi *= lagrange_value(t.nodes, ai % N, xi[d]);
            ai /= N;
          }
          for (int c = 0; c < dim; ++c) f[(size_t)cn[a] * dim + c] += phi * traction[s][c] * jxw;
        }
      }
    }
  }
  return f;
}

int catch_all(const std::function<void()> &fn) {
  try {
    fn();
    return LSMF_OK;
  } catch (const HipError &e) {
    return fail(e.code, e.what());
  } catch (const std::bad_alloc &) {
    return fail(LSMF_ERR_OUT_OF_MEMORY, "host allocation failed");
  } catch (const std::exception &e) {
    return fail(LSMF_ERR_INVALID_ARGUMENT, e.what());
  }
}

// PreconditionMG (V-cycle), PreconditionJacobi(omega) from the fine
// operator's inverse diagonal, or PreconditionIdentity
// (include/mf_elasticity.h:3125-3170)
void apply_preconditioner(lsmf_solver *S, double *z, const double *r) {
  switch (S->opt.preconditioner) {
    case LSMF_PRECONDITIONER_GMG:
      S->mg->precondition(z, r);
      break;
    case LSMF_PRECONDITIONER_JACOBI:
      S->fine.precondition_Jacobi(z, r, S->opt.preconditioner_relaxation);
      break;
    default:
      hip_check(hipMemcpyAsync(z, r, S->n * sizeof(double), hipMemcpyDeviceToDevice, S->s), "identity");
  }
}

void setup_preconditioner(lsmf_solver *S) {
  if (S->opt.preconditioner == LSMF_PRECONDITIONER_GMG)
    S->mg->setup(S->u_tot.p, S->ta.p, S->tb.p);
  else if (S->opt.preconditioner == LSMF_PRECONDITIONER_JACOBI)
    S->fine.compute_diagonal();
}

// SolverCG's condition-number estimate (include/mf_elasticity.h:3029-3046):
// at most `its` unpreconditioned CG steps on A x = b from x = 0 (fewer when
// the residual reaches tol), the Lanczos tridiagonal of the CG coefficients,
// the ratio of its extreme eigenvalues
double condition_number(lsmf_solver *S, const double *b, int its, double tol) {
  const size_t n = S->n;
  hipStream_t s = S->s;
  Reducer &red = *S->red;
  double *x = S->du.p, *r = S->r.p, *p = S->p.p, *Ap = S->Ap.p;
  hipLaunchKernelGGL(k_fill<double>, dim3(nblk(n)), dim3(256), 0, s, x, 0.0, n);
  hip_check(hipMemcpyAsync(r, b, n * sizeof(double), hipMemcpyDeviceToDevice, s), "cond copy");
  hip_check(hipMemcpyAsync(p, b, n * sizeof(double), hipMemcpyDeviceToDevice, s), "cond copy");
  double rr = red.dot(r, r, n);
  const double r0 = std::sqrt(rr);
  std::vector<double> alpha, beta;
  for (int k = 0; k < its && r0 > tol; ++k) {
    S->fine.vmult(Ap, p);
    const double pAp = red.dot(p, Ap, n);
    if (!(pAp > 0)) break;
    const double a = rr / pAp;
    alpha.push_back(a);
    const double rr_new = red.cg_update(x, r, (const double *)p, (const double *)Ap, a, n);
    if (std::sqrt(rr_new) <= tol) break;
    const double bcoef = rr_new / rr;
    beta.push_back(bcoef);
    rr = rr_new;
    launch_xpby<double>(s, p, (const double *)r, bcoef, n);
  }
  const size_t m = alpha.size();
  if (m < 2) return m == 1 ? 1.0 : 0.0;
  std::vector<double> diag(m), off(m - 1);
  diag[0] = 1.0 / alpha[0];
  for (size_t k = 1; k < m; ++k) {
    diag[k] = 1.0 / alpha[k] + beta[k - 1] / alpha[k - 1];
    off[k - 1] = std::sqrt(beta[k - 1]) / alpha[k - 1];
  }
  double lmin, lmax;
  tridiag_extremes(diag, off, lmin, lmax);
  return lmin > 0 ? lmax / lmin : 0.0;
}

// FP64 PCG (SolverCG::solve with the chosen preconditioner).  Returns the
// iteration count on convergence; throws LSMF_ERR_NOT_CONVERGED when the
// iteration limit is reached (also with max_it = 0 and |b| > tol) or p.Ap
// is zero or not finite.
int pcg(lsmf_solver *S, double *x, const double *b, double tol, int max_it, double *res_out, int *its_out = nullptr) {
  const size_t n = S->n;
  hipStream_t s = S->s;
  Reducer &red = *S->red;
  hipLaunchKernelGGL(k_fill<double>, dim3(nblk(n)), dim3(256), 0, s, x, 0.0, n);
  hip_check(hipMemcpyAsync(S->r.p, b, n * sizeof(double), hipMemcpyDeviceToDevice, s), "pcg copy");
  double rn = std::sqrt(red.dot(S->r.p, S->r.p, n));
  if (rn <= tol) {
    *res_out = rn;
    if (its_out) *its_out = 0;
    return 0;
  }
  apply_preconditioner(S, S->z.p, S->r.p);
  hip_check(hipMemcpyAsync(S->p.p, S->z.p, n * sizeof(double), hipMemcpyDeviceToDevice, s), "pcg copy");
  double rz = red.dot(S->r.p, S->z.p, n);
  int it = 0;
  while (it < max_it) {
    ++it;
    S->fine.vmult(S->Ap.p, S->p.p);
    const double pAp = red.dot(S->p.p, S->Ap.p, n);
    // breakdown: p.Ap zero or not finite (deal.II's ExcDivideByZero).  A
    // tiny negative p.Ap from round-off late in a long solve is continued,
    // as SolverCG does: the reference's 3D Q1 Cook output
    // (tests/cook_3d_lin_mf.output) takes such steps
    if (pAp == 0.0 || !std::isfinite(pAp)) {
      *res_out = rn;
      if (its_out) *its_out = it;
      throw HipError(LSMF_ERR_NOT_CONVERGED, "CG: breakdown (p.Ap = " + std::to_string(pAp) + " at iteration " +
                                                 std::to_string(it) + ")");
    }
    const double alpha = rz / pAp;
    const double rr = red.cg_update(x, S->r.p, (const double *)S->p.p, (const double *)S->Ap.p, alpha, n);
    rn = std::sqrt(rr);
    if (rn <= tol) break;
    apply_preconditioner(S, S->z.p, S->r.p);
    const double rz_new = red.dot(S->r.p, S->z.p, n);
    const double beta = rz_new / rz;
    rz = rz_new;
    launch_xpby<double>(s, S->p.p, (const double *)S->z.p, beta, n);
  }
  *res_out = rn;
  if (its_out) *its_out = it;
  if (!(rn <= tol))
    throw HipError(LSMF_ERR_NOT_CONVERGED, "CG: no convergence in " + std::to_string(it) + " iterations (|r| = " +
                                               std::to_string(rn) + ", tolerance " + std::to_string(tol) + ")");
  return it;
}

void set_state(lsmf_solver *S) {
  S->fine.set_linearization(S->u_tot.p);
  S->fine.cache();
  setup_preconditioner(S);
}

}  // namespace

extern "C" {

lsmf_status lsmf_solver_options_default(lsmf_solver_options *o) {
  if (!o) return fail(LSMF_ERR_INVALID_ARGUMENT, "options: null");
  std::memset(o, 0, sizeof(*o));
  o->dim = 3;
  o->degree = 2;
  o->n_q_1d = 3;
  o->n_cells_coarse[0] = o->n_cells_coarse[1] = o->n_cells_coarse[2] = 2;
  o->n_levels = 3;
  o->geometry = LSMF_GEOMETRY_BOX;
  o->hi[0] = o->hi[1] = o->hi[2] = 1.0;
  o->scale = 1.0;
  o->caching = LSMF_CACHING_TENSOR4;
  o->mg_precision = LSMF_FP32;
  o->mg_caching = LSMF_CACHING_TENSOR4;
  o->smoother_degree = 5;
  o->smoothing_range = 20.0;
  o->eig_cg_iterations = 30;
  o->coarse_reduction = 1e-3;
  o->coarse_eig_iterations = 0;
  o->tol_lin = 1e-6;
  o->max_lin_iterations = 1000;
  o->max_newton = 10;
  o->tol_f = 1e-9;
  o->tol_u = 1e-6;
  o->dt = 0.1;
  o->t_end = 1.0;
  o->coarse_solver = LSMF_COARSE_CHEBYSHEV;
  o->use_graph = 1;
  o->preconditioner = LSMF_PRECONDITIONER_GMG;
  o->preconditioner_relaxation = 0.65;
  o->formulation = 1;
  return LSMF_OK;
}

lsmf_status lsmf_solver_create(const lsmf_solver_options *o, const double *mu, const double *lambda,
                               lsmf_solver **out) {
  if (!o || !mu || !lambda || !out) return fail(LSMF_ERR_INVALID_ARGUMENT, "solver_create: null argument");
  if ((o->dim != 2 && o->dim != 3) || o->n_levels < 1 || o->n_levels > 16)
    return fail(LSMF_ERR_INVALID_ARGUMENT, "solver_create: bad dim or n_levels");
  if (o->mg_precision != LSMF_FP32 && o->mg_precision != LSMF_FP64)
    return fail(LSMF_ERR_INVALID_ARGUMENT, "solver_create: bad mg_precision");
  if (o->preconditioner < LSMF_PRECONDITIONER_GMG || o->preconditioner > LSMF_PRECONDITIONER_NONE)
    return fail(LSMF_ERR_INVALID_ARGUMENT, "solver_create: bad preconditioner");
  if (o->dt <= 0.0) return fail(LSMF_ERR_INVALID_ARGUMENT, "solver_create: dt must be positive");
  *out = nullptr;
  lsmf_solver *S = new (std::nothrow) lsmf_solver;
  if (!S) return fail(LSMF_ERR_OUT_OF_MEMORY, "solver_create: host allocation");
  const int rc = catch_all([&] {
    S->opt = *o;
    hip_check(hipStreamCreate(&S->s), "hipStreamCreate");
    const int L = o->n_levels;
    S->geoms.resize(L);
    for (int l = 0; l < L; ++l) {
      S->geoms[l].reset(new LevelGeom);
      build_level_geom(*o, l, *S->geoms[l]);
    }
    LevelGeom &fg = *S->geoms[L - 1];
    fg.mu.assign(mu, mu + fg.mesh->n_cells);
    fg.lam.assign(lambda, lambda + fg.mesh->n_cells);
    for (int l = L - 1; l > 0; --l) {
      coarsen_material(o->dim, S->geoms[l]->mesh->n_cells_dir, S->geoms[l]->mu, S->geoms[l - 1]->mu);
      coarsen_material(o->dim, S->geoms[l]->mesh->n_cells_dir, S->geoms[l]->lam, S->geoms[l - 1]->lam);
    }
    S->n = fg.n_dofs;
    lsmf_mesh_desc md{o->dim, o->degree, fg.mesh->n_cells, fg.mesh->n_nodes, fg.mesh->cell_nodes, fg.mesh->coords,
                      fg.cons.data(), fg.mesh->cell_colour};
    S->fine.initialize(md, o->n_q_1d, o->caching, S->s);
    S->fine.set_formulation(o->formulation);
    S->fine.set_material(fg.mu.data(), fg.lam.data());
    if (o->preconditioner != LSMF_PRECONDITIONER_GMG)
      ;  // no level operators
    else if (o->mg_precision == LSMF_FP32)
      S->mg.reset(new GMG<float>(*o, S->geoms, &S->fine, S->s));
    else
      S->mg.reset(new GMG<double>(*o, S->geoms, &S->fine, S->s));
    S->red.reset(new Reducer(S->s));
    const size_t n = S->n;
    for (DeviceArray<double> *v : {&S->u_n, &S->delta, &S->u_tot, &S->rhs, &S->du, &S->g, &S->Ag, &S->neumann, &S->r,
                                   &S->z, &S->p, &S->Ap})
      v->resize(n);
    size_t tmax = 0;
    for (int l = 1; l < L; ++l) tmax = std::max(tmax, S->geoms[l]->n_dofs * 2);
    S->ta.resize(std::max<size_t>(tmax, 1));
    S->tb.resize(std::max<size_t>(tmax, 1));
    hip_check(hipMemset(S->u_n.p, 0, n * sizeof(double)), "memset");
    // constrained DoFs of the finest level and their increments
    for (size_t i = 0; i < n; ++i)
      if (fg.cons[i]) S->cons_list.push_back((int32_t)i);
    S->cons_inc.assign(S->cons_list.size(), 0.0);
    for (int s = 0; s < 2 * o->dim; ++s) {
      if (!o->dirichlet_components[s]) continue;
      const int cnt = lsmf_mesh_boundary_nodes(fg.mesh, s, nullptr);
      std::vector<int32_t> nodes(cnt);
      lsmf_mesh_boundary_nodes(fg.mesh, s, nodes.data());
      for (int nd : nodes)
        for (int c = 0; c < o->dim; ++c)
          if ((o->dirichlet_components[s] >> c) & 1u) {
            const size_t dof = (size_t)nd * o->dim + c;
            const auto it = std::lower_bound(S->cons_list.begin(), S->cons_list.end(), (int32_t)dof);
            S->cons_inc[it - S->cons_list.begin()] = o->dirichlet_increment[s][c];
            if (o->dirichlet_increment[s][c] != 0.0) S->has_increment = true;
          }
    }
    S->d_cons_list.resize(S->cons_list.size());
    S->d_cons_inc.resize(S->cons_inc.size());
    if (!S->cons_list.empty()) {
      S->d_cons_list.upload(S->cons_list.data(), S->cons_list.size());
      S->d_cons_inc.upload(S->cons_inc.data(), S->cons_inc.size());
    }
    std::vector<double> f = neumann_vector(*o, o->neumann_traction, o->neumann_sides, fg);
    S->neumann.upload(f.data(), n);
    S->has_neumann = o->neumann_sides != 0;
  });
  if (rc) {
    delete S;
    return rc;
  }
  *out = S;
  return LSMF_OK;
}

lsmf_status lsmf_solver_destroy(lsmf_solver *s) {
  delete s;
  return LSMF_OK;
}

size_t lsmf_solver_n_dofs(const lsmf_solver *s, int level) {
  if (!s) return 0;
  if (level < 0) return s->n;
  if (level >= (int)s->geoms.size()) return 0;
  return s->geoms[level]->n_dofs;
}

lsmf_op *lsmf_solver_fine_operator(lsmf_solver *s) { return s ? s->fine.handle() : nullptr; }

lsmf_status lsmf_solver_set_state(lsmf_solver *S, const double *u_host) {
  if (!S || !u_host) return fail(LSMF_ERR_INVALID_ARGUMENT, "set_state: null argument");
  return catch_all([&] {
    S->u_tot.upload(u_host, S->n);
    set_state(S);
    hip_check(hipStreamSynchronize(S->s), "set_state");
  });
}

lsmf_status lsmf_solver_precondition(lsmf_solver *S, double *z, const double *r) {
  if (!S || !z || !r) return fail(LSMF_ERR_INVALID_ARGUMENT, "precondition: null argument");
  return catch_all([&] {
    apply_preconditioner(S, z, r);
    hip_check(hipStreamSynchronize(S->s), "precondition");
  });
}

lsmf_status lsmf_solver_cg(lsmf_solver *S, double *x, const double *b, double tol, int max_it, int *its,
                           double *res) {
  if (!S || !x || !b || !its || !res) return fail(LSMF_ERR_INVALID_ARGUMENT, "cg: null argument");
  return catch_all([&] {
    pcg(S, x, b, tol, max_it, res, its);
  });
}

lsmf_status lsmf_solver_level_info(const lsmf_solver *S, int level, double *lmax, double *lmin, int *degree) {
  if (!S || !S->mg || !lmax || !lmin || !degree || level < 0 || level >= S->mg->n_levels())
    return fail(LSMF_ERR_INVALID_ARGUMENT, "level_info: bad argument");
  S->mg->info(level, *lmax, *lmin, *degree);
  return LSMF_OK;
}

lsmf_status lsmf_solver_transfer(lsmf_solver *S, int level, int mode, double *dst, const double *src) {
  if (!S || !S->mg || !dst || !src || level < 1 || level >= (int)S->geoms.size() || mode < 0 || mode > 2)
    return fail(LSMF_ERR_INVALID_ARGUMENT, "transfer: bad argument");
  return catch_all([&] { S->mg->transfer64(level, mode, dst, src); });
}

lsmf_status lsmf_solver_get_solution(const lsmf_solver *S, double *u_host) {
  if (!S || !u_host) return fail(LSMF_ERR_INVALID_ARGUMENT, "get_solution: null argument");
  return catch_all([&] {
    hip_check(hipStreamSynchronize(S->s), "sync");
    hip_check(hipMemcpy(u_host, S->u_n.p, S->n * sizeof(double), hipMemcpyDeviceToHost), "get_solution");
  });
}

// solve_nonlinear_timestep (include/mf_elasticity.h:2289-2559) for the
// load step at time t
static void newton_step(lsmf_solver *S, double t, int step, lsmf_step_stats &st) {
  const lsmf_solver_options &o = S->opt;
  const size_t n = S->n;
  hipStream_t s = S->s;
  Reducer &red = *S->red;
  const int ncons = (int)S->cons_list.size();
  const auto t_step = Clock::now();
  st = lsmf_step_stats{};
  hipLaunchKernelGGL(k_fill<double>, dim3(nblk(n)), dim3(256), 0, s, S->delta.p, 0.0, n);
  double res0 = 0, upd0 = 0, upd_prev = 0;
  bool converged = false;
  int it = 0;
  for (; it < o.max_newton; ++it) {
    auto t0 = Clock::now();
    // set_total_solution + cache at the new linearisation point
    hipLaunchKernelGGL(k_add<double>, dim3(nblk(n)), dim3(256), 0, s, S->u_tot.p, (const double *)S->u_n.p,
                       (const double *)S->delta.p, n);
    S->fine.set_linearization(S->u_tot.p);
    S->fine.cache();
    st.time_setup_s += seconds_since(t0);
    st.time_setup_cache_s += seconds_since(t0);
    t0 = Clock::now();
    // assemble_system: residual + Neumann + lifting (-K g) at it 0
    abi_check(lsmf_residual(S->fine.handle(), S->rhs.p, s), "residual");
    if (S->has_neumann)
      hipLaunchKernelGGL(k_axpby<double>, dim3(nblk(n)), dim3(256), 0, s, S->rhs.p, (const double *)S->neumann.p,
                         S->neumann_absolute ? 1.0 : t, 1.0, n);
    const bool lift = (it == 0 && S->has_increment);
    if (lift) {
      hipLaunchKernelGGL(k_fill<double>, dim3(nblk(n)), dim3(256), 0, s, S->g.p, 0.0, n);
      hipLaunchKernelGGL(k_set_list<double>, dim3(nblk(ncons)), dim3(256), 0, s, S->g.p, S->d_cons_list.p,
                         (const double *)S->d_cons_inc.p, 0.0, ncons);
      S->fine.vmult(S->Ag.p, S->g.p, LSMF_MASK_PLAIN);
      hipLaunchKernelGGL(k_axpby<double>, dim3(nblk(n)), dim3(256), 0, s, S->rhs.p, (const double *)S->Ag.p, -1.0,
                         1.0, n);
    }
    if (ncons)
      hipLaunchKernelGGL(k_set_list<double>, dim3(nblk(ncons)), dim3(256), 0, s, S->rhs.p, S->d_cons_list.p,
                         (const double *)nullptr, 0.0, ncons);
    const double res = std::sqrt(red.dot(S->rhs.p, S->rhs.p, n));
    st.time_residual_s += seconds_since(t0);
    if (it == 0) {
      res0 = res;
      st.residual_0 = res;
    }
    st.residual_final = res;
    // check_convergence (include/mf_elasticity.h:2245-2284)
    if (it > 2) {
      const double rn = res0 > 0 ? res / res0 : res;
      const double un = upd0 > 0 ? upd_prev / upd0 : upd_prev;
      if (res <= o.tol_f_abs || upd_prev <= o.tol_u_abs || (rn <= o.tol_f && un <= o.tol_u)) converged = true;
    }
    if (converged) break;
    if (o.cond_number_cg_iterations > 0) st.cond_number = condition_number(S, S->rhs.p, o.cond_number_cg_iterations, o.tol_lin * res);
    t0 = Clock::now();
    setup_preconditioner(S);
    hip_check(hipStreamSynchronize(s), "setup sync");
    st.time_setup_s += seconds_since(t0);
    if (S->opt.preconditioner == LSMF_PRECONDITIONER_GMG)
      st.time_setup_cache_s += S->mg->setup_cache_seconds();
    else if (S->opt.preconditioner == LSMF_PRECONDITIONER_JACOBI)
      st.time_setup_cache_s += seconds_since(t0);
    t0 = Clock::now();
    double lin_res = 0;
    const int lin = pcg(S, S->du.p, S->rhs.p, o.tol_lin * res, o.max_lin_iterations, &lin_res);
    st.time_linear_s += seconds_since(t0);
    st.linear_iterations += lin;
    st.max_linear_iterations = std::max(st.max_linear_iterations, lin);
    if (ncons)
      hipLaunchKernelGGL(k_set_list<double>, dim3(nblk(ncons)), dim3(256), 0, s, S->du.p, S->d_cons_list.p,
                         (const double *)nullptr, 0.0, ncons);
    const double upd = std::sqrt(red.dot(S->du.p, S->du.p, n));
    if (it == 0) upd0 = upd;
    upd_prev = upd;
    if (o.verbose)
      std::fprintf(stderr, "[lsmf] step %d it %d |R| %.6e CG %d |du| %.6e\n", step + 1, it, res, lin, upd);
    if (lift)  // constraints.distribute: increment on constrained DoFs
      hipLaunchKernelGGL(k_set_list<double>, dim3(nblk(ncons)), dim3(256), 0, s, S->du.p, S->d_cons_list.p,
                         (const double *)S->d_cons_inc.p, 0.0, ncons);
    hipLaunchKernelGGL(k_axpby<double>, dim3(nblk(n)), dim3(256), 0, s, S->delta.p, (const double *)S->du.p, 1.0, 1.0,
                       n);
  }
  if (!converged) throw HipError(LSMF_ERR_NOT_CONVERGED, "No convergence in nonlinear solver!");
  hipLaunchKernelGGL(k_axpby<double>, dim3(nblk(n)), dim3(256), 0, s, S->u_n.p, (const double *)S->delta.p, 1.0, 1.0,
                     n);
  hip_check(hipStreamSynchronize(s), "step sync");
  st.newton_iterations = it;
  st.time_total_s = seconds_since(t_step);
}

lsmf_status lsmf_solver_step(lsmf_solver *S, lsmf_step_stats *stats, double *t_out, int *done) {
  if (!S || !done) return fail(LSMF_ERR_INVALID_ARGUMENT, "step: null argument");
  *done = 0;
  return catch_all([&] {
    // Time::increment accumulates dt (include/mf_elasticity.h:614-653)
    const double t = S->t_prev + S->opt.dt;
    if (!(t <= S->opt.t_end)) {
      *done = 1;
      return;
    }
    lsmf_step_stats st{};
    newton_step(S, t, S->n_steps_done, st);
    S->t_prev = t;
    ++S->n_steps_done;
    if (stats) *stats = st;
    if (t_out) *t_out = t;
  });
}

lsmf_status lsmf_solver_set_dirichlet_increment(lsmf_solver *S, const double *inc) {
  if (!S || !inc) return fail(LSMF_ERR_INVALID_ARGUMENT, "set_dirichlet_increment: null argument");
  return catch_all([&] {
    S->has_increment = false;
    for (size_t k = 0; k < S->cons_list.size(); ++k) {
      S->cons_inc[k] = inc[S->cons_list[k]];
      if (S->cons_inc[k] != 0.0) S->has_increment = true;
    }
    if (!S->cons_list.empty()) {
      hip_check(hipStreamSynchronize(S->s), "sync");
      S->d_cons_inc.upload(S->cons_inc.data(), S->cons_inc.size());
    }
  });
}

lsmf_status lsmf_solver_set_traction(lsmf_solver *S, const double traction[6][3]) {
  if (!S || !traction) return fail(LSMF_ERR_INVALID_ARGUMENT, "set_traction: null argument");
  return catch_all([&] {
    double tr[6][3] = {};
    unsigned sides = 0;
    for (int s = 0; s < 2 * S->opt.dim; ++s)
      for (int c = 0; c < S->opt.dim; ++c) {
        tr[s][c] = traction[s][c];
        if (tr[s][c] != 0.0) sides |= 1u << s;
      }
    std::vector<double> f = neumann_vector(S->opt, tr, sides, *S->geoms.back());
    hip_check(hipStreamSynchronize(S->s), "sync");
    S->neumann.upload(f.data(), S->n);
    S->has_neumann = sides != 0;
    S->neumann_absolute = true;
  });
}

// Solid::run (include/mf_elasticity.h:1187-1222)
lsmf_status lsmf_solver_run(lsmf_solver *S, int max_steps, lsmf_step_stats *stats, int *n_steps) {
  if (!S || !n_steps) return fail(LSMF_ERR_INVALID_ARGUMENT, "run: null argument");
  *n_steps = 0;
  for (int k = 0; k < max_steps; ++k) {
    int done = 0;
    const int rc = lsmf_solver_step(S, stats ? stats + k : nullptr, nullptr, &done);
    if (rc) return rc;
    if (done) break;
    ++*n_steps;
  }
  return LSMF_OK;
}

}  // extern "C"
