// lsmf_launch.hpp -- host launchers for one (dim, N) instantiation; included
// by the lsmf_inst_*.hip units, each of which instantiates one LaunchTable.
#pragma once

#include "lsmf_error.hpp"
#include "lsmf_kernels.hpp"
#include "lsmf_op.hpp"

namespace lsmf {

template <int DIM, int N> struct Launch {
  using CS = CellShape<DIM, N>;

  // wide tiles exist for 3D Q4 and the variants that run on tensor2 /
  // tensor4 operators (their products, residual and residual VJP)
  template <int VAR> static constexpr bool wide_ok() {
    return DIM == 3 && (N == 5 || N == 3) &&
           (VAR == V_TENSOR2 || VAR == V_TENSOR4 || VAR == V_RESIDUAL || VAR == V_ADJOINT);
  }

  // one colour's fused cell kernel: n tiles (tiled layouts) or n cells, nv
  // vectors (grid y)
  template <typename T, int VAR, bool ADD, bool CONSTR, unsigned PH>
  static int launch_vmult(const lsmf_op *op, const VmultParams<T> &P, int n, unsigned nv, hipStream_t s) {
    if (op->tile_kind == kTileWide) {
      if constexpr (wide_ok<VAR>())
        hipLaunchKernelGGL((k_vmult<DIM, N, T, VAR, ADD, CONSTR, PH, kTileWide>), dim3(n, nv),
                           dim3(tile_threads<kTileWide>()), 0, s, P);
      else
        return LSMF_ERR_UNSUPPORTED;
    } else if (op->tile_kind == 1) {
      hipLaunchKernelGGL((k_vmult<DIM, N, T, VAR, ADD, CONSTR, PH, 1>), dim3(n, nv), dim3(kWG), 0, s, P);
    } else {
      hipLaunchKernelGGL((k_vmult<DIM, N, T, VAR, ADD, CONSTR, PH, 0>), dim3(grid_for(n), nv), dim3(kWG), 0, s, P);
    }
    return 0;
  }

  template <typename T> static VmultParams<T> params(const lsmf_op *op, int k) {
    VmultParams<T> P{};
    P.idx = op->d_idx;
    P.cache = static_cast<const T *>(op->d_cache);
    P.fs = op->field_stride;
    P.tab = static_cast<const T *>(op->d_tab);
    P.u = static_cast<const T *>(op->d_u);
    P.coords = op->d_coords;
    P.mu = static_cast<const T *>(op->d_mu);
    P.lambda = static_cast<const T *>(op->d_lambda);
    const std::vector<int> &ptr = op->tiled ? op->tile_ptr : op->colour_ptr;
    P.cell_begin = ptr[k];
    P.cell_end = ptr[k + 1];
    P.formulation = op->formulation;
    P.curved = op->curved;
    P.tnodes = op->d_tile_nodes;
    P.tinfo = op->d_tile_info;
    P.hang = op->d_hang;
    P.hcons = op->d_hcons;
    P.hbuf = static_cast<T *>(op->d_hbuf);
    P.hstride = (size_t)op->n_hang_inst * DIM;
    P.hptr = op->d_hptr;
    P.dcell = op->d_dcell;
    P.deptr = op->d_deptr;
    P.dtptr = op->d_dtptr;
    P.dta = op->d_dta;
    P.dtw = static_cast<const T *>(op->d_dtw);
    P.dbuf = static_cast<T *>(op->d_dbuf);
    P.hmaster = op->d_hmaster;
    P.hw = static_cast<const T *>(op->d_hw);
    P.dmu = op->adj_dmu;
    P.dlambda = op->adj_dlambda;
    P.perm = op->d_perm;
    using EO = EvenOdd<N, T>;
    static_assert(sizeof(P.sf) / sizeof(P.sf[0]) >= 2 * EO::PK, "1D matrices do not fit the kernel arguments");
    // even-odd forms of S (rows r < R) and D, from the plain tables
    for (int m = 0; m < 2; ++m) {
      const double *M = op->h_sf + m * N * N;
      T *pk = P.sf + m * EO::PK;
      for (int r = 0; r < EO::R; ++r) {
        for (int k = 0; k < EO::H; ++k) {
          pk[r * EO::H + k] = T(0.5 * (M[r * N + k] + M[r * N + N - 1 - k]));
          pk[EO::R * EO::H + r * EO::H + k] = T(0.5 * (M[r * N + k] - M[r * N + N - 1 - k]));
        }
        pk[2 * EO::R * EO::H + r] = EO::ODD ? T(M[r * N + EO::H]) : T(0);
      }
    }
    return P;
  }

  static int grid_for(int n) { return (n + CS::CPW - 1) / CS::CPW; }
  // with node constraints every write adds (no first-touch stores): zero first
  template <typename T> static void zero_if_constrained(const lsmf_op *op, T *dst, bool add, hipStream_t s) {
    if (op->n_node_cons > 0 && !add)
      hipLaunchKernelGGL(k_zero<T>, dim3((unsigned)((op->n_dofs + 255) / 256)), dim3(256), 0, s, dst, op->n_dofs);
  }
  // DoFs that no cell touches keep whatever dst held: zero them (set mode)
  template <typename T> static void zero_orphans(const lsmf_op *op, T *dst, hipStream_t s) {
    if (op->n_orphan > 0)
      hipLaunchKernelGGL(k_zero_list<T>, dim3((op->n_orphan + 255) / 256), dim3(256), 0, s, op->d_orphan,
                         op->n_orphan, dst);
  }
  static int n_colours(const lsmf_op *op) {
    return (int)(op->tiled ? op->tile_ptr.size() : op->colour_ptr.size()) - 1;
  }

  // the masters' share of the constrained nodes' results, after the colours
  // (nv vectors at dst + b * vstride)
  template <typename T, bool CONSTR> static void hang_gather(const lsmf_op *op, T *dst, size_t vstride, unsigned nv,
                                                              hipStream_t s) {
    if (op->n_hang_masters == 0) return;
    hipLaunchKernelGGL((k_hang_gather<DIM, T, CONSTR>), dim3((op->n_hang_masters + 255) / 256, nv), dim3(256), 0, s,
                       dst, vstride, static_cast<const T *>(op->d_hbuf), (size_t)op->n_hang_inst * DIM,
                       op->d_hg_node, op->d_hg_ptr, op->d_hg_inst, static_cast<const T *>(op->d_hg_w),
                       op->n_hang_masters);
  }

  // one launch per colour; the tiled layout launches one workgroup per tile
  template <typename T, int VAR, bool ADD, bool CONSTR, unsigned PH>
  static int run_colours(const lsmf_op *op, T *dst, const T *src, hipStream_t s) {
    const int nc = n_colours(op);
    for (int k = 0; k < nc; ++k) {
      VmultParams<T> P = params<T>(op, k);
      P.src = src;
      P.dst = dst;
      const int n = P.cell_end - P.cell_begin;
      if (n <= 0) continue;
      if (op->prof) op->prof->before(s);
      const int rc = launch_vmult<T, VAR, ADD, CONSTR, PH>(op, P, n, 1, s);
      if (op->prof) op->prof->after(s);
      if (rc) return rc;
    }
    if (op->n_node_cons > 0) hang_gather<T, CONSTR>(op, dst, 0, 1, s);
    return 0;
  }

  template <typename T, int VAR>
  static int vmult_t(const lsmf_op *op, void *dst_, const void *src_, bool add, bool constr, unsigned ph,
                     hipStream_t s) {
    T *dst = static_cast<T *>(dst_);
    const T *src = static_cast<const T *>(src_);
    constexpr unsigned FULL = PH_SF | PH_QD;
    zero_if_constrained(op, dst, add, s);
    if (!add) zero_orphans(op, dst, s);
    int rc;
    if (constr) {
      if (ph == FULL) {
        if (add) rc = run_colours<T, VAR, true, true, FULL>(op, dst, src, s);
        else rc = run_colours<T, VAR, false, true, FULL>(op, dst, src, s);
      } else if (ph == PH_SF) {
        rc = run_colours<T, VAR, false, true, PH_SF>(op, dst, src, s);
      } else {
        rc = run_colours<T, VAR, false, true, 0u>(op, dst, src, s);
      }
      if (op->n_cons > 0)
        hipLaunchKernelGGL(k_constrained_identity<T>, dim3((op->n_cons + 255) / 256), dim3(256), 0, s, op->d_cons,
                           op->n_cons, src, dst, add ? 1 : 0);
    } else {
      if (add) rc = run_colours<T, VAR, true, false, FULL>(op, dst, src, s);
      else rc = run_colours<T, VAR, false, false, FULL>(op, dst, src, s);
    }
    if (rc) return rc;
    return hipGetLastError() == hipSuccess ? 0 : LSMF_ERR_HIP;
  }

  // nvec products in set mode with constraints, vector b at b * n_dofs:
  // one launch per colour for all vectors (grid y = vector), the same code
  // path per vector as vmult_t
  template <typename T, int VAR> static int vmult_batched_t(const lsmf_op *op, T *dst, const T *src, int nvec,
                                                            hipStream_t s) {
    constexpr unsigned FULL = PH_SF | PH_QD;
    // grid y holds the vectors; with node constraints, as many as d_hbuf holds
    const int kMaxY = op->n_node_cons > 0 ? kHangBatch : 65535;
    const size_t n = op->n_dofs;
    const int nc = n_colours(op);
    for (int v0 = 0; v0 < nvec; v0 += kMaxY) {
      const unsigned nv = (unsigned)std::min(nvec - v0, kMaxY);
      T *d = dst + (size_t)v0 * n;
      const T *x = src + (size_t)v0 * n;
      if (op->n_node_cons > 0)
        hipLaunchKernelGGL(k_zero<T>, dim3((unsigned)((nv * n + 255) / 256)), dim3(256), 0, s, d, nv * n);
      if (op->n_orphan > 0)
        hipLaunchKernelGGL(k_zero_list_batch<T>, dim3((op->n_orphan + 255) / 256, nv), dim3(256), 0, s,
                           op->d_orphan, op->n_orphan, d, n);
      for (int k = 0; k < nc; ++k) {
        VmultParams<T> P = params<T>(op, k);
        P.src = x;
        P.dst = d;
        P.vstride = n;
        const int nb = P.cell_end - P.cell_begin;
        if (nb <= 0) continue;
        const int rc = launch_vmult<T, VAR, false, true, FULL>(op, P, nb, nv, s);
        if (rc) return rc;
      }
      if (op->n_node_cons > 0) hang_gather<T, true>(op, d, n, nv, s);
      if (op->n_cons > 0)
        hipLaunchKernelGGL(k_constrained_identity_batch<T>, dim3((op->n_cons + 255) / 256, nv), dim3(256), 0, s,
                           op->d_cons, op->n_cons, x, d, n);
    }
    return hipGetLastError() == hipSuccess ? 0 : LSMF_ERR_HIP;
  }

  static int vmult_batched(const lsmf_op *op, void *dst, const void *src, int nvec, hipStream_t s) {
    const bool f64 = op->precision == LSMF_FP64;
#define LSMF_BATCH(V)                                                                          \
  return f64 ? vmult_batched_t<double, V>(op, (double *)dst, (const double *)src, nvec, s) \
             : vmult_batched_t<float, V>(op, (float *)dst, (const float *)src, nvec, s)
    switch (op->caching) {
      case LSMF_CACHING_TENSOR4: LSMF_BATCH(V_TENSOR4);
      case LSMF_CACHING_TENSOR2: LSMF_BATCH(V_TENSOR2);
      case LSMF_CACHING_SCALAR: LSMF_BATCH(V_SCALAR);
      case LSMF_CACHING_TENSOR4_NS: LSMF_BATCH(V_TENSOR4_NS);
      case LSMF_CACHING_SCALAR_REFERENTIAL: LSMF_BATCH(V_SCALAR_REF);
      default:
        return fail(LSMF_ERR_UNKNOWN_CACHING, "vmult_batched: caching variant not available on the GPU");
    }
#undef LSMF_BATCH
  }

  static int vmult(const lsmf_op *op, void *dst, const void *src, bool add, bool constr, unsigned ph,
                   hipStream_t s) {
    const bool f64 = op->precision == LSMF_FP64;
    switch (op->caching) {
      case LSMF_CACHING_TENSOR4:
        return f64 ? vmult_t<double, V_TENSOR4>(op, dst, src, add, constr, ph, s)
                   : vmult_t<float, V_TENSOR4>(op, dst, src, add, constr, ph, s);
      case LSMF_CACHING_TENSOR2:
        return f64 ? vmult_t<double, V_TENSOR2>(op, dst, src, add, constr, ph, s)
                   : vmult_t<float, V_TENSOR2>(op, dst, src, add, constr, ph, s);
      case LSMF_CACHING_SCALAR:
        return f64 ? vmult_t<double, V_SCALAR>(op, dst, src, add, constr, ph, s)
                   : vmult_t<float, V_SCALAR>(op, dst, src, add, constr, ph, s);
      case LSMF_CACHING_TENSOR4_NS:
        return f64 ? vmult_t<double, V_TENSOR4_NS>(op, dst, src, add, constr, ph, s)
                   : vmult_t<float, V_TENSOR4_NS>(op, dst, src, add, constr, ph, s);
      case LSMF_CACHING_SCALAR_REFERENTIAL:
        return f64 ? vmult_t<double, V_SCALAR_REF>(op, dst, src, add, constr, ph, s)
                   : vmult_t<float, V_SCALAR_REF>(op, dst, src, add, constr, ph, s);
      default:
        return fail(LSMF_ERR_UNKNOWN_CACHING, "vmult: caching variant not available on the GPU");
    }
  }

  template <typename T, int VAR> static int cache_t(const lsmf_op *op, int *bad, hipStream_t s) {
    CacheParams P{};
    P.idx = op->d_idx;
    P.tab = op->d_tab64;
    P.u = op->d_u64;
    P.coords = op->d_coords;
    P.mu = op->d_mu64;
    P.lambda = op->d_lambda64;
    P.cache = op->d_cache;
    P.fs = op->field_stride;
    P.n_cells = op->n_cells;
    P.formulation = op->formulation;
    P.curved = op->curved;
    P.bad = bad;
    hipLaunchKernelGGL((k_cache<DIM, N, T, VAR>), dim3(grid_for(op->n_cells)), dim3(kWG), 0, s, P);
    return hipGetLastError() == hipSuccess ? 0 : LSMF_ERR_HIP;
  }

  static int cache(const lsmf_op *op, int *bad, hipStream_t s) {
    const bool f64 = op->precision == LSMF_FP64;
    switch (op->caching) {
      case LSMF_CACHING_TENSOR4:
        return f64 ? cache_t<double, V_TENSOR4>(op, bad, s) : cache_t<float, V_TENSOR4>(op, bad, s);
      case LSMF_CACHING_TENSOR2:
        return f64 ? cache_t<double, V_TENSOR2>(op, bad, s) : cache_t<float, V_TENSOR2>(op, bad, s);
      case LSMF_CACHING_SCALAR:
        return f64 ? cache_t<double, V_SCALAR>(op, bad, s) : cache_t<float, V_SCALAR>(op, bad, s);
      case LSMF_CACHING_TENSOR4_NS:
        return f64 ? cache_t<double, V_TENSOR4_NS>(op, bad, s) : cache_t<float, V_TENSOR4_NS>(op, bad, s);
      case LSMF_CACHING_SCALAR_REFERENTIAL:
        return f64 ? cache_t<double, V_SCALAR_REF>(op, bad, s) : cache_t<float, V_SCALAR_REF>(op, bad, s);
      default:
        return fail(LSMF_ERR_UNKNOWN_CACHING, "cache: caching variant not available on the GPU");
    }
  }

  template <typename T, int VAR> static int diagonal_t(const lsmf_op *op, hipStream_t s) {
    // with node constraints every cell adds (no first-touch bits): start
    // from zero; cells with constrained nodes leave per-entry results that
    // the gather below adds per node
    if (op->n_node_cons > 0)
      hipLaunchKernelGGL(k_zero<T>, dim3((unsigned)((op->n_dofs + 255) / 256)), dim3(256), 0, s,
                         static_cast<T *>(op->d_diag), op->n_dofs);
    const int nc = n_colours(op);
    for (int k = 0; k < nc; ++k) {
      VmultParams<T> P = params<T>(op, k);
      P.dst = static_cast<T *>(op->d_diag);
      // scalar_referential shares the scalar kernel but integrates on the Q1
      // vertex map on any mesh (include/mf_nh_operator.h:1255-1400)
      if (op->caching == LSMF_CACHING_SCALAR_REFERENTIAL) P.curved = 0;
      const int n = P.cell_end - P.cell_begin;
      if (n <= 0) continue;
      if (op->tile_kind == kTileWide) {
        if constexpr (wide_ok<VAR>())
          hipLaunchKernelGGL((k_diagonal<DIM, N, T, VAR, kTileWide>), dim3(n), dim3(kWG), 0, s, P);
        else
          return LSMF_ERR_UNSUPPORTED;
      } else if (op->tile_kind == 1) {
        hipLaunchKernelGGL((k_diagonal<DIM, N, T, VAR, 1>), dim3(n), dim3(kWG), 0, s, P);
      } else {
        hipLaunchKernelGGL((k_diagonal<DIM, N, T, VAR, 0>), dim3(grid_for(n)), dim3(kWG), 0, s, P);
      }
    }
    T *diag = static_cast<T *>(op->d_diag);
    if (op->n_diag_nodes > 0)
      hipLaunchKernelGGL((k_hang_gather<DIM, T, true>), dim3((op->n_diag_nodes + 255) / 256), dim3(256), 0, s, diag,
                         (size_t)0, static_cast<const T *>(op->d_dbuf), (size_t)0, op->d_dg_node, op->d_dg_ptr,
                         op->d_dg_ent, static_cast<const T *>(nullptr), op->n_diag_nodes);
    zero_orphans(op, diag, s);
    if (op->n_cons > 0)
      hipLaunchKernelGGL(k_set_one<T>, dim3((op->n_cons + 255) / 256), dim3(256), 0, s, op->d_cons, op->n_cons, diag);
    const size_t n = op->n_dofs;
    hipLaunchKernelGGL(k_zero<int>, dim3(1), dim3(64), 0, s, op->d_bad, (size_t)1);
    hipLaunchKernelGGL(k_invert<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, diag,
                       static_cast<T *>(op->d_inv_diag), n, op->d_bad);
    return hipGetLastError() == hipSuccess ? 0 : LSMF_ERR_HIP;
  }


  static int diagonal(const lsmf_op *op, hipStream_t s) {
    const bool f64 = op->precision == LSMF_FP64;
    switch (op->caching) {
      case LSMF_CACHING_TENSOR4:
        return f64 ? diagonal_t<double, V_TENSOR4>(op, s) : diagonal_t<float, V_TENSOR4>(op, s);
      case LSMF_CACHING_TENSOR2:
        return f64 ? diagonal_t<double, V_TENSOR2>(op, s) : diagonal_t<float, V_TENSOR2>(op, s);
      case LSMF_CACHING_SCALAR:
      case LSMF_CACHING_SCALAR_REFERENTIAL:  // c is field 0 of both caches; F from u
        return f64 ? diagonal_t<double, V_SCALAR>(op, s) : diagonal_t<float, V_SCALAR>(op, s);
      case LSMF_CACHING_TENSOR4_NS:
        return f64 ? diagonal_t<double, V_TENSOR4_NS>(op, s) : diagonal_t<float, V_TENSOR4_NS>(op, s);
      default:
        return fail(LSMF_ERR_UNKNOWN_CACHING, "diagonal: caching variant not available on the GPU");
    }
  }

  static int residual(const lsmf_op *op, double *rhs, hipStream_t s) {
    if (op->precision != LSMF_FP64) return fail(LSMF_ERR_INVALID_ARGUMENT, "residual needs an FP64 operator");
    constexpr unsigned FULL = PH_SF | PH_QD;
    zero_if_constrained(op, rhs, false, s);
    zero_orphans(op, rhs, s);
    const int rc = run_colours<double, V_RESIDUAL, false, false, FULL>(op, rhs, nullptr, s);
    if (rc) return rc;
    return hipGetLastError() == hipSuccess ? 0 : LSMF_ERR_HIP;
  }

  static int residual_vjp(const lsmf_op *op, double *du, const double *w, hipStream_t s) {
    if (op->precision != LSMF_FP64) return fail(LSMF_ERR_INVALID_ARGUMENT, "residual VJP needs an FP64 operator");
    constexpr unsigned FULL = PH_SF | PH_QD;
    zero_if_constrained(op, du, false, s);
    zero_orphans(op, du, s);
    const int rc = run_colours<double, V_ADJOINT, false, false, FULL>(op, du, w, s);
    if (rc) return rc;
    return hipGetLastError() == hipSuccess ? 0 : LSMF_ERR_HIP;
  }

  static const LaunchTable *table() {
    static const LaunchTable t{&vmult, &vmult_batched, &cache, &diagonal, &residual, &residual_vjp};
    return &t;
  }
};

}  // namespace lsmf

#define LSMF_INSTANTIATE(D, NN)                                                          \
  namespace lsmf {                                                                       \
  const LaunchTable *launch_table_##D##_##NN() { return Launch<D, NN>::table(); }        \
  }
