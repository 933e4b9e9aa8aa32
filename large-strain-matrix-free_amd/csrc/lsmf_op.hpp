// lsmf_op.hpp -- the operator object behind the opaque lsmf_op handle, and
// the per-(dim, N) launch table the instantiation units fill in.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "lsmf.h"

namespace lsmf {
// per-launch event timing of the operator's cell kernels (the reference's
// TimerOutput / LIKWID regions around vmult, include/mf_nh_operator.h:108-119)
struct Profiler {
  bool on = false;
  std::vector<hipEvent_t> start, stop;
  size_t used = 0;
  void before(hipStream_t s) {
    if (!on) return;
    if (used == start.size()) {
      hipEvent_t a, b;
      if (hipEventCreate(&a) != hipSuccess) return;
      if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return;
      }
      start.push_back(a);
      stop.push_back(b);
    }
    (void)hipEventRecord(start[used], s);
  }
  void after(hipStream_t s) {
    if (!on || used >= start.size()) return;
    (void)hipEventRecord(stop[used], s);
    ++used;
  }
  ~Profiler() {
    for (auto e : start) (void)hipEventDestroy(e);
    for (auto e : stop) (void)hipEventDestroy(e);
  }
};
}  // namespace lsmf

// vectors per k_hang_gather pass of a batched product (the size of d_hbuf)
constexpr int kHangBatch = 8;

struct lsmf_op {
  lsmf::Profiler *prof = nullptr;
  // shape
  int dim = 0, degree = 0, nq = 0, N = 0, nqp = 0;
  int precision = LSMF_FP64, caching = LSMF_CACHING_TENSOR4;
  int formulation = 1;  // material formulation (include/material.h:82-94)
  bool curved = false;  // some support point off its cell's Q1 map: Eulerian Q_p current metric
  int n_cells = 0, n_nodes = 0;
  size_t n_dofs = 0;
  int n_fields = 0;
  size_t field_stride = 0;
  // colouring (host): sorted cells [colour_ptr[k], colour_ptr[k+1])
  std::vector<int> colour_ptr;
  std::vector<int> perm;      // sorted position -> mesh cell
  std::vector<int> inv_perm;  // mesh cell -> sorted position
  // device data
  uint32_t *d_idx = nullptr;        // [n_cells][nqp] node | constraint bits | first-touch bit
  double *d_coords = nullptr;       // [n_nodes][dim]
  void *d_tab = nullptr;            // S, D, St, Dt, qp, qw in operator precision
  double *d_tab64 = nullptr;        // same tables in FP64 (cache kernel)
  double h_sf[2 * 25] = {};         // S[q][k] then D[q][k] (N x N each), kernel-argument copy
  double *d_mu64 = nullptr, *d_lambda64 = nullptr;  // per sorted cell
  void *d_mu = nullptr, *d_lambda = nullptr;        // operator precision
  double *d_u64 = nullptr;          // linearisation state copy
  void *d_u = nullptr;              // operator precision copy (scalar variant)
  void *d_cache = nullptr;
  void *d_diag = nullptr, *d_inv_diag = nullptr;
  int32_t *d_cons = nullptr;        // constrained DoF list
  int n_cons = 0;
  int32_t *d_orphan = nullptr;      // DoFs of nodes that no cell references (zero rows)
  int n_orphan = 0;
  int *d_bad = nullptr;
  int32_t *d_perm = nullptr;        // sorted position -> mesh cell
  double *adj_dmu = nullptr, *adj_dlambda = nullptr;  // outputs of the running residual VJP
  // tiled layout (structured meshes): tiles sorted by colour, tiles of colour
  // k are [tile_ptr[k], tile_ptr[k+1]); each tile's cells are consecutive
  // sorted cells
  bool tiled = false;
  int tile_kind = 0;  // TILE argument of the cell kernels: 0 cells, 1 tiles, 2 wide tiles
  std::vector<int> tile_ptr;
  uint32_t *d_tile_nodes = nullptr;  // [n_tiles][TN] node | constraint bits | first-touch bit, or kNoNode
  int32_t *d_tile_info = nullptr;    // [n_tiles][2] first sorted cell, extents ex | ey << 8 | ez << 16
  // node (hanging) constraints: cells layout, all-add scatter after a zero
  // pass.  Instances
  // (cell, constrained node) leave their results in d_hbuf; the masters
  // gather them after the colours (d_hg_*: per master its instances and
  // weights, in instance order), so masters do not enter the cell colouring.
  int n_node_cons = 0;
  int32_t *d_hang = nullptr, *d_hptr = nullptr, *d_hcons = nullptr;
  uint32_t *d_hmaster = nullptr;
  void *d_hw = nullptr;
  int n_hang_inst = 0, n_hang_masters = 0;
  uint32_t *d_hg_node = nullptr;  // master node | its Dirichlet bits
  int32_t *d_hg_ptr = nullptr, *d_hg_inst = nullptr;
  void *d_hg_w = nullptr;
  void *d_hbuf = nullptr;  // kHangBatch vectors of n_hang_inst * dim values
  // closed-form diagonal with node constraints (VmultParams::dcell ...):
  // cells with constrained nodes evaluate one entry per node they couple;
  // d_dg_*: per node its entries, in entry order, for the gather
  int32_t *d_dcell = nullptr, *d_deptr = nullptr, *d_dtptr = nullptr, *d_dta = nullptr;
  void *d_dtw = nullptr, *d_dbuf = nullptr;
  int n_diag_nodes = 0;
  uint32_t *d_dg_node = nullptr;
  int32_t *d_dg_ptr = nullptr, *d_dg_ent = nullptr;
  bool material_set = false, linearization_set = false, cached = false, diag_ok = false;
  size_t bytes_idx = 0, bytes_cache = 0;
};

namespace lsmf {

size_t elem_size(const lsmf_op *op);

// launch entry points of one (dim, N) instantiation unit
struct LaunchTable {
  int (*vmult)(const lsmf_op *, void *dst, const void *src, bool add, bool constr, unsigned ph, hipStream_t);
  int (*vmult_batched)(const lsmf_op *, void *dst, const void *src, int nvec, hipStream_t);
  int (*cache)(const lsmf_op *, int *bad, hipStream_t);  // bad: det F <= 0 counter (zeroed by the caller)
  int (*diagonal)(const lsmf_op *, hipStream_t);
  int (*residual)(const lsmf_op *, double *rhs, hipStream_t);
  int (*residual_vjp)(const lsmf_op *, double *du, const double *w, hipStream_t);
};

const LaunchTable *launch_table(int dim, int N);
int n_cache_fields(int dim, int caching);

}  // namespace lsmf
