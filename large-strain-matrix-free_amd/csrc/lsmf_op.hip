// lsmf_op.hip -- C ABI of the operator (include/lsmf.h): creation,
// colouring, index tables, material, linearisation, cache, diagonal, vmult,
// residual.  Host logic only; kernels live in lsmf_kernels.hpp and are
// instantiated per (dim, N) in lsmf_inst_*.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "lsmf.h"
#include "lsmf_error.hpp"
#include "lsmf_fe.hpp"
#include "lsmf_kernels.hpp"
#include "lsmf_op.hpp"

namespace lsmf {

std::string &last_error_string() {
  static thread_local std::string s;
  return s;
}

const LaunchTable *launch_table_2_2();
const LaunchTable *launch_table_2_3();
const LaunchTable *launch_table_2_5();
const LaunchTable *launch_table_3_2();
const LaunchTable *launch_table_3_3();
const LaunchTable *launch_table_3_5();

const LaunchTable *launch_table(int dim, int N) {
  if (dim == 2 && N == 2) return launch_table_2_2();
  if (dim == 2 && N == 3) return launch_table_2_3();
  if (dim == 2 && N == 5) return launch_table_2_5();
  if (dim == 3 && N == 2) return launch_table_3_2();
  if (dim == 3 && N == 3) return launch_table_3_3();
  if (dim == 3 && N == 5) return launch_table_3_5();
  return nullptr;
}

size_t elem_size(const lsmf_op *op) { return op->precision == LSMF_FP64 ? sizeof(double) : sizeof(float); }

int n_cache_fields(int dim, int caching) {
  const int nv = dim == 3 ? 6 : 3;
  switch (caching) {
    case LSMF_CACHING_TENSOR4: return dim * dim + nv + nv * (nv + 1) / 2;
    case LSMF_CACHING_TENSOR2: return dim * dim + nv + 2;
    case LSMF_CACHING_TENSOR4_NS: return dim * dim + dim * dim * (dim * dim + 1) / 2;
    case LSMF_CACHING_SCALAR: return 1;
    case LSMF_CACHING_SCALAR_REFERENTIAL: return 1 + dim;
    default: return -1;
  }
}

namespace {

#define LSMF_HIP_TRY(expr, what)                                                                   \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess)                                                                          \
      return fail(e_ == hipErrorOutOfMemory ? LSMF_ERR_OUT_OF_MEMORY : LSMF_ERR_HIP,               \
                  std::string(what) + ": " + hipGetErrorString(e_));                               \
  } while (0)

template <typename T> int dalloc(T **p, size_t n, const char *what) {
  *p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void **)p, n * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(e == hipErrorOutOfMemory ? LSMF_ERR_OUT_OF_MEMORY : LSMF_ERR_HIP,
                std::string("hipMalloc(") + what + "): " + hipGetErrorString(e));
  }
  return 0;
}

void free_op(lsmf_op *op) {
  void *ptrs[] = {op->d_idx,  op->d_coords,    op->d_tab,  op->d_tab64, op->d_mu64,     op->d_lambda64,
                  op->d_mu,   op->d_lambda,    op->d_cache, op->d_diag, op->d_inv_diag, op->d_cons,
                  op->d_bad,  op->d_perm,      op->d_tile_nodes, op->d_tile_info, op->d_orphan, op->d_hang, op->d_hptr,
                  op->d_hmaster, op->d_hw, op->d_hcons, op->d_hg_node, op->d_hg_ptr, op->d_hg_inst, op->d_hg_w,
                  op->d_hbuf, op->d_dcell, op->d_deptr, op->d_dtptr, op->d_dta, op->d_dtw, op->d_dbuf,
                  op->d_dg_node, op->d_dg_ptr, op->d_dg_ent};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (op->d_u64) (void)hipFree(op->d_u64);
  if (op->d_u && op->d_u != op->d_u64) (void)hipFree(op->d_u);
}

__global__ void k_zero_int(int *p) { *p = 0; }

// The lexicographic structured numbering of lsmf_mesh_create:
// node(ix, iy, iz) = ix + Nx (iy + Ny iz), cell (cx, cy, cz) = cx + nx (cy + ny cz),
// local node a = a0 + N (a1 + N a2) at (cx p + a0, cy p + a1, cz p + a2).
// Returns the cells per direction when cell_nodes is exactly that.
bool structured_dims(const lsmf_mesh_desc *m, int dim, int p, int nqp, long long n[3]) {
  const int N = p + 1;
  const int32_t *c0 = m->cell_nodes;
  if (c0[0] != 0 || c0[1] != 1) return false;
  const long long Nx = c0[N];
  if (Nx < N || (Nx - 1) % p) return false;
  long long Ny, Nz = 1;
  if (dim == 3) {
    const long long NxNy = c0[N * N];
    if (NxNy % Nx) return false;
    Ny = NxNy / Nx;
    if (m->n_nodes % (Nx * Ny)) return false;
    Nz = m->n_nodes / (Nx * Ny);
  } else {
    if (m->n_nodes % Nx) return false;
    Ny = m->n_nodes / Nx;
  }
  if ((Ny - 1) % p || (Nz - 1) % p || Ny < N || (dim == 3 && Nz < N)) return false;
  n[0] = (Nx - 1) / p;
  n[1] = (Ny - 1) / p;
  n[2] = dim == 3 ? (Nz - 1) / p : 1;
  if (n[0] * n[1] * n[2] != m->n_cells) return false;
  for (long long cz = 0; cz < n[2]; ++cz)
    for (long long cy = 0; cy < n[1]; ++cy)
      for (long long cx = 0; cx < n[0]; ++cx) {
        const int32_t *cn = c0 + (size_t)(cx + n[0] * (cy + n[1] * cz)) * nqp;
        for (int a = 0; a < nqp; ++a) {
          const int a0 = a % N, a1 = (a / N) % N, a2 = dim == 3 ? a / (N * N) : 0;
          if (cn[a] != (cx * p + a0) + Nx * ((cy * p + a1) + Ny * (cz * p + a2))) return false;
        }
      }
  return true;
}

// Cell grid of a logically structured mesh: cell_at[cx + n0 (cy + n1 cz)]
// is the mesh cell at grid position (cx, cy, cz), recovered from the cells'
// shared faces whatever the node and cell numbering (a box renumbered by
// deal.II's Cuthill-McKee, include/mf_elasticity.h:1610-1708, tiles like
// one in lsmf_mesh_create's lexicographic numbering, which is checked
// first without any search).  Every cell must be in standard orientation:
// the neighbour across local face +d has this cell across its local face
// -d.  Returns false for anything else (the cells layout then serves it).
bool cell_grid(const lsmf_mesh_desc *m, int dim, int p, int nqp, long long n[3], std::vector<int32_t> &cell_at) {
  const int nc = m->n_cells;
  if (structured_dims(m, dim, p, nqp, n)) {
    cell_at.resize(nc);
    for (int c = 0; c < nc; ++c) cell_at[c] = c;
    return true;
  }
  const int N = p + 1, nf = 2 * dim, ncorner = 1 << (dim - 1);
  // local index of corner k of face f = 2 d + s
  auto corner = [&](int f, int k) {
    const int d = f / 2, s = f % 2;
    int bits[3] = {0, 0, 0}, b = 0;
    for (int e = 0; e < dim; ++e) bits[e] = e == d ? s : (k >> b++) & 1;
    int a = 0, stride = 1;
    for (int e = 0; e < dim; ++e, stride *= N) a += bits[e] * p * stride;
    return a;
  };
  struct Face {
    int32_t v[4];
    int64_t id;  // cell * nf + face
  };
  std::vector<Face> faces((size_t)nc * nf);
  for (int c = 0; c < nc; ++c)
    for (int f = 0; f < nf; ++f) {
      Face &F = faces[(size_t)c * nf + f];
      F.v[2] = F.v[3] = -1;
      for (int k = 0; k < ncorner; ++k) F.v[k] = m->cell_nodes[(size_t)c * nqp + corner(f, k)];
      std::sort(F.v, F.v + ncorner);
      F.id = (int64_t)c * nf + f;
    }
  std::sort(faces.begin(), faces.end(), [](const Face &a, const Face &b) {
    return std::lexicographical_compare(a.v, a.v + 4, b.v, b.v + 4);
  });
  std::vector<int32_t> nbr((size_t)nc * nf, -1);
  for (size_t i = 0; i < faces.size();) {
    size_t j = i + 1;
    while (j < faces.size() && std::equal(faces[i].v, faces[i].v + 4, faces[j].v)) ++j;
    if (j - i > 2) return false;
    if (j - i == 2) {
      const int64_t a = faces[i].id, b = faces[i + 1].id;
      const int fa = (int)(a % nf), fb = (int)(b % nf);
      if ((fa ^ 1) != fb) return false;  // not face +d against face -d
      nbr[a] = (int32_t)(b / nf);
      nbr[b] = (int32_t)(a / nf);
    }
    i = j;
  }
  // the one cell without -x, -y, -z neighbours is the grid's origin
  int origin = -1;
  for (int c = 0; c < nc; ++c) {
    bool o = true;
    for (int d = 0; d < dim; ++d) o = o && nbr[(size_t)c * nf + 2 * d] < 0;
    if (o) {
      if (origin >= 0) return false;
      origin = c;
    }
  }
  if (origin < 0) return false;
  n[0] = n[1] = n[2] = 1;
  for (int d = 0; d < dim; ++d) {
    long long len = 1;
    for (int c = origin; nbr[(size_t)c * nf + 2 * d + 1] >= 0; c = nbr[(size_t)c * nf + 2 * d + 1])
      if (++len > nc) return false;
    n[d] = len;
  }
  if (n[0] * n[1] * n[2] != nc) return false;
  cell_at.assign(nc, -1);
  std::vector<uint8_t> seen(nc, 0);
  auto at = [&](long long x, long long y, long long z) -> int32_t & { return cell_at[x + n[0] * (y + n[1] * z)]; };
  for (long long z = 0; z < n[2]; ++z)
    for (long long y = 0; y < n[1]; ++y)
      for (long long x = 0; x < n[0]; ++x) {
        int32_t c = x > 0   ? nbr[(size_t)at(x - 1, y, z) * nf + 1]
                    : y > 0 ? nbr[(size_t)at(x, y - 1, z) * nf + 3]
                    : z > 0 ? nbr[(size_t)at(x, y, z - 1) * nf + 5]
                            : origin;
        if (c < 0 || seen[c]) return false;
        seen[c] = 1;
        at(x, y, z) = c;
      }
  // every face link agrees with the grid
  for (long long z = 0; z < n[2]; ++z)
    for (long long y = 0; y < n[1]; ++y)
      for (long long x = 0; x < n[0]; ++x) {
        const long long g[3] = {x, y, z};
        const int32_t c = at(x, y, z);
        for (int d = 0; d < dim; ++d)
          for (int s = 0; s < 2; ++s) {
            long long h[3] = {g[0], g[1], g[2]};
            h[d] += s ? 1 : -1;
            const int32_t expect = (h[d] < 0 || h[d] >= n[d]) ? -1 : at(h[0], h[1], h[2]);
            if (nbr[(size_t)c * nf + 2 * d + s] != expect) return false;
          }
      }
  return true;
}

// Some support point lies off its cell's Q1 (multi)linear map of the
// vertices (beyond round-off relative to the cell size): the current
// configuration then needs the Eulerian Q_p metric of all support points.
bool support_points_curved(const lsmf_mesh_desc *m, int dim, int p, int nqp) {
  const int N = p + 1, NV = 1 << dim;
  const std::vector<double> g = lobatto_points(p);
  std::vector<double> wq1((size_t)nqp * NV);  // Q1 weight of vertex v at support point a
  std::vector<int> corner(NV);
  for (int v = 0; v < NV; ++v) {
    int a = 0, stride = 1;
    for (int d = 0; d < dim; ++d) {
      a += ((v >> d) & 1) * p * stride;
      stride *= N;
    }
    corner[v] = a;
  }
  for (int a = 0; a < nqp; ++a) {
    int rem = a;
    double xi[3] = {0, 0, 0};
    for (int d = 0; d < dim; ++d) {
      xi[d] = g[rem % N];
      rem /= N;
    }
    for (int v = 0; v < NV; ++v) {
      double w = 1.0;
      for (int d = 0; d < dim; ++d) w *= ((v >> d) & 1) ? xi[d] : 1.0 - xi[d];
      wq1[(size_t)a * NV + v] = w;
    }
  }
  for (int c = 0; c < m->n_cells; ++c) {
    const int32_t *cn = m->cell_nodes + (size_t)c * nqp;
    double X[8][3], h = 0.0;
    for (int v = 0; v < NV; ++v)
      for (int i = 0; i < dim; ++i) {
        X[v][i] = m->coords[(size_t)cn[corner[v]] * dim + i];
        h = std::max(h, std::fabs(X[v][i] - X[0][i]));
      }
    for (int a = 0; a < nqp; ++a)
      for (int i = 0; i < dim; ++i) {
        double x = 0.0;
        for (int v = 0; v < NV; ++v) x += wq1[(size_t)a * NV + v] * X[v][i];
        if (std::fabs(x - m->coords[(size_t)cn[a] * dim + i]) > 1e-10 * h) return true;
      }
  }
  return false;
}

struct TileDims {
  int tx, ty, tz, nx, ny, nz, tn;
};
template <int D, int NN, int WG> TileDims tile_dims_of() {
  using TS = TileShape<D, NN, WG>;
  return {TS::TX, TS::TY, TS::TZ, TS::NX, TS::NY, TS::NZ, TS::TN};
}
bool tile_dims(int dim, int N, int kind, TileDims &t) {
  if (dim == 2 && N == 2) t = tile_dims_of<2, 2, kWG>();
  else if (dim == 2 && N == 3) t = tile_dims_of<2, 3, kWG>();
  else if (dim == 2 && N == 5) t = tile_dims_of<2, 5, kWG>();
  else if (dim == 3 && N == 2) t = tile_dims_of<3, 2, kWG>();
  else if (dim == 3 && N == 3 && kind == kTileWide) t = tile_dims_of<3, 3, 2 * kWG>();
  else if (dim == 3 && N == 3) t = tile_dims_of<3, 3, kWG>();
  else if (dim == 3 && N == 5 && kind == kTileWide) t = tile_dims_of<3, 5, 2 * kWG>();
  else if (dim == 3 && N == 5) t = tile_dims_of<3, 5, kWG>();
  else return false;
  return true;
}

// Tiles of a logically structured mesh (cell grid cell_at, n cells per
// direction), coloured by the parity of their position so that no two tiles
// of one colour share a node.  Fills perm (tile order, the cells of a tile
// consecutive in slot order), tile_ptr, the tile node table (node ids from
// the cells' own node lists; first-touch bit: the node's first tile in
// colour order) and tile info.  Returns kNotTileable when cells that share
// a tile node position disagree about its node (the cells layout then
// serves the mesh).
constexpr int kNotTileable = 1;
int build_tiles(lsmf_op *op, const lsmf_mesh_desc *mesh, const long long n[3], const std::vector<int32_t> &cell_at,
                std::vector<uint32_t> &tnodes, std::vector<int32_t> &tinfo) {
  const int dim = op->dim, p = op->degree, N = p + 1, nqp = op->nqp;
  TileDims td;
  if (!tile_dims(dim, op->N, op->tile_kind, td)) return fail(LSMF_ERR_UNSUPPORTED, "tiles: no tile shape");
  const long long TXn = (n[0] + td.tx - 1) / td.tx, TYn = (n[1] + td.ty - 1) / td.ty,
                  TZn = (n[2] + td.tz - 1) / td.tz;
  const long long n_tiles = TXn * TYn * TZn;
  const int n_col = 1 << dim;
  std::vector<long long> order;
  order.reserve(n_tiles);
  op->tile_ptr.assign(n_col + 1, 0);
  for (int col = 0; col < n_col; ++col) {
    for (long long bz = 0; bz < TZn; ++bz)
      for (long long by = 0; by < TYn; ++by)
        for (long long bx = 0; bx < TXn; ++bx)
          if (((bx & 1) | ((by & 1) << 1) | ((bz & 1) << 2)) == col) order.push_back(bx + TXn * (by + TYn * bz));
    op->tile_ptr[col + 1] = (int)order.size();
  }
  op->perm.assign(op->n_cells, 0);
  op->inv_perm.assign(op->n_cells, 0);
  tinfo.assign((size_t)n_tiles * 2, 0);
  tnodes.assign((size_t)n_tiles * td.tn, kNoNode);
  std::vector<int> min_col(mesh->n_nodes, 1 << 30), stamp(mesh->n_nodes, -1);
  int pos = 0;
  for (int col = 0; col < n_col; ++col)
    for (int ti = op->tile_ptr[col]; ti < op->tile_ptr[col + 1]; ++ti) {
      const long long b = order[ti];
      const long long bx = b % TXn, by = (b / TXn) % TYn, bz = b / (TXn * TYn);
      const int ex = (int)std::min<long long>(td.tx, n[0] - bx * td.tx);
      const int ey = (int)std::min<long long>(td.ty, n[1] - by * td.ty);
      const int ez = (int)std::min<long long>(td.tz, n[2] - bz * td.tz);
      tinfo[(size_t)ti * 2] = pos;
      tinfo[(size_t)ti * 2 + 1] = ex | (ey << 8) | (ez << 16);
      auto tile_cell = [&](int sx, int sy, int sz) {
        return cell_at[(bx * td.tx + sx) + n[0] * ((by * td.ty + sy) + n[1] * (bz * td.tz + sz))];
      };
      for (int sz = 0; sz < ez; ++sz)
        for (int sy = 0; sy < ey; ++sy)
          for (int sx = 0; sx < ex; ++sx) {
            const int cell = tile_cell(sx, sy, sz);
            op->perm[pos] = cell;
            op->inv_perm[cell] = pos;
            ++pos;
          }
      for (int l = 0; l < td.tn; ++l) {
        const int X = l % td.nx, Y = (l / td.nx) % td.ny, Z = l / (td.nx * td.ny);
        if (X > ex * p || Y > ey * p || Z > ez * p) continue;  // beyond a ragged tile
        // the node from every cell of the tile that holds this position
        // (the kernels sum those cells' contributions into it)
        const int xs[2] = {X == 0 ? 0 : (X - 1) / p, std::min(X / p, ex - 1)};
        const int ys[2] = {Y == 0 ? 0 : (Y - 1) / p, std::min(Y / p, ey - 1)};
        const int zs[2] = {Z == 0 ? 0 : (Z - 1) / p, std::min(Z / p, ez - 1)};
        long long node = -1;
        for (int cz = zs[0]; cz <= zs[1]; ++cz)
          for (int cy = ys[0]; cy <= ys[1]; ++cy)
            for (int cx = xs[0]; cx <= xs[1]; ++cx) {
              const int a = (X - cx * p) + N * ((Y - cy * p) + (dim == 3 ? N * (Z - cz * p) : 0));
              const long long nd = mesh->cell_nodes[(size_t)tile_cell(cx, cy, cz) * nqp + a];
              if (node >= 0 && nd != node) return kNotTileable;
              node = nd;
            }
        if (stamp[node] == col) return kNotTileable;  // two tiles of one colour share a node
        stamp[node] = col;
        min_col[node] = std::min(min_col[node], col);
        tnodes[(size_t)ti * td.tn + l] = (uint32_t)node;
      }
    }
  if (pos != op->n_cells) return fail(LSMF_ERR_INVALID_ARGUMENT, "tiles: cell count mismatch");
  for (int col = 0; col < n_col; ++col)
    for (int ti = op->tile_ptr[col]; ti < op->tile_ptr[col + 1]; ++ti)
      for (int l = 0; l < td.tn; ++l) {
        uint32_t &code = tnodes[(size_t)ti * td.tn + l];
        if (code == kNoNode) continue;
        const uint32_t node = code;
        if (mesh->constrained)
          for (int comp = 0; comp < dim; ++comp)
            if (mesh->constrained[(size_t)node * dim + comp]) code |= 1u << (kConsShift + comp);
        if (min_col[node] == col) code |= kFirstBit;
      }
  return LSMF_OK;
}

__global__ void k_convert_f64_f32(const double *__restrict__ a, float *__restrict__ b, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = float(a[i]);
}

}  // namespace
}  // namespace lsmf

namespace lsmf {
namespace {
// cache(): zero the det F <= 0 counter bad_dev, run the cache kernel, read
// the count back (synchronising the stream)
lsmf_status cache_with_counter(lsmf_op *op, int *bad_dev, hipStream_t s, const char *what) {
  if (!op) return fail(LSMF_ERR_INVALID_ARGUMENT, std::string(what) + ": null operator");
  if (!op->material_set) return fail(LSMF_ERR_INVALID_ARGUMENT, std::string(what) + ": set_material was not called");
  if (!op->linearization_set)
    return fail(LSMF_ERR_INVALID_ARGUMENT, std::string(what) + ": set_linearization was not called");
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, s, bad_dev);
  int rc = launch_table(op->dim, op->N)->cache(op, bad_dev, s);
  if (rc) return fail(rc, std::string(what) + ": kernel launch failed");
  int bad = 0;
  LSMF_HIP_TRY(hipMemcpyAsync(&bad, bad_dev, sizeof(int), hipMemcpyDeviceToHost, s), "cache: read det F count");
  LSMF_HIP_TRY(hipStreamSynchronize(s), "cache: synchronize");
  if (bad > 0) {
    op->cached = false;
    return fail(LSMF_ERR_NONPOSITIVE_DET_F,
                "det_F is not positive at " + std::to_string(bad) + " quadrature points");
  }
  op->cached = true;
  op->diag_ok = false;
  return LSMF_OK;
}

// Scratch of lsmf_cache_noalloc: a file-scope device array (each GPU holds
// its own copy for the life of the process), reached through the device
// address hipGetSymbolAddress returns for the current device -- in host
// code the array's name is only a stand-in symbol.  kNoAllocScratchInts
// ints: [0] counts the quadrature points with det F <= 0.
constexpr int kNoAllocScratchInts = 64;
__device__ __align__(256) int scratch[kNoAllocScratchInts];

std::mutex g_scratch_mu;
std::vector<int *> g_scratch_ptr;  // per device ordinal, nullptr until looked up

lsmf_status scratch_for_current_device(int **out) {
  int dev = -1;
  LSMF_HIP_TRY(hipGetDevice(&dev), "cache_noalloc: current device");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    LSMF_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "cache_noalloc: scratch address");
    if (!p) return fail(LSMF_ERR_HIP, "cache_noalloc: null scratch address");
    g_scratch_ptr[dev] = static_cast<int *>(p);
  }
  *out = g_scratch_ptr[dev];
  return LSMF_OK;
}
}  // namespace
}  // namespace lsmf

using namespace lsmf;

extern "C" {

const char *lsmf_last_error(void) { return last_error_string().c_str(); }

lsmf_status lsmf_op_create(const lsmf_mesh_desc *mesh, const lsmf_op_options *opt, lsmf_op **out) {
  if (!mesh || !opt || !out) return fail(LSMF_ERR_INVALID_ARGUMENT, "lsmf_op_create: null argument");
  *out = nullptr;
  const int dim = mesh->dim, p = mesh->degree, nq = opt->n_q_1d;
  if (dim != 2 && dim != 3) return fail(LSMF_ERR_INVALID_ARGUMENT, "lsmf_op_create: dim must be 2 or 3");
  if (opt->precision != LSMF_FP64 && opt->precision != LSMF_FP32)
    return fail(LSMF_ERR_INVALID_ARGUMENT, "lsmf_op_create: unknown precision");
  if (n_cache_fields(dim, opt->caching) < 0)
    return fail(LSMF_ERR_UNKNOWN_CACHING, "Unknown caching");
  if (nq != p + 1 || !launch_table(dim, nq))
    return fail(LSMF_ERR_UNSUPPORTED, "lsmf_op_create: (dim, degree, n_q_1d) = (" + std::to_string(dim) + ", " +
                                          std::to_string(p) + ", " + std::to_string(nq) + ") is not instantiated");
  if (mesh->n_cells <= 0 || mesh->n_nodes <= 0 || !mesh->cell_nodes || !mesh->coords)
    return fail(LSMF_ERR_INVALID_ARGUMENT, "lsmf_op_create: empty mesh");
  if (mesh->n_nodes >= (1 << 28)) return fail(LSMF_ERR_INVALID_ARGUMENT, "lsmf_op_create: more than 2^28 nodes");
  const int N = nq;
  int nqp = 1;
  for (int d = 0; d < dim; ++d) nqp *= N;
  const int nc = mesh->n_cells;
  for (long long i = 0; i < (long long)nc * nqp; ++i)
    if (mesh->cell_nodes[i] < 0 || mesh->cell_nodes[i] >= mesh->n_nodes)
      return fail(LSMF_ERR_INVALID_ARGUMENT, "lsmf_op_create: cell_nodes entry out of range");

  lsmf_op *op = new (std::nothrow) lsmf_op;
  if (!op) return fail(LSMF_ERR_OUT_OF_MEMORY, "lsmf_op_create: host allocation");
  op->dim = dim;
  op->degree = p;
  op->nq = nq;
  op->N = N;
  op->nqp = nqp;
  op->precision = opt->precision;
  op->caching = opt->caching;
  op->n_cells = nc;
  op->n_nodes = mesh->n_nodes;
  op->n_dofs = (size_t)mesh->n_nodes * dim;
  op->n_fields = n_cache_fields(dim, opt->caching);
  op->field_stride = (size_t)nc * nqp;

  // --- node (hanging) constraints: validated, closed (masters are free)
  const int nnc = mesh->n_node_constraints;
  std::vector<int> hang_of;
  if (nnc < 0 || (nnc > 0 && (!mesh->nc_nodes || !mesh->nc_ptr || !mesh->nc_masters || !mesh->nc_weights))) {
    delete op;
    return fail(LSMF_ERR_INVALID_ARGUMENT, "lsmf_op_create: bad node constraint arrays");
  }
  if (nnc > 0) {
    hang_of.assign(mesh->n_nodes, -1);
    if (mesh->nc_ptr[0] != 0) {
      delete op;
      return fail(LSMF_ERR_INVALID_ARGUMENT, "lsmf_op_create: nc_ptr[0] must be 0");
    }
    for (int k = 0; k < nnc; ++k) {
      const int n = mesh->nc_nodes[k];
      if (n < 0 || n >= mesh->n_nodes || hang_of[n] >= 0 || mesh->nc_ptr[k + 1] < mesh->nc_ptr[k]) {
        delete op;
        return fail(LSMF_ERR_INVALID_ARGUMENT, "lsmf_op_create: bad constrained node or nc_ptr");
      }
      hang_of[n] = k;
      if (mesh->constrained)
        for (int c = 0; c < dim; ++c)
          if (mesh->constrained[(size_t)n * dim + c]) {
            delete op;
            return fail(LSMF_ERR_INVALID_ARGUMENT, "lsmf_op_create: a node is both constrained and Dirichlet");
          }
    }
    for (int j = 0; j < mesh->nc_ptr[nnc]; ++j) {
      const int m = mesh->nc_masters[j];
      if (m < 0 || m >= mesh->n_nodes || hang_of[m] >= 0) {
        delete op;
        return fail(LSMF_ERR_INVALID_ARGUMENT, "lsmf_op_create: masters must be unconstrained nodes");
      }
    }
  }
  op->n_node_cons = nnc;
  // nodes cell c writes in its colour: its own nodes (the masters of its
  // constrained ones receive their share after the colours)
  auto write_nodes = [&](int c, std::vector<int> &w) {
    w.clear();
    for (int a = 0; a < nqp; ++a) w.push_back(mesh->cell_nodes[(size_t)c * nqp + a]);
  };
  std::vector<int> wbuf;

  // --- colouring (given, checked below, or greedy first fit)
  std::vector<int> colour(nc, 0);
  if (mesh->cell_colour) {
    for (int c = 0; c < nc; ++c) {
      colour[c] = mesh->cell_colour[c];
      if (colour[c] < 0 || colour[c] > 63) {
        delete op;
        return fail(LSMF_ERR_INVALID_ARGUMENT, "lsmf_op_create: colour out of [0, 63]");
      }
    }
  } else {
    std::vector<uint64_t> used(mesh->n_nodes, 0);
    for (int c = 0; c < nc; ++c) {
      write_nodes(c, wbuf);
      uint64_t m = 0;
      for (int n : wbuf) m |= used[n];
      int col = 0;
      while (col < 64 && (m >> col) & 1ull) ++col;
      if (col == 64) {
        delete op;
        return fail(LSMF_ERR_INVALID_ARGUMENT, "lsmf_op_create: greedy colouring needs more than 64 colours");
      }
      colour[c] = col;
      for (int n : wbuf) used[n] |= (1ull << col);
    }
  }
  std::vector<int> min_col;
  int n_col = 0;
  for (int c = 0; c < nc; ++c) n_col = std::max(n_col, colour[c] + 1);
  op->colour_ptr.assign(n_col + 1, 0);
  for (int c = 0; c < nc; ++c) op->colour_ptr[colour[c] + 1]++;
  for (int k = 0; k < n_col; ++k) op->colour_ptr[k + 1] += op->colour_ptr[k];
  op->perm.assign(nc, 0);
  op->inv_perm.assign(nc, 0);
  {
    std::vector<int> fill(op->colour_ptr.begin(), op->colour_ptr.end() - 1);
    for (int c = 0; c < nc; ++c) {
      const int pos = fill[colour[c]]++;
      op->perm[pos] = c;
      op->inv_perm[c] = pos;
    }
  }
  // colour conflict check: no node written by two cells of one colour; min
  // colour per node
  std::vector<int> stamp(mesh->n_nodes, -1), stamp_cell(mesh->n_nodes, -1);
  min_col.assign(mesh->n_nodes, 1 << 30);
  for (int k = 0; k < n_col; ++k)
    for (int pos = op->colour_ptr[k]; pos < op->colour_ptr[k + 1]; ++pos) {
      const int c = op->perm[pos];
      write_nodes(c, wbuf);
      for (int nd : wbuf) {
        if (stamp[nd] == k && stamp_cell[nd] != c) {
          delete op;
          return fail(LSMF_ERR_INVALID_ARGUMENT, "lsmf_op_create: two cells of one colour share a node");
        }
        stamp[nd] = k;
        stamp_cell[nd] = c;
        min_col[nd] = std::min(min_col[nd], k);
      }
    }
  op->curved = p > 1 && support_points_curved(mesh, dim, p, nqp);
  // --- cell order of the tiled layout (structured meshes) replaces the
  // colour order; the per-cell index table then carries no first-touch bits
  std::vector<uint32_t> tnodes;
  std::vector<int32_t> tinfo;
  long long sn[3] = {1, 1, 1};
  std::vector<int32_t> cell_at;
  op->tiled = opt->layout == LSMF_LAYOUT_AUTO && nnc == 0 && cell_grid(mesh, dim, p, nqp, sn, cell_at);
  // wide (512-thread) tiles for 3D Q2 / Q4 tensor2 and tensor4 products
  op->tile_kind = !op->tiled ? 0
                  : (dim == 3 && (op->N == 5 || op->N == 3) &&
                     (op->caching == LSMF_CACHING_TENSOR2 || op->caching == LSMF_CACHING_TENSOR4))
                      ? kTileWide
                      : 1;
  if (op->tiled) {
    const std::vector<int> perm0 = op->perm, inv0 = op->inv_perm;
    const int rc = build_tiles(op, mesh, sn, cell_at, tnodes, tinfo);
    if (rc == kNotTileable) {  // the colour order of the cells layout stays
      op->tiled = false;
      op->tile_kind = 0;
      op->perm = perm0;
      op->inv_perm = inv0;
      op->tile_ptr.clear();
      tnodes.clear();
      tinfo.clear();
    } else if (rc) {
      delete op;
      return rc;
    }
  }
  // --- index table (sorted cells): node | constraint bits | first-touch bit
  std::vector<uint32_t> idx((size_t)nc * nqp);
  for (int pos = 0; pos < nc; ++pos) {
    const int c = op->perm[pos];
    for (int a = 0; a < nqp; ++a) {
      const int nd = mesh->cell_nodes[(size_t)c * nqp + a];
      uint32_t code = (uint32_t)nd;
      if (mesh->constrained)
        for (int comp = 0; comp < dim; ++comp)
          if (mesh->constrained[(size_t)nd * dim + comp]) code |= 1u << (kConsShift + comp);
      if (!op->tiled && nnc == 0 && colour[c] == min_col[nd]) code |= kFirstBit;
      idx[(size_t)pos * nqp + a] = code;
    }
  }
  std::vector<int32_t> cons;
  if (mesh->constrained)
    for (size_t i = 0; i < op->n_dofs; ++i)
      if (mesh->constrained[i]) cons.push_back((int32_t)i);
  // constrained nodes' rows act as identity, like Dirichlet rows
  for (int k = 0; k < nnc; ++k)
    for (int c = 0; c < dim; ++c) cons.push_back(mesh->nc_nodes[k] * dim + c);
  op->n_cons = (int)cons.size();
  // DoFs of nodes that no cell references: their rows are zero, and the
  // first-touch stores never write them
  std::vector<int32_t> orphan;
  {
    std::vector<uint8_t> used(mesh->n_nodes, 0);
    for (size_t i = 0; i < (size_t)nc * nqp; ++i) used[mesh->cell_nodes[i]] = 1;
    for (int nd = 0; nd < mesh->n_nodes; ++nd)
      if (!used[nd])
        for (int c = 0; c < dim; ++c) orphan.push_back(nd * dim + c);
  }
  op->n_orphan = (int)orphan.size();

  // --- node-constraint tables: instances and their masters (products), and
  // the per-entry tables of the closed-form diagonal
  std::vector<int32_t> hang_tab, hptr;
  std::vector<uint32_t> hmaster;
  std::vector<double> hw;
  std::vector<int32_t> dcell, deptr, dtptr, dta, dg_ptr, dg_ent, dent_node;
  std::vector<double> dtw;
  std::vector<uint32_t> dg_node;
  std::vector<int32_t> hcons, hg_ptr, hg_inst;
  std::vector<uint32_t> hg_node;
  std::vector<double> hg_w;
  if (nnc > 0) {
    // instances in sorted-cell order
    hang_tab.assign((size_t)nc * nqp, -1);
    for (int pos = 0; pos < nc; ++pos)
      for (int a = 0; a < nqp; ++a) {
        const int k = hang_of[mesh->cell_nodes[(size_t)op->perm[pos] * nqp + a]];
        if (k < 0) continue;
        hang_tab[(size_t)pos * nqp + a] = (int32_t)hcons.size();
        hcons.push_back(k);
      }
    hptr.assign(mesh->nc_ptr, mesh->nc_ptr + nnc + 1);
    for (int j = 0; j < mesh->nc_ptr[nnc]; ++j) {
      const int m = mesh->nc_masters[j];
      uint32_t code = (uint32_t)m;
      if (mesh->constrained)
        for (int c = 0; c < dim; ++c)
          if (mesh->constrained[(size_t)m * dim + c]) code |= 1u << (kConsShift + c);
      hmaster.push_back(code);
      hw.push_back(mesh->nc_weights[j]);
    }
    // per master: (instance, weight) in instance order
    std::vector<std::array<int64_t, 3>> ent;  // master, instance, constraint entry
    for (int i = 0; i < (int)hcons.size(); ++i)
      for (int j = mesh->nc_ptr[hcons[i]]; j < mesh->nc_ptr[hcons[i] + 1]; ++j)
        ent.push_back({mesh->nc_masters[j], i, j});
    std::sort(ent.begin(), ent.end());
    for (size_t e = 0; e < ent.size(); ++e) {
      if (e == 0 || ent[e][0] != ent[e - 1][0]) {
        hg_ptr.push_back((int32_t)e);
        hg_node.push_back(hmaster[ent[e][2]]);
      }
      hg_inst.push_back((int32_t)ent[e][1]);
      hg_w.push_back(hw[ent[e][2]]);
    }
    hg_ptr.push_back((int32_t)ent.size());
    op->n_hang_inst = (int)hcons.size();
    op->n_hang_masters = (int)hg_node.size();
    // closed-form diagonal: per cell with constrained nodes, one entry per
    // node it couples (first appearance over the local nodes), each the
    // weighted sum of local basis functions that is that node's constrained
    // basis function on the cell; then per node its entries
    dcell.assign(nc, -1);
    deptr.push_back(0);
    dtptr.push_back(0);
    std::vector<int> eg;                               // entry -> node
    std::vector<std::vector<std::pair<int, double>>> et;  // entry -> (local node, weight)
    for (int pos = 0; pos < nc; ++pos) {
      const int c = op->perm[pos];
      bool any = false;
      for (int a = 0; a < nqp && !any; ++a) any = hang_of[mesh->cell_nodes[(size_t)c * nqp + a]] >= 0;
      if (!any) continue;
      dcell[pos] = (int32_t)(deptr.size() - 1);
      eg.clear();
      et.clear();
      auto add = [&](int g, int a, double w) {
        size_t k = 0;
        while (k < eg.size() && eg[k] != g) ++k;
        if (k == eg.size()) {
          eg.push_back(g);
          et.emplace_back();
        }
        et[k].push_back({a, w});
      };
      for (int a = 0; a < nqp; ++a) {
        const int n = mesh->cell_nodes[(size_t)c * nqp + a];
        const int k = hang_of[n];
        if (k < 0) add(n, a, 1.0);
        else
          for (int j = mesh->nc_ptr[k]; j < mesh->nc_ptr[k + 1]; ++j) add(mesh->nc_masters[j], a, mesh->nc_weights[j]);
      }
      for (size_t k = 0; k < eg.size(); ++k) {
        dent_node.push_back(eg[k]);
        for (const auto &tw : et[k]) {
          dta.push_back(tw.first);
          dtw.push_back(tw.second);
        }
        dtptr.push_back((int32_t)dta.size());
      }
      deptr.push_back((int32_t)dent_node.size());
    }
    std::vector<std::pair<int, int>> ne;  // (node, entry)
    for (int e = 0; e < (int)dent_node.size(); ++e) ne.push_back({dent_node[e], e});
    std::sort(ne.begin(), ne.end());
    for (size_t k = 0; k < ne.size(); ++k) {
      if (k == 0 || ne[k].first != ne[k - 1].first) {
        const int m = ne[k].first;
        uint32_t code = (uint32_t)m;
        if (mesh->constrained)
          for (int c = 0; c < dim; ++c)
            if (mesh->constrained[(size_t)m * dim + c]) code |= 1u << (kConsShift + c);
        dg_ptr.push_back((int32_t)k);
        dg_node.push_back(code);
      }
      dg_ent.push_back(ne[k].second);
    }
    dg_ptr.push_back((int32_t)ne.size());
    op->n_diag_nodes = (int)dg_node.size();
  }

  // --- 1D tables: S, D, St, Dt, qp, qw, and the collocation derivative on
  // the Gauss points Dc[q][k] = l_k'(x_q) (l_k: Lagrange basis on the Gauss
  // points; scalar_referential)
  Tables1D t1;
  t1.init(p, nq);
  std::vector<double> tab(5 * N * N + 2 * N);
  for (int q = 0; q < N; ++q)
    for (int k = 0; k < N; ++k) tab[4 * N * N + 2 * N + q * N + k] = lagrange_derivative(t1.qpts, k, t1.qpts[q]);
  for (int q = 0; q < N; ++q)
    for (int a = 0; a < N; ++a) {
      tab[q * N + a] = t1.S[q * N + a];
      tab[N * N + q * N + a] = t1.D[q * N + a];
      tab[2 * N * N + a * N + q] = t1.S[q * N + a];
      tab[3 * N * N + a * N + q] = t1.D[q * N + a];
    }
  for (int q = 0; q < N; ++q) {
    tab[4 * N * N + q] = t1.qpts[q];
    tab[4 * N * N + N + q] = t1.qw[q];
  }
  for (int i = 0; i < 2 * N * N; ++i) op->h_sf[i] = tab[i];

  const size_t es = elem_size(op);
  int rc = 0;
  if ((rc = dalloc(&op->d_idx, idx.size(), "index table")) ||
      (rc = dalloc(&op->d_coords, (size_t)mesh->n_nodes * dim, "coordinates")) ||
      (rc = dalloc(&op->d_tab64, tab.size(), "tables")) ||
      (rc = dalloc((char **)&op->d_tab, tab.size() * es, "tables")) ||
      (rc = dalloc(&op->d_mu64, (size_t)nc, "mu")) || (rc = dalloc(&op->d_lambda64, (size_t)nc, "lambda")) ||
      (rc = dalloc((char **)&op->d_mu, (size_t)nc * es, "mu")) ||
      (rc = dalloc((char **)&op->d_lambda, (size_t)nc * es, "lambda")) ||
      (rc = dalloc(&op->d_u64, op->n_dofs, "displacement")) ||
      (rc = dalloc((char **)&op->d_cache, (size_t)op->n_fields * op->field_stride * es, "cache")) ||
      (rc = dalloc((char **)&op->d_diag, op->n_dofs * es, "diagonal")) ||
      (rc = dalloc((char **)&op->d_inv_diag, op->n_dofs * es, "inverse diagonal")) ||
      (rc = dalloc(&op->d_cons, cons.size(), "constrained list")) || (rc = dalloc(&op->d_bad, 1, "counter")) ||
      (rc = dalloc(&op->d_orphan, orphan.size(), "unreferenced DoFs")) ||
      (rc = dalloc(&op->d_perm, (size_t)nc, "cell permutation")) ||
      (rc = dalloc(&op->d_tile_nodes, tnodes.size(), "tile nodes")) ||
      (rc = dalloc(&op->d_tile_info, tinfo.size(), "tile info")) ||
      (nnc > 0 && ((rc = dalloc(&op->d_hang, hang_tab.size(), "constraint table")) ||
                   (rc = dalloc(&op->d_hptr, hptr.size(), "constraint offsets")) ||
                   (rc = dalloc(&op->d_hmaster, hmaster.size(), "constraint masters")) ||
                   (rc = dalloc((char **)&op->d_hw, hw.size() * es, "constraint weights")) ||
                   (rc = dalloc(&op->d_hcons, hcons.size(), "constraint instances")) ||
                   (rc = dalloc(&op->d_hg_node, hg_node.size(), "master nodes")) ||
                   (rc = dalloc(&op->d_hg_ptr, hg_ptr.size(), "master offsets")) ||
                   (rc = dalloc(&op->d_hg_inst, hg_inst.size(), "master instances")) ||
                   (rc = dalloc((char **)&op->d_hg_w, hg_w.size() * es, "master weights")) ||
                   (rc = dalloc((char **)&op->d_hbuf, (size_t)kHangBatch * hcons.size() * dim * es,
                                "constrained-node results")) ||
                   (rc = dalloc(&op->d_dcell, dcell.size(), "diagonal cells")) ||
                   (rc = dalloc(&op->d_deptr, deptr.size(), "diagonal entries")) ||
                   (rc = dalloc(&op->d_dtptr, dtptr.size(), "diagonal terms")) ||
                   (rc = dalloc(&op->d_dta, dta.size(), "diagonal local nodes")) ||
                   (rc = dalloc((char **)&op->d_dtw, dtw.size() * es, "diagonal weights")) ||
                   (rc = dalloc((char **)&op->d_dbuf, dent_node.size() * dim * es, "diagonal entry results")) ||
                   (rc = dalloc(&op->d_dg_node, dg_node.size(), "diagonal nodes")) ||
                   (rc = dalloc(&op->d_dg_ptr, dg_ptr.size(), "diagonal node offsets")) ||
                   (rc = dalloc(&op->d_dg_ent, dg_ent.size(), "diagonal node entries"))))) {
    free_op(op);
    delete op;
    return rc;
  }
  if (op->precision == LSMF_FP64) op->d_u = op->d_u64;
  else if ((rc = dalloc((char **)&op->d_u, op->n_dofs * es, "displacement (fp32)"))) {
    free_op(op);
    delete op;
    return rc;
  }
  op->bytes_idx = idx.size() * sizeof(uint32_t);
  op->bytes_cache = (size_t)op->n_fields * op->field_stride * es;
  hipError_t e = hipSuccess;
  e = hipMemcpy(op->d_idx, idx.data(), idx.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(op->d_coords, mesh->coords, (size_t)mesh->n_nodes * dim * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(op->d_tab64, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    if (op->precision == LSMF_FP64) {
      e = hipMemcpy(op->d_tab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice);
    } else {
      std::vector<float> tf(tab.begin(), tab.end());
      e = hipMemcpy(op->d_tab, tf.data(), tf.size() * sizeof(float), hipMemcpyHostToDevice);
    }
  }
  if (e == hipSuccess && !cons.empty())
    e = hipMemcpy(op->d_cons, cons.data(), cons.size() * sizeof(int32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess && !orphan.empty())
    e = hipMemcpy(op->d_orphan, orphan.data(), orphan.size() * sizeof(int32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(op->d_perm, op->perm.data(), (size_t)nc * sizeof(int32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess && !tnodes.empty())
    e = hipMemcpy(op->d_tile_nodes, tnodes.data(), tnodes.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess && !tinfo.empty())
    e = hipMemcpy(op->d_tile_info, tinfo.data(), tinfo.size() * sizeof(int32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess && nnc > 0) {
    e = hipMemcpy(op->d_hang, hang_tab.data(), hang_tab.size() * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(op->d_hptr, hptr.data(), hptr.size() * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(op->d_hmaster, hmaster.data(), hmaster.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      if (op->precision == LSMF_FP64) {
        e = hipMemcpy(op->d_hw, hw.data(), hw.size() * sizeof(double), hipMemcpyHostToDevice);
      } else {
        std::vector<float> hf(hw.begin(), hw.end());
        e = hipMemcpy(op->d_hw, hf.data(), hf.size() * sizeof(float), hipMemcpyHostToDevice);
      }
    }
    if (e == hipSuccess && !hcons.empty())
      e = hipMemcpy(op->d_hcons, hcons.data(), hcons.size() * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess && !hg_node.empty())
      e = hipMemcpy(op->d_hg_node, hg_node.data(), hg_node.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(op->d_hg_ptr, hg_ptr.data(), hg_ptr.size() * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess && !hg_inst.empty())
      e = hipMemcpy(op->d_hg_inst, hg_inst.data(), hg_inst.size() * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess && !hg_w.empty()) {
      if (op->precision == LSMF_FP64) {
        e = hipMemcpy(op->d_hg_w, hg_w.data(), hg_w.size() * sizeof(double), hipMemcpyHostToDevice);
      } else {
        std::vector<float> wf(hg_w.begin(), hg_w.end());
        e = hipMemcpy(op->d_hg_w, wf.data(), wf.size() * sizeof(float), hipMemcpyHostToDevice);
      }
    }
    auto up = [&](void *dst, const void *src, size_t bytes) {
      if (e == hipSuccess && bytes > 0) e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
    };
    up(op->d_dcell, dcell.data(), dcell.size() * sizeof(int32_t));
    up(op->d_deptr, deptr.data(), deptr.size() * sizeof(int32_t));
    up(op->d_dtptr, dtptr.data(), dtptr.size() * sizeof(int32_t));
    up(op->d_dta, dta.data(), dta.size() * sizeof(int32_t));
    up(op->d_dg_node, dg_node.data(), dg_node.size() * sizeof(uint32_t));
    up(op->d_dg_ptr, dg_ptr.data(), dg_ptr.size() * sizeof(int32_t));
    up(op->d_dg_ent, dg_ent.data(), dg_ent.size() * sizeof(int32_t));
    if (op->precision == LSMF_FP64) {
      up(op->d_dtw, dtw.data(), dtw.size() * sizeof(double));
    } else {
      std::vector<float> wf(dtw.begin(), dtw.end());
      up(op->d_dtw, wf.data(), wf.size() * sizeof(float));
    }
  }
  op->bytes_idx += tnodes.size() * sizeof(uint32_t) + tinfo.size() * sizeof(int32_t);
  if (e != hipSuccess) {
    free_op(op);
    delete op;
    return fail(LSMF_ERR_HIP, std::string("lsmf_op_create upload: ") + hipGetErrorString(e));
  }
  *out = op;
  return LSMF_OK;
}

lsmf_status lsmf_op_destroy(lsmf_op *op) {
  if (!op) return LSMF_OK;
  delete op->prof;
  free_op(op);
  delete op;
  return LSMF_OK;
}

lsmf_status lsmf_set_material_per_cell(lsmf_op *op, const double *mu, const double *lambda) {
  if (!op || !mu || !lambda) return fail(LSMF_ERR_INVALID_ARGUMENT, "set_material: null argument");
  const int nc = op->n_cells;
  std::vector<double> m(nc), l(nc);
  for (int c = 0; c < nc; ++c) {
    m[op->inv_perm[c]] = mu[c];
    l[op->inv_perm[c]] = lambda[c];
  }
  LSMF_HIP_TRY(hipMemcpy(op->d_mu64, m.data(), nc * sizeof(double), hipMemcpyHostToDevice), "set_material");
  LSMF_HIP_TRY(hipMemcpy(op->d_lambda64, l.data(), nc * sizeof(double), hipMemcpyHostToDevice), "set_material");
  if (op->precision == LSMF_FP64) {
    LSMF_HIP_TRY(hipMemcpy(op->d_mu, m.data(), nc * sizeof(double), hipMemcpyHostToDevice), "set_material");
    LSMF_HIP_TRY(hipMemcpy(op->d_lambda, l.data(), nc * sizeof(double), hipMemcpyHostToDevice), "set_material");
  } else {
    std::vector<float> mf(m.begin(), m.end()), lf(l.begin(), l.end());
    LSMF_HIP_TRY(hipMemcpy(op->d_mu, mf.data(), nc * sizeof(float), hipMemcpyHostToDevice), "set_material");
    LSMF_HIP_TRY(hipMemcpy(op->d_lambda, lf.data(), nc * sizeof(float), hipMemcpyHostToDevice), "set_material");
  }
  op->material_set = true;
  op->cached = false;
  op->diag_ok = false;
  return LSMF_OK;
}

lsmf_status lsmf_set_material_two(lsmf_op *op, const uint8_t *material_id, double mu0, double lambda0, double mu1,
                                  double lambda1) {
  if (!op || !material_id) return fail(LSMF_ERR_INVALID_ARGUMENT, "set_material: null argument");
  std::vector<double> m(op->n_cells), l(op->n_cells);
  for (int c = 0; c < op->n_cells; ++c) {
    m[c] = material_id[c] ? mu1 : mu0;
    l[c] = material_id[c] ? lambda1 : lambda0;
  }
  return lsmf_set_material_per_cell(op, m.data(), l.data());
}

lsmf_status lsmf_set_formulation(lsmf_op *op, int formulation) {
  if (!op) return fail(LSMF_ERR_INVALID_ARGUMENT, "set_formulation: null operator");
  if (formulation != 0 && formulation != 1)
    return fail(LSMF_ERR_INVALID_ARGUMENT, "set_formulation: Unknown material formulation");
  if (formulation == 0 && op->caching != LSMF_CACHING_SCALAR && op->caching != LSMF_CACHING_TENSOR4)
    return fail(LSMF_ERR_UNSUPPORTED, "set_formulation: formulation 0 needs scalar or tensor4 caching");
  op->formulation = formulation;
  op->cached = false;
  op->diag_ok = false;
  return LSMF_OK;
}

lsmf_status lsmf_set_linearization(lsmf_op *op, const double *u_dev, lsmf_stream stream) {
  if (!op || !u_dev) return fail(LSMF_ERR_INVALID_ARGUMENT, "set_linearization: null argument");
  hipStream_t s = (hipStream_t)stream;
  LSMF_HIP_TRY(hipMemcpyAsync(op->d_u64, u_dev, op->n_dofs * sizeof(double), hipMemcpyDeviceToDevice, s),
               "set_linearization");
  if (op->precision == LSMF_FP32) {
    hipLaunchKernelGGL(k_convert_f64_f32, dim3((unsigned)((op->n_dofs + 255) / 256)), dim3(256), 0, s, op->d_u64,
                       (float *)op->d_u, op->n_dofs);
    LSMF_HIP_TRY(hipGetLastError(), "set_linearization convert");
  }
  op->linearization_set = true;
  op->cached = false;
  op->diag_ok = false;
  return LSMF_OK;
}

lsmf_status lsmf_cache(lsmf_op *op, lsmf_stream stream) {
  if (!op) return fail(LSMF_ERR_INVALID_ARGUMENT, "cache: null operator");
  return cache_with_counter(op, op->d_bad, (hipStream_t)stream, "cache");
}

lsmf_status lsmf_cache_noalloc(lsmf_op *op, lsmf_stream stream) {
  int *scr = nullptr;
  const lsmf_status st = scratch_for_current_device(&scr);
  if (st != LSMF_OK) return st;
  return cache_with_counter(op, scr, (hipStream_t)stream, "cache_noalloc");
}

lsmf_status lsmf_compute_diagonal(lsmf_op *op, lsmf_stream stream) {
  if (!op) return fail(LSMF_ERR_INVALID_ARGUMENT, "compute_diagonal: null operator");
  if (!op->cached) return fail(LSMF_ERR_NOT_CACHED, "compute_diagonal: call cache() first");
  hipStream_t s = (hipStream_t)stream;
  int rc = launch_table(op->dim, op->N)->diagonal(op, s);
  if (rc) return fail(rc, "compute_diagonal: kernel launch failed");
  // entries <= 0 other than the zero rows of unreferenced DoFs
  int bad = 0;
  LSMF_HIP_TRY(hipMemcpyAsync(&bad, op->d_bad, sizeof(int), hipMemcpyDeviceToHost, s),
               "compute_diagonal: read the non-positive count");
  LSMF_HIP_TRY(hipStreamSynchronize(s), "compute_diagonal: synchronize");
  bad -= op->n_orphan;
  if (bad > 0) {
    op->diag_ok = false;
    return fail(LSMF_ERR_NONPOSITIVE_DIAGONAL, "No diagonal entry in a positive definite operator should be zero or "
                                               "negative: " + std::to_string(bad) + " entries");
  }
  op->diag_ok = true;
  return LSMF_OK;
}

lsmf_status lsmf_vmult(lsmf_op *op, void *dst, const void *src, int add, unsigned mask, lsmf_stream stream) {
  if (!op || !dst || !src) return fail(LSMF_ERR_INVALID_ARGUMENT, "vmult: null argument");
  if (dst == src) return fail(LSMF_ERR_INVALID_ARGUMENT, "vmult: dst and src alias");
  if (!op->cached) return fail(LSMF_ERR_NOT_CACHED, "vmult: call cache() first");
  const bool plain = (mask & LSMF_MASK_PLAIN) != 0;
  if (!plain && !(mask & LSMF_MASK_RW)) {
    // no read/write phase: the reference's cell loop neither reads src nor
    // writes dst, and the constrained rows get nothing (:703-705 is under RW)
    if (!add) {
      const size_t bytes = op->n_dofs * elem_size(op);
      hipLaunchKernelGGL(k_zero<unsigned char>, dim3((unsigned)((bytes + 255) / 256)), dim3(256), 0,
                         (hipStream_t)stream, static_cast<unsigned char *>(dst), bytes);
      LSMF_HIP_TRY(hipGetLastError(), "vmult: zero");
    }
    return LSMF_OK;
  }
  unsigned ph = mask & (LSMF_MASK_SF | LSMF_MASK_QD);
  if (!(ph & LSMF_MASK_SF)) ph = 0;  // no sum factorization: gather/scatter only
  if (plain || add) ph = LSMF_MASK_SF | LSMF_MASK_QD;
  const int rc = launch_table(op->dim, op->N)->vmult(op, dst, src, add != 0, !plain, ph, (hipStream_t)stream);
  if (rc) return fail(rc, "vmult: kernel launch failed");
  return LSMF_OK;
}

lsmf_status lsmf_vmult_batched(lsmf_op *op, void *dst, const void *src, int n_vectors, lsmf_stream stream) {
  if (!op || !dst || !src) return fail(LSMF_ERR_INVALID_ARGUMENT, "vmult_batched: null argument");
  if (n_vectors <= 0) return fail(LSMF_ERR_INVALID_ARGUMENT, "vmult_batched: n_vectors must be positive");
  const size_t bytes = (size_t)n_vectors * op->n_dofs * (op->precision == LSMF_FP64 ? 8 : 4);
  const char *d = static_cast<const char *>(dst), *x = static_cast<const char *>(src);
  if (d < x + bytes && x < d + bytes) return fail(LSMF_ERR_INVALID_ARGUMENT, "vmult_batched: dst and src overlap");
  if (!op->cached) return fail(LSMF_ERR_NOT_CACHED, "vmult_batched: call cache() first");
  const int rc = launch_table(op->dim, op->N)->vmult_batched(op, dst, src, n_vectors, (hipStream_t)stream);
  if (rc) return fail(rc, "vmult_batched: kernel launch failed");
  return LSMF_OK;
}

lsmf_status lsmf_get_diagonal(lsmf_op *op, const void **diag_dev, const void **inv_diag_dev) {
  if (!op) return fail(LSMF_ERR_INVALID_ARGUMENT, "get_diagonal: null operator");
  if (!op->diag_ok) return fail(LSMF_ERR_NOT_CACHED, "get_diagonal: call compute_diagonal() first");
  if (diag_dev) *diag_dev = op->d_diag;
  if (inv_diag_dev) *inv_diag_dev = op->d_inv_diag;
  return LSMF_OK;
}

lsmf_status lsmf_precondition_jacobi(lsmf_op *op, void *dst, const void *src, double omega, lsmf_stream stream) {
  if (!op || !dst || !src) return fail(LSMF_ERR_INVALID_ARGUMENT, "precondition_Jacobi: null argument");
  if (!op->diag_ok) return fail(LSMF_ERR_NOT_CACHED, "precondition_Jacobi: call compute_diagonal() first");
  hipStream_t s = (hipStream_t)stream;
  const unsigned g = (unsigned)((op->n_dofs + 255) / 256);
  if (op->precision == LSMF_FP64)
    hipLaunchKernelGGL(k_scale_mul<double>, dim3(g), dim3(256), 0, s, (double *)dst, (const double *)src,
                       (const double *)op->d_inv_diag, omega, op->n_dofs);
  else
    hipLaunchKernelGGL(k_scale_mul<float>, dim3(g), dim3(256), 0, s, (float *)dst, (const float *)src,
                       (const float *)op->d_inv_diag, (float)omega, op->n_dofs);
  LSMF_HIP_TRY(hipGetLastError(), "precondition_Jacobi");
  return LSMF_OK;
}

size_t lsmf_op_size(const lsmf_op *op) { return op ? op->n_dofs : 0; }

size_t lsmf_memory_consumption(const lsmf_op *op) {
  if (!op) return 0;
  const size_t es = elem_size(op);
  size_t b = op->bytes_cache + op->bytes_idx + 2 * (size_t)op->n_cells * es;
  if (op->caching == LSMF_CACHING_SCALAR || op->caching == LSMF_CACHING_SCALAR_REFERENTIAL)
    b += op->n_dofs * es + (size_t)op->n_nodes * op->dim * sizeof(double);
  return b;
}

lsmf_status lsmf_profile_enable(lsmf_op *op, int enable) {
  if (!op) return fail(LSMF_ERR_INVALID_ARGUMENT, "profile_enable: null operator");
  if (!op->prof) op->prof = new (std::nothrow) Profiler;
  if (!op->prof) return fail(LSMF_ERR_OUT_OF_MEMORY, "profile_enable: host allocation");
  op->prof->on = enable != 0;
  op->prof->used = 0;
  return LSMF_OK;
}

lsmf_status lsmf_profile_read(lsmf_op *op, double *total_ms, int *n_launches) {
  if (!op || !total_ms || !n_launches) return fail(LSMF_ERR_INVALID_ARGUMENT, "profile_read: null argument");
  *total_ms = 0.0;
  *n_launches = 0;
  if (!op->prof) return LSMF_OK;
  LSMF_HIP_TRY(hipDeviceSynchronize(), "profile_read: synchronize");
  double sum = 0.0;
  for (size_t i = 0; i < op->prof->used; ++i) {
    float ms = 0.f;
    LSMF_HIP_TRY(hipEventElapsedTime(&ms, op->prof->start[i], op->prof->stop[i]), "profile_read: elapsed");
    sum += ms;
  }
  *total_ms = sum;
  *n_launches = (int)op->prof->used;
  op->prof->used = 0;
  return LSMF_OK;
}

lsmf_status lsmf_residual_vjp(lsmf_op *op, const double *w_dev, double *du_dev, double *dmu, double *dlambda,
                              lsmf_stream stream) {
  if (!op || !w_dev || !du_dev || !dmu || !dlambda) return fail(LSMF_ERR_INVALID_ARGUMENT, "residual_vjp: null argument");
  if (op->precision != LSMF_FP64) return fail(LSMF_ERR_INVALID_ARGUMENT, "residual_vjp: operator must be FP64");
  if ((const void *)w_dev == (const void *)du_dev) return fail(LSMF_ERR_INVALID_ARGUMENT, "residual_vjp: w and du alias");
  if (!op->material_set || !op->linearization_set)
    return fail(LSMF_ERR_INVALID_ARGUMENT, "residual_vjp: material and linearisation must be set");
  if (op->formulation != 1) return fail(LSMF_ERR_UNSUPPORTED, "residual_vjp: formulation 1 only");
  op->adj_dmu = dmu;
  op->adj_dlambda = dlambda;
  const int rc = launch_table(op->dim, op->N)->residual_vjp(op, du_dev, w_dev, (hipStream_t)stream);
  op->adj_dmu = op->adj_dlambda = nullptr;
  if (rc) return fail(rc, "residual_vjp: kernel launch failed");
  return LSMF_OK;
}

lsmf_status lsmf_residual(lsmf_op *op, double *rhs_dev, lsmf_stream stream) {
  if (!op || !rhs_dev) return fail(LSMF_ERR_INVALID_ARGUMENT, "residual: null argument");
  if (op->precision != LSMF_FP64) return fail(LSMF_ERR_INVALID_ARGUMENT, "residual: operator must be FP64");
  if (!op->material_set || !op->linearization_set)
    return fail(LSMF_ERR_INVALID_ARGUMENT, "residual: material and linearisation must be set");
  const int rc = launch_table(op->dim, op->N)->residual(op, rhs_dev, (hipStream_t)stream);
  if (rc) return fail(rc, "residual: kernel launch failed");
  return LSMF_OK;
}

}  // extern "C"
