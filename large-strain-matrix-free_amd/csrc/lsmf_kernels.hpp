// lsmf_kernels.hpp -- HIP kernels of the matrix-free Neo-Hookean tangent for
// gfx950 (MI355X), one fused kernel per cell colour:
//   gather (index table with constraint bits) -> LDS
//   -> sum-factorized gradient (1D matrices, one thread per quadrature point)
//   -> pointwise tangent (scalar / tensor2 / tensor4 caches, FP64 or FP32)
//   -> sum-factorized integration -> colour-ordered scatter (no atomics).
//
// Reference behaviour being reproduced: NeoHookOperator::local_apply_cell and
// do_operation_on_cell (include/mf_nh_operator.h:710-785, 922-1499), cache()
// (:447-592) and compute_diagonal() (:1503-1557), formulation 1.
//
// Thread layout: N = p + 1 = n_q_1d, NQP = N^dim.  Lane r of a cell is both
// the lexicographic support point r (gather / scatter) and the quadrature
// point r (pointwise work), so no lane idles between phases.  A 256-lane
// workgroup holds floor(256 / NQP) cells (Q2-3D: 9 cells, 243 lanes).
//
// Cache layout (structure of arrays, colour-sorted cells):
//   field f of quadrature point r of sorted cell k at  f * fs + k * NQP + r,
// so the NQP * cells_per_workgroup lanes of a workgroup read one contiguous
// run per field.  Jacobian-weighted forms are stored (JxW folded in):
//   tensor4 : K = J_cur^-1 (dim^2), T = tau det(J_ref) w (Voigt, 6),
//             M = (2c S + 2 lambda IxI) det(J_ref) w (Voigt 6x6 upper, 21)
//   tensor2 : K, T, c1 = 2c det(J_ref) w, c2 = 2 lambda det(J_ref) w
//   scalar  : c = mu - 2 lambda ln J   (F, J_ref recomputed per vmult)
// With these, flux = Q K^T where Q = g T + M : sym(g) (or the tensor2 form)
// and g = Grad_xi v K, i.e. exactly the reference's
// Q = [g tau + Jc : sym g] / J integrated with JxW_cur = J det(J_ref) w.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lsmf {

constexpr int kWG = 256;

// Tiled-layout kinds (the TILE template argument of the cell kernels):
// 0 = cells layout, 1 = tiles of one 256-thread workgroup, 2 = wide tiles of
// one 512-thread workgroup.  3D Q2 and Q4 operators with tensor2 or tensor4
// caching use wide tiles.  Q4: four 125-point cells, 4x1x1, gather 425
// nodes for the 256 they own (1.66 touches per node, 17-node rows of 408
// bytes in FP64) where 2x1x1 tiles gather 225 for 128 (1.76 touches, 9-node
// rows of 216 bytes): 64^3 Q4 tensor2 FP64 1.73 -> 1.57 ms, tensor4 2.65 ->
// 2.52 ms.  Q2: 9x2x1 cells (285 nodes for 144 owned, 1.98 touches, rows of
// 19 nodes as in 9x1x1): 128^3 tensor4 FP64 3.94 -> 3.87 ms, FP32 2.08 ->
// 2.01 ms.  Variants with larger register footprints (scalar, tensor4_ns,
// scalar_referential: 150-230 VGPRs in Q4) would run one 512-thread
// workgroup per CU and keep 256 threads.
constexpr int kTileWide = 2;
template <int TILE> __host__ __device__ constexpr int tile_threads() { return TILE == kTileWide ? 2 * kWG : kWG; }
constexpr uint32_t kNodeMask = (1u << 28) - 1u;
constexpr int kConsShift = 28;  // bits 28..30: component c constrained
constexpr uint32_t kFirstBit = 1u << 31;  // this cell writes (not adds) first

// V_RESIDUAL: FP64 internal forces; V_ADJOINT: vector-Jacobian product of
// the residual with respect to u and the per-cell mu, lambda
enum Variant {
  V_SCALAR = 0,
  V_TENSOR2 = 1,
  V_TENSOR4 = 2,
  V_TENSOR4_NS = 3,
  V_SCALAR_REF = 4,
  V_RESIDUAL = 100,
  V_ADJOINT = 101
};

// phase bits (reference MFMask, include/mf_nh_operator.h:73-86)
constexpr unsigned PH_SF = 1u << 3;
constexpr unsigned PH_QD = 1u << 4;

__host__ __device__ constexpr int ipow(int b, int e) { return e == 0 ? 1 : b * ipow(b, e - 1); }

template <int DIM> struct Voigt;
template <> struct Voigt<2> {
  static constexpr int NV = 3;
};
template <> struct Voigt<3> {
  static constexpr int NV = 6;
};
template <int DIM> __host__ __device__ constexpr int voigt_first(int v) {
  return DIM == 3 ? (v < 3 ? v : (v == 5 ? 1 : 0)) : (v < 2 ? v : 0);
}
template <int DIM> __host__ __device__ constexpr int voigt_second(int v) {
  return DIM == 3 ? (v < 3 ? v : (v == 3 ? 1 : 2)) : (v < 2 ? v : 1);
}
template <int DIM> __host__ __device__ constexpr int voigt_index(int i, int j) {
  return i == j ? i : (DIM == 2 ? 2 : (i + j == 1 ? 3 : (i + j == 2 ? 4 : 5)));
}
// packed upper triangle of an NV x NV symmetric matrix
template <int NV> __host__ __device__ constexpr int sym_index(int v, int w) {
  return v <= w ? v * NV - v * (v - 1) / 2 + (w - v) : w * NV - w * (w - 1) / 2 + (v - w);
}

// fields per quadrature point of each cache:
//   tensor4     K = J_cur^-1 (D2), tau (NV), C packed (NV(NV+1)/2)
//   tensor2     K, tau, 2c, 2 lambda
//   tensor4_ns  J_ref^-1 (D2), dP/dF packed by its major symmetry (D2(D2+1)/2)
//   scalar      c
//   scalar_ref  c, u at the point (DIM)
template <int DIM, int VAR> struct CacheFields {
  static constexpr int NV = Voigt<DIM>::NV;
  static constexpr int D2 = DIM * DIM;
  static constexpr int value = VAR == V_TENSOR4      ? D2 + NV + NV * (NV + 1) / 2
                               : VAR == V_TENSOR2    ? D2 + NV + 2
                               : VAR == V_TENSOR4_NS ? D2 + D2 * (D2 + 1) / 2
                               : VAR == V_SCALAR_REF ? 1 + DIM
                               : VAR == V_SCALAR     ? 1
                                                     : 0;
};

// Position of cache field f (of nf) at point cq = sorted cell * NQP + q.
// The fields are stored in groups of G = 16 / sizeof(T) (two FP64 or four
// FP32 fields: one 16-byte vector per point and group, G * fs values per
// group), so that a lane loads G fields with one 16-byte access; the
// nf % G fields left over form one narrower group after them (no padding).
template <typename T> constexpr int kCacheGroup = 16 / (int)sizeof(T);
template <typename T> __host__ __device__ constexpr size_t cache_at(int f, int nf, size_t cq, size_t fs) {
  constexpr int G = kCacheGroup<T>;
  const int full = nf / G * G;
  return f < full ? (size_t)(f / G) * G * fs + (size_t)G * cq + (size_t)(f % G)
                  : (size_t)full * fs + (size_t)(nf - full) * cq + (size_t)(f - full);
}
template <typename T> struct CacheVec;
template <> struct CacheVec<double> { using type = double2; };
template <> struct CacheVec<float> { using type = float4; };
// load the NF fields of point cq: G per 16-byte load, the rest by the
// widest access that fits (FP32: a pair as one 8-byte load)
template <typename T, int NF> __device__ __forceinline__ void load_cache(const T *__restrict__ c, size_t cq, size_t fs,
                                                                         T (&v)[NF]) {
  constexpr int G = kCacheGroup<T>, FULL = NF / G * G, TAIL = NF - FULL;
  using V = typename CacheVec<T>::type;
  const V *cv = reinterpret_cast<const V *>(c) + cq;
#pragma unroll
  for (int f = 0; f < FULL; f += G) {
    const V p = cv[(size_t)(f / G) * fs];
    const T *e = reinterpret_cast<const T *>(&p);
#pragma unroll
    for (int k = 0; k < G; ++k) v[f + k] = e[k];
  }
  const T *ct = c + (size_t)FULL * fs + (size_t)TAIL * cq;
  if constexpr (TAIL == 2 && sizeof(T) == 4) {
    const float2 p = *reinterpret_cast<const float2 *>(ct);
    v[FULL] = p.x;
    v[FULL + 1] = p.y;
  } else {
#pragma unroll
    for (int k = 0; k < TAIL; ++k) v[FULL + k] = ct[k];
  }
}

// 1D tables: S, D, St, Dt (N*N each), Gauss points and weights (N each),
// then the collocation derivative on the Gauss points (N*N, used by
// scalar_referential only)
template <int N> __host__ __device__ constexpr int tab_size(int var) {
  return 4 * N * N + 2 * N + (var == V_SCALAR_REF ? N * N : 0);
}

template <int DIM, int N> struct CellShape {
  static constexpr int NQP = ipow(N, DIM);
  static constexpr int CPW = kWG / NQP;  // cells per workgroup
  static constexpr int NB = DIM * DIM;   // arrays per LDS ping-pong buffer
  static constexpr int NC = 1 << DIM;    // Q1 corners
};

// Tile of the tiled layout: TX x TY x TZ cells of a structured mesh, whose
// (TX p + 1)(TY p + 1)(TZ p + 1) nodes are gathered once into LDS and whose
// contributions are summed there before one write per node.  TX TY TZ = CPW.
template <int DIM, int N, int WG_ = kWG> struct TileShape {
  static constexpr int WG = WG_;
  static constexpr int CPW = WG / ipow(N, DIM);
  // 3D Q2: one row of 9 cells (19 x 3 x 3 nodes) rather than 3 x 3 x 1 (7 x 7 x 3): more node
  // touches (2.5 instead of 2.0 per node) but 19-node rows (456 B in FP64) fill their cache lines
  // where 7-node rows (168 B) span two or three half-used ones
  static constexpr int TX = DIM == 3 ? (N == 2 ? 4 : N == 3 ? 9 : CPW) : (N == 2 ? 8 : N == 3 ? 7 : 5);
  static constexpr int TY = DIM == 3 ? (N == 2 ? 4 : N == 3 ? CPW / 9 : 1) : (N == 2 ? 8 : N == 3 ? 4 : 2);
  static constexpr int TZ = DIM == 3 ? (N == 2 ? 2 : 1) : 1;
  static constexpr int P = N - 1;
  static constexpr int NX = TX * P + 1, NY = TY * P + 1, NZ = DIM == 3 ? TZ * P + 1 : 1;
  static constexpr int TN = NX * NY * NZ;
  static_assert(TX * TY * TZ == CPW, "tile must hold the workgroup's cells");
  static_assert(TN <= WG, "one lane per tile node");
};
constexpr uint32_t kNoNode = 0xFFFFFFFFu;  // tile node outside the mesh

// The DIM values of one node as one access (3D FP32: 12 bytes in one load,
// 3D FP64: 16 + 8 bytes), instead of DIM separate loads each touching three
// times the cache lines per instruction.
template <int DIM, typename T> struct NodeVec {
  T v[DIM];
};
template <int DIM, typename T> __device__ __forceinline__ void load_node(const T *p, T (&v)[DIM]) {
  NodeVec<DIM, T> n;
  __builtin_memcpy(&n, p, sizeof(n));
#pragma unroll
  for (int c = 0; c < DIM; ++c) v[c] = n.v[c];
}
template <int DIM, typename T> __device__ __forceinline__ void store_node(T *p, const T (&v)[DIM]) {
  NodeVec<DIM, T> n;
#pragma unroll
  for (int c = 0; c < DIM; ++c) n.v[c] = v[c];
  __builtin_memcpy(p, &n, sizeof(n));
}

// LDS-only workgroup barrier: waits for this wave's LDS traffic, not for its
// outstanding global loads, so the per-qp cache loads issued at kernel entry
// stay in flight across the sum-factorization stages.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Workgroups are dispatched round-robin over the 8 XCDs; remap so that each
// XCD processes a contiguous range of cells (consecutive workgroups share
// cache lines of the SoA caches and of the vectors at their boundaries).
__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
  const int per = nblocks / 8;
  if (bid >= per * 8) return bid;
  return (bid % 8) * per + bid / 8;
}

template <typename T> struct Tables {
  // S[q][a], D[q][a] (forward), St[a][q] = S[q][a] stored row-major by a, Dt
  // plus Gauss points and weights on [0,1]
  const T *S;  // layout in device memory: S, D, St, Dt each N*N, then qp[N], qw[N]
};

template <int N, int STRIDE, typename T>
__device__ __forceinline__ void load_line(const T *__restrict__ in, int base, T (&x)[N]) {
#pragma unroll
  for (int a = 0; a < N; ++a) x[a] = in[base + a * STRIDE];
}
template <int N, typename T> __device__ __forceinline__ T dotN(const T *__restrict__ row, const T (&x)[N]) {
  T s = row[0] * x[0];
#pragma unroll
  for (int a = 1; a < N; ++a) s = fma(row[a], x[a], s);
  return s;
}

// ------------------------------------------------------------------------
// small tensor helpers (row-major DIM x DIM)
// ------------------------------------------------------------------------
template <int DIM, typename T> __device__ __forceinline__ T det(const T (&A)[DIM][DIM]) {
  if constexpr (DIM == 2) return A[0][0] * A[1][1] - A[0][1] * A[1][0];
  else
    return A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) - A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
           A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
}
// B = A^-1, returns det(A)
template <int DIM, typename T> __device__ __forceinline__ T inverse(const T (&A)[DIM][DIM], T (&B)[DIM][DIM]) {
  if constexpr (DIM == 2) {
    const T d = A[0][0] * A[1][1] - A[0][1] * A[1][0];
    const T id = T(1) / d;
    B[0][0] = A[1][1] * id;
    B[0][1] = -A[0][1] * id;
    B[1][0] = -A[1][0] * id;
    B[1][1] = A[0][0] * id;
    return d;
  } else {
    const T c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
    const T c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
    const T c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
    const T d = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
    const T id = T(1) / d;
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
}

// ------------------------------------------------------------------------
// sum factorization on one cell's LDS slot (lane r, multi-index i0,i1,i2)
// ------------------------------------------------------------------------
template <int DIM, int N, typename T> struct SF {
  static constexpr int NQP = ipow(N, DIM);
  // tab: S, D, St, Dt (N*N each) in LDS

  // nodal values of NCOMP components in A[c*NQP + .] -> gradients G[c][d] in
  // registers; B is scratch.  Ends with data in registers; caller must
  // lds_barrier() before the next writes to A.
  template <int NCOMP>
  static __device__ __forceinline__ void evaluate(const T *tab, T *A, T *B, bool on, int r, int i0, int i1, int i2,
                                                  T (&G)[NCOMP][DIM]) {
    const T *S = tab, *D = tab + N * N;
    // x direction: A -> B (value, derivative)
    if (on) {
      const int base = r - i0;
#pragma unroll
      for (int c = 0; c < NCOMP; ++c) {
        T x[N];
        load_line<N, 1>(A + c * NQP, base, x);
        B[(2 * c) * NQP + r] = dotN<N>(S + i0 * N, x);
        B[(2 * c + 1) * NQP + r] = dotN<N>(D + i0 * N, x);
      }
    }
    lds_barrier();
    if constexpr (DIM == 2) {
      const int base = on ? r - i1 * N : 0;
#pragma unroll
      for (int c = 0; c < NCOMP; ++c) {
        T x0[N], x1[N];
        load_line<N, N>(B + (2 * c) * NQP, base, x0);
        load_line<N, N>(B + (2 * c + 1) * NQP, base, x1);
        G[c][0] = dotN<N>(S + i1 * N, x1);
        G[c][1] = dotN<N>(D + i1 * N, x0);
      }
    } else {
      if (on) {
        const int base = r - i1 * N;
#pragma unroll
        for (int c = 0; c < NCOMP; ++c) {
          T x0[N], x1[N];
          load_line<N, N>(B + (2 * c) * NQP, base, x0);
          load_line<N, N>(B + (2 * c + 1) * NQP, base, x1);
          A[(3 * c + 0) * NQP + r] = dotN<N>(S + i1 * N, x0);  // value in x,y
          A[(3 * c + 1) * NQP + r] = dotN<N>(S + i1 * N, x1);  // d/dx
          A[(3 * c + 2) * NQP + r] = dotN<N>(D + i1 * N, x0);  // d/dy
        }
      }
      lds_barrier();
      const int base = on ? r - i2 * N * N : 0;
#pragma unroll
      for (int c = 0; c < NCOMP; ++c) {
        T y0[N], y1[N], y2[N];
        load_line<N, N * N>(A + (3 * c + 0) * NQP, base, y0);
        load_line<N, N * N>(A + (3 * c + 1) * NQP, base, y1);
        load_line<N, N * N>(A + (3 * c + 2) * NQP, base, y2);
        G[c][0] = dotN<N>(S + i2 * N, y1);
        G[c][1] = dotN<N>(S + i2 * N, y2);
        G[c][2] = dotN<N>(D + i2 * N, y0);
      }
    }
  }

  // fluxes Phi[c][d] (registers) -> out[c] = sum_q sum_d dphi_r/dxi_d Phi_cd.
  // The buffer written first is W (must not be read by anyone after the last
  // barrier), the other is scratch.  3D: W=B, then A, B.  2D: W=A, then B.
  static __device__ __forceinline__ void integrate(const T *tab, T *W, T *O, bool on, int r, int i0, int i1, int i2,
                                                   const T (&Phi)[DIM][DIM], T (&out)[DIM]) {
    const T *St = tab + 2 * N * N, *Dt = tab + 3 * N * N;
    if (on)
#pragma unroll
      for (int c = 0; c < DIM; ++c)
#pragma unroll
        for (int d = 0; d < DIM; ++d) W[(DIM * c + d) * NQP + r] = Phi[c][d];
    lds_barrier();
    if constexpr (DIM == 3) {
      if (on) {
        const int base = r - i2 * N * N;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          T f0[N], f1[N], f2[N];
          load_line<N, N * N>(W + (3 * c + 0) * NQP, base, f0);
          load_line<N, N * N>(W + (3 * c + 1) * NQP, base, f1);
          load_line<N, N * N>(W + (3 * c + 2) * NQP, base, f2);
          O[(3 * c + 0) * NQP + r] = dotN<N>(St + i2 * N, f0);
          O[(3 * c + 1) * NQP + r] = dotN<N>(St + i2 * N, f1);
          O[(3 * c + 2) * NQP + r] = dotN<N>(Dt + i2 * N, f2);
        }
      }
      lds_barrier();
      if (on) {
        const int base = r - i1 * N;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          T r10[N], r01[N], r00[N];
          load_line<N, N>(O + (3 * c + 0) * NQP, base, r10);
          load_line<N, N>(O + (3 * c + 1) * NQP, base, r01);
          load_line<N, N>(O + (3 * c + 2) * NQP, base, r00);
          W[(2 * c + 0) * NQP + r] = dotN<N>(St + i1 * N, r00) + dotN<N>(Dt + i1 * N, r01);
          W[(2 * c + 1) * NQP + r] = dotN<N>(St + i1 * N, r10);
        }
      }
      lds_barrier();
      {
        const int base = on ? r - i0 : 0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          T p0[N], p1[N];
          load_line<N, 1>(W + (2 * c + 0) * NQP, base, p0);
          load_line<N, 1>(W + (2 * c + 1) * NQP, base, p1);
          out[c] = dotN<N>(St + i0 * N, p0) + dotN<N>(Dt + i0 * N, p1);
        }
      }
    } else {
      if (on) {
        const int base = r - i1 * N;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          T fx[N], fy[N];
          load_line<N, N>(W + (2 * c + 0) * NQP, base, fx);
          load_line<N, N>(W + (2 * c + 1) * NQP, base, fy);
          O[(2 * c + 0) * NQP + r] = dotN<N>(Dt + i1 * N, fy);
          O[(2 * c + 1) * NQP + r] = dotN<N>(St + i1 * N, fx);
        }
      }
      lds_barrier();
      {
        const int base = on ? r - i0 : 0;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          T p0[N], p1[N];
          load_line<N, 1>(O + (2 * c + 0) * NQP, base, p0);
          load_line<N, 1>(O + (2 * c + 1) * NQP, base, p1);
          out[c] = dotN<N>(St + i0 * N, p0) + dotN<N>(Dt + i0 * N, p1);
        }
      }
    }
  }
};

// ------------------------------------------------------------------------
// Even-odd form of a 1D matrix M (N x N) that is centrosymmetric (SIG = +1:
// M[N-1-q][N-1-k] = M[q][k], the interpolation S between symmetric point
// sets) or centro-antisymmetric (SIG = -1: the derivative D).  With
// x_e[k] = x[k] + x[N-1-k], x_o[k] = x[k] - x[N-1-k] (k < H = N / 2) and the
// middle value x_m (odd N), M x and M^T x need only the H-column halves
//   E[r][k] = (M[r][k] + M[r][N-1-k]) / 2,  O[r][k] = (M[r][k] - M[r][N-1-k]) / 2
// and the middle column C[r] = M[r][H], rows r < R = H + N % 2: about half
// the multiply-adds and half the constants of the plain products.
// Packed per matrix: E[R][H], O[R][H], C[R] (PK values).
// ------------------------------------------------------------------------
template <int N, typename T> struct EvenOdd {
  static constexpr int H = N / 2, ODD = N % 2, R = H + ODD, PK = R * (2 * H + 1);
  struct Split {
    T e[H], o[H], m;
  };
  static __device__ __forceinline__ Split split(const T (&x)[N]) {
    Split s;
#pragma unroll
    for (int k = 0; k < H; ++k) {
      s.e[k] = x[k] + x[N - 1 - k];
      s.o[k] = x[k] - x[N - 1 - k];
    }
    s.m = ODD ? x[H] : T(0);
    return s;
  }
  static __device__ __forceinline__ T E(const T *pk, int r, int k) { return pk[r * H + k]; }
  static __device__ __forceinline__ T O(const T *pk, int r, int k) { return pk[R * H + r * H + k]; }
  static __device__ __forceinline__ T C(const T *pk, int r) { return pk[2 * R * H + r]; }

  // y = M x
  template <int SIG> static __device__ __forceinline__ void fwd(const T *pk, const Split &s, T (&y)[N]) {
#pragma unroll
    for (int q = 0; q < H; ++q) {
      T ev = ODD ? C(pk, q) * s.m : T(0), od = T(0);
#pragma unroll
      for (int k = 0; k < H; ++k) {
        ev = fma(E(pk, q, k), s.e[k], ev);
        od = fma(O(pk, q, k), s.o[k], od);
      }
      y[q] = ev + od;
      y[N - 1 - q] = SIG > 0 ? ev - od : od - ev;
    }
    if constexpr (ODD) {
      T v = SIG > 0 ? C(pk, H) * s.m : T(0);
#pragma unroll
      for (int k = 0; k < H; ++k) v = SIG > 0 ? fma(E(pk, H, k), s.e[k], v) : fma(O(pk, H, k), s.o[k], v);
      y[H] = v;
    }
  }
  // y = M^T x
  template <int SIG> static __device__ __forceinline__ void tr(const T *pk, const Split &s, T (&y)[N]) {
#pragma unroll
    for (int a = 0; a < H; ++a) {
      T u, w;
      if constexpr (SIG > 0) {
        u = ODD ? E(pk, H, a) * s.m : T(0);  // even part
        w = T(0);                            // odd part
#pragma unroll
        for (int q = 0; q < H; ++q) {
          u = fma(E(pk, q, a), s.e[q], u);
          w = fma(O(pk, q, a), s.o[q], w);
        }
      } else {
        u = T(0);
        w = ODD ? O(pk, H, a) * s.m : T(0);
#pragma unroll
        for (int q = 0; q < H; ++q) {
          u = fma(E(pk, q, a), s.o[q], u);
          w = fma(O(pk, q, a), s.e[q], w);
        }
      }
      y[a] = u + w;
      y[N - 1 - a] = u - w;
    }
    if constexpr (ODD) {
      T v = SIG > 0 ? C(pk, H) * s.m : T(0);
#pragma unroll
      for (int q = 0; q < H; ++q) v = fma(C(pk, q), SIG > 0 ? s.e[q] : s.o[q], v);
      y[H] = v;
    }
  }
};

// ------------------------------------------------------------------------
// 3D sum factorization with one lane per 1D line.  In each of the three
// passes a lane takes one line (slot, component, line of N points along the
// pass direction), reads its values from LDS into registers, applies the 1D
// matrices in even-odd form and writes the results back over the same line.
// The matrices are the same for every lane, so they come from the kernel
// arguments (scalar registers), not from LDS.  Per point and pass that is
// one LDS read and one write per array, where one lane per point reads the
// N line values and N matrix entries for each value it writes.
// Every pass reads and writes only the positions of the lines it owns, so
// the passes work in place: slot s owns one buffer lds + s * SLOT of NB = 9
// arrays of NQP values, array e * 3 + c holding stage value e of component
// c.  sf holds the packed even-odd forms of S, then of D (EvenOdd::PK each).
// ------------------------------------------------------------------------
template <int N, typename T> struct SFLine {
  static constexpr int NQP = N * N * N, NL = N * N, NB = 9, SLOT = NB * NQP;
  using EO = EvenOdd<N, T>;

  // task -> slot s, component c, line l; returns the first array of (s, c)
  static __device__ __forceinline__ T *task_base(T *lds, int task, int &l) {
    const int s = task / (3 * NL), rem = task - s * 3 * NL, c = rem / NL;
    l = rem - c * NL;
    return lds + s * SLOT + c * NQP;
  }
  template <int STRIDE> static __device__ __forceinline__ void load(const T *p, T (&x)[N]) {
#pragma unroll
    for (int k = 0; k < N; ++k) x[k] = p[k * STRIDE];
  }
  template <int STRIDE> static __device__ __forceinline__ void store(T *p, const T (&x)[N]) {
#pragma unroll
    for (int k = 0; k < N; ++k) p[k * STRIDE] = x[k];
  }

  // nodal values in array c (c < 3) of slots s < CPW -> gradients G[c][d]
  // at quadrature point r of this lane's slot (lanes with on), left in
  // array d * 3 + c.  A lane reads only its own point r at the end, so the
  // caller may overwrite point r of any array (integrate) without a barrier.
  template <int CPW, int WG = kWG>
  static __device__ __forceinline__ void evaluate(const T *sf, T *lds, int t, bool on, int slot, int r,
                                                  T (&G)[3][3]) {
    const T *S = sf, *D = sf + EO::PK;
    constexpr int TASKS = CPW * 3 * NL;
    constexpr int E1 = 3 * NQP, E2 = 6 * NQP;  // offsets of stage arrays 1, 2
    // x: e0 = S u, e1 = D u along x; line l = (i1, i2)
#pragma unroll 1
    for (int task = t; task < TASKS; task += WG) {
      int l;
      T *p = task_base(lds, task, l) + l * N;
      T x[N], y[N];
      load<1>(p, x);
      const typename EO::Split sx = EO::split(x);
      EO::template fwd<1>(S, sx, y);
      store<1>(p, y);
      EO::template fwd<-1>(D, sx, y);
      store<1>(p + E1, y);
    }
    lds_barrier();
    // y: e0 <- S e0 (value), e1 <- S e1 (d/dx), e2 = D e0 (d/dy); line (i0, i2)
#pragma unroll 1
    for (int task = t; task < TASKS; task += WG) {
      int l;
      T *p = task_base(lds, task, l) + (l % N) + NL * (l / N);
      T x0[N], x1[N], y[N];
      load<N>(p, x0);
      load<N>(p + E1, x1);
      const typename EO::Split s0 = EO::split(x0), s1 = EO::split(x1);
      EO::template fwd<1>(S, s0, y);
      store<N>(p, y);
      EO::template fwd<-1>(D, s0, y);
      store<N>(p + E2, y);
      EO::template fwd<1>(S, s1, y);
      store<N>(p + E1, y);
    }
    lds_barrier();
    // z: G_c0 = S e1, G_c1 = S e2, G_c2 = D e0 into arrays e0, e1, e2; line (i0, i1)
#pragma unroll 1
    for (int task = t; task < TASKS; task += WG) {
      int l;
      T *p = task_base(lds, task, l) + l;
      T y0[N], y1[N], y2[N], y[N];
      load<NL>(p, y0);
      load<NL>(p + E1, y1);
      load<NL>(p + E2, y2);
      EO::template fwd<1>(S, EO::split(y1), y);
      store<NL>(p, y);
      EO::template fwd<1>(S, EO::split(y2), y);
      store<NL>(p + E1, y);
      EO::template fwd<-1>(D, EO::split(y0), y);
      store<NL>(p + E2, y);
    }
    lds_barrier();
    if (on) {
      const T *g = lds + slot * SLOT + r;
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) G[c][d] = g[(3 * d + c) * NQP];
    }
  }

  // fluxes Phi[c][d] at quadrature point r of this lane's slot -> out[c] =
  // sum_q sum_d dphi_r/dxi_d Phi_cd at support point r (read back from
  // array c after the last barrier).
  template <int CPW, int WG = kWG>
  static __device__ __forceinline__ void integrate(const T *sf, T *lds, int t, bool on, int slot, int r,
                                                   const T (&Phi)[3][3], T (&out)[3]) {
    const T *S = sf, *D = sf + EO::PK;
    constexpr int TASKS = CPW * 3 * NL;
    constexpr int E1 = 3 * NQP, E2 = 6 * NQP;
    if (on) {
      T *w = lds + slot * SLOT + r;
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) w[(3 * d + c) * NQP] = Phi[c][d];
    }
    lds_barrier();
    // z^T: e0 <- S^T e0, e1 <- S^T e1, e2 <- D^T e2; line (i0, i1)
#pragma unroll 1
    for (int task = t; task < TASKS; task += WG) {
      int l;
      T *p = task_base(lds, task, l) + l;
      T f0[N], f1[N], f2[N], y[N];
      load<NL>(p, f0);
      load<NL>(p + E1, f1);
      load<NL>(p + E2, f2);
      EO::template tr<1>(S, EO::split(f0), y);
      store<NL>(p, y);
      EO::template tr<1>(S, EO::split(f1), y);
      store<NL>(p + E1, y);
      EO::template tr<-1>(D, EO::split(f2), y);
      store<NL>(p + E2, y);
    }
    lds_barrier();
    // y^T: e0 <- S^T e2 + D^T e1, e1 <- S^T e0; line (i0, i2)
#pragma unroll 1
    for (int task = t; task < TASKS; task += WG) {
      int l;
      T *p = task_base(lds, task, l) + (l % N) + NL * (l / N);
      T r10[N], r01[N], r00[N], y[N], z[N];
      load<N>(p, r10);
      load<N>(p + E1, r01);
      load<N>(p + E2, r00);
      EO::template tr<1>(S, EO::split(r00), y);
      EO::template tr<-1>(D, EO::split(r01), z);
#pragma unroll
      for (int a = 0; a < N; ++a) y[a] += z[a];
      store<N>(p, y);
      EO::template tr<1>(S, EO::split(r10), y);
      store<N>(p + E1, y);
    }
    lds_barrier();
    // x^T: e0 <- S^T e0 + D^T e1; line (i1, i2)
#pragma unroll 1
    for (int task = t; task < TASKS; task += WG) {
      int l;
      T *p = task_base(lds, task, l) + l * N;
      T p0[N], p1[N], y[N], z[N];
      load<1>(p, p0);
      load<1>(p + E1, p1);
      EO::template tr<1>(S, EO::split(p0), y);
      EO::template tr<-1>(D, EO::split(p1), z);
#pragma unroll
      for (int a = 0; a < N; ++a) y[a] += z[a];
      store<1>(p, y);
    }
    lds_barrier();
    if (on) {
      const T *v = lds + slot * SLOT + r;
#pragma unroll
      for (int c = 0; c < 3; ++c) out[c] = v[c * NQP];
    }
  }
};

// ------------------------------------------------------------------------
// pointwise tangent actions (reference: include/mf_nh_operator.h:1182-1440)
// ------------------------------------------------------------------------
// g = G K
template <int DIM, typename T>
__device__ __forceinline__ void grad_times(const T (&G)[DIM][DIM], const T (&K)[DIM][DIM], T (&g)[DIM][DIM]) {
#pragma unroll
  for (int i = 0; i < DIM; ++i)
#pragma unroll
    for (int j = 0; j < DIM; ++j) {
      T s = G[i][0] * K[0][j];
#pragma unroll
      for (int d = 1; d < DIM; ++d) s = fma(G[i][d], K[d][j], s);
      g[i][j] = s;
    }
}
// Phi = Q K^T
template <int DIM, typename T>
__device__ __forceinline__ void flux_times_kt(const T (&Q)[DIM][DIM], const T (&K)[DIM][DIM], T (&Phi)[DIM][DIM]) {
#pragma unroll
  for (int c = 0; c < DIM; ++c)
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      T s = Q[c][0] * K[d][0];
#pragma unroll
      for (int e = 1; e < DIM; ++e) s = fma(Q[c][e], K[d][e], s);
      Phi[c][d] = s;
    }
}

// tensor4: Q = g T + M : sym(g)   (cached T, M already carry JxW)
template <int DIM, typename T>
__device__ __forceinline__ void qp_tensor4(const T (&G)[DIM][DIM], const T *cch, T (&Phi)[DIM][DIM]) {
  constexpr int NV = Voigt<DIM>::NV;
  T K[DIM][DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d)
#pragma unroll
    for (int j = 0; j < DIM; ++j) K[d][j] = cch[d * DIM + j];
  const T *Tv = cch + DIM * DIM;
  const T *M = Tv + NV;
  T g[DIM][DIM];
  grad_times<DIM>(G, K, g);
  T sv[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const int i = voigt_first<DIM>(v), j = voigt_second<DIM>(v);
    sv[v] = (i == j) ? g[i][i] : g[i][j] + g[j][i];
  }
  T rv[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    T s = M[sym_index<NV>(v, 0)] * sv[0];
#pragma unroll
    for (int w = 1; w < NV; ++w) s = fma(M[sym_index<NV>(v, w)], sv[w], s);
    rv[v] = s;
  }
  T Q[DIM][DIM];
#pragma unroll
  for (int i = 0; i < DIM; ++i)
#pragma unroll
    for (int j = 0; j < DIM; ++j) {
      T s = rv[voigt_index<DIM>(i, j)];
#pragma unroll
      for (int k = 0; k < DIM; ++k) s = fma(g[i][k], Tv[voigt_index<DIM>(k, j)], s);
      Q[i][j] = s;
    }
  flux_times_kt<DIM>(Q, K, Phi);
}

// tensor2: Q = c1 sym(g) + c2 tr(g) I + g T
template <int DIM, typename T>
__device__ __forceinline__ void qp_tensor2(const T (&G)[DIM][DIM], const T *cch, T (&Phi)[DIM][DIM]) {
  constexpr int NV = Voigt<DIM>::NV;
  T K[DIM][DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d)
#pragma unroll
    for (int j = 0; j < DIM; ++j) K[d][j] = cch[d * DIM + j];
  const T *Tv = cch + DIM * DIM;
  const T c1 = cch[DIM * DIM + NV], c2 = cch[DIM * DIM + NV + 1];
  T g[DIM][DIM];
  grad_times<DIM>(G, K, g);
  T tr = g[0][0];
#pragma unroll
  for (int d = 1; d < DIM; ++d) tr += g[d][d];
  const T hc1 = T(0.5) * c1;
  T Q[DIM][DIM];
#pragma unroll
  for (int i = 0; i < DIM; ++i)
#pragma unroll
    for (int j = 0; j < DIM; ++j) {
      T s = hc1 * (g[i][j] + g[j][i]);
      if (i == j) s = fma(c2, tr, s);
#pragma unroll
      for (int k = 0; k < DIM; ++k) s = fma(g[i][k], Tv[voigt_index<DIM>(k, j)], s);
      Q[i][j] = s;
    }
  flux_times_kt<DIM>(Q, K, Phi);
}

// tensor4_ns: P' = A : Grad v with Grad v = G J_ref^-1, flux P' J_ref^-T
// (include/mf_nh_operator.h:1441-1455); A = dP/dF det(J_ref) w is packed by
// its major symmetry over the index pairs p = DIM i + J.
template <int DIM, typename T>
__device__ __forceinline__ void qp_tensor4_ns(const T (&G)[DIM][DIM], const T *cch, T (&Phi)[DIM][DIM]) {
  constexpr int D2 = DIM * DIM;
  T Kr[DIM][DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d)
#pragma unroll
    for (int j = 0; j < DIM; ++j) Kr[d][j] = cch[d * DIM + j];
  const T *Am = cch + D2;
  T H[DIM][DIM];
  grad_times<DIM>(G, Kr, H);
  T P[DIM][DIM];
#pragma unroll
  for (int p = 0; p < D2; ++p) {
    T s = Am[sym_index<D2>(p, 0)] * H[0][0];
#pragma unroll
    for (int q = 1; q < D2; ++q) s = fma(Am[sym_index<D2>(p, q)], H[q / DIM][q % DIM], s);
    P[p / DIM][p % DIM] = s;
  }
  flux_times_kt<DIM>(P, Kr, Phi);
}

// Formulation 0, the volumetric/isochoric split (include/material.h:116-141,
// 171-226, 260-295; include/mf_nh_operator.h:1005-1180), with
// kappa = lambda + 2 mu / 3, b_bar = sbar^2 F F^T (sbar = J^{-1/d}),
// tau_bar = mu b_bar, tau_iso = dev(tau_bar), J dPsi_vol/dJ = kappa/2 (J^2 - 1):
//   tau   = tau_bar + (J dPsi/dJ - tr(tau_bar)/d) I
//   J c:s = a s + b tr(s) I - (2/d) (tr(s) tau_iso + (tau_iso : s) I)
//   a = -2 J dPsi/dJ + (2/d) tr(tau_bar),  b = kappa J^2 - (2/d^2) tr(tau_bar)
template <int DIM, typename T> struct Form0 {
  T tau[DIM][DIM], tiso[DIM][DIM];
  T a, b;
  __device__ __forceinline__ Form0(const T (&F)[DIM][DIM], T J, T sbar, T m, T l) {
    const T kappa = l + T(2) * m / T(3);
    T trtb = T(0);
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
      for (int j = 0; j < DIM; ++j) {
        T s = F[i][0] * F[j][0];
#pragma unroll
        for (int k = 1; k < DIM; ++k) s = fma(F[i][k], F[j][k], s);
        tau[i][j] = m * sbar * sbar * s;
      }
#pragma unroll
    for (int d = 0; d < DIM; ++d) trtb += tau[d][d];
    const T dpsiJ = T(0.5) * kappa * (J - T(1) / J) * J;
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
      for (int j = 0; j < DIM; ++j) tiso[i][j] = tau[i][j] - (i == j ? trtb / T(DIM) : T(0));
#pragma unroll
    for (int d = 0; d < DIM; ++d) tau[d][d] += dpsiJ - trtb / T(DIM);
    a = T(-2) * dpsiJ + T(2) / T(DIM) * trtb;
    b = kappa * J * J - T(2) / T(DIM * DIM) * trtb;
  }
  // Q = J c : sym(g) + g tau
  __device__ __forceinline__ void action(const T (&g)[DIM][DIM], T (&Q)[DIM][DIM]) const {
    T tr = g[0][0];
#pragma unroll
    for (int d = 1; d < DIM; ++d) tr += g[d][d];
    T ts = T(0);  // tau_iso : sym(g)
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
      for (int j = 0; j < DIM; ++j) ts = fma(tiso[i][j], g[i][j], ts);
    const T f = T(2) / T(DIM);
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
      for (int j = 0; j < DIM; ++j) {
        T s = T(0.5) * a * (g[i][j] + g[j][i]) - f * tr * tiso[i][j];
        if (i == j) s += b * tr - f * ts;
#pragma unroll
        for (int k = 0; k < DIM; ++k) s = fma(g[i][k], tau[k][j], s);
        Q[i][j] = s;
      }
  }
};

// tensor4 fields of one quadrature point (the reference's tensor4 cache,
// include/mf_nh_operator.h:548-562, and its formulation-0 counterpart):
// T = tau det(J_ref) w (Voigt), M = C det(J_ref) w packed, C the spatial
// tangent J c acting on sym(g); formulation 1: C = 2 c S + 2 lambda I x I.
template <int DIM, typename T>
__device__ __forceinline__ void tensor4_fields(const T (&F)[DIM][DIM], T J, T m, T l, T dw, int formulation,
                                               T *Tv, T *M) {
  constexpr int NV = Voigt<DIM>::NV;
  if (formulation == 0) {
    const Form0<DIM, T> f0(F, J, T(pow(double(J), -1.0 / DIM)), m, l);
    const T f = T(2) / T(DIM);
#pragma unroll
    for (int v = 0; v < NV; ++v) Tv[v] = f0.tau[voigt_first<DIM>(v)][voigt_second<DIM>(v)] * dw;
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
      for (int w2 = v; w2 < NV; ++w2) {
        const int i = voigt_first<DIM>(v), j = voigt_second<DIM>(v);
        const int k = voigt_first<DIM>(w2), ll = voigt_second<DIM>(w2);
        const T Sijkl = T(0.5) * (T(i == k && j == ll) + T(i == ll && j == k));
        const T dij = T(i == j), dkl = T(k == ll);
        M[sym_index<NV>(v, w2)] =
            (f0.a * Sijkl + f0.b * dij * dkl - f * (f0.tiso[i][j] * dkl + dij * f0.tiso[k][ll])) * dw;
      }
    return;
  }
  const T cs = m - T(2) * l * T(log(double(J)));
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const int i = voigt_first<DIM>(v), j = voigt_second<DIM>(v);
    T s = T(0);
#pragma unroll
    for (int k = 0; k < DIM; ++k) s += F[i][k] * F[j][k];
    Tv[v] = (m * s - (i == j ? cs : T(0))) * dw;
  }
#pragma unroll
  for (int v = 0; v < NV; ++v)
#pragma unroll
    for (int w2 = v; w2 < NV; ++w2) {
      const int i = voigt_first<DIM>(v), j = voigt_second<DIM>(v);
      const int k = voigt_first<DIM>(w2), ll = voigt_second<DIM>(w2);
      const T Sijkl = T(0.5) * (T(i == k && j == ll) + T(i == ll && j == k));
      const T IxI = (i == j && k == ll) ? T(1) : T(0);
      M[sym_index<NV>(v, w2)] = (T(2) * cs * Sijkl + T(2) * l * IxI) * dw;
    }
}

// Q1 reference Jacobian at Gauss point (i0,i1,i2) from the cell's corner
// coordinates X[v][i] (LDS); returns J_ref[i][d] = dX_i / dxi_d.
template <int DIM, typename T, typename TC>
__device__ __forceinline__ void q1_jacobian(const TC *X, const T *qp, int i0, int i1, int i2, T (&J)[DIM][DIM]) {
  const T xi[3] = {qp[i0], qp[i1], DIM == 3 ? qp[i2] : T(0)};
#pragma unroll
  for (int i = 0; i < DIM; ++i)
#pragma unroll
    for (int d = 0; d < DIM; ++d) J[i][d] = T(0);
#pragma unroll
  for (int v = 0; v < (1 << DIM); ++v) {
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      T g = T(1);
#pragma unroll
      for (int e = 0; e < DIM; ++e) {
        const int ve = (v >> e) & 1;
        if (e == d) g *= ve ? T(1) : T(-1);
        else g *= ve ? xi[e] : T(1) - xi[e];
      }
#pragma unroll
      for (int i = 0; i < DIM; ++i) J[i][d] = fma(T(X[v * DIM + i]), g, J[i][d]);
    }
  }
}

// local node index of Q1 corner v
template <int DIM, int N> __host__ __device__ constexpr int corner_node(int v) {
  return ((v & 1) ? N - 1 : 0) + ((v >> 1 & 1) ? (N - 1) * N : 0) + (DIM == 3 && (v >> 2 & 1) ? (N - 1) * N * N : 0);
}
// corner index of local node r, or -1
template <int DIM, int N> __device__ __forceinline__ int node_corner(int i0, int i1, int i2) {
  const bool e0 = (i0 == 0 || i0 == N - 1), e1 = (i1 == 0 || i1 == N - 1), e2 = (DIM == 2) || (i2 == 0 || i2 == N - 1);
  if (!(e0 && e1 && e2)) return -1;
  return (i0 ? 1 : 0) + (i1 ? 2 : 0) + ((DIM == 3 && i2) ? 4 : 0);
}

template <typename Number> struct VmultParams {
  const uint32_t *__restrict__ idx;   // [cells][NQP] sorted cells
  const Number *__restrict__ cache;   // SoA caches
  size_t fs;                          // field stride = n_cells * NQP
  const Number *__restrict__ tab;     // S, D, St, Dt, qp, qw (Number)
  const Number *__restrict__ src;
  Number *__restrict__ dst;
  size_t vstride;                     // batched products: distance between vectors (0: one vector)
  const Number *__restrict__ u;       // scalar / residual: displacement
  const double *__restrict__ coords;  // scalar / residual: node coordinates
  const Number *__restrict__ mu;      // per sorted cell
  const Number *__restrict__ lambda;
  int cell_begin, cell_end;  // sorted cells of one colour, or its tiles (tiled layout)
  int formulation;  // 1, or 0: the volumetric/isochoric split (scalar, tensor4, residual)
  int curved;       // support points off the Q1 maps: Eulerian Q_p metric (scalar variants)
  const uint32_t *__restrict__ tnodes;  // tiled layout: [tile][TN] node codes
  const int32_t *__restrict__ tinfo;    // tiled layout: [tile][2] first cell, extents
  // node (hanging) constraints, cells layout only; hang == nullptr: none.
  // Each (cell, constrained node) pair is an instance: its cell reads the
  // weighted sum of the masters and leaves its result in hbuf, from where
  // k_hang_gather adds it to the masters once all colours are done.
  const int32_t *__restrict__ hang;     // [sorted cell][NQP] instance or -1
  const int32_t *__restrict__ hcons;    // [instance] constraint index
  const int32_t *__restrict__ hptr;     // [n + 1]
  const uint32_t *__restrict__ hmaster; // master node | its Dirichlet bits
  const Number *__restrict__ hw;        // weights
  Number *__restrict__ hbuf;            // [vector][instance][DIM] results
  size_t hstride;                       // values per vector in hbuf
  // closed-form diagonal of cells with constrained nodes: per sorted cell
  // its index among them or -1; per such cell its entries (one per node the
  // cell couples: own unconstrained nodes and masters), per entry the local
  // nodes and weights whose basis functions sum to the node's constrained
  // basis function on the cell; results per entry in dbuf
  const int32_t *__restrict__ dcell;
  const int32_t *__restrict__ deptr;
  const int32_t *__restrict__ dtptr;
  const int32_t *__restrict__ dta;
  const Number *__restrict__ dtw;
  Number *__restrict__ dbuf;
  // adjoint: per-cell outputs in mesh order (perm: sorted -> mesh cell)
  double *__restrict__ dmu;
  double *__restrict__ dlambda;
  const int32_t *__restrict__ perm;
  // even-odd forms of the 1D matrices S and D (EvenOdd<N>::PK values each,
  // N <= 5) for the line-per-lane sum factorization: uniform across lanes,
  // read from the kernel arguments
  Number sf[2 * 15];
};

// Kinematic prologue shared by the scalar vmult and the residual: gathers u
// and the corner coordinates, returns F = I + Grad_xi u J_ref^-1 and
// J_ref^-1, det(J_ref) at this lane's quadrature point.  Leaves A,B free.
// Meshes whose support points do not all lie on their cells' Q1 maps
// ("curved"): the current configuration is the Eulerian Q_p mapping
// x(xi) = sum_a (X_a + u_a) phi_a(xi) (MappingQEulerian of degree p,
// include/mf_elasticity.h:1821-1831), so its Jacobian is Grad_xi X + Grad_xi u,
// while the reference quantities (F, the residual) keep the Q1 vertex map
// (StaticMappingQ1, :1828-1829).  Returns GX = Grad_xi X at this lane's
// quadrature point from the support-point coordinates; A (and B for the
// ping-pong sum factorization) are free again on return.
template <int DIM, int N, typename T, bool LINE, int CPW = CellShape<DIM, N>::CPW, int WG = kWG>
__device__ __forceinline__ void geometry_gradient(const double *__restrict__ coords, const T *sf, const T *tab,
                                                  T *lds, T *A, T *B, bool on, bool active, uint32_t code, int t,
                                                  int slot, int r, int i0, int i1, int i2, T (&GX)[DIM][DIM]) {
  constexpr int NQP = ipow(N, DIM);
  const size_t node = code & kNodeMask;
  if (active)
#pragma unroll
    for (int c = 0; c < DIM; ++c) A[c * NQP + r] = T(coords[node * DIM + c]);
  lds_barrier();
  if constexpr (LINE && DIM == 3)
    SFLine<N, T>::template evaluate<CPW, WG>(sf, lds, t, on, slot, r, GX);
  else
    SF<DIM, N, T>::template evaluate<DIM>(tab, A, B, on, r, i0, i1, i2, GX);
  lds_barrier();
}

// current metric of a curved mesh: K = (GX + Gu)^-1 and the integration
// factor JxW_cur / J = det(GX + Gu) w / J
template <int DIM, typename T>
__device__ __forceinline__ void eulerian_metric(const T (&GX)[DIM][DIM], const T (&Gu)[DIM][DIM], T J, T w,
                                                T (&K)[DIM][DIM], T &dw) {
  T Jc[DIM][DIM];
#pragma unroll
  for (int i = 0; i < DIM; ++i)
#pragma unroll
    for (int d = 0; d < DIM; ++d) Jc[i][d] = GX[i][d] + Gu[i][d];
  dw = inverse<DIM>(Jc, K) * w / J;
}

// LINE: 3D gradient by the line-per-lane sum factorization (P.sf set,
// lds = the workgroup's slot buffers, t = thread, slot = this lane's cell).
// Gu = Grad_xi u (rows: components, columns: xi directions).
template <int DIM, int N, typename T, bool LINE = false, int CPW = CellShape<DIM, N>::CPW, int WG = kWG>
__device__ __forceinline__ void kinematics_prologue(const VmultParams<T> &P, const T *tab, T *A, T *B, T *Xc,
                                                    bool on, bool active, uint32_t code, int r, int i0, int i1, int i2,
                                                    T (&F)[DIM][DIM], T (&Jri)[DIM][DIM], T &detJr, T (&Gu)[DIM][DIM],
                                                    T *lds = nullptr, int t = 0, int slot = 0) {
  constexpr int NQP = ipow(N, DIM);
  const int node = code & kNodeMask;
  if (active) {
#pragma unroll
    for (int c = 0; c < DIM; ++c) A[c * NQP + r] = P.u[(size_t)node * DIM + c];
    const int v = node_corner<DIM, N>(i0, i1, i2);
    if (v >= 0)
#pragma unroll
      for (int c = 0; c < DIM; ++c) Xc[v * DIM + c] = T(P.coords[(size_t)node * DIM + c]);
  }
  lds_barrier();
  if constexpr (LINE && DIM == 3)
    SFLine<N, T>::template evaluate<CPW, WG>(P.sf, lds, t, on, slot, r, Gu);
  else
    SF<DIM, N, T>::template evaluate<DIM>(tab, A, B, on, r, i0, i1, i2, Gu);
  T Jr[DIM][DIM];
  q1_jacobian<DIM, T, T>(Xc, tab + 4 * N * N, i0, i1, i2, Jr);
  detJr = inverse<DIM>(Jr, Jri);
  grad_times<DIM>(Gu, Jri, F);
#pragma unroll
  for (int d = 0; d < DIM; ++d) F[d][d] += T(1);
  lds_barrier();  // all lanes done reading A/B/Xc of this prologue
}

// scalar_referential kinematics (include/mf_nh_operator.h:1276-1310): the
// displacement at the Gauss points comes from the cache (uq), its reference
// gradient from the collocation derivative on the Gauss points (no gather of
// u); J_ref from the corner coordinates.  Leaves A free.
template <int DIM, int N, typename T>
__device__ __forceinline__ void kinematics_collocation(const VmultParams<T> &P, const T *tab, T *A, T *Xc, bool on,
                                                       bool active, uint32_t code, int r, int i0, int i1, int i2,
                                                       const T (&uq)[DIM], T (&F)[DIM][DIM], T (&Jri)[DIM][DIM],
                                                       T &detJr, T (&Gu)[DIM][DIM]) {
  constexpr int NQP = ipow(N, DIM);
  const int node = code & kNodeMask;
  if (on) {
#pragma unroll
    for (int c = 0; c < DIM; ++c) A[c * NQP + r] = uq[c];
  }
  if (active) {
    const int v = node_corner<DIM, N>(i0, i1, i2);
    if (v >= 0)
#pragma unroll
      for (int c = 0; c < DIM; ++c) Xc[v * DIM + c] = T(P.coords[(size_t)node * DIM + c]);
  }
  lds_barrier();
  const T *Dc = tab + 4 * N * N + 2 * N;
  const int ii[3] = {i0, i1, i2};
  const int stride[3] = {1, N, N * N};
#pragma unroll
  for (int c = 0; c < DIM; ++c)
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const int base = r - ii[d] * stride[d];  // in [0, NQP) for every lane
      T s = T(0);
#pragma unroll
      for (int k = 0; k < N; ++k) s = fma(Dc[ii[d] * N + k], A[c * NQP + base + k * stride[d]], s);
      Gu[c][d] = s;
    }
  T Jr[DIM][DIM];
  q1_jacobian<DIM, T, T>(Xc, tab + 4 * N * N, i0, i1, i2, Jr);
  detJr = inverse<DIM>(Jr, Jri);
  grad_times<DIM>(Gu, Jri, F);
#pragma unroll
  for (int d = 0; d < DIM; ++d) F[d][d] += T(1);
  lds_barrier();  // all lanes done reading A/Xc
}

// ------------------------------------------------------------------------
// fused vmult kernel, one launch per colour
//   VAR: V_SCALAR / V_SCALAR_REF / V_TENSOR2 / V_TENSOR4 / V_TENSOR4_NS /
//        V_RESIDUAL / V_ADJOINT
//   ADD: dst += A src for every touch; otherwise the first-touch cell of each
//        DoF stores (no separate zeroing pass)
//   CONSTR: constrained DoFs read as zero and are skipped on write
//   PH: phase bits (SF, QD) for breakdown timings
//   TILE: 0 = cells layout; 1 (2: wide tiles) = tiled layout -- one
//        workgroup per tile of a structured mesh; the tile's nodes are
//        gathered once into LDS and the cells' results are summed there
//        (fixed order) before one read-modify-write per node
// ------------------------------------------------------------------------
// Minimum waves per SIMD of the fused cell kernel (HIP's second
// __launch_bounds__ argument): wide tiles at least 4 (two 512-thread
// workgroups per CU, at most 128 VGPRs); the wide tensor2 kernel 6 (three
// workgroups, at most 80 VGPRs, which its FP64 form needs no more than).
template <int VAR, int TILE> constexpr int vmult_min_waves() {
  return TILE != kTileWide ? 1 : VAR == V_TENSOR2 ? 6 : 4;
}

template <int DIM, int N, typename T, int VAR, bool ADD, bool CONSTR, unsigned PH, int TILE>
__global__ void __launch_bounds__(tile_threads<TILE>(), (vmult_min_waves<VAR, TILE>())) k_vmult(const VmultParams<T> P) {
  using CS = CellShape<DIM, N>;
  // threads and cells of this workgroup: a tile, or CS::CPW cells
  constexpr int WG = tile_threads<TILE>();
  using TS = TileShape<DIM, N, WG>;
  constexpr int NQP = CS::NQP, CPW = WG / NQP, NB = CS::NB, NC = CS::NC;
  constexpr int TN = TILE ? TS::TN : 1;
  constexpr bool RECOMP = (VAR == V_RESIDUAL || VAR == V_ADJOINT);  // no cache, c from ln J in FP64
  constexpr int NF = CacheFields<DIM, (RECOMP ? V_SCALAR : VAR)>::value;
  constexpr bool KIN = (VAR == V_SCALAR || VAR == V_SCALAR_REF || RECOMP);
  // the 1D tables in LDS serve the 2D sum factorization and the kinematics;
  // the 3D cached tangents read their line operators from P.sf
  constexpr int TABN = (DIM == 2 || KIN) ? tab_size<N>(VAR) : 0;
  constexpr int RED = (VAR == V_ADJOINT) ? CPW * 2 * NQP : 0;  // per-cell reductions
  // per-cell LDS buffers: 3D sums in place (SFLine, NB arrays), 2D ping-pongs
  constexpr int SLOTW = (DIM == 3 ? 1 : 2) * NB * NQP;
  __shared__ T lds[CPW * SLOTW + TABN + (KIN ? CPW * NC * DIM : 0) + RED];
  __shared__ T tile_v[TILE ? TN * DIM : 1];         // tile nodes' source values
  __shared__ uint32_t tile_c[TILE && KIN ? TN : 1];  // tile nodes' codes (kinematics)
  T *tab = lds + CPW * SLOTW;
  T *Xall = tab + TABN;
  T *Rall = Xall + (KIN ? CPW * NC * DIM : 0);

  // batched products (all columns of a small matrix in one launch): vector
  // blockIdx.y lives at offset blockIdx.y * vstride of src and dst
  const T *__restrict__ src = P.src + (size_t)blockIdx.y * P.vstride;
  T *__restrict__ dst = P.dst + (size_t)blockIdx.y * P.vstride;
  const int t = threadIdx.x;
  if (t < TABN) tab[t] = P.tab[t];
  const int slot = t / NQP;
  const int r = t - slot * NQP;
  const int i0 = r % N, i1 = (r / N) % N, i2 = (DIM == 3) ? r / (N * N) : 0;
  const int L = xcd_remap(blockIdx.x, gridDim.x);
  const bool on = slot < CPW;
  T *A = lds + (slot < CPW ? slot : 0) * SLOTW;
  T *B = DIM == 3 ? A : A + NB * NQP;  // 2D sum factorization only
  T *Xc = Xall + (slot < CPW ? slot : 0) * NC * DIM;

  // Load order matters: vmcnt is an in-order counter, so consuming a load
  // waits for every load issued before it.  The dependent index -> source
  // gather goes first; the per-qp cache stream is issued after it and stays
  // in flight through the sum-factorization stages (LDS-only barriers).
  // The loads are branch-free (idle lanes read a valid cell of this colour
  // and discard the values) so the wait counts stay static.
  int cell;
  bool active;
  size_t cell_s;
  uint32_t code = 0;
  size_t node = 0;
  // tiled layout: slot -> (sx, sy, sz) in the tile, lane t < TN -> tile node
  const int sx = slot % TS::TX, sy = (slot / TS::TX) % TS::TY, sz = slot / (TS::TX * TS::TY);
  const int tile = P.cell_begin + L;
  uint32_t tcode = 0;
  bool tvalid = false;
  size_t tnode = 0;
  T sv[DIM];
  T dv[DIM];
#pragma unroll
  for (int c = 0; c < DIM; ++c) dv[c] = T(0);
  if constexpr (TILE) {
    const int first = P.tinfo[2 * (size_t)tile], ext = P.tinfo[2 * (size_t)tile + 1];
    const int ex = ext & 0xff, ey = (ext >> 8) & 0xff, ez = ext >> 16;
    active = on && sx < ex && sy < ey && sz < ez;
    cell = first + sx + ex * (sy + ey * sz);
    cell_s = active ? (size_t)cell : (size_t)first;
    tcode = P.tnodes[(size_t)tile * TN + (t < TN ? t : 0)];
    tvalid = t < TN && tcode != kNoNode;
    tnode = tvalid ? (tcode & kNodeMask) : 0;
    if constexpr (VAR != V_RESIDUAL) load_node<DIM, T>(src + tnode * DIM, sv);
    // the node's destination values, read now rather than at the final
    // read-modify-write: no other tile of this colour touches the node, so
    // they do not change before this tile writes them
    if (tvalid && (ADD || !(tcode & kFirstBit))) load_node<DIM, T>(dst + tnode * DIM, dv);
  } else {
    cell = P.cell_begin + L * CPW + slot;
    active = on && (cell < P.cell_end);
    cell_s = active ? (size_t)cell : (size_t)P.cell_begin;
    code = P.idx[cell_s * NQP + (on ? r : 0)];
    node = code & kNodeMask;
    if constexpr (!KIN) {
#pragma unroll
      for (int c = 0; c < DIM; ++c) sv[c] = src[node * DIM + c];
    }
  }
  const int r_s = on ? r : 0;
  __builtin_amdgcn_sched_barrier(0);
  T cch[NF > 0 ? NF : 1];
  if constexpr ((PH & PH_QD) && NF > 0 && !RECOMP) load_cache<T, NF>(P.cache, cell_s * NQP + r_s, P.fs, cch);
  __builtin_amdgcn_sched_barrier(0);

  // tile-local index of this lane's node (tiled layout)
  const int ln = on ? (sx * (N - 1) + i0) + TS::NX * ((sy * (N - 1) + i1) + TS::NY * (sz * (N - 1) + i2)) : 0;
  if constexpr (TILE) {
    if (t < TN) {
      if constexpr (KIN) tile_c[t] = tcode;
      if constexpr (VAR == V_RESIDUAL) {
        // no source values: the tile holds the destination values throughout
#pragma unroll
        for (int c = 0; c < DIM; ++c) tile_v[t * DIM + c] = dv[c];
      } else {
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
          const bool zero = !tvalid || (CONSTR && ((tcode >> (kConsShift + c)) & 1u));
          tile_v[t * DIM + c] = zero ? T(0) : sv[c];
        }
      }
    }
    lds_barrier();
    if constexpr (KIN) {
      code = tile_c[ln];
      node = code & kNodeMask;
    }
  }

  T F[DIM][DIM], Jri[DIM][DIM], detJr = T(1);
  T Gu[DIM][DIM], GX[DIM][DIM];  // Grad_xi u, Grad_xi X (curved meshes)
  if constexpr (VAR == V_SCALAR_REF && (PH & PH_QD)) {
    T uq[DIM];
#pragma unroll
    for (int c = 0; c < DIM; ++c) uq[c] = cch[1 + c];
    // referential: the Q1 metric only, whatever the mesh (:1255-1400)
    kinematics_collocation<DIM, N, T>(P, tab, A, Xc, on, active, code, r, i0, i1, i2, uq, F, Jri, detJr, Gu);
  } else if constexpr (KIN && (PH & PH_QD)) {
    kinematics_prologue<DIM, N, T, true, CPW, WG>(P, tab, A, B, Xc, on, active, code, r, i0, i1, i2, F, Jri, detJr,
                                                  Gu, lds, t, slot);
    if (VAR == V_SCALAR && P.curved)
      geometry_gradient<DIM, N, T, true, CPW, WG>(P.coords, P.sf, tab, lds, A, B, on, active, code, t, slot, r, i0,
                                                  i1, i2, GX);
  }

  T G[DIM][DIM];
  T out[DIM];
  if constexpr (VAR != V_RESIDUAL) {
    // gather src with constraints (read_dof_values, include/mf_nh_operator.h:745-749)
    if constexpr (TILE) {
      if (on)
#pragma unroll
        for (int c = 0; c < DIM; ++c) A[c * NQP + r] = tile_v[ln * DIM + c];
    } else if (active) {
      const int hidx = P.hang ? P.hang[cell_s * NQP + r_s] : -1;
      const int hk = hidx >= 0 ? P.hcons[hidx] : 0;
#pragma unroll
      for (int c = 0; c < DIM; ++c) {
        T v;
        if (hidx >= 0) {
          // a constrained node reads the weighted sum of its masters
          v = T(0);
          for (int j = P.hptr[hk]; j < P.hptr[hk + 1]; ++j) {
            const uint32_t mc = P.hmaster[j];
            if (CONSTR && ((mc >> (kConsShift + c)) & 1u)) continue;
            v = fma(P.hw[j], src[(size_t)(mc & kNodeMask) * DIM + c], v);
          }
        } else {
          if constexpr (KIN) v = src[node * DIM + c];
          else v = sv[c];
          if (CONSTR && ((code >> (kConsShift + c)) & 1u)) v = T(0);
        }
        A[c * NQP + r] = v;
      }
    }
    lds_barrier();
    if constexpr (TILE) {
      // every lane has gathered: the tile's destination values take the
      // source values' place (out of registers until the final write)
      if (t < TN)
#pragma unroll
        for (int c = 0; c < DIM; ++c) tile_v[t * DIM + c] = dv[c];
    }
    if constexpr (!(PH & PH_SF)) {
#pragma unroll
      for (int c = 0; c < DIM; ++c) out[c] = A[c * NQP + r];
    } else {
      if constexpr (DIM == 3)
        SFLine<N, T>::template evaluate<CPW, WG>(P.sf, lds, t, on, slot, r, G);
      else
        SF<DIM, N, T>::template evaluate<DIM>(tab, A, B, on, r, i0, i1, i2, G);
    }
  }

  if constexpr ((PH & PH_SF) || VAR == V_RESIDUAL) {
    T Phi[DIM][DIM];
    if constexpr (!(PH & PH_QD)) {
#pragma unroll
      for (int c = 0; c < DIM; ++c)
#pragma unroll
        for (int d = 0; d < DIM; ++d) Phi[c][d] = G[c][d];
    } else if constexpr (VAR == V_TENSOR4) {
      qp_tensor4<DIM, T>(G, cch, Phi);
    } else if constexpr (VAR == V_TENSOR4_NS) {
      qp_tensor4_ns<DIM, T>(G, cch, Phi);
    } else if constexpr (VAR == V_TENSOR2) {
      qp_tensor2<DIM, T>(G, cch, Phi);
    } else {
      // scalar (include/mf_nh_operator.h:1182-1254), scalar_referential
      // (:1255-1400) or residual (include/mf_elasticity.h:2769-2804);
      // kinematics recomputed here
      const T m = active ? P.mu[cell] : T(0), l = active ? P.lambda[cell] : T(0);
      T Fi[DIM][DIM];
      const T J = inverse<DIM>(F, Fi);
      // quadrature weight from the table (qw after qp)
      const T *qw = tab + 4 * N * N + N;
      const T w = (DIM == 3) ? qw[i0] * qw[i1] * qw[i2] : qw[i0] * qw[i1];
      T dw = detJr * w;
      T K[DIM][DIM];  // J_cur^-1 = J_ref^-1 F^-1
      if (VAR == V_SCALAR && P.curved)
        eulerian_metric<DIM, T>(GX, Gu, J, w, K, dw);
      else
        grad_times<DIM>(Jri, Fi, K);
      if (P.formulation == 0 && VAR != V_ADJOINT) {
        // formulation 0 (include/mf_nh_operator.h:1005-1180): the scalar
        // cache holds J^{-1/d}; the residual evaluates it here
        T sbar;
        if constexpr (RECOMP) sbar = T(pow(double(J), -1.0 / DIM));
        else sbar = active ? cch[0] : T(1);
        const Form0<DIM, T> f0(F, J, sbar, m, l);
        if constexpr (VAR == V_RESIDUAL) {
#pragma unroll
          for (int c = 0; c < DIM; ++c)
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
              T s = f0.tau[c][0] * K[d][0];
#pragma unroll
              for (int e = 1; e < DIM; ++e) s = fma(f0.tau[c][e], K[d][e], s);
              Phi[c][d] = -s * dw;
            }
        } else {
          T g[DIM][DIM], Q[DIM][DIM];
          grad_times<DIM>(G, K, g);
          f0.action(g, Q);
#pragma unroll
          for (int i = 0; i < DIM; ++i)
#pragma unroll
            for (int j = 0; j < DIM; ++j) Q[i][j] *= dw;
          flux_times_kt<DIM>(Q, K, Phi);
        }
      } else {
      T tau[DIM][DIM];
      T cs;
      if constexpr (RECOMP) cs = m - T(2) * l * T(log(double(J)));
      else cs = active ? cch[0] : T(0);
#pragma unroll
      for (int i = 0; i < DIM; ++i)
#pragma unroll
        for (int j = 0; j < DIM; ++j) {
          T s = F[i][0] * F[j][0];
#pragma unroll
          for (int k = 1; k < DIM; ++k) s = fma(F[i][k], F[j][k], s);
          tau[i][j] = m * s - (i == j ? cs : T(0));
        }
      if constexpr (VAR == V_RESIDUAL) {
        // flux = -(tau F^-T) J_ref^-T det(J_ref) w = -tau K^T dw
#pragma unroll
        for (int c = 0; c < DIM; ++c)
#pragma unroll
          for (int d = 0; d < DIM; ++d) {
            T s = tau[c][0] * K[d][0];
#pragma unroll
            for (int e = 1; e < DIM; ++e) s = fma(tau[c][e], K[d][e], s);
            Phi[c][d] = -s * dw;
          }
      } else {
        T g[DIM][DIM];
        grad_times<DIM>(G, K, g);
        T tr = g[0][0];
#pragma unroll
        for (int d = 1; d < DIM; ++d) tr += g[d][d];
        T Q[DIM][DIM];
#pragma unroll
        for (int i = 0; i < DIM; ++i)
#pragma unroll
          for (int j = 0; j < DIM; ++j) {
            T s = cs * (g[i][j] + g[j][i]);
            if (i == j) s = fma(T(2) * l, tr, s);
#pragma unroll
            for (int k = 0; k < DIM; ++k) s = fma(g[i][k], tau[k][j], s);
            Q[i][j] = s * dw;
          }
        flux_times_kt<DIM>(Q, K, Phi);
        if constexpr (VAR == V_ADJOINT) {
          // (dR/du)^T w = -K w, and with tau = mu (b - I) + 2 lambda ln J I:
          // (dR/dmu_c)^T w = -sum_q (b - I) : g dw, (dR/dlambda_c)^T w =
          // -sum_q 2 ln J tr(g) dw, where g = grad_x w.  Recomputed from u.
#pragma unroll
          for (int c = 0; c < DIM; ++c)
#pragma unroll
            for (int d = 0; d < DIM; ++d) Phi[c][d] = -Phi[c][d];
          T bg = T(0);
#pragma unroll
          for (int i = 0; i < DIM; ++i)
#pragma unroll
            for (int j = 0; j < DIM; ++j) {
              // b_ij - delta_ij = (tau_ij + cs delta_ij) / mu - delta_ij
              T s = F[i][0] * F[j][0];
#pragma unroll
              for (int k = 1; k < DIM; ++k) s = fma(F[i][k], F[j][k], s);
              bg = fma(s - (i == j ? T(1) : T(0)), g[i][j], bg);
            }
          if (on) {
            T *Rs = Rall + (slot < CPW ? slot : 0) * 2 * NQP;
            Rs[r] = -bg * dw;
            Rs[NQP + r] = -T(2) * T(log(double(J))) * tr * dw;
          }
        }
      }
      }  // formulation 1
    }
    if constexpr (DIM == 3)
      SFLine<N, T>::template integrate<CPW, WG>(P.sf, lds, t, on, slot, r, Phi, out);
    else
      SF<DIM, N, T>::integrate(tab, A, B, on, r, i0, i1, i2, Phi, out);
    if constexpr (VAR == V_ADJOINT) {
      // the integration's barriers ordered the reduction writes above
      if (active && r == 0) {
        const T *Rs = Rall + slot * 2 * NQP;
        T a = T(0), b = T(0);
        for (int q = 0; q < NQP; ++q) {
          a += Rs[q];
          b += Rs[NQP + q];
        }
        const int mc = P.perm[cell];
        P.dmu[mc] = double(a);
        P.dlambda[mc] = double(b);
      }
    }
  }

  if constexpr (TILE) {
    // sum the cells' results per tile node in LDS (fixed order), then one
    // write per node (distribute_local_to_global, :769-780).  In 3D the
    // results are read where the integration (or, without it, the gather)
    // left them after its last barrier: component c of point rr of slot s
    // at lds[s * SLOTW + c * NQP + rr]; in 2D they are copied out first.
    T *res = lds;
    if constexpr (DIM == 2) {
      lds_barrier();  // every lane is done with the sum-factorization buffers
      if (on)
#pragma unroll
        for (int c = 0; c < DIM; ++c) res[t * DIM + c] = out[c];
      lds_barrier();
    } else {
      (void)out;
    }
    if (tvalid) {
      const int first = P.tinfo[2 * (size_t)tile], ext = P.tinfo[2 * (size_t)tile + 1];
      (void)first;
      const int ex = ext & 0xff, ey = (ext >> 8) & 0xff, ez = ext >> 16;
      const int X = t % TS::NX, Y = (t / TS::NX) % TS::NY, Z = t / (TS::NX * TS::NY);
      constexpr int PP = N - 1;
      T acc[DIM];
#pragma unroll
      for (int c = 0; c < DIM; ++c) acc[c] = T(0);
      const int zlo = Z == 0 ? 0 : (Z - 1) / PP, zhi = min(Z / PP, ez - 1);
      const int ylo = Y == 0 ? 0 : (Y - 1) / PP, yhi = min(Y / PP, ey - 1);
      const int xlo = X == 0 ? 0 : (X - 1) / PP, xhi = min(X / PP, ex - 1);
      for (int cz = zlo; cz <= zhi; ++cz)
        for (int cy = ylo; cy <= yhi; ++cy)
          for (int cx = xlo; cx <= xhi; ++cx) {
            const int rr = (X - cx * PP) + N * ((Y - cy * PP) + N * (Z - cz * PP));
            const int s = cx + TS::TX * (cy + TS::TY * cz);
#pragma unroll
            for (int c = 0; c < DIM; ++c)
              acc[c] += DIM == 3 ? res[s * SLOTW + c * NQP + rr] : res[(s * NQP + rr) * DIM + c];
          }
      T nv[DIM];
#pragma unroll
      for (int c = 0; c < DIM; ++c) nv[c] = tile_v[t * DIM + c] + acc[c];  // destination: 0 on first touch
      if (!CONSTR || !((tcode >> kConsShift) & ((1u << DIM) - 1u))) {
        store_node<DIM, T>(dst + tnode * DIM, nv);
      } else {
#pragma unroll
        for (int c = 0; c < DIM; ++c)
          if (!((tcode >> (kConsShift + c)) & 1u)) dst[tnode * DIM + c] = nv[c];
      }
    }
    return;
  }
  // colour-ordered scatter (distribute_local_to_global, :769-780): all reads
  // of the node's values are issued before the writes (one round trip)
  const int hidx = (P.hang && active) ? P.hang[cell_s * NQP + r_s] : -1;
  if (active && hidx < 0) {
    const bool first = !ADD && (code & kFirstBit);
    T old[DIM];
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      const bool skip = CONSTR && ((code >> (kConsShift + c)) & 1u);
      old[c] = (first || skip) ? T(0) : dst[node * DIM + c];
    }
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      if (CONSTR && ((code >> (kConsShift + c)) & 1u)) continue;
      dst[node * DIM + c] = old[c] + out[c];
    }
  }
  // a constrained node's result waits in hbuf for k_hang_gather
  // (distribute_local_to_global with constraints, :769-780)
  if (hidx >= 0) {
    T *h = P.hbuf + (size_t)blockIdx.y * P.hstride + (size_t)hidx * DIM;
#pragma unroll
    for (int c = 0; c < DIM; ++c) h[c] = out[c];
  }
}

// After all colours: every master adds the weighted results of the
// constrained-node instances that name it, in instance order (fixed, so the
// sums are reproducible), to its components that are not Dirichlet.  One
// thread per master, grid y = vector.  w == nullptr: unit weights (the
// diagonal's per-entry results).
template <int DIM, typename T, bool CONSTR>
__global__ void k_hang_gather(T *__restrict__ dst, size_t vstride, const T *__restrict__ hbuf, size_t hstride,
                              const uint32_t *__restrict__ node, const int32_t *__restrict__ ptr,
                              const int32_t *__restrict__ inst, const T *__restrict__ w, int n) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const uint32_t code = node[q];
  const size_t m = code & kNodeMask;
  const T *hb = hbuf + (size_t)blockIdx.y * hstride;
  T acc[DIM];
#pragma unroll
  for (int c = 0; c < DIM; ++c) acc[c] = T(0);
  for (int e = ptr[q]; e < ptr[q + 1]; ++e) {
    const T we = w ? w[e] : T(1);
    const T *h = hb + (size_t)inst[e] * DIM;
#pragma unroll
    for (int c = 0; c < DIM; ++c) acc[c] = fma(we, h[c], acc[c]);
  }
  T *d = dst + (size_t)blockIdx.y * vstride + m * DIM;
#pragma unroll
  for (int c = 0; c < DIM; ++c)
    if (!CONSTR || !((code >> (kConsShift + c)) & 1u)) d[c] += acc[c];
}

template <typename T> __global__ void k_zero(T *__restrict__ p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = T(0);
}

template <typename T>
__global__ void k_constrained_identity(const int32_t *__restrict__ list, int n, const T *__restrict__ src,
                                       T *__restrict__ dst, int add) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const int k = list[i];
    dst[k] = add ? dst[k] + src[k] : src[k];
  }
}

// ------------------------------------------------------------------------
// cache kernel (include/mf_nh_operator.h:447-592, formulation 1): FP64
// kinematics for both storage precisions; counts det F <= 0.
// ------------------------------------------------------------------------
struct CacheParams {
  const uint32_t *__restrict__ idx;
  const double *__restrict__ tab;  // S, D, St, Dt, qp, qw (double)
  const double *__restrict__ u;
  const double *__restrict__ coords;
  const double *__restrict__ mu;
  const double *__restrict__ lambda;
  void *cache;
  size_t fs;
  int n_cells;
  int formulation;
  int curved;  // Eulerian Q_p metric for the spatial caches (tensor2, tensor4)
  int *bad;
};

template <int DIM, int N, typename T, int VAR> __global__ void __launch_bounds__(kWG) k_cache(const CacheParams P) {
  using CS = CellShape<DIM, N>;
  constexpr int NQP = CS::NQP, CPW = CS::CPW, NB = CS::NB, NC = CS::NC;
  constexpr int NV = Voigt<DIM>::NV;
  constexpr int TABN = 4 * N * N + 2 * N;
  __shared__ double lds[CPW * 2 * NB * NQP + TABN + CPW * NC * DIM];
  double *tab = lds + CPW * 2 * NB * NQP;
  const int t = threadIdx.x;
  if (t < TABN) tab[t] = P.tab[t];
  const int slot = t / NQP, r = t - slot * NQP;
  const int i0 = r % N, i1 = (r / N) % N, i2 = (DIM == 3) ? r / (N * N) : 0;
  const int cell = blockIdx.x * CPW + slot;
  constexpr int D2 = DIM * DIM;
  const bool on = slot < CPW;
  const bool active = on && (cell < P.n_cells);
  double *A = lds + (slot < CPW ? slot : 0) * 2 * NB * NQP;
  double *B = A + NB * NQP;
  double *Xc = tab + TABN + (slot < CPW ? slot : 0) * NC * DIM;
  uint32_t code = 0;
  if (active) code = P.idx[(size_t)cell * NQP + r];
  VmultParams<double> KP{};
  KP.formulation = P.formulation;
  KP.u = P.u;
  KP.coords = P.coords;
  double F[DIM][DIM], Jri[DIM][DIM], detJr, Gu[DIM][DIM], GX[DIM][DIM];
  kinematics_prologue<DIM, N, double>(KP, tab, A, B, Xc, on, active, code, r, i0, i1, i2, F, Jri, detJr, Gu);
  // the spatial caches carry the current metric: Eulerian Q_p on curved meshes
  constexpr bool SPATIAL = (VAR == V_TENSOR2 || VAR == V_TENSOR4);
  if (SPATIAL && P.curved)
    geometry_gradient<DIM, N, double, false>(P.coords, nullptr, tab, lds, A, B, on, active, code, t, slot, r, i0, i1,
                                             i2, GX);
  double uq[DIM];
  if constexpr (VAR == V_SCALAR_REF) {
    // u at this Gauss point: the tensor-product interpolation of the nodal
    // values (phi_reference values, include/mf_nh_operator.h:527-533)
    const int node = code & kNodeMask;
    if (active)
#pragma unroll
      for (int c = 0; c < DIM; ++c) A[c * NQP + r] = P.u[(size_t)node * DIM + c];
    lds_barrier();
    const double *S = tab;
#pragma unroll
    for (int c = 0; c < DIM; ++c) uq[c] = 0.0;
    if (on)
      for (int a = 0; a < NQP; ++a) {
        const int a0 = a % N, a1 = (a / N) % N, a2 = (DIM == 3) ? a / (N * N) : 0;
        double ph = S[i0 * N + a0] * S[i1 * N + a1];
        if (DIM == 3) ph *= S[i2 * N + a2];
#pragma unroll
        for (int c = 0; c < DIM; ++c) uq[c] = fma(ph, A[c * NQP + a], uq[c]);
      }
  }
  if (!active) return;
  const double m = P.mu[cell], l = P.lambda[cell];
  double Fi[DIM][DIM];
  const double J = inverse<DIM>(F, Fi);
  if (!(J > 0.0)) {
    atomicAdd(P.bad, 1);
  }
  const double cs = m - 2.0 * l * log(J);
  const double *qw = tab + 4 * N * N + N;
  const double w = (DIM == 3) ? qw[i0] * qw[i1] * qw[i2] : qw[i0] * qw[i1];
  double dw = detJr * w;
  double K[DIM][DIM];
  if (SPATIAL && P.curved)
    eulerian_metric<DIM, double>(GX, Gu, J, w, K, dw);
  else
    grad_times<DIM>(Jri, Fi, K);
  T *out = static_cast<T *>(P.cache);
  const size_t cq = (size_t)cell * NQP + r;
  constexpr int NFC = CacheFields<DIM, VAR>::value;
  auto put = [&](int f, double v) { out[cache_at<T>(f, NFC, cq, P.fs)] = T(v); };
  if constexpr (VAR == V_SCALAR) {
    // formulation 0 caches J^{-1/d} (include/mf_nh_operator.h:467-487)
    put(0, P.formulation == 0 ? pow(J, -1.0 / DIM) : cs);
  } else if constexpr (VAR == V_SCALAR_REF) {
    put(0, cs);
#pragma unroll
    for (int c = 0; c < DIM; ++c) put(1 + c, uq[c]);
  } else if constexpr (VAR == V_TENSOR4_NS) {
    // dP/dF_(iJ)(kL) = mu d_ik d_JL + c Fi_Jk Fi_Li + 2 lambda Fi_Ji Fi_Lk
    // (include/mf_nh_operator.h:563-582), times det(J_ref) w
#pragma unroll
    for (int d = 0; d < DIM; ++d)
#pragma unroll
      for (int j = 0; j < DIM; ++j) put(d * DIM + j, Jri[d][j]);
#pragma unroll
    for (int p = 0; p < D2; ++p)
#pragma unroll
      for (int q = p; q < D2; ++q) {
        const int i = p / DIM, Jc = p % DIM, k = q / DIM, L = q % DIM;
        const double a = (i == k && Jc == L ? m : 0.0) + cs * Fi[Jc][k] * Fi[L][i] + 2.0 * l * Fi[Jc][i] * Fi[L][k];
        put(D2 + sym_index<D2>(p, q), a * dw);
      }
  } else if constexpr (VAR == V_TENSOR4) {
#pragma unroll
    for (int d = 0; d < DIM; ++d)
#pragma unroll
      for (int j = 0; j < DIM; ++j) put(d * DIM + j, K[d][j]);
    double Tv[NV], M[NV * (NV + 1) / 2];
    tensor4_fields<DIM, double>(F, J, m, l, dw, P.formulation, Tv, M);
#pragma unroll
    for (int v = 0; v < NV; ++v) put(D2 + v, Tv[v]);
#pragma unroll
    for (int k = 0; k < NV * (NV + 1) / 2; ++k) put(D2 + NV + k, M[k]);
  } else {  // tensor2 (include/mf_nh_operator.h:535-547)
#pragma unroll
    for (int d = 0; d < DIM; ++d)
#pragma unroll
      for (int j = 0; j < DIM; ++j) put(d * DIM + j, K[d][j]);
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const int i = voigt_first<DIM>(v), j = voigt_second<DIM>(v);
      double s = 0;
#pragma unroll
      for (int k = 0; k < DIM; ++k) s += F[i][k] * F[j][k];
      const double tau = m * s - (i == j ? cs : 0.0);
      put(DIM * DIM + v, tau * dw);
    }
    put(DIM * DIM + NV, 2.0 * cs * dw);
    put(DIM * DIM + NV + 1, 2.0 * l * dw);
  }
}

// ------------------------------------------------------------------------
// exact diagonal (include/mf_nh_operator.h:862-918 probes unit vectors; the
// same entries in closed form): for local DoF (a, c) with
// gamma = K^T grad_xi phi_a(q):
//   A_(a,c),(a,c) = sum_q gamma.T.gamma + s~^T M s~   (tensor4)
//                 = sum_q gamma.T.gamma + c1/2 (|gamma|^2 + gamma_c^2)
//                          + c2 gamma_c^2               (tensor2 / scalar)
// where s~ is the Voigt vector of sym(e_c (x) gamma) with doubled shears.
// ------------------------------------------------------------------------
template <int DIM, int N, typename T, int VAR, int TILE>
__global__ void __launch_bounds__(kWG) k_diagonal(const VmultParams<T> P) {
  using CS = CellShape<DIM, N>;
  using TS = TileShape<DIM, N, tile_threads<TILE>()>;  // the layout's tiles (this kernel: 256 threads)
  constexpr int NQP = CS::NQP, CPW = CS::CPW, NB = CS::NB, NC = CS::NC;
  constexpr int NV = Voigt<DIM>::NV;
  // the scalar variant evaluates the tensor4 fields per point (either
  // formulation) and shares the tensor4 closed form
  constexpr int NF = (VAR == V_TENSOR2) ? CacheFields<DIM, V_TENSOR2>::value
                     : (VAR == V_TENSOR4_NS) ? CacheFields<DIM, V_TENSOR4_NS>::value
                                             : CacheFields<DIM, V_TENSOR4>::value;
  constexpr int TABN = 4 * N * N + 2 * N;
  constexpr int SFN = (VAR == V_SCALAR) ? CPW * 2 * NB * NQP : 0;
  constexpr int XN = (VAR == V_SCALAR) ? CPW * NC * DIM : 0;
  // the kinematic prologue's scratch (SF ping-pong, corners) is dead once the
  // per-point fields are written, so it shares the fields' region
  constexpr int QDN = CPW * NF * NQP > SFN + XN ? CPW * NF * NQP : SFN + XN;
  // a tile of TC cells is done in ROUNDS rounds of CPW cells (3D Q4 tiles
  // hold four cells, twice this kernel's); the per-lane results of all the
  // tile's cells then live in their own region, with one round in the
  // fields' region once those are dead
  constexpr int TC = TILE ? TS::CPW : CPW;
  constexpr int ROUNDS = (TC + CPW - 1) / CPW;
  constexpr int RESN = ROUNDS > 1 ? TC * NQP * DIM : 0;  // TC cells' results
  static_assert(ROUNDS > 1 || CPW * NQP * DIM <= QDN, "tile results fit the fields' region");
  __shared__ T lds[QDN + TABN + RESN];
  T *qd_all = lds;
  T *res = ROUNDS > 1 ? lds + QDN + TABN : lds;
  T *tab = lds + QDN;
  T *sf_all = lds;
  T *X_all = sf_all + SFN;
  const int t = threadIdx.x;
  if (t < TABN) tab[t] = P.tab[t];
  const int slot = t / NQP, r = t - slot * NQP;
  const int i0 = r % N, i1 = (r / N) % N, i2 = (DIM == 3) ? r / (N * N) : 0;
  const int L = xcd_remap(blockIdx.x, gridDim.x);
  const bool on = slot < CPW;
  const int tile = P.cell_begin + L;
  int ex = 0, ey = 0, ez = 0;
  for (int round = 0; round < ROUNDS; ++round) {
  const int ts = round * CPW + slot;  // this lane's cell within the tile
  int cell;
  bool active;
  if constexpr (TILE) {
    const int first = P.tinfo[2 * (size_t)tile], ext = P.tinfo[2 * (size_t)tile + 1];
    ex = ext & 0xff, ey = (ext >> 8) & 0xff, ez = ext >> 16;
    const int sx = ts % TS::TX, sy = (ts / TS::TX) % TS::TY, sz = ts / (TS::TX * TS::TY);
    active = on && ts < TC && sx < ex && sy < ey && sz < ez;
    cell = first + sx + ex * (sy + ey * sz);
  } else {
    cell = P.cell_begin + L * CPW + slot;
    active = on && (cell < P.cell_end);
  }
  const int sl = slot < CPW ? slot : 0;
  T *qd = qd_all + sl * NF * NQP;
  uint32_t code = 0;
  if (active) code = P.idx[(size_t)cell * NQP + r];
  if constexpr (VAR == V_SCALAR) {
    T *A = sf_all + sl * 2 * NB * NQP;
    T *B = A + NB * NQP;
    T *Xc = X_all + sl * NC * DIM;
    T F[DIM][DIM], Jri[DIM][DIM], detJr, Gu[DIM][DIM], GX[DIM][DIM];
    kinematics_prologue<DIM, N, T>(P, tab, A, B, Xc, on, active, code, r, i0, i1, i2, F, Jri, detJr, Gu);
    if (P.curved)
      geometry_gradient<DIM, N, T, false>(P.coords, P.sf, tab, lds, A, B, on, active, code, t, slot, r, i0, i1, i2,
                                          GX);
    if (active) {
      const T m = P.mu[cell], l = P.lambda[cell];
      T Fi[DIM][DIM], K[DIM][DIM];
      const T J = inverse<DIM>(F, Fi);
      const T *qw = tab + 4 * N * N + N;
      const T w = (DIM == 3) ? qw[i0] * qw[i1] * qw[i2] : qw[i0] * qw[i1];
      T dw = detJr * w;
      if (P.curved)
        eulerian_metric<DIM, T>(GX, Gu, J, w, K, dw);
      else
        grad_times<DIM>(Jri, Fi, K);
      T Tv[NV], M[NV * (NV + 1) / 2];
      tensor4_fields<DIM, T>(F, J, m, l, dw, P.formulation, Tv, M);
#pragma unroll
      for (int d = 0; d < DIM; ++d)
#pragma unroll
        for (int j = 0; j < DIM; ++j) qd[(d * DIM + j) * NQP + r] = K[d][j];
#pragma unroll
      for (int v = 0; v < NV; ++v) qd[(DIM * DIM + v) * NQP + r] = Tv[v];
#pragma unroll
      for (int k = 0; k < NV * (NV + 1) / 2; ++k) qd[(DIM * DIM + NV + k) * NQP + r] = M[k];
    }
  } else {
    if (active) {
      T v[NF];
      load_cache<T, NF>(P.cache, (size_t)cell * NQP + r, P.fs, v);
#pragma unroll
      for (int f = 0; f < NF; ++f) qd[f * NQP + r] = v[f];
    }
  }
  lds_barrier();
  if (!TILE && !active) return;  // the tiled layout's lanes all reach the sum below
  const T *S = tab, *D = tab + N * N;
  // reference gradient of the basis function of local node a at point q
  auto grad_ref = [&](int a, int q, T (&g)[DIM]) {
    const int ai[3] = {a % N, (a / N) % N, (DIM == 3) ? a / (N * N) : 0};
    const int qi[3] = {q % N, (q / N) % N, (DIM == 3) ? q / (N * N) : 0};
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      T v = T(1);
#pragma unroll
      for (int e = 0; e < DIM; ++e) v *= (e == d ? D : S)[qi[e] * N + ai[e]];
      g[d] = v;
    }
  };
  // point q's share of the diagonal entries (one per component) of the
  // basis function with reference gradient gph
  auto point = [&](int q, const T (&gph)[DIM], T (&acc)[DIM]) {
    T gam[DIM];
#pragma unroll
    for (int j = 0; j < DIM; ++j) {
      T s = T(0);
#pragma unroll
      for (int d = 0; d < DIM; ++d) s = fma(gph[d], qd[(d * DIM + j) * NQP + q], s);
      gam[j] = s;
    }
    if constexpr (VAR == V_TENSOR4_NS) {
      // gam = J_ref^-T grad_xi phi_a: A_(a,c),(a,c) = sum_q gam_e A_(c e)(c f) gam_f
      constexpr int D2 = DIM * DIM;
      const T *Am = qd + D2 * NQP;
#pragma unroll
      for (int c = 0; c < DIM; ++c) {
        T s = T(0);
#pragma unroll
        for (int e = 0; e < DIM; ++e)
#pragma unroll
          for (int f = 0; f < DIM; ++f) s = fma(gam[e] * gam[f], Am[sym_index<D2>(c * DIM + e, c * DIM + f) * NQP + q], s);
        acc[c] += s;
      }
      return;
    }
    T geo = T(0);
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
      for (int j = 0; j < DIM; ++j) geo = fma(gam[i] * gam[j], qd[(DIM * DIM + voigt_index<DIM>(i, j)) * NQP + q], geo);
    if constexpr (VAR == V_TENSOR4 || VAR == V_SCALAR) {
      const T *M = qd + (DIM * DIM + NV) * NQP;
#pragma unroll
      for (int c = 0; c < DIM; ++c) {
        T sv[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          const int i = voigt_first<DIM>(v), j = voigt_second<DIM>(v);
          sv[v] = (i == j) ? (i == c ? gam[c] : T(0)) : ((i == c ? gam[j] : T(0)) + (j == c ? gam[i] : T(0)));
        }
        T s = T(0);
#pragma unroll
        for (int v = 0; v < NV; ++v)
#pragma unroll
          for (int w = 0; w < NV; ++w) s = fma(sv[v] * sv[w], M[sym_index<NV>(v, w) * NQP + q], s);
        acc[c] += geo + s;
      }
    } else {
      const T c1 = qd[(DIM * DIM + NV) * NQP + q], c2 = qd[(DIM * DIM + NV + 1) * NQP + q];
      T g2 = T(0);
#pragma unroll
      for (int j = 0; j < DIM; ++j) g2 = fma(gam[j], gam[j], g2);
#pragma unroll
      for (int c = 0; c < DIM; ++c) acc[c] += geo + T(0.5) * c1 * (g2 + gam[c] * gam[c]) + c2 * gam[c] * gam[c];
    }
  };
  if constexpr (!TILE) {
    const int hc = P.dcell ? P.dcell[cell] : -1;
    if (hc >= 0) {
      // a cell with constrained nodes: per entry (a node the cell couples)
      // the basis function is the weighted sum of local ones; results wait
      // in dbuf for the gather that follows the colours
      for (int e = P.deptr[hc] + r; e < P.deptr[hc + 1]; e += NQP) {
        T acc[DIM];
#pragma unroll
        for (int c = 0; c < DIM; ++c) acc[c] = T(0);
        for (int q = 0; q < NQP; ++q) {
          T gph[DIM];
#pragma unroll
          for (int d = 0; d < DIM; ++d) gph[d] = T(0);
          for (int k = P.dtptr[e]; k < P.dtptr[e + 1]; ++k) {
            T g[DIM];
            grad_ref(P.dta[k], q, g);
            const T w = P.dtw[k];
#pragma unroll
            for (int d = 0; d < DIM; ++d) gph[d] = fma(w, g[d], gph[d]);
          }
          point(q, gph, acc);
        }
#pragma unroll
        for (int c = 0; c < DIM; ++c) P.dbuf[(size_t)e * DIM + c] = acc[c];
      }
      return;
    }
  }
  T acc[DIM];
#pragma unroll
  for (int c = 0; c < DIM; ++c) acc[c] = T(0);
  for (int q = 0; q < NQP; ++q) {
    T gph[DIM];
    grad_ref(r, q, gph);
    point(q, gph, acc);
  }
  if constexpr (TILE) {
    lds_barrier();  // every lane is done with the per-point fields
    if (on)
#pragma unroll
      for (int c = 0; c < DIM; ++c)
        if (ts < TC) res[(ts * NQP + r) * DIM + c] = active ? acc[c] : T(0);
    lds_barrier();
  } else {
    const size_t node = code & kNodeMask;
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      if ((code >> (kConsShift + c)) & 1u) continue;
      T *d = P.dst + node * DIM + c;
      if (code & kFirstBit) *d = acc[c];
      else *d += acc[c];
    }
  }
  }  // rounds
  if constexpr (TILE) {
    for (int nd = t; nd < TS::TN; nd += kWG) {
      const uint32_t tcode = P.tnodes[(size_t)tile * TS::TN + nd];
      if (tcode != kNoNode) {
        const int X = nd % TS::NX, Y = (nd / TS::NX) % TS::NY, Z = nd / (TS::NX * TS::NY);
        constexpr int PP = N - 1;
        T sum[DIM];
#pragma unroll
        for (int c = 0; c < DIM; ++c) sum[c] = T(0);
        const int zlo = Z == 0 ? 0 : (Z - 1) / PP, zhi = min(Z / PP, ez - 1);
        const int ylo = Y == 0 ? 0 : (Y - 1) / PP, yhi = min(Y / PP, ey - 1);
        const int xlo = X == 0 ? 0 : (X - 1) / PP, xhi = min(X / PP, ex - 1);
        for (int cz = zlo; cz <= zhi; ++cz)
          for (int cy = ylo; cy <= yhi; ++cy)
            for (int cx = xlo; cx <= xhi; ++cx) {
              const int rr = (X - cx * PP) + N * ((Y - cy * PP) + N * (Z - cz * PP));
              const int lane = (cx + TS::TX * (cy + TS::TY * cz)) * NQP + rr;
#pragma unroll
              for (int c = 0; c < DIM; ++c) sum[c] += res[lane * DIM + c];
            }
        const size_t tnode = tcode & kNodeMask;
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
          if ((tcode >> (kConsShift + c)) & 1u) continue;
          T *d = P.dst + tnode * DIM + c;
          if (tcode & kFirstBit) *d = sum[c];
          else *d += sum[c];
        }
      }
    }
  }
}

// rows of DoFs that no cell touches are zero (the reference zeroes dst and
// the diagonal before its cell loop, include/mf_nh_operator.h:646-706, 1517)
// constrained rows of a batch of vectors (vector blockIdx.y at blockIdx.y * stride)
template <typename T>
__global__ void k_constrained_identity_batch(const int32_t *__restrict__ list, int n, const T *__restrict__ src,
                                             T *__restrict__ dst, size_t stride) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const size_t k = (size_t)blockIdx.y * stride + (size_t)list[i];
    dst[k] = src[k];
  }
}
template <typename T>
__global__ void k_zero_list_batch(const int32_t *__restrict__ list, int n, T *__restrict__ dst, size_t stride) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[(size_t)blockIdx.y * stride + (size_t)list[i]] = T(0);
}

template <typename T> __global__ void k_zero_list(const int32_t *__restrict__ list, int n, T *__restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[list[i]] = T(0);
}

// constrained diagonal entries are one (include/mf_nh_operator.h:1533-1534)
template <typename T> __global__ void k_set_one(const int32_t *__restrict__ list, int n_cons, T *__restrict__ diag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_cons) diag[list[i]] = T(1);
}
// inverse diagonal; entries that are not positive (the zero rows of DoFs no
// cell references, or a tangent that is not positive definite) get inverse 0
// and are counted in *bad (the reference asserts diag > 0 before inverting,
// include/mf_nh_operator.h:1543-1547)
template <typename T>
__global__ void k_invert(const T *__restrict__ diag, T *__restrict__ inv, size_t n, int *__restrict__ bad) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const T d = diag[i];
    const bool ok = d > T(0);
    inv[i] = ok ? T(1) / d : T(0);
    if (!ok) atomicAdd(bad, 1);
  }
}
template <typename T>
__global__ void k_scale_mul(T *__restrict__ dst, const T *__restrict__ a, const T *__restrict__ b, T omega, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = omega * a[i] * b[i];
}

}  // namespace lsmf
