// lsmf_vec.hpp -- vector kernels of the solver: deterministic dot products
// (wave-shuffle + fixed two-stage reduction), fused CG and Chebyshev updates,
// precision conversion, constrained-DoF updates and the tensor-product
// multigrid transfer (one banded 1D operator applied along one direction).
//
// Reference counterparts (all in deal.II, called from
// include/mf_elasticity.h): SolverCG vector updates and inner products
// (:3049-3061, :3175-3180), PreconditionChebyshev (:2081-2113),
// MGTransferMatrixFree prolongate / restrict_and_add / interpolate_to_mg
// (:1917-1942, :2236-2238), copy_to_mg / copy_from_mg (FP64 <-> FP32).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace lsmf {

constexpr int kRedBlocks = 1024;  // fixed grid of the first reduction stage
constexpr int kRedThreads = 256;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// block sum of NV values per thread -> out[blockIdx.x * NV + k]
template <int NV> __device__ __forceinline__ void block_reduce_store(double (&v)[NV], double *out) {
  __shared__ double part[NV][kRedThreads / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const double s = wave_sum(v[k]);
    if (lane == 0) part[k][wid] = s;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double s = 0;
#pragma unroll
    for (int w = 0; w < kRedThreads / 64; ++w) s += part[threadIdx.x][w];
    out[blockIdx.x * NV + threadIdx.x] = s;
  }
}

// stage 2: single block sums kRedBlocks partials (fixed order) -> res[k]
template <int NV> __global__ void k_reduce_final(const double *__restrict__ part, double *__restrict__ res) {
  double v[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = 0;
  for (int b = threadIdx.x; b < kRedBlocks; b += kRedThreads)
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] += part[b * NV + k];
  block_reduce_store<NV>(v, res);  // launched with one block: res[k]
}

// dot products a.b (and optionally c.d) over n entries
template <typename TA, typename TB>
__global__ void __launch_bounds__(kRedThreads) k_dot(const TA *__restrict__ a, const TB *__restrict__ b, size_t n,
                                                     double *__restrict__ part) {
  double v[1] = {0.0};
  for (size_t i = (size_t)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += (size_t)kRedBlocks * kRedThreads)
    v[0] += double(a[i]) * double(b[i]);
  block_reduce_store<1>(v, part);
}

template <typename T>
__global__ void __launch_bounds__(kRedThreads) k_dot2(const T *__restrict__ a, const T *__restrict__ b,
                                                      const T *__restrict__ c, const T *__restrict__ d, size_t n,
                                                      double *__restrict__ part) {
  double v[2] = {0.0, 0.0};
  for (size_t i = (size_t)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += (size_t)kRedBlocks * kRedThreads) {
    v[0] += double(a[i]) * double(b[i]);
    v[1] += double(c[i]) * double(d[i]);
  }
  block_reduce_store<2>(v, part);
}

// CG: x += alpha p ; r -= alpha Ap ; partial r.r
template <typename T>
__global__ void __launch_bounds__(kRedThreads) k_cg_update(T *__restrict__ x, T *__restrict__ r,
                                                           const T *__restrict__ p, const T *__restrict__ Ap,
                                                           double alpha, size_t n, double *__restrict__ part) {
  double v[1] = {0.0};
  for (size_t i = (size_t)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += (size_t)kRedBlocks * kRedThreads) {
    x[i] = x[i] + T(alpha) * p[i];
    const T ri = r[i] - T(alpha) * Ap[i];
    r[i] = ri;
    v[0] += double(ri) * double(ri);
  }
  block_reduce_store<1>(v, part);
}

// Device-resident CG for the Lanczos eigenvalue estimate: its scalars live
// in a small device state, so a step needs no host round trip.  State
// slots (doubles): rz, r0, alpha and beta of the current step, the number
// of recorded steps m and a stop flag (a non-positive p.Ap, or the residual
// below 1e-10 r0, ends the recording, as the host loop's breaks did); the
// reductions land in their own slots.  Same arithmetic as the host loop.
enum EigSlot { kEigRz = 0, kEigR0, kEigAlpha, kEigBeta, kEigM, kEigStop, kEigPAp, kEigRr, kEigRzNew, kEigSlots };

__global__ void k_eig_init(double *__restrict__ st) {
  st[kEigR0] = sqrt(st[kEigR0]);  // holds r.r on entry
  st[kEigM] = 0.0;
  st[kEigStop] = 0.0;
}
// alpha_k = rz / p.Ap
__global__ void k_eig_alpha(double *__restrict__ st, double *__restrict__ alpha, int k) {
  if (st[kEigStop] != 0.0) return;
  const double pAp = st[kEigPAp];
  if (!(pAp > 0)) {
    st[kEigStop] = 1.0;
    return;
  }
  const double a = st[kEigRz] / pAp;
  alpha[k] = a;
  st[kEigAlpha] = a;
  st[kEigM] = k + 1;
}
__global__ void k_eig_check(double *__restrict__ st) {
  if (st[kEigStop] != 0.0) return;
  if (sqrt(st[kEigRr]) <= 1e-10 * st[kEigR0]) st[kEigStop] = 1.0;
}
// beta_k = rz_new / rz
__global__ void k_eig_beta(double *__restrict__ st, double *__restrict__ beta, int k) {
  if (st[kEigStop] != 0.0) return;
  const double b = st[kEigRzNew] / st[kEigRz];
  beta[k] = b;
  st[kEigBeta] = b;
  st[kEigRz] = st[kEigRzNew];
}
// k_cg_update with alpha from the state; no-op once the recording stopped
template <typename T>
__global__ void __launch_bounds__(kRedThreads) k_cg_update_dev(T *__restrict__ x, T *__restrict__ r,
                                                               const T *__restrict__ p, const T *__restrict__ Ap,
                                                               const double *__restrict__ st, size_t n,
                                                               double *__restrict__ part) {
  if (st[kEigStop] != 0.0) return;
  const double alpha = st[kEigAlpha];
  double v[1] = {0.0};
  for (size_t i = (size_t)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += (size_t)kRedBlocks * kRedThreads) {
    x[i] = x[i] + T(alpha) * p[i];
    const T ri = r[i] - T(alpha) * Ap[i];
    r[i] = ri;
    v[0] += double(ri) * double(ri);
  }
  block_reduce_store<1>(v, part);
}
// p = z + beta p with beta from the state; no-op once the recording stopped
template <typename T>
__global__ void k_xpby_dev(T *__restrict__ p, const T *__restrict__ z, const double *__restrict__ st, size_t n) {
  if (st[kEigStop] != 0.0) return;
  const double beta = st[kEigBeta];
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = z[i] + T(beta) * p[i];
}

// Element-wise vector kernels: thread k takes elements [k W, k W + W), W =
// 16 / sizeof(T) when every array is 16-byte aligned (one 16-byte access
// per array and thread), else W = 1; the arithmetic per element is the same
// either way.
template <typename T> constexpr int kVecW = 16 / (int)sizeof(T);
template <typename T, int W> struct alignas(W * sizeof(T)) Lanes {
  T v[W];
};
template <int W, typename T> __device__ __forceinline__ Lanes<T, W> ld(const T *p) {
  return *reinterpret_cast<const Lanes<T, W> *>(p);
}
template <int W, typename T> __device__ __forceinline__ void st(T *p, const Lanes<T, W> &x) {
  *reinterpret_cast<Lanes<T, W> *>(p) = x;
}
inline bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }
template <typename T, int W> inline unsigned vec_blocks(size_t n) {
  return (unsigned)(((n + W - 1) / W + 255) / 256);
}

// p = z + beta p
template <typename T, int W>
__global__ void k_xpby(T *__restrict__ p, const T *__restrict__ z, double beta, size_t n) {
  const size_t i0 = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * W;
  if (i0 + W <= n) {
    Lanes<T, W> pv = ld<W>(p + i0);
    const Lanes<T, W> zv = ld<W>(z + i0);
#pragma unroll
    for (int k = 0; k < W; ++k) pv.v[k] = zv.v[k] + T(beta) * pv.v[k];
    st<W>(p + i0, pv);
  } else {
    for (size_t i = i0; i < n; ++i) p[i] = z[i] + T(beta) * p[i];
  }
}

// y = a x + b y
template <typename T>
__global__ void k_axpby(T *__restrict__ y, const T *__restrict__ x, double a, double b, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = T(a) * x[i] + T(b) * y[i];
}

// z = invD .* r
template <typename T, int W>
__global__ void k_jacobi(T *__restrict__ z, const T *__restrict__ r, const T *__restrict__ invd, size_t n) {
  const size_t i0 = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * W;
  if (i0 + W <= n) {
    const Lanes<T, W> rv = ld<W>(r + i0), iv = ld<W>(invd + i0);
    Lanes<T, W> zv;
#pragma unroll
    for (int k = 0; k < W; ++k) zv.v[k] = iv.v[k] * rv.v[k];
    st<W>(z + i0, zv);
  } else {
    for (size_t i = i0; i < n; ++i) z[i] = invd[i] * r[i];
  }
}

// t = b - t
template <typename T, int W> __global__ void k_residual(T *__restrict__ t, const T *__restrict__ b, size_t n) {
  const size_t i0 = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * W;
  if (i0 + W <= n) {
    Lanes<T, W> tv = ld<W>(t + i0);
    const Lanes<T, W> bv = ld<W>(b + i0);
#pragma unroll
    for (int k = 0; k < W; ++k) tv.v[k] = bv.v[k] - tv.v[k];
    st<W>(t + i0, tv);
  } else {
    for (size_t i = i0; i < n; ++i) t[i] = b[i] - t[i];
  }
}

// Chebyshev first step: d = f2 invD (b - t) (t == nullptr: b), x = x + d or x = d
template <typename T, int W>
__global__ void k_cheb_first(T *__restrict__ x, T *__restrict__ d, const T *__restrict__ b,
                             const T *__restrict__ t, const T *__restrict__ invd, double f2, int x_is_zero,
                             size_t n) {
  const size_t i0 = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * W;
  if (i0 + W <= n) {
    const Lanes<T, W> bv = ld<W>(b + i0), iv = ld<W>(invd + i0);
    Lanes<T, W> tv, xv, dv;
    if (t) tv = ld<W>(t + i0);
    if (!x_is_zero) xv = ld<W>(x + i0);
#pragma unroll
    for (int k = 0; k < W; ++k) {
      const T res = t ? bv.v[k] - tv.v[k] : bv.v[k];
      const T di = T(f2) * iv.v[k] * res;
      dv.v[k] = di;
      xv.v[k] = x_is_zero ? di : xv.v[k] + di;
    }
    st<W>(d + i0, dv);
    st<W>(x + i0, xv);
  } else {
    for (size_t i = i0; i < n; ++i) {
      const T res = t ? b[i] - t[i] : b[i];
      const T di = T(f2) * invd[i] * res;
      d[i] = di;
      x[i] = x_is_zero ? di : x[i] + di;
    }
  }
}

// Chebyshev recurrence (deal.II's vector_updates): d = f1 d + f2 invD (b - t); x += d
template <typename T, int W>
__global__ void k_cheb_update(T *__restrict__ x, T *__restrict__ d, const T *__restrict__ b,
                              const T *__restrict__ t, const T *__restrict__ invd, double f1, double f2, size_t n) {
  const size_t i0 = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * W;
  if (i0 + W <= n) {
    Lanes<T, W> xv = ld<W>(x + i0), dv = ld<W>(d + i0);
    const Lanes<T, W> bv = ld<W>(b + i0), tv = ld<W>(t + i0), iv = ld<W>(invd + i0);
#pragma unroll
    for (int k = 0; k < W; ++k) {
      const T di = T(f1) * dv.v[k] + T(f2) * iv.v[k] * (bv.v[k] - tv.v[k]);
      dv.v[k] = di;
      xv.v[k] += di;
    }
    st<W>(d + i0, dv);
    st<W>(x + i0, xv);
  } else {
    for (size_t i = i0; i < n; ++i) {
      const T di = T(f1) * d[i] + T(f2) * invd[i] * (b[i] - t[i]);
      d[i] = di;
      x[i] += di;
    }
  }
}

// host launchers: the 16-byte form when every array allows it
template <typename T> void launch_xpby(hipStream_t s, T *p, const T *z, double beta, size_t n) {
  if (aligned16(p) && aligned16(z))
    hipLaunchKernelGGL((k_xpby<T, kVecW<T>>), dim3(vec_blocks<T, kVecW<T>>(n)), dim3(256), 0, s, p, z, beta, n);
  else
    hipLaunchKernelGGL((k_xpby<T, 1>), dim3(vec_blocks<T, 1>(n)), dim3(256), 0, s, p, z, beta, n);
}
template <typename T> void launch_jacobi(hipStream_t s, T *z, const T *r, const T *invd, size_t n) {
  if (aligned16(z) && aligned16(r) && aligned16(invd))
    hipLaunchKernelGGL((k_jacobi<T, kVecW<T>>), dim3(vec_blocks<T, kVecW<T>>(n)), dim3(256), 0, s, z, r, invd, n);
  else
    hipLaunchKernelGGL((k_jacobi<T, 1>), dim3(vec_blocks<T, 1>(n)), dim3(256), 0, s, z, r, invd, n);
}
template <typename T> void launch_residual(hipStream_t s, T *t, const T *b, size_t n) {
  if (aligned16(t) && aligned16(b))
    hipLaunchKernelGGL((k_residual<T, kVecW<T>>), dim3(vec_blocks<T, kVecW<T>>(n)), dim3(256), 0, s, t, b, n);
  else
    hipLaunchKernelGGL((k_residual<T, 1>), dim3(vec_blocks<T, 1>(n)), dim3(256), 0, s, t, b, n);
}
template <typename T>
void launch_cheb_first(hipStream_t s, T *x, T *d, const T *b, const T *t, const T *invd, double f2, int x_is_zero,
                       size_t n) {
  if (aligned16(x) && aligned16(d) && aligned16(b) && aligned16(t) && aligned16(invd))
    hipLaunchKernelGGL((k_cheb_first<T, kVecW<T>>), dim3(vec_blocks<T, kVecW<T>>(n)), dim3(256), 0, s, x, d, b, t,
                       invd, f2, x_is_zero, n);
  else
    hipLaunchKernelGGL((k_cheb_first<T, 1>), dim3(vec_blocks<T, 1>(n)), dim3(256), 0, s, x, d, b, t, invd, f2,
                       x_is_zero, n);
}
template <typename T>
void launch_cheb_update(hipStream_t s, T *x, T *d, const T *b, const T *t, const T *invd, double f1, double f2,
                        size_t n) {
  if (aligned16(x) && aligned16(d) && aligned16(b) && aligned16(t) && aligned16(invd))
    hipLaunchKernelGGL((k_cheb_update<T, kVecW<T>>), dim3(vec_blocks<T, kVecW<T>>(n)), dim3(256), 0, s, x, d, b, t,
                       invd, f1, f2, n);
  else
    hipLaunchKernelGGL((k_cheb_update<T, 1>), dim3(vec_blocks<T, 1>(n)), dim3(256), 0, s, x, d, b, t, invd, f1, f2,
                       n);
}

template <typename TI, typename TO>
__global__ void k_convert(const TI *__restrict__ a, TO *__restrict__ b, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = TO(a[i]);
}

template <typename T> __global__ void k_fill(T *__restrict__ a, T v, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// a[list[k]] = values ? values[k] : v
template <typename T>
__global__ void k_set_list(T *__restrict__ a, const int32_t *__restrict__ list, const double *__restrict__ values,
                           double v, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) a[list[k]] = T(values ? values[k] : v);
}

// z = x + y (used for u_total = u_n + delta)
template <typename T>
__global__ void k_add(T *__restrict__ z, const T *__restrict__ x, const T *__restrict__ y, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) z[i] = x[i] + y[i];
}

// deterministic pseudo-random entries in [-1, 1) (SplitMix64 of the index):
// start vector of the eigenvalue estimate
template <typename T> __global__ void k_pseudo_random(T *__restrict__ v, size_t n, uint64_t seed) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    uint64_t z = seed + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    v[i] = T(double(z >> 11) * (2.0 / 9007199254740992.0) - 1.0);
  }
}

// dense FP64 matrix-vector product x = M b (coarse direct solve with the
// inverse coarse matrix); one wave per row, shuffle reduction
template <typename T>
__global__ void k_dense_matvec(const double *__restrict__ M, const T *__restrict__ b, T *__restrict__ x, int n) {
  const int row = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  double s = 0.0;
  for (int j = lane; j < n; j += 64) s += M[(size_t)row * n + j] * double(b[j]);
  s = wave_sum(s);
  if (lane == 0) x[row] = T(s);
}

// the n x n identity, row j = e_j (the probe vectors of the coarse matrix)
__global__ void k_identity(double *__restrict__ E, int n) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)n * n) return;
  E[t] = (t / n == t % n) ? 1.0 : 0.0;
}

__global__ void k_sym_average(double *__restrict__ A, int n) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)n * n) return;
  const int i = (int)(t / n), j = (int)(t % n);
  if (i <= j) return;
  const double a = 0.5 * (A[(size_t)i * n + j] + A[(size_t)j * n + i]);
  A[(size_t)i * n + j] = a;
  A[(size_t)j * n + i] = a;
}

// In-place inverse of a dense SPD matrix (row-major, n x n) by blocked
// Gauss-Jordan elimination without pivoting, kGJB rows and columns per block
// step K = [K0, K0 + bk), three launches per step:
//   k_bgj_diag    W = inverse of the diagonal block A_KK (unblocked
//                 Gauss-Jordan in LDS; flags a pivot that is not positive)
//   k_bgj_panels  Cold = A_:K, Cnew = -A_:K W (rows outside K), R = W A_K:
//   k_bgj_update  A_KK = W, A_Kj = R_j, A_iK = Cnew_i, A_ij -= Cold_i . R_j
// The Schur complements of an SPD matrix are SPD, so every pivot stays
// positive and no row exchanges are needed.  Scratch: W kGJB^2, Cold and
// Cnew n * kGJB (row i at i * kGJB), R kGJB * n (row r at r * n).
constexpr int kGJB = 32;

__global__ void __launch_bounds__(256) k_bgj_diag(const double *__restrict__ A, int n, int K0, int bk,
                                                  double *__restrict__ W, int *__restrict__ bad) {
  __shared__ double M[kGJB * kGJB];
  __shared__ double rk[kGJB], ck[kGJB];
  const int t = threadIdx.x;
  for (int e = t; e < kGJB * kGJB; e += 256) {
    const int i = e / kGJB, j = e % kGJB;
    M[e] = (i < bk && j < bk) ? A[(size_t)(K0 + i) * n + K0 + j] : 0.0;
  }
  __syncthreads();
  for (int k = 0; k < bk; ++k) {
    if (t < bk) {
      rk[t] = M[k * kGJB + t];
      ck[t] = M[t * kGJB + k];
    }
    __syncthreads();
    const double p = rk[k];
    if (t == 0 && !(p > 0.0)) bad[0] = 1;
    const double pinv = 1.0 / p;
    for (int e = t; e < bk * kGJB; e += 256) {
      const int i = e / kGJB, j = e % kGJB;
      if (j >= bk) continue;
      double v;
      if (i == k)
        v = j == k ? pinv : rk[j] * pinv;
      else if (j == k)
        v = -ck[i] * pinv;
      else
        v = M[e] - ck[i] * rk[j] * pinv;
      M[e] = v;
    }
    __syncthreads();
  }
  for (int e = t; e < kGJB * kGJB; e += 256) W[e] = M[e];
}

__global__ void k_bgj_panels(const double *__restrict__ A, int n, int K0, int bk, const double *__restrict__ W,
                             double *__restrict__ Cold, double *__restrict__ Cnew, double *__restrict__ R) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t half = (size_t)n * kGJB;
  if (t < half) {
    const int i = (int)(t / kGJB), c = (int)(t % kGJB);
    if (c >= bk) return;
    const double *a = A + (size_t)i * n + K0;
    Cold[t] = a[c];
    if (i >= K0 && i < K0 + bk) return;
    double s = 0.0;
    for (int m = 0; m < bk; ++m) s = fma(a[m], W[m * kGJB + c], s);
    Cnew[t] = -s;
  } else if (t < 2 * half) {
    const size_t u = t - half;
    const int r = (int)(u / n), j = (int)(u % n);
    if (r >= bk) return;
    double s = 0.0;
    for (int m = 0; m < bk; ++m) s = fma(W[r * kGJB + m], A[(size_t)(K0 + m) * n + j], s);
    R[(size_t)r * n + j] = s;
  }
}

// one workgroup per 16 x 64 tile of A: lane -> column, wave -> 4 rows; the
// tile's Cold rows and R columns are staged in LDS
__global__ void __launch_bounds__(256) k_bgj_update(double *__restrict__ A, int n, int K0, int bk,
                                                    const double *__restrict__ W, const double *__restrict__ Cold,
                                                    const double *__restrict__ Cnew, const double *__restrict__ R) {
  __shared__ double Rs[kGJB * 64];
  __shared__ double Cs[16 * kGJB];
  const int t = threadIdx.x;
  const int j0 = blockIdx.x * 64, i0 = blockIdx.y * 16;
  for (int e = t; e < kGJB * 64; e += 256) {
    const int m = e / 64, jj = e % 64;
    Rs[e] = (m < bk && j0 + jj < n) ? R[(size_t)m * n + j0 + jj] : 0.0;
  }
  for (int e = t; e < 16 * kGJB; e += 256) {
    const int ii = e / kGJB, m = e % kGJB;
    Cs[e] = (m < bk && i0 + ii < n) ? Cold[(size_t)(i0 + ii) * kGJB + m] : 0.0;
  }
  __syncthreads();
  const int jj = t % 64, w = t / 64;
  const int j = j0 + jj;
  if (j >= n) return;
  const bool jK = j >= K0 && j < K0 + bk;
  for (int q = 0; q < 4; ++q) {
    const int ii = w * 4 + q, i = i0 + ii;
    if (i >= n) break;
    const bool iK = i >= K0 && i < K0 + bk;
    double *a = A + (size_t)i * n + j;
    if (iK && jK) {
      *a = W[(i - K0) * kGJB + (j - K0)];
    } else if (iK) {
      *a = Rs[(i - K0) * 64 + jj];
    } else if (jK) {
      *a = Cnew[(size_t)i * kGJB + (j - K0)];
    } else {
      double v = *a;
      for (int m = 0; m < bk; ++m) v = fma(-Cs[ii * kGJB + m], Rs[m * 64 + jj], v);
      *a = v;
    }
  }
}

// One banded 1D operator along direction DIR of a node array of shape
// (n2, n1, n0, C) (x fastest, C components interleaved); the output has
// n_DIR replaced by n_out:
//   out[.., o, ..] = sum_{k < len[o]} w[o * W + k] in[.., start[o] + k, ..]
// (add != 0: out +=, the prolongate-and-add of the V-cycle).  T is the
// level's precision: FP32 levels transfer in FP32 with FP32 weights.  The
// transfer runs k_band_line (x) and k_band_rows (y, z) below; this form,
// one thread per output node with its C components loaded as one access
// per tap, serves x-lines too long to stage in LDS.  Grid: x covers the
// output nodes (i0, i1) flattened, y is i2.  Along x the lanes differ in o
// and loop over the padded width W (zero weights past a row's length).
template <int C, typename T> struct BandNode {
  T v[C];
};
template <typename T, int DIR, int C>
__global__ void __launch_bounds__(256) k_band(const T *__restrict__ in, T *__restrict__ out,
                                              const int32_t *__restrict__ start, const int32_t *__restrict__ len,
                                              const T *__restrict__ w, int W, int n0, int n1, int n_out, int add) {
  // output node (i0, i1, i2): (i0, i1) flattened over grid x, so that short
  // lines (257 nodes) fill whole workgroups; i2 = grid y
  const int m0 = DIR == 0 ? n_out : n0, m1 = DIR == 1 ? n_out : n1;
  const unsigned idx = blockIdx.x * 256u + threadIdx.x;
  if (idx >= (unsigned)m0 * (unsigned)m1) return;
  const int i1 = (int)(idx / (unsigned)m0), i0 = (int)idx - i1 * m0, i2 = (int)blockIdx.y;
  const int o = DIR == 0 ? i0 : (DIR == 1 ? i1 : i2);
  const int s = start[o];
  const int j0 = DIR == 0 ? s : i0, j1 = DIR == 1 ? s : i1, j2 = DIR == 2 ? s : i2;
  const size_t line = (size_t)n0 * C;
  const T *src = in + ((size_t)j2 * n1 + j1) * line + (size_t)j0 * C;
  const size_t stride = DIR == 0 ? (size_t)C : (DIR == 1 ? line : line * n1);
  const T *wr = w + (size_t)o * W;
  const int nk = DIR == 0 ? W : len[o];
  T acc[C];
#pragma unroll
  for (int c = 0; c < C; ++c) acc[c] = T(0);
  for (int k = 0; k < nk; ++k) {
    BandNode<C, T> x;
    __builtin_memcpy(&x, src + (size_t)k * stride, sizeof(x));
    const T wk = wr[k];
#pragma unroll
    for (int c = 0; c < C; ++c) acc[c] = fma(wk, x.v[c], acc[c]);
  }
  T *dst = out + (((size_t)i2 * m1 + i1) * m0 + i0) * C;
  BandNode<C, T> r;
#pragma unroll
  for (int c = 0; c < C; ++c) r.v[c] = acc[c];
  if (add) {
    BandNode<C, T> d;
    __builtin_memcpy(&d, dst, sizeof(d));
#pragma unroll
    for (int c = 0; c < C; ++c) r.v[c] += d.v[c];
  }
  __builtin_memcpy(dst, &r, sizeof(r));
}


// Banded 1D operator along x (DIR 0) with the input line staged in LDS: one
// workgroup per line (i1, i2) reads its n0 nodes once, coalesced, and each
// lane then computes output nodes o = t, t + 256, ... from LDS (the W taps
// from start[o]; rows are padded to W taps inside the line).  Same
// arithmetic as k_band<T, 0, C>.  Dynamic LDS: n0 * C values.
template <typename T, int C>
__global__ void __launch_bounds__(256) k_band_line(const T *__restrict__ in, T *__restrict__ out,
                                                   const int32_t *__restrict__ start, const T *__restrict__ w, int W,
                                                   int n0, int n_out, int add) {
  extern __shared__ unsigned char band_smem[];
  T *buf = reinterpret_cast<T *>(band_smem);
  const size_t line = blockIdx.x;
  const T *src = in + line * (size_t)n0 * C;
  for (int e = threadIdx.x; e < n0 * C; e += 256) buf[e] = src[e];
  __syncthreads();
  T *dst = out + line * (size_t)n_out * C;
  for (int o = threadIdx.x; o < n_out; o += 256) {
    const T *b = buf + (size_t)start[o] * C;
    const T *wr = w + (size_t)o * W;
    T acc[C];
#pragma unroll
    for (int c = 0; c < C; ++c) acc[c] = T(0);
    for (int k = 0; k < W; ++k) {
      const T wk = wr[k];
#pragma unroll
      for (int c = 0; c < C; ++c) acc[c] = fma(wk, b[k * C + c], acc[c]);
    }
    BandNode<C, T> r;
#pragma unroll
    for (int c = 0; c < C; ++c) r.v[c] = acc[c];
    if (add) {
      BandNode<C, T> d;
      __builtin_memcpy(&d, dst + (size_t)o * C, sizeof(d));
#pragma unroll
      for (int c = 0; c < C; ++c) r.v[c] += d.v[c];
    }
    __builtin_memcpy(dst + (size_t)o * C, &r, sizeof(r));
  }
}

// Banded 1D operator along y (DIR 1) or z (DIR 2), R consecutive outputs
// per lane: the lanes of a workgroup cover nodes of the two other
// directions (grid x) and share the outputs o0 = R blockIdx.y, ...,
// o0 + R - 1, so every input row of their bands' union is loaded once
// (coalesced along x) instead of once per output; the band bookkeeping is
// uniform over the workgroup.  Each output sums its taps in the order of
// k_band<T, DIR, C>.
template <typename T, int DIR, int C, int R>
__global__ void __launch_bounds__(256) k_band_rows(const T *__restrict__ in, T *__restrict__ out,
                                                   const int32_t *__restrict__ start,
                                                   const int32_t *__restrict__ len, const T *__restrict__ w, int W,
                                                   int n0, int n1, int n2, int n_out, int add) {
  static_assert(DIR == 1 || DIR == 2, "rows of y or z");
  const unsigned idx = blockIdx.x * 256u + threadIdx.x;
  const int nb = DIR == 1 ? n2 : n1;  // the other direction flattened with x
  if (idx >= (unsigned)n0 * (unsigned)nb) return;
  const int ib = (int)(idx / (unsigned)n0), i0 = (int)idx - ib * n0;
  const int o0 = (int)blockIdx.y * R, nr = min(R, n_out - o0);
  int st[R], ln[R];
  int jlo = start[o0], jhi = jlo;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    st[r] = r < nr ? start[o0 + r] : 0;
    ln[r] = r < nr ? len[o0 + r] : 0;
    if (r < nr) {
      jlo = min(jlo, st[r]);
      jhi = max(jhi, st[r] + ln[r]);
    }
  }
  // input node (i0, j, ib) or (i0, ib, j); element stride between rows
  const size_t row = DIR == 1 ? (size_t)n0 * C : (size_t)n0 * n1 * C;
  const T *src = in + (DIR == 1 ? (size_t)ib * n1 * n0 * C : (size_t)ib * n0 * C) + (size_t)i0 * C;
  T acc[R][C];
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int c = 0; c < C; ++c) acc[r][c] = T(0);
  for (int j = jlo; j < jhi; ++j) {
    BandNode<C, T> x;
    __builtin_memcpy(&x, src + (size_t)j * row, sizeof(x));
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int k = j - st[r];
      if (k >= 0 && k < ln[r]) {
        const T wk = w[(size_t)(o0 + r) * W + k];
#pragma unroll
        for (int c = 0; c < C; ++c) acc[r][c] = fma(wk, x.v[c], acc[r][c]);
      }
    }
  }
  // output node (i0, o, ib) or (i0, ib, o)
  const size_t orow = DIR == 1 ? (size_t)n0 * C : (size_t)n0 * n1 * C;
  T *dst = out + (DIR == 1 ? (size_t)ib * n_out * n0 * C : (size_t)ib * n0 * C) + (size_t)i0 * C;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (r >= nr) break;
    T *d = dst + (size_t)(o0 + r) * orow;
    BandNode<C, T> v;
#pragma unroll
    for (int c = 0; c < C; ++c) v.v[c] = acc[r][c];
    if (add) {
      BandNode<C, T> old;
      __builtin_memcpy(&old, d, sizeof(old));
#pragma unroll
      for (int c = 0; c < C; ++c) v.v[c] += old.v[c];
    }
    __builtin_memcpy(d, &v, sizeof(v));
  }
}

}  // namespace lsmf
