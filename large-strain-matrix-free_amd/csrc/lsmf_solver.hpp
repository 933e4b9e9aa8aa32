// lsmf_solver.hpp -- host C++ driver of the GMG-preconditioned CG / Newton
// solve, written over the C ABI of the operator (include/lsmf.h).
//
// NeoHookOperatorHIP<number> re-exposes the member set of the reference's
// NeoHookOperator<dim, fe_degree, n_q_points_1d, number>
// (include/mf_nh_operator.h:147-296) so the CG / Chebyshev / multigrid code
// below calls it where the reference's deal.II solvers call NeoHookOperator
// (include/mf_elasticity.h:897-941, 2081-2238, 3012-3214).
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "lsmf.h"

namespace lsmf {

struct HipError : std::runtime_error {
  int code;
  HipError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess)
    throw HipError(e == hipErrorOutOfMemory ? LSMF_ERR_OUT_OF_MEMORY : LSMF_ERR_HIP,
                   std::string(what) + ": " + hipGetErrorString(e));
}
inline void abi_check(int rc, const char *what) {
  if (rc != LSMF_OK) throw HipError(rc, std::string(what) + ": " + lsmf_last_error());
}

// owning device buffer
template <typename T> struct DeviceArray {
  T *p = nullptr;
  size_t n = 0;
  DeviceArray() = default;
  explicit DeviceArray(size_t n_) { resize(n_); }
  DeviceArray(const DeviceArray &) = delete;
  DeviceArray &operator=(const DeviceArray &) = delete;
  DeviceArray(DeviceArray &&o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  DeviceArray &operator=(DeviceArray &&o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DeviceArray() { release(); }
  void resize(size_t n_) {
    release();
    n = n_;
    if (n) hip_check(hipMalloc((void **)&p, n * sizeof(T)), "hipMalloc");
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void upload(const T *h, size_t count) {
    hip_check(hipMemcpy(p, h, count * sizeof(T), hipMemcpyHostToDevice), "upload");
  }
};

// The reference's operator interface over one lsmf_op handle.
template <typename number> class NeoHookOperatorHIP {
 public:
  NeoHookOperatorHIP() = default;
  NeoHookOperatorHIP(const NeoHookOperatorHIP &) = delete;
  NeoHookOperatorHIP &operator=(const NeoHookOperatorHIP &) = delete;
  ~NeoHookOperatorHIP() { clear(); }

  // initialize(data_current, data_reference, displacement, caching)
  // (include/mf_nh_operator.h:394-443); the displacement is passed to
  // set_linearization, which copies it.
  void initialize(const lsmf_mesh_desc &mesh, int n_q_1d, int caching, hipStream_t s) {
    clear();
    lsmf_op_options o{n_q_1d, sizeof(number) == 8 ? LSMF_FP64 : LSMF_FP32, caching};
    abi_check(lsmf_op_create(&mesh, &o, &op_), "NeoHookOperator::initialize");
    stream_ = s;
  }
  void set_formulation(int formulation) { abi_check(lsmf_set_formulation(op_, formulation), "set_formulation"); }
  void set_material(const double *mu, const double *lambda) {
    abi_check(lsmf_set_material_per_cell(op_, mu, lambda), "set_material");
  }
  // the reference's set_material(matrix, inclusion): material_id != 0 selects the inclusion
  void set_material(const uint8_t *material_id, double mu0, double lambda0, double mu1, double lambda1) {
    abi_check(lsmf_set_material_two(op_, material_id, mu0, lambda0, mu1, lambda1), "set_material");
  }
  void set_linearization(const double *u_dev) { abi_check(lsmf_set_linearization(op_, u_dev, stream_), "set_linearization"); }
  void cache() { abi_check(lsmf_cache(op_, stream_), "cache"); }
  void compute_diagonal() { abi_check(lsmf_compute_diagonal(op_, stream_), "compute_diagonal"); }
  void vmult(number *dst, const number *src, unsigned mask = LSMF_MASK_DEFAULT) const {
    abi_check(lsmf_vmult(op_, dst, src, 0, mask, stream_), "vmult");
  }
  // n_vectors products at once (vector b at b * m()): the coarse-matrix probing
  void vmult_batched(number *dst, const number *src, int n_vectors) const {
    abi_check(lsmf_vmult_batched(op_, dst, src, n_vectors, stream_), "vmult_batched");
  }
  void vmult_add(number *dst, const number *src) const { abi_check(lsmf_vmult(op_, dst, src, 1, LSMF_MASK_DEFAULT, stream_), "vmult_add"); }
  void Tvmult(number *dst, const number *src) const { vmult(dst, src); }
  void Tvmult_add(number *dst, const number *src) const { vmult_add(dst, src); }
  void precondition_Jacobi(number *dst, const number *src, number omega) const {
    abi_check(lsmf_precondition_jacobi(op_, dst, src, omega, stream_), "precondition_Jacobi");
  }
  const number *get_matrix_diagonal_inverse() const {
    const void *d = nullptr, *inv = nullptr;
    abi_check(lsmf_get_diagonal(op_, &d, &inv), "get_matrix_diagonal_inverse");
    return static_cast<const number *>(inv);
  }
  number el(size_t row, size_t col) const {
    if (row != col) throw HipError(LSMF_ERR_INVALID_ARGUMENT, "el: only diagonal entries are available");
    const void *d = nullptr, *inv = nullptr;
    abi_check(lsmf_get_diagonal(op_, &d, &inv), "el");
    number v;
    hip_check(hipMemcpy(&v, static_cast<const number *>(d) + row, sizeof(number), hipMemcpyDeviceToHost), "el");
    return v;
  }
  size_t m() const { return lsmf_op_size(op_); }
  size_t n() const { return lsmf_op_size(op_); }
  size_t memory_consumption() const { return lsmf_memory_consumption(op_); }
  void clear() {
    if (op_) lsmf_op_destroy(op_);
    op_ = nullptr;
  }
  lsmf_op *handle() const { return op_; }
  hipStream_t stream() const { return stream_; }

 private:
  lsmf_op *op_ = nullptr;
  hipStream_t stream_ = nullptr;
};

// deterministic inner products (fixed two-stage reduction)
class Reducer {
 public:
  explicit Reducer(hipStream_t s);
  template <typename TA, typename TB> double dot(const TA *a, const TB *b, size_t n);
  template <typename T> void dot2(const T *a, const T *b, const T *c, const T *d, size_t n, double &r0, double &r1);
  template <typename T> double cg_update(T *x, T *r, const T *p, const T *Ap, double alpha, size_t n);
  // the same reductions with the result left in device memory (no host sync)
  template <typename TA, typename TB> void dot_to(const TA *a, const TB *b, size_t n, double *dst);
  template <typename T> void cg_update_dev(T *x, T *r, const T *p, const T *Ap, const double *st, size_t n,
                                           double *dst);

 private:
  double read(int nv, double *out);
  hipStream_t s_;
  DeviceArray<double> part_, res_;
};

}  // namespace lsmf
