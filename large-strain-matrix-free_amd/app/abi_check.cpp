// abi_check.cpp -- a C++ caller of the C ABI (include/lsmf.h) with no
// Python or torch in the process: the call sequence the reference's
// NeoHookOperator users make (include/mf_elasticity.h:1833-1836, 2032-2033,
// 2109-2110, 3146-3180), driven on one operator read from a file.
//
//   lsmf_abi_check in.bin out.bin
//
// in.bin (little-endian): int32 dim, degree, n_cells, n_nodes, caching,
// precision, then int32 cell_nodes[n_cells * (degree+1)^dim], float64
// coords[n_nodes * dim], uint8 constrained[n_nodes * dim], float64
// mu[n_cells], lambda[n_cells], u[n_dofs], src[n_dofs], add0[n_dofs],
// float64 omega.
// out.bin: float64 vmult[n_dofs] (dst = A src), vmult_add[n_dofs]
// (add0 + A src), diagonal[n_dofs], inverse diagonal[n_dofs],
// jacobi[n_dofs] (omega D^-1 src), el00, then int64 m, n and
// memory_consumption.  FP32 operators compute in FP32; every value is
// written as float64.
//
// Sequence: lsmf_op_create (initialize) -> lsmf_set_material_per_cell ->
// lsmf_set_linearization -> lsmf_cache -> lsmf_compute_diagonal ->
// lsmf_vmult (set and add) -> lsmf_get_diagonal (get_matrix_diagonal_inverse,
// el) -> lsmf_precondition_jacobi -> lsmf_op_size / lsmf_memory_consumption
// -> lsmf_op_destroy.  Exit status 0 on success, 1 with the library's
// message on stderr otherwise.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "lsmf.h"

namespace {

void check(lsmf_status st, const char *what) {
  if (st != LSMF_OK) throw std::runtime_error(std::string(what) + ": " + lsmf_last_error());
}
void hip(hipError_t e, const char *what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <typename V> void read(std::ifstream &f, V *dst, size_t n) {
  f.read(reinterpret_cast<char *>(dst), (std::streamsize)(n * sizeof(V)));
  if (!f) throw std::runtime_error("input file too short");
}

// a device vector in the operator's precision, filled from / read back to FP64
struct DevVec {
  void *p = nullptr;
  size_t n = 0, es = 8;
  DevVec(size_t n_, size_t es_) : n(n_), es(es_) { hip(hipMalloc(&p, n * es), "hipMalloc"); }
  ~DevVec() { (void)hipFree(p); }
  void upload(const std::vector<double> &h) {
    if (es == 8) {
      hip(hipMemcpy(p, h.data(), n * 8, hipMemcpyHostToDevice), "upload");
    } else {
      std::vector<float> f(h.begin(), h.end());
      hip(hipMemcpy(p, f.data(), n * 4, hipMemcpyHostToDevice), "upload");
    }
  }
  std::vector<double> download(const void *src = nullptr) const {
    std::vector<double> h(n);
    const void *from = src ? src : p;
    if (es == 8) {
      hip(hipMemcpy(h.data(), from, n * 8, hipMemcpyDeviceToHost), "download");
    } else {
      std::vector<float> f(n);
      hip(hipMemcpy(f.data(), from, n * 4, hipMemcpyDeviceToHost), "download");
      for (size_t i = 0; i < n; ++i) h[i] = f[i];
    }
    return h;
  }
};

int run(const char *in_path, const char *out_path) {
  std::ifstream f(in_path, std::ios::binary);
  if (!f) throw std::runtime_error(std::string("cannot open ") + in_path);
  int32_t hdr[6];
  read(f, hdr, 6);
  const int dim = hdr[0], degree = hdr[1], n_cells = hdr[2], n_nodes = hdr[3], caching = hdr[4], precision = hdr[5];
  int npc = 1;
  for (int d = 0; d < dim; ++d) npc *= degree + 1;
  const size_t n_dofs = (size_t)n_nodes * dim;
  std::vector<int32_t> cell_nodes((size_t)n_cells * npc);
  std::vector<double> coords(n_dofs), mu(n_cells), lambda(n_cells), u(n_dofs), src(n_dofs), add0(n_dofs);
  std::vector<uint8_t> cons(n_dofs);
  double omega = 0;
  read(f, cell_nodes.data(), cell_nodes.size());
  read(f, coords.data(), coords.size());
  read(f, cons.data(), cons.size());
  read(f, mu.data(), mu.size());
  read(f, lambda.data(), lambda.size());
  read(f, u.data(), u.size());
  read(f, src.data(), src.size());
  read(f, add0.data(), add0.size());
  read(f, &omega, 1);

  hipStream_t s;
  hip(hipStreamCreate(&s), "hipStreamCreate");
  // initialize (include/mf_nh_operator.h:394-443)
  lsmf_mesh_desc m{};
  m.dim = dim;
  m.degree = degree;
  m.n_cells = n_cells;
  m.n_nodes = n_nodes;
  m.cell_nodes = cell_nodes.data();
  m.coords = coords.data();
  m.constrained = cons.data();
  m.cell_colour = nullptr;
  lsmf_op_options o{degree + 1, precision, caching, LSMF_LAYOUT_AUTO};
  lsmf_op *op = nullptr;
  check(lsmf_op_create(&m, &o, &op), "lsmf_op_create");
  const size_t es = precision == LSMF_FP64 ? 8 : 4;
  std::vector<double> out;
  try {
    check(lsmf_set_material_per_cell(op, mu.data(), lambda.data()), "set_material");
    DevVec du(n_dofs, 8);
    du.upload(u);
    check(lsmf_set_linearization(op, static_cast<const double *>(du.p), s), "set_linearization");
    check(lsmf_cache(op, s), "cache");
    check(lsmf_compute_diagonal(op, s), "compute_diagonal");
    DevVec x(n_dofs, es), y(n_dofs, es), z(n_dofs, es), jac(n_dofs, es);
    x.upload(src);
    z.upload(add0);
    check(lsmf_vmult(op, y.p, x.p, 0, LSMF_MASK_DEFAULT, s), "vmult");
    check(lsmf_vmult(op, z.p, x.p, 1, LSMF_MASK_DEFAULT, s), "vmult_add");
    const void *diag = nullptr, *inv = nullptr;
    check(lsmf_get_diagonal(op, &diag, &inv), "get_diagonal");
    check(lsmf_precondition_jacobi(op, jac.p, x.p, omega, s), "precondition_Jacobi");
    hip(hipStreamSynchronize(s), "synchronize");
    const std::vector<double> vy = y.download(), vz = z.download(), vd = y.download(diag), vi = y.download(inv),
                              vj = jac.download();
    for (const auto *v : {&vy, &vz, &vd, &vi, &vj}) out.insert(out.end(), v->begin(), v->end());
    out.push_back(vd[0]);  // el(0, 0)
    const int64_t mn[3] = {(int64_t)lsmf_op_size(op), (int64_t)lsmf_op_size(op), (int64_t)lsmf_memory_consumption(op)};
    std::ofstream g(out_path, std::ios::binary);
    g.write(reinterpret_cast<const char *>(out.data()), (std::streamsize)(out.size() * sizeof(double)));
    g.write(reinterpret_cast<const char *>(mn), sizeof(mn));
    if (!g) throw std::runtime_error(std::string("cannot write ") + out_path);
  } catch (...) {
    lsmf_op_destroy(op);
    (void)hipStreamDestroy(s);
    throw;
  }
  check(lsmf_op_destroy(op), "lsmf_op_destroy");
  hip(hipStreamDestroy(s), "hipStreamDestroy");
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 3) {
    std::cerr << "usage: lsmf_abi_check in.bin out.bin" << std::endl;
    return 2;
  }
  try {
    return run(argv[1], argv[2]);
  } catch (const std::exception &e) {
    std::cerr << "lsmf_abi_check: " << e.what() << std::endl;
    return 1;
  }
}
