// lsmf_mesh.cpp -- structured meshes of the unit square / cube and of the
// Cook membrane, host only.
//
// Replaces, for the geometries this library runs, the reference's
// Solid::make_grid (include/mf_elasticity.h:1339-1398: subdivided_hyper_rectangle,
// grid_y_transform at :1252-1270, GridTools::scale) together with the DoF
// distribution of Solid::system_setup.  Numbering is our own canonical one:
//   node(ix, iy, iz) = ix + Nx * (iy + Ny * iz),  Nx = nx * p + 1, ...
//   dof = dim * node + component
//   cells lexicographic (x fastest); local nodes lexicographic (x fastest)
// Support points sit at the Gauss-Lobatto points of each cell.
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "lsmf.h"
#include "lsmf_error.hpp"
#include "lsmf_fe.hpp"

namespace {

// grid_y_transform (include/mf_elasticity.h:1252-1270)
void cook_transform(double x, double &y) {
  const double y_upper = 44.0 + (16.0 / 48.0) * x;
  const double y_lower = 0.0 + (44.0 / 48.0) * x;
  const double theta = y / 44.0;
  y = (1.0 - theta) * y_lower + theta * y_upper;
}

}  // namespace

extern "C" {

lsmf_status lsmf_mesh_create(int dim, int degree, const int *n_cells_dir, const double *lo, const double *hi,
                             int geometry, double scale, lsmf_mesh **out) {
  if (!out || !n_cells_dir || !lo || !hi || (dim != 2 && dim != 3) || degree < 1 || degree > 8)
    return lsmf::fail(LSMF_ERR_INVALID_ARGUMENT, "lsmf_mesh_create: bad arguments");
  if (geometry != LSMF_GEOMETRY_BOX && geometry != LSMF_GEOMETRY_COOK)
    return lsmf::fail(LSMF_ERR_INVALID_ARGUMENT, "lsmf_mesh_create: unknown geometry");
  int n[3] = {1, 1, 1};
  for (int d = 0; d < dim; ++d) {
    n[d] = n_cells_dir[d];
    if (n[d] < 1) return lsmf::fail(LSMF_ERR_INVALID_ARGUMENT, "lsmf_mesh_create: n_cells_dir < 1");
  }
  const int p = degree;
  int nn[3] = {1, 1, 1};
  for (int d = 0; d < dim; ++d) nn[d] = n[d] * p + 1;
  const long long n_nodes = (long long)nn[0] * nn[1] * nn[2];
  const long long n_cells = (long long)n[0] * n[1] * n[2];
  if (n_nodes >= (1ll << 28)) return lsmf::fail(LSMF_ERR_INVALID_ARGUMENT, "lsmf_mesh_create: more than 2^28 nodes");
  int nd1 = 1;
  for (int d = 0; d < dim; ++d) nd1 *= (p + 1);

  lsmf_mesh *m = new (std::nothrow) lsmf_mesh;
  if (!m) return lsmf::fail(LSMF_ERR_OUT_OF_MEMORY, "lsmf_mesh_create: host allocation");
  std::memset(m, 0, sizeof(*m));
  m->dim = dim;
  m->degree = p;
  for (int d = 0; d < 3; ++d) m->n_cells_dir[d] = (d < dim) ? n[d] : 1;
  m->n_cells = (int)n_cells;
  m->n_nodes = (int)n_nodes;
  m->cell_nodes = (int32_t *)std::malloc(sizeof(int32_t) * n_cells * nd1);
  m->coords = (double *)std::malloc(sizeof(double) * n_nodes * dim);
  m->cell_colour = (int32_t *)std::malloc(sizeof(int32_t) * n_cells);
  if (!m->cell_nodes || !m->coords || !m->cell_colour) {
    lsmf_mesh_destroy(m);
    return lsmf::fail(LSMF_ERR_OUT_OF_MEMORY, "lsmf_mesh_create: host allocation");
  }
  const std::vector<double> gl = lsmf::lobatto_points(p);
  // 1D node coordinates per direction
  std::vector<double> x1[3];
  for (int d = 0; d < dim; ++d) {
    x1[d].resize(nn[d]);
    const double h = (hi[d] - lo[d]) / n[d];
    for (int i = 0; i < nn[d]; ++i) {
      const int k = (i / p < n[d]) ? i / p : n[d] - 1;
      const int j = i - k * p;
      x1[d][i] = lo[d] + (k + gl[j]) * h;
    }
  }
  for (int iz = 0; iz < nn[2]; ++iz)
    for (int iy = 0; iy < nn[1]; ++iy)
      for (int ix = 0; ix < nn[0]; ++ix) {
        const long long node = ix + (long long)nn[0] * (iy + (long long)nn[1] * iz);
        double X[3] = {x1[0][ix], x1[1][iy], dim == 3 ? x1[2][iz] : 0.0};
        if (geometry == LSMF_GEOMETRY_COOK) cook_transform(X[0], X[1]);
        for (int d = 0; d < dim; ++d) m->coords[node * dim + d] = X[d] * scale;
      }
  for (int cz = 0; cz < n[2]; ++cz)
    for (int cy = 0; cy < n[1]; ++cy)
      for (int cx = 0; cx < n[0]; ++cx) {
        const long long cell = cx + (long long)n[0] * (cy + (long long)n[1] * cz);
        m->cell_colour[cell] = (cx & 1) + 2 * (cy & 1) + (dim == 3 ? 4 * (cz & 1) : 0);
        int a = 0;
        for (int az = 0; az < (dim == 3 ? p + 1 : 1); ++az)
          for (int ay = 0; ay <= p; ++ay)
            for (int ax = 0; ax <= p; ++ax, ++a) {
              const long long ix = (long long)cx * p + ax, iy = (long long)cy * p + ay,
                              iz = (long long)cz * p + az;
              m->cell_nodes[cell * nd1 + a] = (int32_t)(ix + nn[0] * (iy + (long long)nn[1] * iz));
            }
      }
  *out = m;
  return LSMF_OK;
}

lsmf_status lsmf_mesh_destroy(lsmf_mesh *mesh) {
  if (!mesh) return LSMF_OK;
  std::free(mesh->cell_nodes);
  std::free(mesh->coords);
  std::free(mesh->cell_colour);
  delete mesh;
  return LSMF_OK;
}

int lsmf_mesh_boundary_faces(const lsmf_mesh *m, int side, int32_t *cells, int32_t *faces) {
  if (!m || side < 0 || side >= 2 * m->dim) return lsmf::fail(LSMF_ERR_INVALID_ARGUMENT, "boundary_faces: bad side");
  const int d = side / 2, hi = side % 2;
  const int *n = m->n_cells_dir;
  int count = 0;
  for (int cz = 0; cz < n[2]; ++cz)
    for (int cy = 0; cy < n[1]; ++cy)
      for (int cx = 0; cx < n[0]; ++cx) {
        const int c[3] = {cx, cy, cz};
        if (c[d] != (hi ? n[d] - 1 : 0)) continue;
        if (cells) cells[count] = cx + n[0] * (cy + n[1] * cz);
        if (faces) faces[count] = side;
        ++count;
      }
  return count;
}

int lsmf_mesh_boundary_nodes(const lsmf_mesh *m, int side, int32_t *nodes) {
  if (!m || side < 0 || side >= 2 * m->dim) return lsmf::fail(LSMF_ERR_INVALID_ARGUMENT, "boundary_nodes: bad side");
  const int d = side / 2, hi = side % 2, p = m->degree;
  int nn[3] = {1, 1, 1};
  for (int e = 0; e < m->dim; ++e) nn[e] = m->n_cells_dir[e] * p + 1;
  int count = 0;
  for (int iz = 0; iz < nn[2]; ++iz)
    for (int iy = 0; iy < nn[1]; ++iy)
      for (int ix = 0; ix < nn[0]; ++ix) {
        const int i[3] = {ix, iy, iz};
        if (i[d] != (hi ? nn[d] - 1 : 0)) continue;
        if (nodes) nodes[count] = ix + nn[0] * (iy + nn[1] * iz);
        ++count;
      }
  return count;
}

}  // extern "C"
