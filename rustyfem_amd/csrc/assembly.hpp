// Element assembly of linear elastic stiffness matrices on one MI355X:
// replaces the element loop and the serial SystemMatrix::add scatter of
// src/fem/assembly.rs:315-401 for the continuum elements
//   C3D8  trilinear hexahedron, 2x2x2 Gauss, volumetric B-bar
//         (src/elements/continuum3d.rs:39-84, :117-121);
//   C3D4  linear tetrahedron, one point of weight 1/6
//         (src/elements/solid_higher.rs:109, :137-181, :453-463);
//   CPS4  bilinear quadrilateral, plane stress, 2x2 Gauss
//   CPE4  bilinear quadrilateral, plane strain, planar B-bar
//         (src/elements/continuum2d.rs:24-27, :35-80, :119-123)
// with the B matrices, B-bar and D B / B^T D B products of
// src/elements/common.rs:142-202, :278-297 and the isotropic D of
// src/materials/elastic.rs:55-75.
//
// The pattern and, for every stored entry, the list of element-matrix
// entries that add to it (in the order of the reference's element loop and
// its (a, b) loops) are built once on the host; the values are computed on
// the GPU: one lane per node pair of an element computes that block of
// 0.5 (K_e + K_e^T), and one thread per stored entry sums its list from zero.
#pragma once
#include <cstdint>
#include <memory>
#include <vector>

namespace rf {

enum class ElemType : int32_t { C3D8 = 1, C3D4 = 2, CPS4 = 3, CPE4 = 4 };

class GpuAssembly {
 public:
  virtual ~GpuAssembly() = default;
  virtual int64_t n() const = 0;
  virtual int64_t nnz() const = 0;
  // CSC of the symmetric matrix (both triangles, rows sorted per column)
  virtual const std::vector<int64_t>& indptr() const = 0;
  virtual const std::vector<int64_t>& indices() const = 0;
  // Values for per-element Young's modulus, Poisson's ratio and thickness
  // (host arrays of nelem; thickness only for the plane elements, nullptr =
  // 1).  data: nnz doubles on the host or the device.  Returns -1, or the
  // lowest element whose Jacobian determinant fails the reference's test.
  virtual int64_t values(const double* young, const double* nu, const double* thickness, double* data,
                         bool on_device) = 0;
  // Binds a second output layout: vmap[p] (one per stored entry, -1 = not
  // used) is the position in an array of nout values that entry p goes to;
  // every output position takes at most one entry.  `owner` names what the
  // layout was bound for (bound_owner()).
  virtual void bind(const std::vector<int64_t>& vmap, int64_t nout, uint64_t owner) = 0;
  virtual uint64_t bound_owner() const = 0;
  // values() into the bound layout, a device array of nout owned by the
  // assembly (bound_values()).
  virtual int64_t values_bound(const double* young, const double* nu, const double* thickness) = 0;
  virtual const double* bound_values() const = 0;
  // [0] element kernel ms, [1] gather kernel ms, [2] host build s (pattern
  // and lists), [3] contributions, [4] element-matrix doubles per element,
  // [5] bound gather ms
  virtual void stats(double* out, int len) const = 0;
};

// conn: nelem x nodes-per-element node indices; coords nn x 3 (plane
// elements use x and y); eq: nn x dofs-per-node equation numbers (-1: no
// equation), n equations.
std::unique_ptr<GpuAssembly> make_gpu_assembly(int64_t nelem, ElemType etype, const int64_t* conn, int64_t nn,
                                               const double* coords, const int64_t* eq, int64_t n);

// nodes per element and DOFs per node of a type (0, 0 for an unknown code)
void elem_shape(int32_t etype, int* nodes, int* dofs);

}  // namespace rf
