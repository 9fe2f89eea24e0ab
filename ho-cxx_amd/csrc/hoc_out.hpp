// hoc_out.hpp - node-averaged output and sample points on the device (hoc_out.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "hoc_kernels.hpp"

namespace hoc {

// Element-to-node extrapolation with the boundary values on boundary faces
// (save_smooth_vtk, src/HO_2D.cpp:3413-3506).
struct SmoothArgs {
  int N, K, L;                  // elements, solution points and geometry nodes per direction
  double sgb[kMaxKd][4];         // Lagrange basis of the K solution points at the L geometry nodes
  int slot[16];                 // Gmsh node slot of tensor position (i, j) at [j L + i]
  const double *w, *psi, *vel;  // [N][K2], [N][K2], [N][K2][2]
  const int32_t* nb;            // [N][4] neighbour, or -1-edge on a boundary face
  const double *bvort, *bnvel, *bpvel;  // [E][K] boundary vorticity, normal and tangential velocity (edge order)
  const double* bmet;           // [E][K][4] x_xi, x_eta, y_xi, y_eta at the owner face's flux points (local order)
  double* elemout;              // [N][L2][4] vorticity, psi, u, v at the element's geometry nodes
};
hipError_t launch_smooth_elem(const SmoothArgs& A, hipStream_t s);
// out[node][4]: the mean of the elements' values at each node, accumulated in element order
// (src/HO_2D.cpp:3511-3528); incidences of node nd are ninc[nptr[nd] .. nptr[nd+1]) = el L2 + l.
hipError_t launch_smooth_nodes(int n_nodes, const int32_t* nptr, const int32_t* ninc, const double* elemout,
                               double* out, hipStream_t s);
// out[p][3] = vorticity, u, v at sample point p: sum over the containing element's nodes of the
// shape function value times the node-averaged field (src/HO_2D.cpp:3593-3612).
hipError_t launch_sample(int n, int L2, const int32_t* sel, const double* shape, const int32_t* elem_nodes,
                         const double* nodes, double* out, hipStream_t s);

}  // namespace hoc
