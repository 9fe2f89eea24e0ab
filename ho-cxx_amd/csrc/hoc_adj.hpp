// hoc_adj.hpp - kernels of the discrete adjoint of the RK stage (hoc_adj.hip).
//
// The stage map omega -> k is applied transposed by probing its linearisation (k_stage in its
// frozen-switch mode) with coloured unit seeds: a seed puts 1 at local dof m of every element
// of colour c, and when no two elements of one colour lie in the dependence ring of any output
// element, the product y = J seed holds, at every element e, the column of its one
// colour-c ring member.  Then (J^T lam)[i][m] = sum over the elements e whose ring holds i of
// lam_e . y_e, a gather over i's ring of the per-element products p_e = lam_e . y_e.
#pragma once

#include "hoc_kernels.hpp"

namespace hoc {

// x = unit seed: 1 at local dof m of every element of colour c, 0 elsewhere
hipError_t launch_adj_seed(int N, int K2, const int32_t* color, int c, int m, double* x, hipStream_t s);
// p[e] = sum_r lam[e][r] y[e][r]
hipError_t launch_adj_dot(int N, int K2, const double* lam, const double* y, double* p, hipStream_t s);
// out[i][m] = sum over e in ring(i) of p[e], for the elements i of colour c (ring as CSR)
hipError_t launch_adj_gather(int N, int K2, const int32_t* color, int c, int m, const int32_t* rptr,
                             const int32_t* ridx, const double* p, double* out, hipStream_t s);
// boundary-edge seed: 1 at edge flux point j of every edge owned through face f, 0 elsewhere
hipError_t launch_adj_edge_seed(int E, int K, const int8_t* own_f, int f, int j, double* x, hipStream_t s);
// out[e][j] += p[owner element of e] for the edges owned through face f
hipError_t launch_adj_edge_gather(int E, int K, const int32_t* own_el, const int8_t* own_f, int f, int j,
                                  const double* p, double* out, hipStream_t s);
// y = a x + b y
hipError_t launch_adj_axpby(size_t n, double a, const double* x, double b, double* y, hipStream_t s);
// mode 0: y = x / (w_j w_i J);  mode 1: y += w_j w_i J x   (quadrature weights w, Jacobian J)
hipError_t launch_adj_wj(int N, int K, const double* wts, const double* jac, const double* x, double* y, int mode,
                         hipStream_t s);

}  // namespace hoc
