// hoc_adj.hpp - kernels of the discrete adjoint of the RK stage (hoc_adj.hip).
//
// The stage map omega -> k is applied transposed by probing its linearisation (k_stage in its
// frozen-switch mode) with coloured unit seeds: probe m of colour c puts 1 at local dof m of
// every element of colour c, and when no two elements of one colour lie in the dependence ring
// of any output element, the product y = J seed holds, at every element e, the column of its
// one colour-c ring member.  Then (J^T lam)[i][m] = sum over the elements e whose ring holds i
// of lam_e . y_e, a gather over i's ring of the per-element products p_e = lam_e . y_e.  The
// K^2 probes of a colour run as one batch (probe b at offset b * the array size).
#pragma once

#include "hoc_kernels.hpp"

namespace hoc {

// K^2 probes of colour c: x[b] = unit seed at local dof b of the colour-c elements ([K2][N][K2])
// and tr[b] its face traces in the trace-plane layout ([K2][4][N][2K])
hipError_t launch_adj_seed_batch(int N, int K, const int32_t* color, int c, const Tab& T, double* x, double* tr,
                                 hipStream_t s);
// p[b][e] = sum_r lam[e][r] y[b][e][r], b < nb
hipError_t launch_adj_dot_batch(int N, int K2, int nb, const double* lam, const double* y, double* p, hipStream_t s);
// out[i][ooff + m] (row stride ostride) = sum over e in ring(i) of p[m][e], for the nmem elements
// i of a colour (members; ring as CSR)
hipError_t launch_adj_gather_batch(int N, int K2, int nmem, const int32_t* members, const int32_t* rptr,
                                   const int32_t* ridx, const double* p, double* out, int ostride, int ooff,
                                   hipStream_t s);
// K^2 probes of colour c on velocity component comp of U ([N][2][K2]): U[b] = unit seed at local dof b
// of component comp of the colour-c elements, tru[b] its contravariant normal traces ([4][N][K])
hipError_t launch_adj_seed_u_batch(int N, int K, const int32_t* color, int c, int comp, const Tab& T, double* U,
                                   double* tru, hipStream_t s);
// Assembly of the (state-independent) velocity map psi -> U from K^2 probes of colour c
// (dU [K2][N][2K2]): Mv[e][slot][m][o] = dU[m][e][o] for the slot (0 self, 1 + f neighbour across
// face f) holding the colour-c member of e's ring; and of the wall flux psi -> diff of each
// boundary edge from its owner element (dd [K2][E][K]): Md[g][j][m] = dd[m][g][j].
hipError_t launch_adj_assemble_vel(int N, int E, int K, const int32_t* color, int c, const int32_t* nb,
                                   const int32_t* own_el, const double* dU, const double* dd, double* Mv, double* Md,
                                   hipStream_t s);
// psibar[i][m] = sum over e in {i, face neighbours of i} of Mv[e][slot of i][m] . Ubar[e]
//              + sum over the boundary edges g of i of Md[g][.][m] . diffbar[g]
hipError_t launch_adj_vel_transpose(int N, int K, const int32_t* nb, const int8_t* nbf, const double* Mv,
                                    const double* Md, const double* Ubar, const double* diffbar, double* psibar,
                                    hipStream_t s);
// pvbar[g][j] += sign_g idr diffbar[g][j] on no-slip edges (diff = (vjump + sign pvel) idr there)
hipError_t launch_adj_wall(int E, int K, const int8_t* no_slip, const double* esign, double idr,
                           const double* diffbar, double* pvbar, hipStream_t s);
// 4K boundary-edge probes: x[f*K + j] = 1 at edge flux point j of every edge owned through face f
hipError_t launch_adj_edge_seed_batch(int E, int K, const int8_t* own_f, double* x, hipStream_t s);
// out[e][j] += p[own_f[e]*K + j][owner element of e]
hipError_t launch_adj_edge_gather_batch(int N, int E, int K, const int32_t* own_el, const int8_t* own_f,
                                        const double* p, double* out, hipStream_t s);
// y = a x + b y
hipError_t launch_adj_axpby(size_t n, double a, const double* x, double b, double* y, hipStream_t s);
// mode 0: y = x / (w_j w_i J);  mode 1: y += w_j w_i J x   (quadrature weights w, Jacobian J)
hipError_t launch_adj_wj(int N, int K, const double* wts, const double* jac, const double* x, double* y, int mode,
                         hipStream_t s);

}  // namespace hoc
