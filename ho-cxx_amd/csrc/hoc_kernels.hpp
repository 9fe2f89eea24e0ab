// hoc_kernels.hpp - gfx950 kernels of the HO-CXX time-step hot path.
//
// Mapping: one thread per element, 64 elements (one wavefront) per workgroup.
// Element-major fields ([el][j][i], K*K values per element) are moved between HBM
// and registers through an LDS tile so that every global access is coalesced.
// Neighbour data are read from face-major trace arrays [face][el][value|derivative][K].
// Face numbering ijp: 0 west (xi=-1), 1 east, 2 south (eta=-1), 3 north.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace hoc {

constexpr int kWG = 64;      // elements (threads) per workgroup
constexpr int kMaxKd = 8;

// 1-D FR tables, passed by value (kernel argument space), in the kernel's working
// precision R (double everywhere; float for the mixed-precision preconditioner).
template <class R>
struct TabT {
  R bb[kMaxKd][2];      // l_m(-1), l_m(+1)
  R bgb[kMaxKd][2];     // l_m'(-1), l_m'(+1)
  R D[kMaxKd][kMaxKd];  // D[m][k] = l_m'(xi_k)
  R rad[kMaxKd][2];     // g_L, g_R at xi_k
  R grad[kMaxKd][2];    // g_L', g_R' at xi_k
  R gp[2];              // g'(-1), g'(+1)
  R w[kMaxKd];          // Gauss weights
  R x[kMaxKd];          // Gauss points (bilinear coarse-space weights)
};
using Tab = TabT<double>;

// Bilinear (Q1) weight of element corner a (0 SW, 1 SE, 2 NE, 3 NW of the element's
// reference square) at solution point (i along xi, j along eta).
template <class R>
__device__ __forceinline__ R q1w(const TabT<R>& T, int a, int j, int i) {
  const R xi = T.x[i], et = T.x[j];
  const R sx = (a == 1 || a == 2) ? R(1) + xi : R(1) - xi;
  const R sy = (a >= 2) ? R(1) + et : R(1) - et;
  return R(0.25) * sx * sy;
}

// Static per-mesh device arrays (metrics in precision R; the volume Jacobian and the
// Cartesian derivatives are only used by the FP64 stage kernels).
template <class R>
struct GeoT {
  int N, E;
  const R *jac, *xxi, *xeta, *yxi, *yeta;  // [N][K2]
  const R *Ga, *Gb, *Gc;                   // g22/J, g12/J, g11/J at sps
  const R *fa, *fb;                        // [N][4][K]  A/fjac, B/fjac
  const R *fA, *fB, *fjac, *fAn;           // [N][4][K]  raw face metrics
  const R* gel;                            // [N][3]     Ga, Gb, Gc of affine elements
  const R* gface;                          // [N][4][2]  fa, fb of affine elements
  const R* gjm;                            // [N][5]     xxi, xeta, yxi, yeta, jac of affine elements
  const int32_t* nb;                       // [N][4] neighbour or -1-edge
  const int8_t* nbf;                       // [N][4] neighbour's face
  const int8_t *sw_psi, *sw_diff;          // [E] 0 Dirichlet / 1 Neumann
  const int8_t* no_slip;                   // [E]
  const R* esign;                          // [E] 2*(ijp%2)-1 of the owner face
  const int32_t* inst;                     // [N] ensemble instance of each element, or null
  const int* iact = nullptr;               // [instances] active flags: a Poisson-operator block whose
                                           // elements all belong to one inactive instance returns at once
};
using Geo = GeoT<double>;

// Hybrid operation (problem_type 11): per-edge kind and the open-boundary values at the
// start and end of the outer step, blended per stage (src/HO_2D.cpp:2104-2191).
struct HybridDev {
  const uint8_t* kind;       // [E] 0 wall, 1 open, 2 inlet, 3 outlet
  const double* inlet_nvel;  // [E]
  const double* esign;       // [E]
  const int8_t* sw_diff;     // [E]
  const double *vn0, *vn1, *vp0, *vp1, *vw0, *vw1;  // [E][K] normal/tangential velocity, vorticity
  const double *w0, *w1, *wgt;                      // [N][K2] solution vorticity targets and weights
};

// Per-instance parameters of an ensemble (null pointers: one instance, scalar arguments).
struct InstDev {
  const double* re_inv;      // [n_inst] 1/Re
  const double* dt;          // [n_inst] time step
  const double* inv_dt_re;   // [n_inst] 1/(dt Re^-1), wall-vorticity flux scale
  const double* ramp;        // [n_inst] time factor of the boundary values at the stage time
  const int32_t* edge_inst;  // [E] instance of each boundary edge
};

// Time-dependent boundary values [E][K] (edge flux-point order).
struct BCDev {
  double *psi, *diff, *nvel, *pvel, *vjump, *vort, *adv;
  const double *psi0, *psi_r, *diff0, *diff_r, *nvel0, *nvel_r, *pvel0, *pvel_r;
};

// ------------------------------------------------------------------------------------------
// Tile helpers: NV values of type V per element, block of kWG elements.  A tile starts at
// element e0 (a multiple of kWG), so with a 16-byte aligned buffer it is moved in 16-byte
// chunks; the LDS tile must be declared __align__(16).
// ------------------------------------------------------------------------------------------
template <int NV, class V>
__device__ __forceinline__ void tile_load(const V* __restrict__ g, V* lds, int e0, int N) {
  constexpr int P = 16 / sizeof(V);
  const int tid = threadIdx.x;
  const int nel = min(kWG, N - e0);
  const int cnt = nel * NV;
  const V* src = g + (size_t)e0 * NV;
  int done = 0;
  if ((reinterpret_cast<uintptr_t>(src) & 15) == 0) {
    const int nch = cnt / P;
    const uint4* s4 = reinterpret_cast<const uint4*>(src);
    uint4* l4 = reinterpret_cast<uint4*>(lds);
#pragma unroll 4
    for (int i = tid; i < nch; i += kWG) l4[i] = s4[i];
    done = nch * P;
  }
  for (int i = done + tid; i < cnt; i += kWG) lds[i] = src[i];
  __syncthreads();
}
template <int NV, class V>
__device__ __forceinline__ void tile_store(V* __restrict__ g, V* lds, int e0, int N) {
  constexpr int P = 16 / sizeof(V);
  __syncthreads();
  const int tid = threadIdx.x;
  const int nel = min(kWG, N - e0);
  const int cnt = nel * NV;
  V* dst = g + (size_t)e0 * NV;
  int done = 0;
  if ((reinterpret_cast<uintptr_t>(dst) & 15) == 0) {
    const int nch = cnt / P;
    uint4* d4 = reinterpret_cast<uint4*>(dst);
    const uint4* l4 = reinterpret_cast<const uint4*>(lds);
#pragma unroll 4
    for (int i = tid; i < nch; i += kWG) d4[i] = l4[i];
    done = nch * P;
  }
  for (int i = done + tid; i < cnt; i += kWG) dst[i] = lds[i];
  __syncthreads();
}

}  // namespace hoc
