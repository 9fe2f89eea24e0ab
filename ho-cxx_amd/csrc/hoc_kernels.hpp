// hoc_kernels.hpp - gfx950 kernels of the HO-CXX time-step hot path.
//
// Mapping: one thread per element, 64 elements (one wavefront) per workgroup.
// Element-major fields ([el][j][i], K*K doubles per element) are moved between HBM
// and registers through an LDS tile so that every global access is coalesced.
// Neighbour data are read from per-face trace arrays [el][face][...][K].
// Face numbering ijp: 0 west (xi=-1), 1 east, 2 south (eta=-1), 3 north.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace hoc {

constexpr int kWG = 64;      // elements (threads) per workgroup
constexpr int kMaxKd = 8;

// 1-D FR tables, passed by value (kernel argument space).
struct Tab {
  double bb[kMaxKd][2];      // l_m(-1), l_m(+1)
  double bgb[kMaxKd][2];     // l_m'(-1), l_m'(+1)
  double D[kMaxKd][kMaxKd];  // D[m][k] = l_m'(xi_k)
  double rad[kMaxKd][2];     // g_L, g_R at xi_k
  double grad[kMaxKd][2];    // g_L', g_R' at xi_k
  double gp[2];              // g'(-1), g'(+1)
  double w[kMaxKd];          // Gauss weights
  double x[kMaxKd];          // Gauss points (bilinear coarse-space weights)
};

// Bilinear (Q1) weight of element corner a (0 SW, 1 SE, 2 NE, 3 NW of the element's
// reference square) at solution point (i along xi, j along eta).
__device__ __forceinline__ double q1w(const Tab& T, int a, int j, int i) {
  const double xi = T.x[i], et = T.x[j];
  const double sx = (a == 1 || a == 2) ? 1. + xi : 1. - xi;
  const double sy = (a >= 2) ? 1. + et : 1. - et;
  return 0.25 * sx * sy;
}

// Static per-mesh device arrays.
struct Geo {
  int N, E;
  const double *jac, *xxi, *xeta, *yxi, *yeta;  // [N][K2]
  const double *Ga, *Gb, *Gc;                   // g22/J, g12/J, g11/J at sps
  const double *fa, *fb;                        // [N][4][K]  A/fjac, B/fjac
  const double *fA, *fB, *fjac, *fAn;           // [N][4][K]  raw face metrics
  const double* gel;                            // [N][3]     Ga, Gb, Gc of affine elements
  const double* gface;                          // [N][4][2]  fa, fb of affine elements
  const int32_t* nb;                            // [N][4] neighbour or -1-edge
  const int8_t* nbf;                            // [N][4] neighbour's face
  const int8_t *sw_psi, *sw_diff;               // [E] 0 Dirichlet / 1 Neumann
  const int8_t* no_slip;                        // [E]
  const double* esign;                          // [E] 2*(ijp%2)-1 of the owner face
};

// Time-dependent boundary values [E][K] (edge flux-point order).
struct BCDev {
  double *psi, *diff, *nvel, *pvel, *vjump, *vort, *adv;
  const double *psi0, *psi_r, *diff0, *diff_r, *nvel0, *nvel_r, *pvel0, *pvel_r;
};

// ------------------------------------------------------------------------------------------
// Tile helpers: NV doubles per element, block of kWG elements.
// ------------------------------------------------------------------------------------------
template <int NV>
__device__ __forceinline__ void tile_load(const double* __restrict__ g, double* lds, int e0, int N) {
  const int tid = threadIdx.x;
  const int nel = min(kWG, N - e0);
  const int cnt = nel * NV;
  const double* src = g + (size_t)e0 * NV;
#pragma unroll 4
  for (int i = tid; i < cnt; i += kWG) lds[i] = src[i];
  __syncthreads();
}
template <int NV>
__device__ __forceinline__ void tile_store(double* __restrict__ g, double* lds, int e0, int N) {
  __syncthreads();
  const int tid = threadIdx.x;
  const int nel = min(kWG, N - e0);
  const int cnt = nel * NV;
  double* dst = g + (size_t)e0 * NV;
#pragma unroll 4
  for (int i = tid; i < cnt; i += kWG) dst[i] = lds[i];
  __syncthreads();
}

}  // namespace hoc
