// hoc_adj.hip - seeds, per-element products and transposed gathers of the discrete adjoint.
#include "hoc_adj.hpp"

namespace hoc {

namespace {

constexpr int kTB = 256;

inline unsigned grid(size_t n) { return (unsigned)((n + kTB - 1) / kTB); }

// seeds: one thread per (value, probe)
__global__ void k_adj_seed(size_t D, int K2, const int32_t* __restrict__ color, int c, double* __restrict__ x) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (i >= D) return;
  const size_t e = i / K2;
  x[(size_t)b * D + i] = (color[e] == c && (int)(i - e * K2) == b) ? 1. : 0.;
}

// traces of the seeds (the traces of src/HO_2D.cpp:2695-2713 of a unit value at (j0, i0)):
// faces 0/1 (xi = -1/+1) carry l_i0 and l_i0' at eta point j0, faces 2/3 l_j0 and l_j0' at xi point i0
__global__ void k_adj_seed_tr(int N, int K, const int32_t* __restrict__ color, int c, Tab T, double* __restrict__ tr) {
  const size_t TR = (size_t)4 * N * 2 * K;
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;  // plane f, element e, slot kk
  const int b = blockIdx.y;
  if (q >= TR) return;
  const int f = (int)(q / ((size_t)N * 2 * K));
  const size_t rem = q - (size_t)f * N * 2 * K;
  const int e = (int)(rem / (2 * K)), kk = (int)(rem - (size_t)e * 2 * K);
  const int j0 = b / K, i0 = b % K;
  const int pos = f < 2 ? j0 : i0, idx = f < 2 ? i0 : j0, ib = f & 1;
  const int k = kk < K ? kk : kk - K;
  double v = 0.;
  if (color[e] == c && k == pos) v = kk < K ? T.bb[idx][ib] : T.bgb[idx][ib];
  tr[(size_t)b * TR + q] = v;
}

// 64 elements per block: the products are formed with coalesced reads into an LDS tile, then
// each thread sums its element's row
constexpr int kDotEl = 64;
constexpr int kMaxK2 = 36;
__global__ __launch_bounds__(kDotEl) void k_adj_dot(int N, int K2, const double* __restrict__ lam,
                                                   const double* __restrict__ y, double* __restrict__ p) {
  __shared__ double prod[kDotEl * kMaxK2];
  const int e0 = blockIdx.x * kDotEl;
  const int b = blockIdx.y;
  const int nel = min(kDotEl, N - e0);
  const size_t o = (size_t)e0 * K2;
  const double* yb = y + (size_t)b * N * K2;
  for (int v = threadIdx.x; v < nel * K2; v += kDotEl) prod[v] = lam[o + v] * yb[o + v];
  __syncthreads();
  if (threadIdx.x >= nel) return;
  double s = 0.;
  for (int r = 0; r < K2; ++r) s += prod[threadIdx.x * K2 + r];
  p[(size_t)b * N + e0 + threadIdx.x] = s;
}

// one thread per (member of the colour, dof): members = the colour's element list
__global__ void k_adj_gather(int N, int K2, int nmem, const int32_t* __restrict__ members,
                             const int32_t* __restrict__ rptr, const int32_t* __restrict__ ridx,
                             const double* __restrict__ p, double* __restrict__ out, int ostride, int ooff) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)nmem * K2) return;
  const int k = (int)(t / K2), m = (int)(t - (size_t)k * K2);
  const int e = members[k];
  const double* pm = p + (size_t)m * N;
  double s = 0.;
  for (int j = rptr[e]; j < rptr[e + 1]; ++j) s += pm[ridx[j]];
  out[(size_t)e * ostride + ooff + m] = s;
}

// U seeds and their normal traces (k_velocity's: faces 0/1 from U1 along xi at eta point j,
// faces 2/3 from U2 along eta at xi point i)
__global__ void k_adj_seed_u(int N, int K, const int32_t* __restrict__ color, int c, int comp, Tab T,
                             double* __restrict__ U, double* __restrict__ tru) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  const int K2 = K * K, j0 = b / K, i0 = b % K;
  const size_t NU = (size_t)N * 2 * K2, TU = (size_t)4 * N * K;
  if (t < NU) {  // U values
    const int e = (int)(t / (2 * K2)), o = (int)(t - (size_t)e * 2 * K2);
    U[(size_t)b * NU + t] = (color[e] == c && o == comp * K2 + b) ? 1. : 0.;
  } else if (t < NU + TU) {  // normal traces
    const size_t q = t - NU;
    const int f = (int)(q / ((size_t)N * K));
    const size_t rem = q - (size_t)f * N * K;
    const int e = (int)(rem / K), k = (int)(rem - (size_t)e * K);
    const bool face_on = color[e] == c && ((f < 2) == (comp == 0));
    const int pos = f < 2 ? j0 : i0, idx = f < 2 ? i0 : j0, ib = f & 1;
    tru[(size_t)b * TU + q] = (face_on && k == pos) ? T.bb[idx][ib] : 0.;
  }
}

__global__ void k_adj_assemble_vel(int N, int K, const int32_t* __restrict__ color, int c,
                                   const int32_t* __restrict__ nb, const double* __restrict__ dU,
                                   double* __restrict__ Mv) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = blockIdx.y;
  if (e >= N) return;
  const int K2 = K * K, O = 2 * K2;
  for (int sl = 0; sl < 5; ++sl) {
    const int i = sl == 0 ? e : nb[(size_t)e * 4 + sl - 1];
    if (i < 0 || color[i] != c) continue;
    const double* src = dU + ((size_t)m * N + e) * O;
    double* dst = Mv + (((size_t)e * 5 + sl) * K2 + m) * O;
    for (int o = 0; o < O; ++o) dst[o] = src[o];
  }
}

__global__ void k_adj_assemble_diff(int N, int E, int K, const int32_t* __restrict__ color, int c,
                                    const int32_t* __restrict__ own_el, const double* __restrict__ dd,
                                    double* __restrict__ Md) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = blockIdx.y;
  if (g >= E || color[own_el[g]] != c) return;
  const int K2 = K * K;
  for (int j = 0; j < K; ++j) Md[((size_t)g * K + j) * K2 + m] = dd[(size_t)m * E * K + (size_t)g * K + j];
}

// one wavefront per element i: lane l takes outputs o = l, l + 64, ... of each contributing block
// (contiguous rows), accumulating per dof m; the per-m partials are summed across the wave in LDS
template <int K>
__global__ __launch_bounds__(64) void k_adj_vel_transpose(int N, const int32_t* __restrict__ nb,
                                                         const int8_t* __restrict__ nbf, const double* __restrict__ Mv,
                                                         const double* __restrict__ Md,
                                                         const double* __restrict__ Ubar,
                                                         const double* __restrict__ diffbar,
                                                         double* __restrict__ psibar) {
  constexpr int K2 = K * K, O = 2 * K2;
  __shared__ double part[64 * K2];
  const int i = blockIdx.x, l = threadIdx.x;
  double acc[K2];
#pragma unroll
  for (int m = 0; m < K2; ++m) acc[m] = 0.;
  for (int sl = 0; sl < 5; ++sl) {
    int e = i, es = 0;
    if (sl > 0) {
      e = nb[(size_t)i * 4 + sl - 1];
      if (e < 0) continue;
      es = 1 + nbf[(size_t)i * 4 + sl - 1];
    }
    const double* a = Mv + ((size_t)e * 5 + es) * K2 * O;
    const double* u = Ubar + (size_t)e * O;
    for (int o = l; o < O; o += 64) {
      const double uo = u[o];
#pragma unroll
      for (int m = 0; m < K2; ++m) acc[m] += a[(size_t)m * O + o] * uo;
    }
  }
#pragma unroll
  for (int m = 0; m < K2; ++m) part[l * K2 + m] = acc[m];
  __syncthreads();
  if (l < K2) {
    double s = 0.;
    for (int q = 0; q < 64; ++q) s += part[q * K2 + l];
    for (int f = 0; f < 4; ++f) {  // boundary edges of i: their wall flux
      const int e = nb[(size_t)i * 4 + f];
      if (e >= 0) continue;
      const int g = -1 - e;
      for (int j = 0; j < K; ++j) s += Md[((size_t)g * K + j) * K2 + l] * diffbar[(size_t)g * K + j];
    }
    psibar[(size_t)i * K2 + l] = s;
  }
}

__global__ void k_adj_wall(int E, int K, const int8_t* __restrict__ no_slip, const double* __restrict__ esign,
                           double idr, const double* __restrict__ diffbar, double* __restrict__ pvbar) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= E * K) return;
  const int g = i / K;
  if (no_slip[g]) pvbar[i] += esign[g] * idr * diffbar[i];
}

__global__ void k_adj_edge_seed(int E, int K, const int8_t* __restrict__ own_f, double* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;  // b = f K + j
  if (i >= E * K) return;
  const int e = i / K;
  x[(size_t)b * E * K + i] = (own_f[e] == b / K && i - e * K == b % K) ? 1. : 0.;
}

__global__ void k_adj_edge_gather(int N, int E, int K, const int32_t* __restrict__ own_el,
                                  const int8_t* __restrict__ own_f, const double* __restrict__ p,
                                  double* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  const int f = own_f[e], el = own_el[e];
  for (int j = 0; j < K; ++j) out[(size_t)e * K + j] += p[(size_t)(f * K + j) * N + el];
}

__global__ void k_adj_axpby(size_t n, double a, const double* __restrict__ x, double b, double* __restrict__ y) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i] + (b == 0. ? 0. : b * y[i]);
}

struct Wts {
  double w[kMaxKd];
};

__global__ void k_adj_wj(size_t D, int K, Wts W, const double* __restrict__ jac, const double* __restrict__ x,
                         double* __restrict__ y, int mode) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= D) return;
  const int K2 = K * K;
  const int v = (int)(i % K2);
  const double wj = W.w[v / K] * W.w[v % K] * jac[i];
  if (mode == 0) y[i] = x[i] / wj;
  else y[i] += wj * x[i];
}

}  // namespace

hipError_t launch_adj_seed_batch(int N, int K, const int32_t* color, int c, const Tab& T, double* x, double* tr,
                                 hipStream_t s) {
  const int K2 = K * K;
  const size_t D = (size_t)N * K2;
  if (D == 0) return hipSuccess;
  if (K > kMaxKd) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_adj_seed, dim3(grid(D), K2), dim3(kTB), 0, s, D, K2, color, c, x);
  hipLaunchKernelGGL(k_adj_seed_tr, dim3(grid((size_t)4 * N * 2 * K), K2), dim3(kTB), 0, s, N, K, color, c, T, tr);
  return hipGetLastError();
}

hipError_t launch_adj_dot_batch(int N, int K2, int nb, const double* lam, const double* y, double* p, hipStream_t s) {
  if (N == 0 || nb == 0) return hipSuccess;
  if (K2 > kMaxK2) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_adj_dot, dim3((N + kDotEl - 1) / kDotEl, nb), dim3(kDotEl), 0, s, N, K2, lam, y, p);
  return hipGetLastError();
}

hipError_t launch_adj_gather_batch(int N, int K2, int nmem, const int32_t* members, const int32_t* rptr,
                                   const int32_t* ridx, const double* p, double* out, int ostride, int ooff,
                                   hipStream_t s) {
  if (nmem == 0) return hipSuccess;
  hipLaunchKernelGGL(k_adj_gather, dim3(grid((size_t)nmem * K2)), dim3(kTB), 0, s, N, K2, nmem, members, rptr, ridx,
                     p, out, ostride, ooff);
  return hipGetLastError();
}

hipError_t launch_adj_seed_u_batch(int N, int K, const int32_t* color, int c, int comp, const Tab& T, double* U,
                                   double* tru, hipStream_t s) {
  if (N == 0) return hipSuccess;
  if (K > kMaxKd || comp < 0 || comp > 1) return hipErrorInvalidValue;
  const size_t n = (size_t)N * 2 * K * K + (size_t)4 * N * K;
  hipLaunchKernelGGL(k_adj_seed_u, dim3(grid(n), K * K), dim3(kTB), 0, s, N, K, color, c, comp, T, U, tru);
  return hipGetLastError();
}

hipError_t launch_adj_assemble_vel(int N, int E, int K, const int32_t* color, int c, const int32_t* nb,
                                   const int32_t* own_el, const double* dU, const double* dd, double* Mv, double* Md,
                                   hipStream_t s) {
  if (N == 0) return hipSuccess;
  hipLaunchKernelGGL(k_adj_assemble_vel, dim3(grid(N), K * K), dim3(kTB), 0, s, N, K, color, c, nb, dU, Mv);
  if (E > 0)
    hipLaunchKernelGGL(k_adj_assemble_diff, dim3(grid(E), K * K), dim3(kTB), 0, s, N, E, K, color, c, own_el, dd, Md);
  return hipGetLastError();
}

hipError_t launch_adj_vel_transpose(int N, int K, const int32_t* nb, const int8_t* nbf, const double* Mv,
                                    const double* Md, const double* Ubar, const double* diffbar, double* psibar,
                                    hipStream_t s) {
  if (N == 0) return hipSuccess;
  switch (K) {
#define VT(KK) \
  case KK: hipLaunchKernelGGL(k_adj_vel_transpose<KK>, dim3(N), dim3(64), 0, s, N, nb, nbf, Mv, Md, Ubar, diffbar, psibar); break;
    VT(2) VT(3) VT(4) VT(5) VT(6)
#undef VT
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_adj_wall(int E, int K, const int8_t* no_slip, const double* esign, double idr,
                           const double* diffbar, double* pvbar, hipStream_t s) {
  if (E == 0) return hipSuccess;
  hipLaunchKernelGGL(k_adj_wall, dim3(grid((size_t)E * K)), dim3(kTB), 0, s, E, K, no_slip, esign, idr, diffbar,
                     pvbar);
  return hipGetLastError();
}

hipError_t launch_adj_edge_seed_batch(int E, int K, const int8_t* own_f, double* x, hipStream_t s) {
  if (E == 0) return hipSuccess;
  hipLaunchKernelGGL(k_adj_edge_seed, dim3(grid((size_t)E * K), 4 * K), dim3(kTB), 0, s, E, K, own_f, x);
  return hipGetLastError();
}

hipError_t launch_adj_edge_gather_batch(int N, int E, int K, const int32_t* own_el, const int8_t* own_f,
                                        const double* p, double* out, hipStream_t s) {
  if (E == 0) return hipSuccess;
  hipLaunchKernelGGL(k_adj_edge_gather, dim3(grid(E)), dim3(kTB), 0, s, N, E, K, own_el, own_f, p, out);
  return hipGetLastError();
}

hipError_t launch_adj_axpby(size_t n, double a, const double* x, double b, double* y, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_adj_axpby, dim3(grid(n)), dim3(kTB), 0, s, n, a, x, b, y);
  return hipGetLastError();
}

hipError_t launch_adj_wj(int N, int K, const double* wts, const double* jac, const double* x, double* y, int mode,
                         hipStream_t s) {
  const size_t D = (size_t)N * K * K;
  if (D == 0) return hipSuccess;
  if (K > kMaxKd) return hipErrorInvalidValue;
  Wts W{};
  for (int k = 0; k < K; ++k) W.w[k] = wts[k];
  hipLaunchKernelGGL(k_adj_wj, dim3(grid(D)), dim3(kTB), 0, s, D, K, W, jac, x, y, mode);
  return hipGetLastError();
}

}  // namespace hoc
