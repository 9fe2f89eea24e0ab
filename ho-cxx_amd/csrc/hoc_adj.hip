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
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (e >= N) return;
  const int j0 = b / K, i0 = b % K;
  const bool on = color[e] == c;
  double* t = tr + (size_t)b * 4 * N * 2 * K;
  for (int f = 0; f < 4; ++f) {
    double* q = t + ((size_t)f * N + e) * 2 * K;
    const int pos = f < 2 ? j0 : i0, idx = f < 2 ? i0 : j0, ib = f & 1;
    for (int k = 0; k < K; ++k) {
      q[k] = (on && k == pos) ? T.bb[idx][ib] : 0.;
      q[K + k] = (on && k == pos) ? T.bgb[idx][ib] : 0.;
    }
  }
}

__global__ void k_adj_dot(int N, int K2, const double* __restrict__ lam, const double* __restrict__ y,
                          double* __restrict__ p) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (e >= N) return;
  const size_t o = (size_t)e * K2;
  const double* yb = y + (size_t)b * N * K2;
  double s = 0.;
  for (int r = 0; r < K2; ++r) s += lam[o + r] * yb[o + r];
  p[(size_t)b * N + e] = s;
}

__global__ void k_adj_gather(int N, int K2, const int32_t* __restrict__ color, int c,
                             const int32_t* __restrict__ rptr, const int32_t* __restrict__ ridx,
                             const double* __restrict__ p, double* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= N || color[e] != c) return;
  for (int m = 0; m < K2; ++m) {
    const double* pm = p + (size_t)m * N;
    double s = 0.;
    for (int j = rptr[e]; j < rptr[e + 1]; ++j) s += pm[ridx[j]];
    out[(size_t)e * K2 + m] = s;
  }
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
  hipLaunchKernelGGL(k_adj_seed_tr, dim3(grid(N), K2), dim3(kTB), 0, s, N, K, color, c, T, tr);
  return hipGetLastError();
}

hipError_t launch_adj_dot_batch(int N, int K2, int nb, const double* lam, const double* y, double* p, hipStream_t s) {
  if (N == 0 || nb == 0) return hipSuccess;
  hipLaunchKernelGGL(k_adj_dot, dim3(grid(N), nb), dim3(kTB), 0, s, N, K2, lam, y, p);
  return hipGetLastError();
}

hipError_t launch_adj_gather_batch(int N, int K2, const int32_t* color, int c, const int32_t* rptr,
                                   const int32_t* ridx, const double* p, double* out, hipStream_t s) {
  if (N == 0) return hipSuccess;
  hipLaunchKernelGGL(k_adj_gather, dim3(grid(N)), dim3(kTB), 0, s, N, K2, color, c, rptr, ridx, p, out);
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
