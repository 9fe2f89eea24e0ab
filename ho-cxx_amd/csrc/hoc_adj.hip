// hoc_adj.hip - seeds, per-element products and transposed gathers of the discrete adjoint.
#include "hoc_adj.hpp"

namespace hoc {

namespace {

constexpr int kTB = 256;

__global__ void k_adj_seed(size_t D, int K2, const int32_t* __restrict__ color, int c, int m, double* __restrict__ x) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= D) return;
  const size_t e = i / K2;
  x[i] = (color[e] == c && (int)(i - e * K2) == m) ? 1. : 0.;
}

// one thread per element
__global__ void k_adj_dot(int N, int K2, const double* __restrict__ lam, const double* __restrict__ y,
                          double* __restrict__ p) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= N) return;
  const size_t b = (size_t)e * K2;
  double s = 0.;
  for (int r = 0; r < K2; ++r) s += lam[b + r] * y[b + r];
  p[e] = s;
}

__global__ void k_adj_gather(int N, int K2, const int32_t* __restrict__ color, int c, int m,
                             const int32_t* __restrict__ rptr, const int32_t* __restrict__ ridx,
                             const double* __restrict__ p, double* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= N || color[e] != c) return;
  double s = 0.;
  for (int j = rptr[e]; j < rptr[e + 1]; ++j) s += p[ridx[j]];
  out[(size_t)e * K2 + m] = s;
}

__global__ void k_adj_edge_seed(int E, int K, const int8_t* __restrict__ own_f, int f, int j, double* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= E * K) return;
  const int e = i / K;
  x[i] = (own_f[e] == f && i - e * K == j) ? 1. : 0.;
}

__global__ void k_adj_edge_gather(int E, int K, const int32_t* __restrict__ own_el, const int8_t* __restrict__ own_f,
                                  int f, int j, const double* __restrict__ p, double* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E || own_f[e] != f) return;
  out[(size_t)e * K + j] += p[own_el[e]];
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

inline unsigned grid(size_t n) { return (unsigned)((n + kTB - 1) / kTB); }

}  // namespace

hipError_t launch_adj_seed(int N, int K2, const int32_t* color, int c, int m, double* x, hipStream_t s) {
  const size_t D = (size_t)N * K2;
  if (D == 0) return hipSuccess;
  hipLaunchKernelGGL(k_adj_seed, dim3(grid(D)), dim3(kTB), 0, s, D, K2, color, c, m, x);
  return hipGetLastError();
}

hipError_t launch_adj_dot(int N, int K2, const double* lam, const double* y, double* p, hipStream_t s) {
  if (N == 0) return hipSuccess;
  hipLaunchKernelGGL(k_adj_dot, dim3(grid(N)), dim3(kTB), 0, s, N, K2, lam, y, p);
  return hipGetLastError();
}

hipError_t launch_adj_gather(int N, int K2, const int32_t* color, int c, int m, const int32_t* rptr,
                             const int32_t* ridx, const double* p, double* out, hipStream_t s) {
  if (N == 0) return hipSuccess;
  hipLaunchKernelGGL(k_adj_gather, dim3(grid(N)), dim3(kTB), 0, s, N, K2, color, c, m, rptr, ridx, p, out);
  return hipGetLastError();
}

hipError_t launch_adj_edge_seed(int E, int K, const int8_t* own_f, int f, int j, double* x, hipStream_t s) {
  if (E == 0) return hipSuccess;
  hipLaunchKernelGGL(k_adj_edge_seed, dim3(grid((size_t)E * K)), dim3(kTB), 0, s, E, K, own_f, f, j, x);
  return hipGetLastError();
}

hipError_t launch_adj_edge_gather(int E, int K, const int32_t* own_el, const int8_t* own_f, int f, int j,
                                  const double* p, double* out, hipStream_t s) {
  if (E == 0) return hipSuccess;
  hipLaunchKernelGGL(k_adj_edge_gather, dim3(grid(E)), dim3(kTB), 0, s, E, K, own_el, own_f, f, j, p, out);
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
