// hoc_la.hip - vector, reduction, CSR and probing kernels of the Poisson solver.
#include "hoc_la.hpp"

#include <algorithm>
#include <type_traits>

namespace hoc {

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block (256 threads) sum of two values; result valid in thread 0.
__device__ __forceinline__ void block_sum2(double& a, double& b) {
  __shared__ double sa[4], sb[4];
  a = wsum(a); b = wsum(b);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) { sa[w] = a; sb[w] = b; }
  __syncthreads();
  if (threadIdx.x == 0) {
    a = sa[0] + sa[1] + sa[2] + sa[3];
    b = sb[0] + sb[1] + sb[2] + sb[3];
  }
}

// Position within its element (j K + i) of a grid-stride loop's flat dof index v
// (element-major, K*K per element), advanced without a 64-bit division per step; the
// inverse weight 1/(w_j w_i) of that position is A.winv[m].
struct DofSlot {
  int m, step, K2;
  size_t el, estep;  // the element of the dof
  __device__ DofSlot(const VecArgs& A, size_t v0, size_t stride) {
    K2 = A.K * A.K;
    m = (int)(v0 % (size_t)K2);
    step = (int)(stride % (size_t)K2);
    el = v0 / (size_t)K2;
    estep = stride / (size_t)K2;
  }
  __device__ void next() {
    m += step;
    el += estep;
    if (m >= K2) { m -= K2; ++el; }
  }
};

// Elementwise Krylov updates with fused partial norms.  Scalars come from S.
template <class R>
__global__ void k_vec(int op, VecArgs A, size_t n, R* __restrict__ a, R* __restrict__ b, const R* __restrict__ c,
                      const R* __restrict__ d, R* __restrict__ e, const double* __restrict__ S,
                      const int* __restrict__ F, double* __restrict__ part) {
  if (F && !F[F_ACTIVE]) return;
  if (op >= VOP_BICG_X && F && F[F_HALF]) return;
  double p0 = 0., p1 = 0.;
  const size_t v0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (size_t)gridDim.x * blockDim.x;
  DofSlot ds(A, v0, stride);
  for (size_t v = v0; v < n; v += stride, ds.next()) {
    switch (op) {
      case VOP_PCG_UPD: {  // x += alpha p; r -= alpha q      a=x b=r c=p d=q
        if (A.ialpha && A.ialpha[A.einst[ds.el]] == 0.) break;  // a stopped ensemble instance
        const R al = (R)(A.ialpha ? A.ialpha[A.einst[ds.el]] : S[S_ALPHA]);
        a[v] += al * c[v];
        const R r = b[v] - al * d[v];
        b[v] = r;
        const double ru = r * A.winv[ds.m];
        p0 += (double)r * r; p1 += ru * ru;
      } break;
      case VOP_BICG_P: {   // p = r + beta (p - omega v)      a=p c=r d=v
        a[v] = c[v] + (R)S[S_BETA] * (a[v] - (R)S[S_OMEGA] * d[v]);
      } break;
      case VOP_BICG_S: {   // s = r - alpha v (in r)          b=r d=v
        const R s = b[v] - (R)S[S_ALPHA] * d[v];
        b[v] = s;
        const double su = s * A.winv[ds.m];
        p0 += (double)s * s; p1 += su * su;
      } break;
      case VOP_BICG_X: {   // x += alpha ph + omega sh; r = s - omega t   a=x b=r(s) c=ph d=sh e=t
        a[v] += (R)S[S_ALPHA] * c[v] + (R)S[S_OMEGA] * d[v];
        const R r = b[v] - (R)S[S_OMEGA] * e[v];
        b[v] = r;
        const double ru = r * A.winv[ds.m];
        p0 += (double)r * r; p1 += ru * ru;
      } break;
      case VOP_DOT2: {     // partials of a.c and b.d (read only)
        p0 += (double)a[v] * c[v]; p1 += (double)b[v] * d[v];
      } break;
      case VOP_COPY: {     // a = c
        a[v] = c[v];
      } break;
    }
  }
  if (part) {
    block_sum2(p0, p1);
    if (threadIdx.x == 0) { part[blockIdx.x] = p0; part[gridDim.x + blockIdx.x] = p1; }
  }
}

// x += alpha ph when BiCGSTAB converged at the half step.
template <class R>
__global__ void k_bicg_half(size_t n, R* __restrict__ x, const R* __restrict__ ph, const double* __restrict__ S,
                            int* __restrict__ F) {
  if (!F[F_ACTIVE] || !F[F_HALF]) return;
  for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (size_t)gridDim.x * blockDim.x)
    x[v] += (R)S[S_ALPHA] * ph[v];
}

template <class R>
__global__ void k_zero_if_flag(size_t n, R* __restrict__ x, const int* __restrict__ F) {
  if (!F[F_ZERO]) return;
  for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (size_t)gridDim.x * blockDim.x) x[v] = 0;
}

// Scalar logic of the Krylov solvers.  pa: np partials of q0 then np of q1.  With mirror
// set (pinned host memory), thread 0 finally publishes seq * 4 + near * 2 + (active flag) there
// with a system-scope store, whether or not the phase ran, so the host can follow the iterations
// without a copy (the host reads only this word, so the store needs no ordering); near: still
// active with the residual within 8x of the tolerance, so the next iteration may be the last.
__device__ void finalize_phase(int phase, double a, double b, double* __restrict__ S, int* __restrict__ F,
                               int maxiter, const double* __restrict__ irr, const double* __restrict__ ibb,
                               int ninst);
__global__ void k_finalize(int phase, const double* __restrict__ pa, int np, double* __restrict__ S,
                           int* __restrict__ F, int maxiter, int* __restrict__ mirror, int seq,
                           const double* __restrict__ irr, const double* __restrict__ ibb, int ninst) {
  const bool live = phase == PH_INIT || phase == PH_BB || F[F_ACTIVE];
  const bool half = live && phase >= PH_OMEGA && F[F_HALF];  // BiCGSTAB converged at the half step
  if (live && !half) {  // uniform over the block
    double a = 0., b = 0.;
    for (int i = threadIdx.x; i < np; i += blockDim.x) { a += pa[i]; b += pa[np + i]; }
    block_sum2(a, b);
    if (threadIdx.x == 0) finalize_phase(phase, a, b, S, F, maxiter >= 0 ? maxiter : F[F_MAXIT], irr, ibb, ninst);
  } else if (half && phase == PH_RNORM && threadIdx.x == 0) {
    F[F_ACTIVE] = 0;
  }
  if (mirror && threadIdx.x == 0) {
    const int sq = seq >= 0 ? seq : ++F[F_SEQ];
    const int act = F[F_ACTIVE] ? 1 : 0;
    const int near = act && (irr ? S[S_IMAX] <= 64. : S[S_RR] <= 64. * S[S_TOL2] * S[S_BB]);  // within 8x of the tolerance
    __hip_atomic_store(mirror, sq * 4 + near * 2 + act, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}
__global__ void k_set_int(int* __restrict__ p, int v) { *p = v; }
__device__ void finalize_phase(int phase, double a, double b, double* __restrict__ S, int* __restrict__ F,
                               int maxiter, const double* __restrict__ irr, const double* __restrict__ ibb,
                               int ninst) {
  const double tol2bb = S[S_TOL2] * S[S_BB];
  // the convergence test of PH_INIT / PH_UPD: the joint residual, or every instance's own
  bool above = b > tol2bb;
  if (irr && (phase == PH_INIT || phase == PH_UPD)) {
    double bmax = 0., rmax = 0.;
    for (int i = 0; i < ninst; ++i) bmax = fmax(bmax, ibb[i]);
    for (int i = 0; i < ninst; ++i) {
      const double d = S[S_TOL2] * (ibb[i] > 0. ? ibb[i] : bmax);
      rmax = fmax(rmax, d > 0. ? irr[i] / d : (irr[i] > 0. ? HUGE_VAL : 0.));
    }
    S[S_IMAX] = rmax;
    above = rmax > 1. || !isfinite(rmax);
  }
  switch (phase) {
    case PH_BB: S[S_BB] = a; break;
    case PH_INIT: {
      S[S_RR] = b;
      const double tiny = 2. * 2.220446049250313e-16;
      const bool zero = S[S_BB] < tiny * tiny;
      F[F_ZERO] = zero;
      F[F_ITER] = 0;
      F[F_HALF] = 0;
      F[F_ACTIVE] = (!zero && above && maxiter > 0) ? 1 : 0;
      S[S_RZ] = 0.; S[S_RHO] = 1.; S[S_ALPHA] = 1.; S[S_OMEGA] = 1.; S[S_BETA] = 0.;
    } break;
    case PH_RZ: {  // breakdown (z = 0 or an indefinite preconditioner) ends the solve
      if (!(a > 0.) || !isfinite(a)) { F[F_ACTIVE] = 0; break; }
      S[S_BETA] = F[F_ITER] > 0 ? a / S[S_RZ] : 0.;
      S[S_RZ] = a;
    } break;
    case PH_PQ: {
      const double al = S[S_RZ] / a;
      if (!(a > 0.) || !isfinite(al)) { F[F_ACTIVE] = 0; break; }
      S[S_PQ] = a;
      S[S_ALPHA] = al;
    } break;
    case PH_UPD: {
      S[S_RR] = b;
      F[F_ITER] += 1;
      F[F_ACTIVE] = (above && F[F_ITER] < maxiter && isfinite(b)) ? 1 : 0;
    } break;
    case PH_RHO: {
      if (a == 0. || !isfinite(a)) { F[F_ACTIVE] = 0; break; }
      S[S_BETA] = F[F_ITER] > 0 ? (a / S[S_RHO]) * (S[S_ALPHA] / S[S_OMEGA]) : 0.;
      S[S_RHO] = a;
    } break;
    case PH_ALPHA: {
      const double al = S[S_RHO] / a;
      if (a == 0. || !isfinite(al)) { F[F_ACTIVE] = 0; break; }
      S[S_ALPHA] = al;
    } break;
    case PH_SNORM: {
      S[S_RR] = b;
      if (b <= tol2bb) { F[F_HALF] = 1; F[F_ITER] += 1; }
    } break;
    case PH_OMEGA: {  // a = t.s, b = t.t
      const double om = b != 0. ? a / b : 0.;
      S[S_OMEGA] = isfinite(om) ? om : 0.;
    } break;
    case PH_RNORM: {
      S[S_RR] = b;
      F[F_ITER] += 1;
      F[F_ACTIVE] = (b > tol2bb && F[F_ITER] < maxiter && S[S_OMEGA] != 0. && isfinite(b)) ? 1 : 0;
    } break;
  }
}

// ---- Sparse levels -------------------------------------------------------------------------
// mode 0: y = A x;  1: y = f - A x;  2: y = f + A x;  3: y = x + M (f - A x) (spai0 step);
//      4: xio = M f and y = f - A (M f) (spai0 step from zero fused with its residual; x unused).
// Every mode writes only row i of its outputs, and no output may alias an input read at
// other rows (x, or f in mode 4): the relaxation is a true Jacobi-type step, independent of
// the order in which rows complete.
template <class R>
__device__ __forceinline__ void sparse_out(int mode, int i, R s, const R* __restrict__ x, const R* __restrict__ f,
                                           R* __restrict__ y, const R* __restrict__ M, R* __restrict__ xio) {
  if (mode == 0) y[i] = s;
  else if (mode == 1) y[i] = f[i] - s;
  else if (mode == 2) y[i] = f[i] + s;
  else if (mode == 3) y[i] = x[i] + M[i] * (f[i] - s);
  else { y[i] = f[i] - s; xio[i] = M[i] * f[i]; }
}

// CSR, one thread per row (short rows).
template <class R>
__global__ void k_csr(int mode, int n, const int32_t* __restrict__ ptr, const int32_t* __restrict__ col,
                      const R* __restrict__ val, const R* __restrict__ x, const R* __restrict__ f, R* __restrict__ y,
                      const R* __restrict__ M, R* __restrict__ xio, const int* __restrict__ F) {
  if (F && !F[F_ACTIVE]) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  R s = 0;
  if (mode == 4)
    for (int j = ptr[i]; j < ptr[i + 1]; ++j) s += val[j] * (M[col[j]] * f[col[j]]);
  else
    for (int j = ptr[i]; j < ptr[i + 1]; ++j) s += val[j] * x[col[j]];
  sparse_out(mode, i, s, x, f, y, M, xio);
}

// CSR with LPR lanes per row (rows of many nonzeros and uneven length): the lanes of a
// group stride over the row together, so a wavefront's loads are contiguous.
template <int LPR, class R>
__global__ void k_csr_vec(int mode, int n, const int32_t* __restrict__ ptr, const int32_t* __restrict__ col,
                          const R* __restrict__ val, const R* __restrict__ x, const R* __restrict__ f,
                          R* __restrict__ y, const R* __restrict__ M, R* __restrict__ xio, const int* __restrict__ F) {
  if (F && !F[F_ACTIVE]) return;
  const int lane = threadIdx.x % LPR;
  const int i = (blockIdx.x * blockDim.x + threadIdx.x) / LPR;
  R s = 0;
  if (i < n) {
    if (mode == 4)
      for (int j = ptr[i] + lane; j < ptr[i + 1]; j += LPR) s += val[j] * (M[col[j]] * f[col[j]]);
    else
      for (int j = ptr[i] + lane; j < ptr[i + 1]; j += LPR) s += val[j] * x[col[j]];
  }
#pragma unroll
  for (int off = LPR / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, LPR);
  if (i >= n || lane != 0) return;
  sparse_out(mode, i, s, x, f, y, M, xio);
}

// ELL (rows of nearly equal length, e.g. the Q1 vertex operator) with L lanes per row: slot k
// of row i is at ((k / L) n + i) L + k % L, so the L threads of a row, and consecutive rows,
// read consecutive words, and each thread walks w / L slots.  Padding slots hold a column of
// the same row and value 0.
template <int L, class R>
__global__ void k_ell(int mode, int n, int w, const int32_t* __restrict__ col, const R* __restrict__ val,
                      const R* __restrict__ x, const R* __restrict__ f, R* __restrict__ y, const R* __restrict__ M,
                      R* __restrict__ xio, const int* __restrict__ F) {
  if (F && !F[F_ACTIVE]) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = t / L, lane = t % L;
  R s = 0;
  if (i < n) {
    const int groups = w / L;
    if (mode == 4) {
#pragma unroll 4
      for (int g = 0; g < groups; ++g) {
        const size_t q = ((size_t)g * n + i) * L + lane;
        const int c = col[q];
        s += val[q] * (M[c] * f[c]);
      }
    } else {
#pragma unroll 4
      for (int g = 0; g < groups; ++g) {
        const size_t q = ((size_t)g * n + i) * L + lane;
        s += val[q] * x[col[q]];
      }
    }
  }
#pragma unroll
  for (int off = L / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, L);
  if (i >= n || lane != 0) return;
  sparse_out(mode, i, s, x, f, y, M, xio);
}

// x = M f (spai0 from a zero guess)
template <class R>
__global__ void k_scale(int n, const R* __restrict__ M, const R* __restrict__ f, R* __restrict__ x,
                        const int* __restrict__ F) {
  if (F && !F[F_ACTIVE]) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = M[i] * f[i];
}

// Coarsest level: x = Ainv f (dense, row-major), one wavefront per row.
template <class R>
__global__ void k_dense(int n, const R* __restrict__ Ainv, const R* __restrict__ f, R* __restrict__ x,
                        const int* __restrict__ F) {
  if (F && !F[F_ACTIVE]) return;
  const int row = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  R s = 0;
  for (int j = lane; j < n; j += 64) s += Ainv[(size_t)row * n + j] * f[j];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) x[row] = s;
}

// y = -y / (w_j w_i) at every solution point (el K^2 + j K + i): As = -W A turned back into A.
struct Weights6 { double w[6]; };
__global__ void k_unweight(size_t n, int K, Weights6 W, double* __restrict__ y) {
  const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const int m = (int)(v % (size_t)(K * K));
  y[v] = -y[v] / (W.w[m / K] * W.w[m % K]);
}

// ---- probing ---------------------------------------------------------------------------------
// x[el][m] = (color[el] == c && (m == rs || rs < 0)) ? 1 : 0
__global__ void k_probe_set(int N, int K2, const int32_t* __restrict__ color, int c, int rs, double* __restrict__ x) {
  const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= (size_t)N * K2) return;
  const int el = (int)(v / K2), m = (int)(v % K2);
  x[v] = (color[el] == c && (rs < 0 || m == rs)) ? 1. : 0.;
}
// diag[el][rs] = y[el][rs] for elements of colour c
__global__ void k_probe_diag(int N, int K2, const int32_t* __restrict__ color, int c, int rs, const double* __restrict__ y,
                             double* __restrict__ diag) {
  const int el = blockIdx.x * blockDim.x + threadIdx.x;
  if (el >= N || color[el] != c) return;
  diag[(size_t)el * K2 + rs] = y[(size_t)el * K2 + rs];
}
// ---- Q1 coarse space ---------------------------------------------------------------------------
// rc[v] = sum over the (element, corner) incidences of vertex v of ec[element][corner]
template <class R>
__global__ void k_q1_gather(int nv, const int32_t* __restrict__ vptr, const int32_t* __restrict__ vinc,
                            const R* __restrict__ ec, R* __restrict__ rc, const int* __restrict__ F) {
  if (F && !F[F_ACTIVE]) return;
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  R s = 0;
  for (int k = vptr[v]; k < vptr[v + 1]; ++k) s += ec[vinc[k]];
  rc[v] = s;
}
// For a probe y = As x (x = unit vector at local dof rs of every element of one colour),
// M[f][a][b] += (sum_ij q1w(a,ij) y[f][ij]) * q1w(b, rs): the Galerkin block between the
// corners of f and the corners of f's probed neighbour.
__global__ void k_q1_probe_accum(int N, int K, Tab T, int rs, const double* __restrict__ y, double* __restrict__ M) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= N) return;
  const int K2 = K * K;
  double t[4] = {0., 0., 0., 0.};
  for (int m = 0; m < K2; ++m) {
    const double v = y[(size_t)f * K2 + m];
    for (int a = 0; a < 4; ++a) t[a] += q1w(T, a, m / K, m % K) * v;
  }
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) M[(size_t)f * 16 + a * 4 + b] += t[a] * q1w(T, b, rs / K, rs % K);
}

// ---- mixed precision ---------------------------------------------------------------------------
// FP64 Krylov method with the FP32 V-cycle as its preconditioner: the V-cycle's input
// r32 = (float) r64, and its output z64 = (double) z32 with partials of r64 . z64 (slot 0; slot 1
// zero) when part is set.
__global__ void k_demote(size_t n, const double* __restrict__ r64, float* __restrict__ r32, const int* __restrict__ F) {
  if (F && !F[F_ACTIVE]) return;
  for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (size_t)gridDim.x * blockDim.x)
    r32[v] = (float)r64[v];
}
__global__ void k_promote_dot(size_t n, const float* __restrict__ z32, double* __restrict__ z64,
                              const double* __restrict__ r64, double* __restrict__ part, const int* __restrict__ F) {
  if (F && !F[F_ACTIVE]) return;
  double p0 = 0., p1 = 0.;
  for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (size_t)gridDim.x * blockDim.x) {
    const double z = (double)z32[v];
    z64[v] = z;
    if (part) p0 += r64[v] * z;
  }
  if (part) {
    block_sum2(p0, p1);
    if (threadIdx.x == 0) { part[blockIdx.x] = p0; part[gridDim.x + blockIdx.x] = p1; }
  }
}

// ---- ensembles ---------------------------------------------------------------------------------
__global__ void k_inst_norms(VecArgs A, const int32_t* __restrict__ off, int K2, const double* __restrict__ b,
                             double* __restrict__ out) {
  const size_t v0 = (size_t)off[blockIdx.x] * K2, v1 = (size_t)off[blockIdx.x + 1] * K2;
  double p0 = 0.;
  DofSlot ds(A, v0 + threadIdx.x, blockDim.x);
  for (size_t v = v0 + threadIdx.x; v < v1; v += blockDim.x, ds.next()) {
    const double u = b[v] * A.winv[ds.m];
    p0 += u * u;
  }
  double p1 = 0.;
  block_sum2(p0, p1);
  if (threadIdx.x == 0) out[blockIdx.x] = p0;
}

// ---- launchers -------------------------------------------------------------------------------
static int grid_for(size_t n, int bs) {
  size_t g = (n + bs - 1) / bs;
  if (g > 4096) g = 4096;
  if (g < 1) g = 1;
  return (int)g;
}

int vec_grid(size_t n) { return grid_for(n, 256); }

template <class R>
void vec(hipStream_t s, int op, const VecArgs& A, size_t n, R* a, R* b, const R* c, const R* d, R* e, const double* S,
         const int* F, double* part) {
  hipLaunchKernelGGL(k_vec<R>, dim3(vec_grid(n)), dim3(256), 0, s, op, A, n, a, b, c, d, e, S, F, part);
}
template <class R>
void bicg_half(hipStream_t s, size_t n, R* x, const R* ph, const double* S, int* F) {
  hipLaunchKernelGGL(k_bicg_half<R>, dim3(vec_grid(n)), dim3(256), 0, s, n, x, ph, S, F);
}
__global__ void k_psi_predict(size_t n, double f, int extrap, double* __restrict__ psi, double* __restrict__ prev) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double a = psi[i];
    if (extrap) psi[i] = a + f * (a - prev[i]);
    prev[i] = a;
  }
}
struct Comb5 {
  double a[8];
  const double* x[8];
};
__global__ void k_psi_combine(size_t n, int nterm, Comb5 C, double* __restrict__ psi, double* __restrict__ save) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double old = psi[i];
    double v = 0.;
    for (int j = 0; j < nterm; ++j) v += C.a[j] * (C.x[j] == psi ? old : C.x[j][i]);
    psi[i] = v;
    if (save) save[i] = old;
  }
}
// The RK4 guess history at the first stage of a step, in one pass (T1 = psi on entry, the
// first stage solution of the new step; S1..S4 the previous step's stage solutions, Q4 its fourth
// stage difference): psi = 4 T1 - 2 S3 - S4 - Q4/2 (qok) or 4 T1 + S1 - S2 - 2 S3 - S4; then
// Q1 = 2 S2 - 2 S1, Q2 = 2 S3 - 2 S1, Q3 = S4 - S1, Q4 = 6 T1 + 2 S1 - 2 S2 - 4 S3 - 2 S4, S1 = T1.
__global__ void k_rk4_history(size_t n, int qok, double* __restrict__ psi, double* __restrict__ s1,
                              const double* __restrict__ s2, const double* __restrict__ s3,
                              const double* __restrict__ s4, double* __restrict__ q1, double* __restrict__ q2,
                              double* __restrict__ q3, double* __restrict__ q4) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double t = psi[i], a = s1[i], b = s2[i], c = s3[i], d = s4[i];
    double g = 4. * t;
    if (qok) {
      g += -2. * c;
      g += -1. * d;
      g += -0.5 * q4[i];
    } else {
      g += 1. * a;
      g += -1. * b;
      g += -2. * c;
      g += -1. * d;
    }
    psi[i] = g;
    double v = 2. * b;
    v += -2. * a;
    q1[i] = v;
    v = 2. * c;
    v += -2. * a;
    q2[i] = v;
    v = 6. * t;
    v += 2. * a;
    v += -2. * b;
    v += -4. * c;
    v += -2. * d;
    q4[i] = v;
    v = 1. * d;
    v += -1. * a;
    q3[i] = v;
    s1[i] = t;
  }
}
void rk4_history(hipStream_t s, size_t n, int qok, double* psi, double* const* S, double* const* Q) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_rk4_history, dim3(vec_grid(n)), dim3(256), 0, s, n, qok, psi, S[0], S[1], S[2], S[3], Q[0],
                     Q[1], Q[2], Q[3]);
}
void psi_combine(hipStream_t s, size_t n, int nterm, const double* a, const double* const* x, double* psi,
                 double* save) {
  if (n == 0 || nterm < 1 || nterm > 8) return;
  Comb5 C{};
  for (int j = 0; j < nterm; ++j) { C.a[j] = a[j]; C.x[j] = x[j]; }
  hipLaunchKernelGGL(k_psi_combine, dim3(vec_grid(n)), dim3(256), 0, s, n, nterm, C, psi, save);
}
// ---- projection basis kernels -----------------------------------------------------------------
// Every pass runs over the dof segments of a ProjSeg (one per ensemble instance): the blocks of
// segment s are [bbeg[s], bbeg[s+1]) and stride over its dofs [sbeg[s], sbeg[s+1]); reductions
// leave per-block partials that k_proj_fin sums per segment in block order (deterministic).
struct SegRange {
  size_t i0, i1, step;
  int s, lb;
};
__device__ __forceinline__ SegRange seg_range(const ProjSeg& S) {
  const int s = S.blk_seg[blockIdx.x];
  const int lb = blockIdx.x - S.bbeg[s], nbs = S.bbeg[s + 1] - S.bbeg[s];
  return {(size_t)S.sbeg[s] + (size_t)lb * blockDim.x + threadIdx.x, (size_t)S.sbeg[s + 1], (size_t)nbs * blockDim.x,
          s, lb};
}
void proj_seg_layout(const std::vector<int64_t>& sbeg, std::vector<int32_t>& bbeg, std::vector<int32_t>& blk_seg) {
  const int nseg = (int)sbeg.size() - 1;
  bbeg.assign(nseg + 1, 0);
  blk_seg.clear();
  for (int s = 0; s < nseg; ++s) {  // about 8 dofs per thread, at least one block per segment
    const int64_t n = sbeg[s + 1] - sbeg[s];
    const int nb = (int)std::max<int64_t>(1, (n + 2047) / 2048);
    bbeg[s + 1] = bbeg[s] + nb;
    for (int k = 0; k < nb; ++k) blk_seg.push_back(s);
  }
}

// The projection kernels take the basis size M as a template argument, so that the M + 2 loads
// of one dof are issued back to back (one streaming pass over the basis at full bandwidth);
// proj_dispatch maps the run-time m in [lo, kProjMax] to the instance.
template <int M, int LO, class Fn>
void proj_dispatch_impl(int m, Fn& fn) {
  if constexpr (M >= LO) {
    if (m == M) fn(std::integral_constant<int, M>{});
    else proj_dispatch_impl<M - 1, LO>(m, fn);
  }
}
template <int LO, class Fn>
void proj_dispatch(int m, Fn fn) { proj_dispatch_impl<kProjMax, LO>(m, fn); }

// Sum over the block (256 threads) of NV per-thread values; thread k < NV writes sum k to out[k].
template <int NV>
__device__ __forceinline__ void block_sums(double (&v)[NV], double* __restrict__ out) {
  __shared__ double red[4][NV];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const double w = wsum(v[k]);
    if (lane == 0) red[wv][k] = w;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    const int k = threadIdx.x;
    out[k] = red[0][k] + red[1][k] + red[2][k] + red[3][k];
  }
}
// part[block][0, M) = Y_j . t,  part[block][M, 2M) = Y_j . b over the block's segment
template <int M>
__global__ __launch_bounds__(256) void k_proj_dots2(size_t n, ProjSeg S, const double* __restrict__ Y,
                                                    const double* __restrict__ t, const double* __restrict__ b,
                                                    double* __restrict__ part) {
  double s[2 * M];
#pragma unroll
  for (int k = 0; k < 2 * M; ++k) s[k] = 0.;
  const SegRange R = seg_range(S);
  for (size_t i = R.i0; i < R.i1; i += R.step) {
    const double tv = t[i], bv = b[i];
#pragma unroll
    for (int j = 0; j < M; ++j) {
      const double y = Y[(size_t)j * n + i];
      s[j] += y * tv;
      s[M + j] += y * bv;
    }
  }
  block_sums<2 * M>(s, part + (size_t)blockIdx.x * 2 * M);
}
// k_proj_dots2 fused with the Gram column of the newest basis vector: Y_j . t, Y_j . b and
// Y_j . ax (j < M; ax the image of the newest vector, whose Gram column k_proj_gcol would give
// from the same loads of Y); partials at part + b 3M.
template <int M>
__global__ __launch_bounds__(256) void k_proj_dots3(size_t n, ProjSeg S, const double* __restrict__ Y,
                                                    const double* __restrict__ ax, const double* __restrict__ t,
                                                    const double* __restrict__ b, double* __restrict__ part) {
  double s[3 * M];
#pragma unroll
  for (int k = 0; k < 3 * M; ++k) s[k] = 0.;
  const SegRange R = seg_range(S);
  for (size_t i = R.i0; i < R.i1; i += R.step) {
    const double tv = t[i], bv = b[i], av = ax[i];
#pragma unroll
    for (int j = 0; j < M; ++j) {
      const double y = Y[(size_t)j * n + i];
      s[j] += y * tv;
      s[M + j] += y * bv;
      s[2 * M + j] += y * av;
    }
  }
  block_sums<3 * M>(s, part + (size_t)blockIdx.x * 3 * M);
}
// out[s * ostride + k] = sum of part[b * nv + k] over the blocks b of segment s (grid: nv x nseg),
// strided partial sums, then a fixed-order tree
__global__ __launch_bounds__(256) void k_proj_fin(ProjSeg S, int nv, const double* __restrict__ part,
                                                  double* __restrict__ out, int ostride) {
  const int k = blockIdx.x, sg = blockIdx.y;
  double s = 0.;
  for (int b = S.bbeg[sg] + threadIdx.x; b < S.bbeg[sg + 1]; b += 256) s += part[(size_t)b * nv + k];
  s = wsum(s);
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[(size_t)sg * ostride + k] = red[0] + red[1] + red[2] + red[3];
}
// out += sum_j c_j X_j, c of the block's segment at c + s cstride
template <int M>
__global__ __launch_bounds__(256) void k_proj_combine(size_t n, ProjSeg S, const double* __restrict__ X,
                                                      const double* __restrict__ c, int cstride,
                                                      double* __restrict__ out, double* __restrict__ v) {
  const SegRange R = seg_range(S);
  double cj[M];
#pragma unroll
  for (int j = 0; j < M; ++j) cj[j] = c[(size_t)R.s * cstride + j];
  if (v) {  // uniform: also v = -sum_j e_j X_j with e = c + M, from the same loads of X
    double ej[M];
#pragma unroll
    for (int j = 0; j < M; ++j) ej[j] = c[(size_t)R.s * cstride + M + j];
    for (size_t i = R.i0; i < R.i1; i += R.step) {
      double o = out[i], w = 0.;
#pragma unroll
      for (int j = 0; j < M; ++j) {
        const double xj = X[(size_t)j * n + i];
        o += cj[j] * xj;
        w -= ej[j] * xj;
      }
      out[i] = o;
      v[i] = w;
    }
    return;
  }
  for (size_t i = R.i0; i < R.i1; i += R.step) {
    double o = out[i];
#pragma unroll
    for (int j = 0; j < M; ++j) o += cj[j] * X[(size_t)j * n + i];
    out[i] = o;
  }
}
// The append of proj_ortho when k_proj_combine has left v = -sum_j e_j X_j before the solve (e
// for a basis of m vectors): x = psi + v (x holds v on entry), with n0's partials and wcol as
// k_proj_ortho writes them; reads 3 vectors instead of m + 2.
__global__ __launch_bounds__(256) void k_proj_append(size_t n, ProjSeg S, int m, int mr,
                                                     const double* __restrict__ e, int estride,
                                                     const double* __restrict__ psi, const double* __restrict__ b,
                                                     double* __restrict__ x, double* __restrict__ part,
                                                     double* __restrict__ wcol, int wstride, int ldw) {
  const SegRange R = seg_range(S);
  double s[1] = {0.};
  if (R.lb == 0 && (int)threadIdx.x < ldw) {
    const int t = threadIdx.x;
    wcol[(size_t)R.s * wstride + t] = t < m ? e[(size_t)R.s * estride + t] : t == m ? 1. : 0.;
  }
  for (size_t i = R.i0; i < R.i1; i += R.step) {
    const double p = psi[i], bv = b[i];
    x[i] = p + x[i];
    s[0] += (mr ? bv : p) * bv;
  }
  block_sums<1>(s, part + blockIdx.x);
}
// Gram-Schmidt step appending a new solution psi, with the coefficients e = G^-1 Y . b computed
// before its solve (per segment, at e + s estride): x = psi - sum_j e_j X_j, and the per-block
// partials of psi's squared norm (mr ? b . b : psi . b; the scale of the in-span test).  Its
// image As x is then applied exactly and k_proj_gcol forms its Gram column.  The first block of
// each segment writes psi's coefficients in the basis with x appended: wcol + s wstride =
// (e_0 .. e_{M-1}, 1, 0 ..), ldw values.
template <int M>
__global__ __launch_bounds__(256) void k_proj_ortho(size_t n, ProjSeg S, int mr, const double* __restrict__ X,
                                                    const double* __restrict__ e, int estride,
                                                    const double* __restrict__ psi, const double* __restrict__ b,
                                                    double* __restrict__ x, double* __restrict__ part,
                                                    double* __restrict__ wcol, int wstride, int ldw) {
  const SegRange R = seg_range(S);
  double ej[M > 0 ? M : 1], s[1] = {0.};
#pragma unroll
  for (int j = 0; j < M; ++j) ej[j] = e[(size_t)R.s * estride + j];
  if (R.lb == 0 && (int)threadIdx.x < ldw) {
    const int t = threadIdx.x;
    wcol[(size_t)R.s * wstride + t] = t < M ? e[(size_t)R.s * estride + t] : t == M ? 1. : 0.;
  }
  for (size_t i = R.i0; i < R.i1; i += R.step) {
    const double p = psi[i], bv = b[i];
    double xv = p;
#pragma unroll
    for (int j = 0; j < M; ++j) xv -= ej[j] * X[(size_t)j * n + i];
    x[i] = xv;
    s[0] += (mr ? bv : p) * bv;
  }
  block_sums<1>(s, part + blockIdx.x);
}
// Gram column of an appended vector x from its image ax = As x: Y_i . ax (i < M), then
// (mr ? ax : x) . ax; per-block partials.
template <int M>
__global__ __launch_bounds__(256) void k_proj_gcol(size_t n, ProjSeg S, int mr, const double* __restrict__ Y,
                                                   const double* __restrict__ x, const double* __restrict__ ax,
                                                   double* __restrict__ part) {
  const SegRange R = seg_range(S);
  double s[M + 1];
#pragma unroll
  for (int k = 0; k < M + 1; ++k) s[k] = 0.;
  for (size_t i = R.i0; i < R.i1; i += R.step) {
    const double a = ax[i];
#pragma unroll
    for (int j = 0; j < M; ++j) s[j] += Y[(size_t)j * n + i] * a;
    s[M] += (mr ? a : x[i]) * a;
  }
  block_sums<M + 1>(s, part + (size_t)blockIdx.x * (M + 1));
}
// Restart of a full basis of M vectors: slots k < kk become sum_j C[k M + j] X_j (C of the
// block's segment at C + s cstride), and the same for AX.  In place: each dof's M values are
// read before the same thread writes its kk values.
template <int M>
__global__ __launch_bounds__(256) void k_proj_reseed(size_t n, ProjSeg S, int kk, const double* __restrict__ C,
                                                     int cstride, double* X, double* AX) {
  const SegRange R = seg_range(S);
  __shared__ double Cs[kProjMax * kProjMax];
  for (int t = threadIdx.x; t < kk * M; t += blockDim.x) Cs[t] = C[(size_t)R.s * cstride + t];
  __syncthreads();
  for (size_t i = R.i0; i < R.i1; i += R.step) {
    double xv[M], av[M];
#pragma unroll
    for (int j = 0; j < M; ++j) {
      xv[j] = X[(size_t)j * n + i];
      av[j] = AX[(size_t)j * n + i];
    }
    for (int k = 0; k < kk; ++k) {
      double sx = 0., sa = 0.;
#pragma unroll
      for (int j = 0; j < M; ++j) {
        sx += Cs[k * M + j] * xv[j];
        sa += Cs[k * M + j] * av[j];
      }
      X[(size_t)k * n + i] = sx;
      AX[(size_t)k * n + i] = sa;
    }
  }
}
// The small system of a basis of m vectors, one per segment (grid: nseg; G + s gstride,
// v + s vstride): column j of G (at G + j ldg) holds G_ij (i < j), G_jj, and in row j + 1 the
// squared norm of the solution the vector came from.  A vector with G_jj <= 1e-24 of that norm
// (left in the span of the others), or whose Cholesky pivot falls below 1e-8 G_jj, is left out.
// v[0, m) and v[m, 2m) become G^-1 v on the kept vectors and 0 on the others: Cholesky by one
// wavefront, lane i holding row i of L in registers (m <= kProjMax <= 64); each column of L is
// broadcast by lane shuffles, and L goes to LDS once for the back substitution, which reads it
// by columns.
__global__ __launch_bounds__(64) void k_proj_solve(int m, const double* __restrict__ Gall, int ldg, int gstride,
                                                   double* __restrict__ vall, int vstride) {
  constexpr int MX = kProjMax;
  static_assert(MX <= 64, "one lane per row");
  const double* G = Gall + (size_t)blockIdx.x * gstride;
  double* v = vall + (size_t)blockIdx.x * vstride;
  __shared__ double Ls[MX][MX + 1];
  const int i = threadIdx.x;
  const bool row = i < m;
  double a[MX];  // row i of the lower triangle: G_ik (k <= i), the upper half of column i
#pragma unroll
  for (int k = 0; k < MX; ++k) a[k] = (row && k <= i) ? G[(size_t)i * ldg + k] : 0.;
  const double gd = row ? G[(size_t)i * ldg + i] : 0.;
  int ok = row && gd > 1e-24 * fabs(G[(size_t)i * ldg + i + 1]) && gd > 0.;
  double y0 = row ? v[i] : 0., y1 = row ? v[m + i] : 0.;
#pragma unroll
  for (int j = 0; j < MX; ++j) {  // right-looking: column j, then the trailing rows
    const double d = __shfl(a[j], j);
    const bool okj = j < m && __shfl(ok, j) && d > 1e-8 * __shfl(gd, j);  // the same on every lane
    if (okj) {
      const double l = sqrt(d);
      if (i == j) a[j] = l;
      else if (i > j && ok) a[j] /= l;
#pragma unroll
      for (int k = j + 1; k < MX; ++k) {
        const double lkj = __shfl(a[j], k);
        const int okk = __shfl(ok, k);
        if (k < m && k <= i && i > j && ok && okk) a[k] -= a[j] * lkj;
      }
    } else if (i == j) {
      ok = 0;
    }
  }
#pragma unroll
  for (int j = 0; j < MX; ++j) {  // L y = v
    const double ljj = __shfl(a[j], j);
    const double y0j = __shfl(y0, j) / ljj, y1j = __shfl(y1, j) / ljj;
    if (j < m && __shfl(ok, j)) {
      if (i == j) {
        y0 = y0j;
        y1 = y1j;
      } else if (i > j && ok) {
        y0 -= a[j] * y0j;
        y1 -= a[j] * y1j;
      }
    }
  }
  if (row) {
#pragma unroll
    for (int k = 0; k < MX; ++k) Ls[i][k] = a[k];
  }
  __syncthreads();
  for (int j = m - 1; j >= 0; --j) {  // L^T x = y
    if (!__shfl(ok, j)) continue;
    const double ljj = Ls[j][j];
    const double x0 = __shfl(y0, j) / ljj, x1 = __shfl(y1, j) / ljj;
    if (i == j) {
      y0 = x0;
      y1 = x1;
    } else if (i < j && ok) {
      const double lji = Ls[j][i];
      y0 -= lji * x0;
      y1 -= lji * x1;
    }
  }
  if (row) {
    v[i] = ok ? y0 : 0.;
    v[m + i] = ok ? y1 : 0.;
  }
}
// Gram matrix G_ij = Y_i . AX_j of a basis of M vectors over each segment: per block, chunks of
// 32 dofs of both sets staged in LDS, thread t summing pairs t, t + 256, ..; per-block partials
// part[block][M * M].
template <int M>
__global__ __launch_bounds__(256) void k_proj_gram(size_t n, ProjSeg S, const double* __restrict__ Y,
                                                   const double* __restrict__ AX, double* __restrict__ part) {
  constexpr int CH = 32, NP = (M * M + 255) / 256;
  __shared__ double ys[M][CH + 1], as[M][CH + 1];  // (+1: rows on distinct banks)
  const int t = threadIdx.x;
  const int sg = S.blk_seg[blockIdx.x];
  const int lb = blockIdx.x - S.bbeg[sg], nbs = S.bbeg[sg + 1] - S.bbeg[sg];
  const size_t d0 = S.sbeg[sg], d1 = S.sbeg[sg + 1];
  double s[NP];
#pragma unroll
  for (int k = 0; k < NP; ++k) s[k] = 0.;
  for (size_t c0 = d0 + (size_t)lb * CH; c0 < d1; c0 += (size_t)nbs * CH) {
    for (int q = t; q < M * CH; q += 256) {
      const int v = q / CH, d = q % CH;
      const size_t i = c0 + d;
      ys[v][d] = i < d1 ? Y[(size_t)v * n + i] : 0.;
      as[v][d] = i < d1 ? AX[(size_t)v * n + i] : 0.;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      const int pr = t + 256 * k;
      if (pr < M * M) {
        const int pi = pr / M, pj = pr % M;
        for (int d = 0; d < CH; ++d) s[k] += ys[pi][d] * as[pj][d];
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < NP; ++k)
    if (t + 256 * k < M * M) part[(size_t)blockIdx.x * M * M + t + 256 * k] = s[k];
}
void proj_dots2(hipStream_t s, size_t n, const ProjSeg& S, int m, const double* Y, const double* t, const double* b,
                double* part, double* out, int ostride) {
  if (m <= 0 || m > kProjMax || n == 0) return;
  proj_dispatch<1>(m, [&](auto Mc) {
    constexpr int M = decltype(Mc)::value;
    hipLaunchKernelGGL(k_proj_dots2<M>, dim3(S.nb), dim3(256), 0, s, n, S, Y, t, b, part);
  });
  hipLaunchKernelGGL(k_proj_fin, dim3(2 * m, S.nseg), dim3(256), 0, s, S, 2 * m, part, out, ostride);
}
void proj_dots3(hipStream_t s, size_t n, const ProjSeg& S, int m, const double* Y, const double* ax, const double* t,
                const double* b, double* part, double* out, int ostride, double* gcol, int gstride) {
  if (m <= 0 || m > kProjMax || n == 0) return;
  proj_dispatch<1>(m, [&](auto Mc) {
    constexpr int M = decltype(Mc)::value;
    hipLaunchKernelGGL(k_proj_dots3<M>, dim3(S.nb), dim3(256), 0, s, n, S, Y, ax, t, b, part);
  });
  hipLaunchKernelGGL(k_proj_fin, dim3(2 * m, S.nseg), dim3(256), 0, s, S, 3 * m, part, out, ostride);
  hipLaunchKernelGGL(k_proj_fin, dim3(m, S.nseg), dim3(256), 0, s, S, 3 * m, part + 2 * m, gcol, gstride);
}
void proj_solve(hipStream_t s, const ProjSeg& S, int m, const double* G, int ldg, int gstride, double* v,
                int vstride) {
  if (m <= 0 || m > kProjMax) return;
  hipLaunchKernelGGL(k_proj_solve, dim3(S.nseg), dim3(64), 0, s, m, G, ldg, gstride, v, vstride);
}
void proj_combine(hipStream_t s, size_t n, const ProjSeg& S, int m, const double* X, const double* c, int cstride,
                  double* out, double* v) {
  if (m <= 0 || m > kProjMax || n == 0 || (v && 2 * m > cstride)) return;
  proj_dispatch<1>(m, [&](auto Mc) {
    constexpr int M = decltype(Mc)::value;
    hipLaunchKernelGGL(k_proj_combine<M>, dim3(S.nb), dim3(256), 0, s, n, S, X, c, cstride, out, v);
  });
}
void proj_append(hipStream_t s, size_t n, const ProjSeg& S, int m, int mr, const double* e, int estride,
                 const double* psi, const double* b, double* x, double* part, double* n0, int n0stride,
                 double* wcol, int wstride, int ldw) {
  if (m < 0 || m > kProjMax || n == 0 || ldw > 256) return;
  hipLaunchKernelGGL(k_proj_append, dim3(S.nb), dim3(256), 0, s, n, S, m, mr, e, estride, psi, b, x, part, wcol,
                     wstride, ldw);
  hipLaunchKernelGGL(k_proj_fin, dim3(1, S.nseg), dim3(256), 0, s, S, 1, part, n0, n0stride);
}
void proj_ortho(hipStream_t s, size_t n, const ProjSeg& S, int m, int mr, const double* X, const double* e,
                int estride, const double* psi, const double* b, double* x, double* part, double* n0, int n0stride,
                double* wcol, int wstride, int ldw) {
  if (m < 0 || m > kProjMax || n == 0 || ldw > 256) return;
  proj_dispatch<0>(m, [&](auto Mc) {
    constexpr int M = decltype(Mc)::value;
    hipLaunchKernelGGL(k_proj_ortho<M>, dim3(S.nb), dim3(256), 0, s, n, S, mr, X, e, estride, psi, b, x, part, wcol,
                       wstride, ldw);
  });
  hipLaunchKernelGGL(k_proj_fin, dim3(1, S.nseg), dim3(256), 0, s, S, 1, part, n0, n0stride);
}
void proj_gcol(hipStream_t s, size_t n, const ProjSeg& S, int m, int mr, const double* Y, const double* x,
               const double* ax, double* part, double* gcol, int gstride) {
  if (m < 0 || m > kProjMax || n == 0) return;
  proj_dispatch<0>(m, [&](auto Mc) {
    constexpr int M = decltype(Mc)::value;
    hipLaunchKernelGGL(k_proj_gcol<M>, dim3(S.nb), dim3(256), 0, s, n, S, mr, Y, x, ax, part);
    hipLaunchKernelGGL(k_proj_fin, dim3(M + 1, S.nseg), dim3(256), 0, s, S, M + 1, part, gcol, gstride);
  });
}
void proj_gram(hipStream_t s, size_t n, const ProjSeg& S, int m, const double* Y, const double* AX, double* part,
               double* out) {
  if (m <= 0 || m > kProjMax || n == 0) return;
  proj_dispatch<1>(m, [&](auto Mc) {
    constexpr int M = decltype(Mc)::value;
    hipLaunchKernelGGL(k_proj_gram<M>, dim3(S.nb), dim3(256), 0, s, n, S, Y, AX, part);
  });
  hipLaunchKernelGGL(k_proj_fin, dim3(m * m, S.nseg), dim3(256), 0, s, S, m * m, part, out, m * m);
}
void proj_reseed(hipStream_t s, size_t n, const ProjSeg& S, int m, int kk, const double* C, int cstride, double* X,
                 double* AX) {
  if (m <= 0 || m > kProjMax || kk <= 0 || kk > m || n == 0) return;
  proj_dispatch<1>(m, [&](auto Mc) {
    constexpr int M = decltype(Mc)::value;
    hipLaunchKernelGGL(k_proj_reseed<M>, dim3(S.nb), dim3(256), 0, s, n, S, kk, C, cstride, X, AX);
  });
}
// The per-segment sum of one partial per block (k_proj_fin with nv = 1), leaving out[s] as it is
// for a segment whose flag in iact (may be null) is 0.
__global__ __launch_bounds__(256) void k_seg_fin(ProjSeg S, const double* __restrict__ part, double* __restrict__ out,
                                                 const int* __restrict__ iact) {
  const int sg = blockIdx.y;
  if (iact && !iact[sg]) return;
  double s = 0.;
  for (int b = S.bbeg[sg] + threadIdx.x; b < S.bbeg[sg + 1]; b += 256) s += part[b];
  s = wsum(s);
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[sg] = red[0] + red[1] + red[2] + red[3];
}
__global__ __launch_bounds__(256) void k_seg_norm2(ProjSeg S, VecArgs A, const double* __restrict__ r,
                                                   const int* __restrict__ F, double* __restrict__ part,
                                                   const int* __restrict__ iact) {
  if (F && !F[F_ACTIVE]) return;
  const SegRange R = seg_range(S);
  if (iact && !iact[R.s]) return;  // a stopped instance: its sum is left as it was
  DofSlot ds(A, R.i0, R.step);
  double p[1] = {0.};
  for (size_t v = R.i0; v < R.i1; v += R.step, ds.next()) {
    const double u = r[v] * A.winv[ds.m];
    p[0] += u * u;
  }
  block_sums<1>(p, part + blockIdx.x);
}
void seg_norm2(hipStream_t s, const ProjSeg& S, const VecArgs& A, const double* r, const int* F, double* part,
               double* out, const int* iact) {
  hipLaunchKernelGGL(k_seg_norm2, dim3(S.nb), dim3(256), 0, s, S, A, r, F, part, iact);
  hipLaunchKernelGGL(k_seg_fin, dim3(1, S.nseg), dim3(256), 0, s, S, part, out, iact);
}
template <class RB>
__global__ __launch_bounds__(256) void k_seg_dot(ProjSeg S, const double* __restrict__ a, const RB* __restrict__ b,
                                                 const int* __restrict__ F, double* __restrict__ part,
                                                 const int* __restrict__ iact) {
  if (F && !F[F_ACTIVE]) return;
  const SegRange R = seg_range(S);
  if (iact && !iact[R.s]) return;
  double p[1] = {0.};
  for (size_t v = R.i0; v < R.i1; v += R.step) p[0] += a[v] * (double)b[v];
  block_sums<1>(p, part + blockIdx.x);
}
template <class RB>
void seg_dot(hipStream_t s, const ProjSeg& S, const double* a, const RB* b, const int* F, double* part, double* out,
             const int* iact) {
  hipLaunchKernelGGL(k_seg_dot<RB>, dim3(S.nb), dim3(256), 0, s, S, a, b, F, part, iact);
  hipLaunchKernelGGL(k_seg_fin, dim3(1, S.nseg), dim3(256), 0, s, S, part, out, iact);
}
template void seg_dot<double>(hipStream_t, const ProjSeg&, const double*, const double*, const int*, double*, double*,
                              const int*);
template void seg_dot<float>(hipStream_t, const ProjSeg&, const double*, const float*, const int*, double*, double*,
                             const int*);

// One block: thread t handles instances t, t + 256, ..; the block then reduces "any active", the
// largest residual ratio of the active instances and the sum of the residual norms.
__global__ __launch_bounds__(256) void k_finalize_ens(int phase, EnsKry E, double* __restrict__ S,
                                                      int* __restrict__ F, int maxiter, int* __restrict__ mirror,
                                                      int seq) {
  const bool live = phase == ENS_INIT || F[F_ACTIVE];
  __shared__ int sany[256];
  __shared__ double smax[256], ssum[256];
  int any = 0;
  double rmax = 0., rsum = 0.;
  if (live) {
    const double tol2 = S[S_TOL2];
    double bmax = 0.;
    if (phase == ENS_INIT || phase == ENS_UPD)
      for (int i = 0; i < E.n; ++i) bmax = fmax(bmax, E.bb[i]);
    for (int i = threadIdx.x; i < E.n; i += blockDim.x) {
      const double v = E.in[i];
      int a = E.act[i];
      switch (phase) {
        case ENS_INIT: {
          const double d = tol2 * (E.bb[i] > 0. ? E.bb[i] : bmax);
          a = F[F_ACTIVE] && (d > 0. ? v > d : v > 0.);
          E.rz[i] = 0.;
          E.alpha[i] = 0.;
          E.beta[i] = 0.;
        } break;
        case ENS_RZ: {
          double be = 0.;
          if (a) {
            if (!(v > 0.) || !isfinite(v)) a = 0;
            else { be = F[F_ITER] > 0 ? v / E.rz[i] : 0.; E.rz[i] = v; }
          }
          E.beta[i] = a ? be : 0.;
        } break;
        case ENS_PQ: {
          double al = 0.;
          if (a) {
            al = E.rz[i] / v;
            if (!(v > 0.) || !isfinite(al)) { a = 0; al = 0.; }
          }
          E.alpha[i] = al;
        } break;
        case ENS_UPD: {
          const double d = tol2 * (E.bb[i] > 0. ? E.bb[i] : bmax);
          const double ratio = d > 0. ? v / d : (v > 0. ? HUGE_VAL : 0.);
          if (a) a = ratio > 1. && isfinite(v);
          if (a) rmax = fmax(rmax, ratio);
          rsum += v;
          if (!a) E.alpha[i] = 0.;
        } break;
      }
      E.act[i] = a;
      any |= a;
    }
  }
  sany[threadIdx.x] = any;
  smax[threadIdx.x] = rmax;
  ssum[threadIdx.x] = rsum;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sany[threadIdx.x] |= sany[threadIdx.x + w];
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
      ssum[threadIdx.x] += ssum[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  if (live) {
    if (phase == ENS_UPD) {
      F[F_ITER] += 1;
      S[S_RR] = ssum[0];
      S[S_IMAX] = smax[0];
      F[F_ACTIVE] = (sany[0] && F[F_ITER] < (maxiter >= 0 ? maxiter : F[F_MAXIT])) ? 1 : 0;
    } else if (phase == ENS_INIT) {
      F[F_ACTIVE] = F[F_ACTIVE] && sany[0];
    } else {
      F[F_ACTIVE] = sany[0] ? 1 : 0;
    }
  }
  if (mirror) {
    const int sq = seq >= 0 ? seq : ++F[F_SEQ];
    const int act = F[F_ACTIVE] ? 1 : 0;
    const int near = act && S[S_IMAX] <= 64.;
    __hip_atomic_store(mirror, sq * 4 + near * 2 + act, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}
void finalize_ens(hipStream_t s, int phase, const EnsKry& E, double* S, int* F, int maxiter, int* mirror, int seq) {
  hipLaunchKernelGGL(k_finalize_ens, dim3(1), dim3(256), 0, s, phase, E, S, F, maxiter, mirror, seq);
}
void psi_predict(hipStream_t s, size_t n, double f, int extrap, double* psi, double* prev) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_psi_predict, dim3(vec_grid(n)), dim3(256), 0, s, n, f, extrap, psi, prev);
}
template <class R>
void zero_if_flag(hipStream_t s, size_t n, R* x, const int* F) {
  hipLaunchKernelGGL(k_zero_if_flag<R>, dim3(vec_grid(n)), dim3(256), 0, s, n, x, F);
}
void set_int(hipStream_t s, int* p, int v) { hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, s, p, v); }
void finalize(hipStream_t s, int phase, const double* part, int np, double* S, int* F, int maxiter, int* mirror,
              int seq, const double* irr, const double* ibb, int ninst) {
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, s, phase, part, np, S, F, maxiter, mirror, seq, irr, ibb,
                     ninst);
}
template <class R>
void csr(hipStream_t s, int mode, int n, const int32_t* ptr, const int32_t* col, const R* val, const R* x, const R* f,
         R* y, const R* M, R* xio, const int* F, int lpr) {
  const size_t threads = (size_t)n * lpr;
  const dim3 grid((unsigned)((threads + 255) / 256));
  switch (lpr) {
    case 1: hipLaunchKernelGGL(k_csr<R>, dim3((n + 255) / 256), dim3(256), 0, s, mode, n, ptr, col, val, x, f, y, M, xio, F); break;
    case 2: hipLaunchKernelGGL((k_csr_vec<2, R>), grid, dim3(256), 0, s, mode, n, ptr, col, val, x, f, y, M, xio, F); break;
    case 4: hipLaunchKernelGGL((k_csr_vec<4, R>), grid, dim3(256), 0, s, mode, n, ptr, col, val, x, f, y, M, xio, F); break;
    case 8: hipLaunchKernelGGL((k_csr_vec<8, R>), grid, dim3(256), 0, s, mode, n, ptr, col, val, x, f, y, M, xio, F); break;
    default: hipLaunchKernelGGL((k_csr_vec<16, R>), grid, dim3(256), 0, s, mode, n, ptr, col, val, x, f, y, M, xio, F); break;
  }
}
template <class R>
void ell(hipStream_t s, int mode, int n, int w, int lanes, const int32_t* col, const R* val, const R* x, const R* f,
         R* y, const R* M, R* xio, const int* F) {
  if (lanes == 4)
    hipLaunchKernelGGL((k_ell<4, R>), dim3((unsigned)(((size_t)n * 4 + 255) / 256)), dim3(256), 0, s, mode, n, w, col,
                       val, x, f, y, M, xio, F);
  else
    hipLaunchKernelGGL((k_ell<1, R>), dim3((n + 255) / 256), dim3(256), 0, s, mode, n, w, col, val, x, f, y, M, xio, F);
}
int csr_lanes(int n, size_t nnz) {
  const double avg = n ? (double)nnz / n : 0.;
  return avg >= 24 ? 16 : avg >= 12 ? 8 : avg >= 6 ? 4 : avg >= 3 ? 2 : 1;
}
template <class R>
void scale(hipStream_t s, int n, const R* M, const R* f, R* x, const int* F) {
  hipLaunchKernelGGL(k_scale<R>, dim3((n + 255) / 256), dim3(256), 0, s, n, M, f, x, F);
}
template <class R>
void dense(hipStream_t s, int n, const R* Ainv, const R* f, R* x, const int* F) {
  hipLaunchKernelGGL(k_dense<R>, dim3((n + 3) / 4), dim3(256), 0, s, n, Ainv, f, x, F);
}
template <class R>
void q1_gather(hipStream_t s, int nv, const int32_t* vptr, const int32_t* vinc, const R* ec, R* rc, const int* F) {
  hipLaunchKernelGGL(k_q1_gather<R>, dim3((nv + 255) / 256), dim3(256), 0, s, nv, vptr, vinc, ec, rc, F);
}
void demote(hipStream_t s, size_t n, const double* r64, float* r32, const int* F) {
  hipLaunchKernelGGL(k_demote, dim3(vec_grid(n)), dim3(256), 0, s, n, r64, r32, F);
}
void promote_dot(hipStream_t s, size_t n, const float* z32, double* z64, const double* r64, double* part,
                 const int* F) {
  hipLaunchKernelGGL(k_promote_dot, dim3(vec_grid(n)), dim3(256), 0, s, n, z32, z64, r64, part, F);
}
void inst_norms(hipStream_t s, const VecArgs& A, int n_inst, const int32_t* off, int K2, const double* b, double* out) {
  hipLaunchKernelGGL(k_inst_norms, dim3(n_inst), dim3(256), 0, s, A, off, K2, b, out);
}
void unweight(hipStream_t s, size_t n, int K, const double* w, double* y) {
  Weights6 W{};
  for (int i = 0; i < K && i < 6; ++i) W.w[i] = w[i];
  if (n) hipLaunchKernelGGL(k_unweight, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, K, W, y);
}
void probe_set(hipStream_t s, int N, int K2, const int32_t* color, int c, int rs, double* x) {
  size_t n = (size_t)N * K2;
  hipLaunchKernelGGL(k_probe_set, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, N, K2, color, c, rs, x);
}
void probe_diag(hipStream_t s, int N, int K2, const int32_t* color, int c, int rs, const double* y, double* diag) {
  hipLaunchKernelGGL(k_probe_diag, dim3((N + 255) / 256), dim3(256), 0, s, N, K2, color, c, rs, y, diag);
}
void q1_probe_accum(hipStream_t s, int N, int K, const Tab& T, int rs, const double* y, double* M) {
  hipLaunchKernelGGL(k_q1_probe_accum, dim3((N + 255) / 256), dim3(256), 0, s, N, K, T, rs, y, M);
}

#define HOC_LA_INST(R)                                                                                                 \
  template void vec<R>(hipStream_t, int, const VecArgs&, size_t, R*, R*, const R*, const R*, R*, const double*,        \
                       const int*, double*);                                                                           \
  template void bicg_half<R>(hipStream_t, size_t, R*, const R*, const double*, int*);                                 \
  template void zero_if_flag<R>(hipStream_t, size_t, R*, const int*);                                                 \
  template void csr<R>(hipStream_t, int, int, const int32_t*, const int32_t*, const R*, const R*, const R*, R*,        \
                       const R*, R*, const int*, int);                                                                 \
  template void ell<R>(hipStream_t, int, int, int, int, const int32_t*, const R*, const R*, const R*, R*, const R*,  \
                       R*, const int*);                                                                        \
  template void scale<R>(hipStream_t, int, const R*, const R*, R*, const int*);                                       \
  template void dense<R>(hipStream_t, int, const R*, const R*, R*, const int*);                                       \
  template void q1_gather<R>(hipStream_t, int, const int32_t*, const int32_t*, const R*, R*, const int*);
HOC_LA_INST(double)
HOC_LA_INST(float)
#undef HOC_LA_INST

}  // namespace hoc
