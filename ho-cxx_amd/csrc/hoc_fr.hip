// hoc_fr.hip - flux-reconstruction stage kernels for gfx950.
//
// Second-order FR operator (the Poisson matrix of src/HO_2D.cpp:1533-1804 applied
// matrix-free, and the diffusion term of src/HO_2D.cpp:2666-2831), velocity from the
// stream function (src/HO_2D.cpp:3174-3366), advection (src/HO_2D.cpp:2245-2645) and
// the RK stage update (src/HO_2D.cpp:1465-1531, 1833-1836).
#include "hoc_fr.hpp"

namespace hoc {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ int rev_flag(int ijp, int ijpm) { return (ijp + ijpm - ijp / 2 - ijpm / 2 + 1) & 1; }
__device__ __forceinline__ int face_sign_i(int ijp, int ijpm) { return ((ijp + ijpm) & 1) ? 1 : -1; }
__device__ __forceinline__ double face_sign(int ijp, int ijpm) { return (double)face_sign_i(ijp, ijpm); }
// element-local flux point k of face ijp -> boundary-edge flux point (src/HO_2D.cpp:2229)
template <int K>
__device__ __forceinline__ int edge_fp(int ijp, int k) { return (ijp == 0 || ijp == 3) ? K - 1 - k : k; }

// Value and normal-derivative traces on the four faces (src/HO_2D.cpp:2695-2713).
template <int K, class R>
__device__ __forceinline__ void traces(const TabT<R>& T, const R (&u)[K * K], R (&bv)[4][K], R (&bg)[4][K]) {
#pragma unroll
  for (int rc = 0; rc < K; ++rc)
#pragma unroll
    for (int ib = 0; ib < 2; ++ib) {
      R s = 0, g = 0, s2 = 0, g2 = 0;
#pragma unroll
      for (int m = 0; m < K; ++m) {
        s += u[rc * K + m] * T.bb[m][ib];
        g += u[rc * K + m] * T.bgb[m][ib];
        s2 += u[m * K + rc] * T.bb[m][ib];
        g2 += u[m * K + rc] * T.bgb[m][ib];
      }
      bv[ib][rc] = s; bg[ib][rc] = g;
      bv[2 + ib][rc] = s2; bg[2 + ib][rc] = g2;
    }
}

// Face traces are stored face-major: plane f holds [N][2][K] (values, then normal
// derivatives) of every element's face f, so a wavefront that reads face f' of its
// neighbours (adjacent elements) reads contiguous memory.
template <int K, class R>
__device__ __forceinline__ void store_traces(R* __restrict__ tr, int N, int el, const R (&bv)[4][K], const R (&bg)[4][K]) {
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    R* t = tr + ((size_t)f * N + el) * 2 * K;
#pragma unroll
    for (int k = 0; k < K; ++k) { t[k] = bv[f][k]; t[K + k] = bg[f][k]; }
  }
}

// Common value and common normal flux on an interior face (src/HO_2D.cpp:2833-2870 applied
// to both sides, then the average of :2742-2768).
template <int K, class R>
__device__ __forceinline__ void interior_face(const TabT<R>& T, int ib, int ibn, int rev, R sgn,
                                              const R (&bv)[K], const R (&bg)[K],
                                              const R (&vn)[K], const R (&gn)[K],
                                              const R (&fa)[K], const R (&fb)[K],
                                              const R (&fan)[K], const R (&fbn)[K],
                                              R (&cv)[K], R (&cg)[K]) {
  R com[K], comn[K], F[K], Fn[K];
#pragma unroll
  for (int k = 0; k < K; ++k) com[k] = R(0.5) * (bv[k] + (rev ? vn[K - 1 - k] : vn[k]));
#pragma unroll
  for (int m = 0; m < K; ++m) comn[m] = rev ? com[K - 1 - m] : com[m];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    R cr = 0, crn = 0;
#pragma unroll
    for (int d = 0; d < K; ++d) { cr += com[d] * T.D[d][k]; crn += comn[d] * T.D[d][k]; }
    const R dv = bg[k] + (com[k] - bv[k]) * T.gp[ib];
    const R dvn = gn[k] + (comn[k] - vn[k]) * T.gp[ibn];
    F[k] = dv * fa[k] - cr * fb[k];
    Fn[k] = dvn * fan[k] - crn * fbn[k];
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    cv[k] = com[k];
    cg[k] = R(0.5) * (F[k] + sgn * (rev ? Fn[K - 1 - k] : Fn[k]));
  }
}

// Boundary face (src/HO_2D.cpp:2872-2927).  bc: edge values (edge flux-point order) or
// nullptr for zero values; bscale multiplies the Neumann values.
template <int K, class R>
__device__ __forceinline__ void boundary_face(const TabT<R>& T, const GeoT<R>& G, int el, int ijp, int sw,
                                              const R* __restrict__ bc, R bscale,
                                              const R (&bv)[K], const R (&bg)[K],
                                              const R (&fa)[K], const R (&fb)[K],
                                              R (&cv)[K], R (&cg)[K]) {
  const int ib = ijp & 1;
  if (sw == 0) {  // Dirichlet: common value = BC
    R com[K];
#pragma unroll
    for (int m = 0; m < K; ++m) com[m] = bc ? bc[edge_fp<K>(ijp, m)] : R(0);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      R cr = 0;
#pragma unroll
      for (int d = 0; d < K; ++d) cr += com[d] * T.D[d][k];
      const R dv = bg[k] + (com[k] - bv[k]) * T.gp[ib];
      cv[k] = com[k];
      cg[k] = dv * fa[k] - cr * fb[k];
    }
  } else {  // Neumann: common flux = BC * |g|, common value from the K x K system
    const size_t q = ((size_t)el * 4 + ijp) * K;
    R M[K][K], r[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const R A = G.fA[q + k], B = G.fB[q + k], J = G.fjac[q + k];
      cg[k] = (bc ? bscale * bc[edge_fp<K>(ijp, k)] : R(0)) * G.fAn[q + k];
      r[k] = J * cg[k] - A * (bg[k] - bv[k] * T.gp[ib]);
#pragma unroll
      for (int i = 0; i < K; ++i) M[k][i] = -B * T.D[i][k];
      M[k][k] += A * T.gp[ib];
    }
    // Gauss elimination without pivoting (src/MathHelpers.hpp:53-74)
#pragma unroll
    for (int k = 0; k < K - 1; ++k)
#pragma unroll
      for (int i = k + 1; i < K; ++i) {
        const R c = M[i][k] / M[k][k];
        r[i] -= c * r[k];
#pragma unroll
        for (int j = k + 1; j < K; ++j) M[i][j] -= c * M[k][j];
      }
#pragma unroll
    for (int i = K - 1; i >= 0; --i) {
      R c = 0;
#pragma unroll
      for (int j = i + 1; j < K; ++j) c += M[i][j] * cv[j];
      cv[i] = (r[i] - c) / M[i][i];
    }
  }
}

// Gathers one face's neighbour trace pair (value, derivative) and neighbour metrics.
template <int K, bool AFF = false, class R>
__device__ __forceinline__ void load_nbr(const GeoT<R>& G, const R* __restrict__ tr, int nb, int fn,
                                         R (&vn)[K], R (&gn)[K], R (&fan)[K], R (&fbn)[K]) {
  const R* t = tr + ((size_t)fn * G.N + nb) * 2 * K;
  if (AFF) {
    const R a = G.gface[((size_t)nb * 4 + fn) * 2], b = G.gface[((size_t)nb * 4 + fn) * 2 + 1];
#pragma unroll
    for (int k = 0; k < K; ++k) { vn[k] = t[k]; gn[k] = t[K + k]; fan[k] = a; fbn[k] = b; }
  } else {
    const size_t q = ((size_t)nb * 4 + fn) * K;
#pragma unroll
    for (int k = 0; k < K; ++k) { vn[k] = t[k]; gn[k] = t[K + k]; fan[k] = G.fa[q + k]; fbn[k] = G.fb[q + k]; }
  }
}

// All four faces of the second-order operator.  mode: 0 Poisson (psi switches, bc values
// or zero), 1 diffusion (diffusion switches).
template <int K, bool AFF = false, class R>
__device__ __forceinline__ void faces2(const TabT<R>& T, const GeoT<R>& G, int el, const R* __restrict__ tr,
                                       const int8_t* __restrict__ sw, const R* __restrict__ bcv, bool signed_neumann,
                                       const R (&bv)[4][K], const R (&bg)[4][K],
                                       R (&cv)[4][K], R (&cg)[4][K]) {
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    const int nb = G.nb[(size_t)el * 4 + f];
    R fa[K], fb[K];
    if (AFF) {
      const R a = G.gface[((size_t)el * 4 + f) * 2], b = G.gface[((size_t)el * 4 + f) * 2 + 1];
#pragma unroll
      for (int k = 0; k < K; ++k) { fa[k] = a; fb[k] = b; }
    } else {
      const size_t q = ((size_t)el * 4 + f) * K;
#pragma unroll
      for (int k = 0; k < K; ++k) { fa[k] = G.fa[q + k]; fb[k] = G.fb[q + k]; }
    }
    if (nb >= 0) {
      const int fn = G.nbf[(size_t)el * 4 + f];
      R vn[K], gn[K], fan[K], fbn[K];
      load_nbr<K, AFF>(G, tr, nb, fn, vn, gn, fan, fbn);
      interior_face<K>(T, f & 1, fn & 1, rev_flag(f, fn), R(face_sign_i(f, fn)), bv[f], bg[f], vn, gn, fa, fb, fan, fbn,
                       cv[f], cg[f]);
    } else {
      const int e = -1 - nb;
      const R* b = bcv ? bcv + (size_t)e * K : nullptr;
      const R sc = signed_neumann ? G.esign[e] : R(1);
      boundary_face<K>(T, G, el, f, sw[e], b, sc, bv[f], bg[f], fa, fb, cv[f], cg[f]);
    }
  }
}

// Contravariant corrected gradient (f~, g~) and its corrected divergence (src/HO_2D.cpp:2771-2816).
template <int K, class R>
__device__ __forceinline__ void second_order_core(const TabT<R>& T, const R (&u)[K * K], const R (&bv)[4][K],
                                                  const R (&cv)[4][K], const R (&cg)[4][K],
                                                  const R* ga, const R* gb, const R* gc,
                                                  R (&out)[K * K]) {
  R ft[K * K], gt[K * K];
#pragma unroll
  for (int ky = 0; ky < K; ++ky)
#pragma unroll
    for (int kx = 0; kx < K; ++kx) {
      R dx = 0, de = 0;
#pragma unroll
      for (int d = 0; d < K; ++d) { dx += u[ky * K + d] * T.D[d][kx]; de += u[d * K + kx] * T.D[d][ky]; }
      dx += (cv[0][ky] - bv[0][ky]) * T.grad[kx][0] + (cv[1][ky] - bv[1][ky]) * T.grad[kx][1];
      de += (cv[2][kx] - bv[2][kx]) * T.grad[ky][0] + (cv[3][kx] - bv[3][kx]) * T.grad[ky][1];
      const int v = ky * K + kx;
      ft[v] = dx * ga[v] - de * gb[v];
      gt[v] = de * gc[v] - dx * gb[v];
    }
  R fB[2][K], gB[2][K];
#pragma unroll
  for (int ib = 0; ib < 2; ++ib)
#pragma unroll
    for (int k = 0; k < K; ++k) {
      R a = cg[ib][k], b = cg[2 + ib][k];
#pragma unroll
      for (int d = 0; d < K; ++d) { a -= ft[k * K + d] * T.bb[d][ib]; b -= gt[d * K + k] * T.bb[d][ib]; }
      fB[ib][k] = a; gB[ib][k] = b;
    }
#pragma unroll
  for (int ky = 0; ky < K; ++ky)
#pragma unroll
    for (int kx = 0; kx < K; ++kx) {
      R r = 0;
#pragma unroll
      for (int d = 0; d < K; ++d) r += ft[ky * K + d] * T.D[d][kx];
      r += fB[0][ky] * T.grad[kx][0] + fB[1][ky] * T.grad[kx][1];
#pragma unroll
      for (int d = 0; d < K; ++d) r += gt[d * K + kx] * T.D[d][ky];
      r += gB[0][kx] * T.grad[ky][0] + gB[1][kx] * T.grad[ky][1];
      out[ky * K + kx] = r;
    }
}

// ==========================================================================================
// Trace kernel with fused vector pre-operations producing the traced field u:
//   PRE_COPY   u = x
//   PRE_CHEB0  u = r * dinv * coef (coef = 1/theta); writes x = d = u   (Chebyshev step from zero)
//   PRE_PROL   u = x + Pc xc (bilinear from the corner vertices); writes x = u
//   PRE_PUPD   u = z + beta * p; writes p = u                     (PCG direction; beta =
//              scal[scal_idx], or with scal_idx < 0 scal[einst[element]]: per ensemble instance)
// einst, iact (may be null): a block inside one ensemble instance whose flag in iact is 0
// returns at once.
// RA: the type of a (R, or float for the FP64 direction update of the mixed-precision solve,
// whose V-cycle leaves z in FP32).
// ==========================================================================================
template <int K, int PRE, class R, class RA = R>
__global__ __launch_bounds__(kWG) void k_trace(int N, TabT<R> T, R* __restrict__ x, const RA* __restrict__ a,
                                                const R* __restrict__ b, R* __restrict__ d, R coef,
                                                const double* __restrict__ scal, int scal_idx, const int* __restrict__ active,
                                                R* __restrict__ tr, const int32_t* __restrict__ ev,
                                                const int32_t* __restrict__ einst, const int* __restrict__ iact) {
  if (active && !*active) return;
  constexpr int K2 = K * K;
  __shared__ __align__(16) R lds[kWG * K2];  // one tile: two inputs are staged in turn
  const int e0 = blockIdx.x * kWG;
  if (iact) {  // a block inside one stopped ensemble instance
    const int i0 = einst[e0];
    if (i0 == einst[min(e0 + kWG, N) - 1] && !iact[i0]) return;
  }
  const int el = e0 + threadIdx.x;
  const bool live = el < N;
  R u[K2];
  R beta = 0;
  if (PRE == PRE_PUPD) beta = (R)(scal_idx >= 0 ? scal[scal_idx] : live ? scal[einst[el]] : 0.);
  R xc[4] = {0, 0, 0, 0};
  if (PRE == PRE_PROL && live)
#pragma unroll
    for (int c = 0; c < 4; ++c) xc[c] = (R)a[ev[(size_t)el * 4 + c]];
  RA* la = reinterpret_cast<RA*>(lds);  // sizeof(RA) <= sizeof(R): a fits the tile
  if constexpr (PRE == PRE_COPY || PRE == PRE_PROL) tile_load<K2>((const R*)x, lds, e0, N);
  else tile_load<K2>(a, la, e0, N);
#pragma unroll
  for (int m = 0; m < K2; ++m) {
    const int li = threadIdx.x * K2 + m;
    R v = (PRE == PRE_COPY || PRE == PRE_PROL) ? lds[li] : (R)la[li];
    if (PRE == PRE_PROL) {
      const int j = m / K, i = m % K;
      v += q1w(T, 0, j, i) * xc[0] + q1w(T, 1, j, i) * xc[1] + q1w(T, 2, j, i) * xc[2] + q1w(T, 3, j, i) * xc[3];
    }
    u[m] = v;
  }
  if (PRE == PRE_CHEB0 || PRE == PRE_PUPD) {  // second input
    __syncthreads();
    tile_load<K2>(PRE == PRE_CHEB0 ? b : (const R*)x, lds, e0, N);
#pragma unroll
    for (int m = 0; m < K2; ++m) {
      const R w = lds[threadIdx.x * K2 + m];
      u[m] = PRE == PRE_CHEB0 ? u[m] * w * coef : u[m] + beta * w;
    }
  }
  if (PRE == PRE_CHEB0 || PRE == PRE_PROL || PRE == PRE_PUPD) {
    __syncthreads();
#pragma unroll
    for (int m = 0; m < K2; ++m) lds[threadIdx.x * K2 + m] = u[m];
    if (PRE == PRE_CHEB0) { tile_store<K2>(x, lds, e0, N); tile_store<K2>(d, lds, e0, N); }
    else tile_store<K2>(x, lds, e0, N);
  }
  if (!live) return;
  R bv[4][K], bg[4][K];
  traces<K>(T, u, bv, bg);
  store_traces<K>(tr, N, el, bv, bg);
}

// ==========================================================================================
// Poisson operator As = -W * L (W = w_i w_j) applied matrix-free, with fused epilogues:
//   EPI_STORE     y = As x; partial sums of x.y                    (PCG / probing)
//   EPI_RESID     y = r - As x; partial sums of y.y and (y/W).(y/W) (initial residual)
//   EPI_CHEB      res = r - As x; e = c1 d + c2 res dinv; x += e; y = e unless y is null;
//                 optional partial r.x
//   EPI_CHEBF     res = r - As x; e = c2 res dinv (c2 = 1/theta); x += e; y = e unless null
//                 (first step from x != 0)
//   (dinv = 1/diag(As), the Jacobi scaling of the Chebyshev smoother; d and y are distinct
//   direction buffers, and the last step of a sweep passes y = null)
//   EPI_RESTRICT  ec[el][a] = sum_ij q1w(a,ij) (r - As x)_ij        (Pc^T residual, per element)
// ==========================================================================================
template <int K, int EPI, bool AFF, class R>
// Affine FP32 instances with K <= 5 are held to 128 VGPRs (4 waves per SIMD; their LDS tile
// allows more), which they fit without spilling; the others keep the compiler's allocation.
__global__ __launch_bounds__(kWG, (AFF && sizeof(R) == 4 && K <= 5) ? 4 : 1) void k_lap(int N, TabT<R> T, GeoT<R> G, R* __restrict__ x, const R* __restrict__ tr,
                                              R* __restrict__ y, const R* __restrict__ r,
                                              const R* __restrict__ dinv, R* __restrict__ d,
                                              R c1, R c2, double* __restrict__ part,
                                              const int* __restrict__ active, R* __restrict__ tr_out,
                                              int* __restrict__ ran) {
  const bool go = !active || *active;
  if (ran && blockIdx.x == 0 && threadIdx.x == 0) *ran = go ? 1 : 0;  // timed launches: did it work?
  if (!go) return;
  constexpr int K2 = K * K;
  __shared__ __align__(16) R lds[kWG * K2];
  const int e0 = blockIdx.x * kWG;
  if (G.iact) {  // a block inside one stopped ensemble instance
    const int i0 = G.inst[e0];
    if (i0 == G.inst[min(e0 + kWG, N) - 1] && !G.iact[i0]) return;
  }
  const int el = e0 + threadIdx.x;
  const bool live = el < N;
  R u[K2], ga[K2], gb[K2], gc[K2];
  tile_load<K2>(x, lds, e0, N);
#pragma unroll
  for (int m = 0; m < K2; ++m) u[m] = live ? lds[threadIdx.x * K2 + m] : R(0);  // tail threads: no stale LDS
  if (AFF) {  // metrics constant per element
    const R a = live ? G.gel[(size_t)el * 3] : R(0), b = live ? G.gel[(size_t)el * 3 + 1] : R(0),
            c = live ? G.gel[(size_t)el * 3 + 2] : R(0);
#pragma unroll
    for (int m = 0; m < K2; ++m) { ga[m] = a; gb[m] = b; gc[m] = c; }
  } else {
    __syncthreads();
    tile_load<K2>(G.Ga, lds, e0, N);
#pragma unroll
    for (int m = 0; m < K2; ++m) ga[m] = lds[threadIdx.x * K2 + m];
    __syncthreads();
    tile_load<K2>(G.Gb, lds, e0, N);
#pragma unroll
    for (int m = 0; m < K2; ++m) gb[m] = lds[threadIdx.x * K2 + m];
    __syncthreads();
    tile_load<K2>(G.Gc, lds, e0, N);
#pragma unroll
    for (int m = 0; m < K2; ++m) gc[m] = lds[threadIdx.x * K2 + m];
  }
  R out[K2];
  if (live) {
    R bv[4][K], bg[4][K], cv[4][K], cg[4][K];
    traces<K>(T, u, bv, bg);
    faces2<K, AFF>(T, G, el, tr, G.sw_psi, static_cast<const R*>(nullptr), false, bv, bg, cv, cg);
    second_order_core<K>(T, u, bv, cv, cg, ga, gb, gc, out);
#pragma unroll
    for (int j = 0; j < K; ++j)
#pragma unroll
      for (int i = 0; i < K; ++i) out[j * K + i] *= -(T.w[j] * T.w[i]);
  } else {
#pragma unroll
    for (int m = 0; m < K2; ++m) out[m] = 0;
  }
  double p0 = 0., p1 = 0.;
  if (EPI == EPI_STORE) {
#pragma unroll
    for (int m = 0; m < K2; ++m) p0 += u[m] * out[m];
    __syncthreads();
#pragma unroll
    for (int m = 0; m < K2; ++m) lds[threadIdx.x * K2 + m] = out[m];
    tile_store<K2>(y, lds, e0, N);
  } else {
    __syncthreads();
    tile_load<K2>(r, lds, e0, N);
    R res[K2];
#pragma unroll
    for (int m = 0; m < K2; ++m) res[m] = lds[threadIdx.x * K2 + m] - out[m];
    if (EPI == EPI_RESID) {
#pragma unroll
      for (int j = 0; j < K; ++j)
#pragma unroll
        for (int i = 0; i < K; ++i) {
          const double v = res[j * K + i], uw = v / ((double)T.w[j] * T.w[i]);
          p0 += v * v; p1 += uw * uw;
        }
      if (!live) { p0 = 0.; p1 = 0.; }
      __syncthreads();
#pragma unroll
      for (int m = 0; m < K2; ++m) lds[threadIdx.x * K2 + m] = res[m];
      tile_store<K2>(y, lds, e0, N);
    } else if (EPI == EPI_RESTRICT) {
      R s[4] = {0, 0, 0, 0};
#pragma unroll
      for (int j = 0; j < K; ++j)
#pragma unroll
        for (int i = 0; i < K; ++i)
#pragma unroll
          for (int c = 0; c < 4; ++c) s[c] += q1w(T, c, j, i) * res[j * K + i];
      if (live)
#pragma unroll
        for (int c = 0; c < 4; ++c) y[(size_t)el * 4 + c] = s[c];
    } else {  // Chebyshev updates
      __syncthreads();
      tile_load<K2>(dinv, lds, e0, N);
      R dd[K2];
#pragma unroll
      for (int m = 0; m < K2; ++m) dd[m] = res[m] * lds[threadIdx.x * K2 + m];
      if (EPI == EPI_CHEB) {
        __syncthreads();
        tile_load<K2>(d, lds, e0, N);
#pragma unroll
        for (int m = 0; m < K2; ++m) dd[m] = c1 * lds[threadIdx.x * K2 + m] + c2 * dd[m];
      } else {
#pragma unroll
        for (int m = 0; m < K2; ++m) dd[m] *= c2;
      }
      if (y) {  // uniform: the direction for the next step of the sweep
        __syncthreads();
#pragma unroll
        for (int m = 0; m < K2; ++m) lds[threadIdx.x * K2 + m] = dd[m];
        tile_store<K2>(y, lds, e0, N);
      }
#pragma unroll
      for (int m = 0; m < K2; ++m) u[m] += dd[m];
#pragma unroll
      for (int m = 0; m < K2; ++m) lds[threadIdx.x * K2 + m] = u[m];
      tile_store<K2>(x, lds, e0, N);
      if (tr_out && live) {  // traces of the updated iterate for the next operator application
        R bv[4][K], bg[4][K];
        traces<K>(T, u, bv, bg);
        store_traces<K>(tr_out, N, el, bv, bg);
      }
      if (part) {  // r . x (PCG's r.z after the last smoothing step)
        tile_load<K2>(r, lds, e0, N);
#pragma unroll
        for (int m = 0; m < K2; ++m) p0 += lds[threadIdx.x * K2 + m] * u[m];
        if (!live) p0 = 0.;
      }
    }
  }
  if (part) {
    p0 = wave_sum(p0);
    p1 = wave_sum(p1);
    if (threadIdx.x == 0) { part[blockIdx.x] = p0; part[gridDim.x + blockIdx.x] = p1; }
  }
}

// ==========================================================================================
// Poisson right-hand side (src/HO_2D.cpp:1881-1900) in the scaled form
//   bs = W (omega J + L(0; BC_psi)),   with partial sums of (bs/W)^2 = ||b||^2.
// ==========================================================================================
template <int K>
__global__ __launch_bounds__(kWG) void k_rhs(int N, Tab T, Geo G, const double* __restrict__ omega,
                                              const double* __restrict__ bcpsi, double* __restrict__ bs,
                                              double* __restrict__ part) {
  constexpr int K2 = K * K;
  __shared__ __align__(16) double lds[kWG * K2];
  const int e0 = blockIdx.x * kWG;
  const int el = e0 + threadIdx.x;
  const bool live = el < N;
  double w[K2], jac[K2];
  tile_load<K2>(omega, lds, e0, N);
#pragma unroll
  for (int m = 0; m < K2; ++m) w[m] = lds[threadIdx.x * K2 + m];
  __syncthreads();
  tile_load<K2>(G.jac, lds, e0, N);
#pragma unroll
  for (int m = 0; m < K2; ++m) jac[m] = lds[threadIdx.x * K2 + m];
  double out[K2];
#pragma unroll
  for (int m = 0; m < K2; ++m) out[m] = 0.;
  bool bnd = false;
  if (live)
#pragma unroll
    for (int f = 0; f < 4; ++f) bnd |= G.nb[(size_t)el * 4 + f] < 0;
  if (bnd) {
    double z[K2], bv[4][K], bg[4][K], cv[4][K], cg[4][K];
#pragma unroll
    for (int m = 0; m < K2; ++m) z[m] = 0.;
#pragma unroll
    for (int f = 0; f < 4; ++f)
#pragma unroll
      for (int k = 0; k < K; ++k) { bv[f][k] = 0.; bg[f][k] = 0.; cv[f][k] = 0.; cg[f][k] = 0.; }
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      const int nb = G.nb[(size_t)el * 4 + f];
      if (nb >= 0) continue;  // a zero field gives zero common values on interior faces
      const int e = -1 - nb;
      double fa[K], fb[K];
      const size_t q = ((size_t)el * 4 + f) * K;
#pragma unroll
      for (int k = 0; k < K; ++k) { fa[k] = G.fa[q + k]; fb[k] = G.fb[q + k]; }
      // Neumann values enter with the owner face's outward sign: boundary_source . BC_Poisson
      // (src/HO_2D.cpp:1660-1714, sign (2t-1) of :1685) equals L(0; sign * BC) on such faces
      boundary_face<K>(T, G, el, f, G.sw_psi[e], bcpsi + (size_t)e * K, G.esign[e], bv[f], bg[f], fa, fb, cv[f], cg[f]);
    }
    const size_t base = (size_t)el * K2;
    second_order_core<K>(T, z, bv, cv, cg, G.Ga + base, G.Gb + base, G.Gc + base, out);
  }
  double p0 = 0.;
#pragma unroll
  for (int j = 0; j < K; ++j)
#pragma unroll
    for (int i = 0; i < K; ++i) {
      const int m = j * K + i;
      const double b = -w[m] * jac[m] - out[m];  // unscaled reference RHS
      p0 += b * b;
      out[m] = -(T.w[j] * T.w[i]) * b;
    }
  if (!live) p0 = 0.;
  __syncthreads();
#pragma unroll
  for (int m = 0; m < K2; ++m) lds[threadIdx.x * K2 + m] = out[m];
  tile_store<K2>(bs, lds, e0, N);
  p0 = wave_sum(p0);
  if (threadIdx.x == 0) part[blockIdx.x] = p0;
}

// ==========================================================================================
// Velocity from the stream function (src/HO_2D.cpp:3174-3366) with the no-slip wall
// vorticity flux (update_diffusion_BC, src/HO_2D.cpp:2647-2664), fused with the traces the
// stage kernel needs from neighbours: omega value/derivative traces and the contravariant
// normal-velocity trace.  AFF: metrics constant per element (G.gel, G.gface, G.gjm).
// ==========================================================================================
// Affine instances with K <= 5 are held to 2 waves per SIMD (256 registers; a few spill to
// scratch), which beats one wave with no spills.
template <int K, bool AFF, bool BATCH = false>
__global__ __launch_bounds__(kWG, (AFF && K <= 5) ? 2 : 1) void k_velocity(int N, Tab T, Geo G, BCDev B, const double* __restrict__ psi,
                                                   const double* __restrict__ trpsi, const double* __restrict__ omega,
                                                   double* __restrict__ vel, double* __restrict__ U,
                                                   double* __restrict__ trw, double* __restrict__ tru,
                                                   double inv_dt_re, InstDev I, VelBatch vb) {
  constexpr int K2 = K * K;
  if constexpr (BATCH) {  // probe blockIdx.y of a batch
    const size_t b = blockIdx.y;
    psi += b * vb.psi; trpsi += b * vb.trpsi; omega += b * vb.omega; vel += b * vb.vel; U += b * vb.U;
    trw += b * vb.trw; tru += b * vb.tru; B.vjump += b * vb.vjump; B.diff += b * vb.diff; B.pvel += b * vb.pvel;
  }
  __shared__ __align__(16) double lds[kWG * K2 * 2];
  const int e0 = blockIdx.x * kWG;
  const int el = e0 + threadIdx.x;
  const bool live = el < N;
  double u[K2];
  tile_load<K2>(psi, lds, e0, N);
#pragma unroll
  for (int m = 0; m < K2; ++m) u[m] = lds[threadIdx.x * K2 + m];
  double vx[K2], vy[K2], U1[K2], U2[K2];
  if (live) {
    double bv[4][K], bg[4][K], cv[4][K], cg[4][K];
    traces<K>(T, u, bv, bg);
    faces2<K, AFF>(T, G, el, trpsi, G.sw_psi, B.psi, true, bv, bg, cv, cg);
    // boundary side effects
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      const int nb = G.nb[(size_t)el * 4 + f];
      if (nb >= 0) continue;
      const int e = -1 - nb;
      const size_t q = ((size_t)el * 4 + f) * K;
      if (G.sw_psi[e] == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) B.vjump[(size_t)e * K + edge_fp<K>(f, k)] = cg[f][k] / G.fAn[q + k];
      } else {
        const double vs = 1. - 2. * (f / 2);
#pragma unroll
        for (int k = 0; k < K; ++k) {
          double s = 0.;
#pragma unroll
          for (int d = 0; d < K; ++d) s += vs * cv[f][d] * T.D[d][k];
          B.nvel[(size_t)e * K + edge_fp<K>(f, k)] = s / G.fAn[q + k];
        }
      }
      if (G.no_slip[e]) {
        const double sg = G.esign[e];
#pragma unroll
        for (int m = 0; m < K; ++m) {
          const size_t i = (size_t)e * K + m;
          const double vj = B.vjump[i] + sg * B.pvel[i];
          B.vjump[i] = vj;
          B.diff[i] = vj * (I.inv_dt_re ? I.inv_dt_re[G.inst[el]] : inv_dt_re);
        }
      }
    }
    const size_t base = (size_t)el * K2;
    double ea = 0., eb = 0., ec = 0., exx = 0., exe = 0., eyx = 0., eye = 0., ej = 0.;
    if (AFF) {
      ea = G.gel[(size_t)el * 3]; eb = G.gel[(size_t)el * 3 + 1]; ec = G.gel[(size_t)el * 3 + 2];
      exx = G.gjm[(size_t)el * 5]; exe = G.gjm[(size_t)el * 5 + 1]; eyx = G.gjm[(size_t)el * 5 + 2];
      eye = G.gjm[(size_t)el * 5 + 3]; ej = G.gjm[(size_t)el * 5 + 4];
    }
    double ft[K2], gt[K2];
#pragma unroll
    for (int ky = 0; ky < K; ++ky)
#pragma unroll
      for (int kx = 0; kx < K; ++kx) {
        double dx = 0., de = 0.;
#pragma unroll
        for (int d = 0; d < K; ++d) { dx += u[ky * K + d] * T.D[d][kx]; de += u[d * K + kx] * T.D[d][ky]; }
        dx += (cv[0][ky] - bv[0][ky]) * T.grad[kx][0] + (cv[1][ky] - bv[1][ky]) * T.grad[kx][1];
        de += (cv[2][kx] - bv[2][kx]) * T.grad[ky][0] + (cv[3][kx] - bv[3][kx]) * T.grad[ky][1];
        const int v = ky * K + kx;
        const double ga = AFF ? ea : G.Ga[base + v], gb = AFF ? eb : G.Gb[base + v], gc = AFF ? ec : G.Gc[base + v];
        ft[v] = dx * ga - de * gb;
        gt[v] = de * gc - dx * gb;
      }
    double fB[2][K], gB[2][K];
#pragma unroll
    for (int ib = 0; ib < 2; ++ib)
#pragma unroll
      for (int k = 0; k < K; ++k) {
        double a = cg[ib][k], b = cg[2 + ib][k];
#pragma unroll
        for (int d = 0; d < K; ++d) { a -= ft[k * K + d] * T.bb[d][ib]; b -= gt[d * K + k] * T.bb[d][ib]; }
        fB[ib][k] = a; gB[ib][k] = b;
      }
#pragma unroll
    for (int ky = 0; ky < K; ++ky)
#pragma unroll
      for (int kx = 0; kx < K; ++kx) {
        const int v = ky * K + kx;
        const double tx = ft[v] + fB[0][ky] * T.rad[kx][0] + fB[1][ky] * T.rad[kx][1];
        const double ty = gt[v] + gB[0][kx] * T.rad[ky][0] + gB[1][kx] * T.rad[ky][1];
        const double xxi = AFF ? exx : G.xxi[base + v], xet = AFF ? exe : G.xeta[base + v];
        const double yxi = AFF ? eyx : G.yxi[base + v], yet = AFF ? eye : G.yeta[base + v];
        const double J = AFF ? ej : G.jac[base + v];
        const double dpdx = (tx * xxi + ty * xet) / J;
        const double dpdy = (tx * yxi + ty * yet) / J;
        vx[v] = dpdy;    // u =  dpsi/dy
        vy[v] = -dpdx;   // v = -dpsi/dx
        U1[v] = vx[v] * yet - vy[v] * xet;    // J (xi_x u + xi_y v)
        U2[v] = -vx[v] * yxi + vy[v] * xxi;   // J (eta_x u + eta_y v)
      }
  } else {
#pragma unroll
    for (int m = 0; m < K2; ++m) { vx[m] = vy[m] = U1[m] = U2[m] = 0.; }
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < K2; ++m) { lds[threadIdx.x * 2 * K2 + 2 * m] = vx[m]; lds[threadIdx.x * 2 * K2 + 2 * m + 1] = vy[m]; }
  if (vel) tile_store<2 * K2>(vel, lds, e0, N);  // (uniform) intermediate RK stages skip it
  else __syncthreads();
#pragma unroll
  for (int m = 0; m < K2; ++m) { lds[threadIdx.x * 2 * K2 + m] = U1[m]; lds[threadIdx.x * 2 * K2 + K2 + m] = U2[m]; }
  tile_store<2 * K2>(U, lds, e0, N);
  // omega traces for the stage kernel
  double w[K2];
  tile_load<K2>(omega, lds, e0, N);
#pragma unroll
  for (int m = 0; m < K2; ++m) w[m] = lds[threadIdx.x * K2 + m];
  if (!live) return;
  double bv[4][K], bg[4][K];
  traces<K>(T, w, bv, bg);
  store_traces<K>(trw, N, el, bv, bg);
  // contravariant normal-velocity traces (src/HO_2D.cpp:2332, 2489-2503), face-major [4][N][K]
#pragma unroll
  for (int rc = 0; rc < K; ++rc)
#pragma unroll
    for (int ib = 0; ib < 2; ++ib) {
      double s = 0., s2 = 0.;
#pragma unroll
      for (int m = 0; m < K; ++m) { s += U1[rc * K + m] * T.bb[m][ib]; s2 += U2[m * K + rc] * T.bb[m][ib]; }
      tru[((size_t)ib * N + el) * K + rc] = s;
      tru[((size_t)(2 + ib) * N + el) * K + rc] = s2;
    }
}

// ==========================================================================================
// Stage kernel: diffusion (src/HO_2D.cpp:2666-2831, BC_vorticity at :2726-2727),
// update_advection_BC (:2219-2243), advection type 1 (:2245-2437) or 2 (:2439-2645),
// k = dt (adv + diff) (:1833-1836) and the RK update (:1465-1531).  AFF: metrics constant per
// element (G.gel, G.gface, G.gjm).
// MODE 0: the stage.  MODE 1 (advection type 1): the stage, recording the upwind switch of
// every interior flux point in so.sig: sign(dw) sign(fo - s fn) where |dw| > 1e-4, else 0, so
// that the upwind term 1/2 s_f |a| dw equals 1/2 s_f sig (fo - s fn).  MODE 2 (advection type
// 1): the stage map with the switches frozen at so.sig, which is bilinear in (U, omega) and
// linear in the boundary values: diffusion of omega/trw with the boundary values in B, advection
// of so.omega_adv/so.trw_adv by U/tru; boundary arrays are not written.  Called with (U, domega)
// or (dU, omega) and homogeneous boundary values it gives the stage's tangent.
// ==========================================================================================
template <int K, int ADV, bool AFF, int MODE = 0>
__global__ __launch_bounds__(kWG) void k_stage(int N, Tab T, Geo G, BCDev B, const double* __restrict__ omega,
                                                const double* __restrict__ U, const double* __restrict__ trw,
                                                const double* __restrict__ tru, double re_inv, double dt,
                                                InstDev I, StageOut so) {
  constexpr int K2 = K * K;
  if constexpr (MODE == 2) {  // probe blockIdx.y of a batch
    const size_t b = blockIdx.y;
    omega += b * so.bs_w; U += b * so.bs_U; trw += b * so.bs_trw; tru += b * so.bs_tru;
    so.omega_adv += b * so.bs_wa; so.trw_adv += b * so.bs_trwa; so.k += b * so.bs_k; B.diff += b * so.bs_diff;
  }
  __shared__ __align__(16) double lds[kWG * K2 * 2];
  const int e0 = blockIdx.x * kWG;
  const int el = e0 + threadIdx.x;
  const bool live = el < N;
  double w[K2], U1[K2], U2[K2];
  tile_load<K2>(omega, lds, e0, N);
#pragma unroll
  for (int m = 0; m < K2; ++m) w[m] = lds[threadIdx.x * K2 + m];
  __syncthreads();
  tile_load<2 * K2>(U, lds, e0, N);
#pragma unroll
  for (int m = 0; m < K2; ++m) { U1[m] = lds[threadIdx.x * 2 * K2 + m]; U2[m] = lds[threadIdx.x * 2 * K2 + K2 + m]; }
  double wa[MODE == 2 ? K2 : 1];  // MODE 2: the advected vorticity
  if constexpr (MODE == 2) {
    __syncthreads();
    tile_load<K2>(so.omega_adv, lds, e0, N);
#pragma unroll
    for (int m = 0; m < K2; ++m) wa[m] = lds[threadIdx.x * K2 + m];
  }
  auto wadv = [&](int m) -> double {
    if constexpr (MODE == 2) return wa[m]; else return w[m];
  };
  const double* trwa = MODE == 2 ? so.trw_adv : trw;
  double kk[K2];
  if (live) {
    const size_t base = (size_t)el * K2;
    // ---- diffusion (mode 2 with so.adv_only: zero diffusion inputs, skipped)
    const bool do_diff = MODE != 2 || !so.adv_only;
    double bv[4][K], bg[4][K], cv[4][K], cg[4][K];
    if (do_diff) {
      traces<K>(T, w, bv, bg);
      faces2<K, AFF>(T, G, el, trw, G.sw_diff, B.diff, false, bv, bg, cv, cg);
    } else {
#pragma unroll
      for (int f = 0; f < 4; ++f)
#pragma unroll
        for (int k = 0; k < K; ++k) bv[f][k] = bg[f][k] = cv[f][k] = cg[f][k] = 0.;
    }
    double bva[MODE == 2 ? 4 : 1][K];
    if constexpr (MODE == 2) {
      double bga[4][K];
      traces<K>(T, wa, bva, bga);
    }
    auto bvadv = [&](int f, int k) -> double {
      if constexpr (MODE == 2) return bva[f][k]; else return bv[f][k];
    };
    double dif[K2];
    double ej = 0.;
    if (AFF) ej = G.gjm[(size_t)el * 5 + 4];
    if (!do_diff) {
#pragma unroll
      for (int m = 0; m < K2; ++m) dif[m] = 0.;
    } else if (AFF) {
      double ga[K2], gb[K2], gc[K2];
      const double a = G.gel[(size_t)el * 3], b = G.gel[(size_t)el * 3 + 1], c = G.gel[(size_t)el * 3 + 2];
#pragma unroll
      for (int m = 0; m < K2; ++m) { ga[m] = a; gb[m] = b; gc[m] = c; }
      second_order_core<K>(T, w, bv, cv, cg, ga, gb, gc, dif);
    } else {
      second_order_core<K>(T, w, bv, cv, cg, G.Ga + base, G.Gb + base, G.Gc + base, dif);
    }
    // ---- boundary vorticity and advective boundary flux
    double bca[4][K];
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      const int nb = G.nb[(size_t)el * 4 + f];
      if (nb >= 0) continue;
      const int e = -1 - nb;
      const size_t q = ((size_t)el * 4 + f) * K;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int fp = edge_fp<K>(f, k);
        const double vort = cv[f][k];
        const double a = vort * B.nvel[(size_t)e * K + fp] * G.fAn[q + k];
        if constexpr (MODE != 2) {
          B.vort[(size_t)e * K + fp] = vort;
          B.adv[(size_t)e * K + fp] = a;
        }
        bca[f][k] = a;
      }
    }
    // ---- advection
    double upw[4][K], bu[4][K], cvf[4][K];
#pragma unroll
    for (int rc = 0; rc < K; ++rc)
#pragma unroll
      for (int ib = 0; ib < 2; ++ib) {
        double s = 0., s2 = 0.;
#pragma unroll
        for (int m = 0; m < K; ++m) { s += U1[rc * K + m] * T.bb[m][ib]; s2 += U2[m * K + rc] * T.bb[m][ib]; }
        bu[ib][rc] = s;
        bu[2 + ib][rc] = s2;
      }
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      const int nb = G.nb[(size_t)el * 4 + f];
      if (nb < 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) { upw[f][k] = bca[f][k]; cvf[f][k] = 0.; }
        if (ADV == 2) {
          const int e = -1 - nb;
          const size_t q = ((size_t)el * 4 + f) * K;
#pragma unroll
          for (int k = 0; k < K; ++k) cvf[f][k] = B.nvel[(size_t)e * K + edge_fp<K>(f, k)] * G.fAn[q + k];
        }
        continue;
      }
      const int fn = G.nbf[(size_t)el * 4 + f];
      const int rev = rev_flag(f, fn);
      const double sgn = face_sign(f, fn);
      const double* tw = trwa + ((size_t)fn * N + nb) * 2 * K;
      const double* tuu = tru + ((size_t)fn * N + nb) * K;
      double wn[K], un[K];
#pragma unroll
      for (int k = 0; k < K; ++k) { wn[k] = tw[k]; un[k] = tuu[k]; }
      if (ADV == 1) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const int kn = rev ? K - 1 - k : k;
          const double wnk = rev ? wn[K - 1 - k] : wn[k];
          const double unk = rev ? un[K - 1 - k] : un[k];
          (void)kn;
          const double fo = bu[f][k] * bvadv(f, k);
          const double fnn = unk * wnk;
          double up = 0.5 * (fo + sgn * fnn);
          const double dw = bvadv(f, k) - wnk;
          if constexpr (MODE == 2) {
            const double sg = (double)so.sig[((size_t)el * 4 + f) * K + k];
            up += 0.5 * ((f & 1) ? 1. : -1.) * sg * (fo - sgn * fnn);
          } else {
            if (fabs(dw) > 1.e-4) {
              const double a = (fo - sgn * fnn) / dw;
              up += 0.5 * ((f & 1) ? 1. : -1.) * fabs(a) * dw;
            }
            if constexpr (MODE == 1) {
              const double df = fo - sgn * fnn;
              const int sd = dw > 0. ? 1 : -1;
              const int sf = df > 0. ? 1 : (df < 0. ? -1 : 0);
              so.sig[((size_t)el * 4 + f) * K + k] = (int8_t)(fabs(dw) > 1.e-4 ? sd * sf : 0);
            }
          }
          upw[f][k] = up;
        }
      } else {
        const double t = (double)(f & 1);
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const double wnk = rev ? wn[K - 1 - k] : wn[k];
          const double unk = rev ? un[K - 1 - k] : un[k];
          const double c = 0.5 * (bu[f][k] + sgn * unk);
          cvf[f][k] = c;
          if (c > 0.) upw[f][k] = c * (t * bvadv(f, k) + (1. - t) * wnk);
          else upw[f][k] = c * ((1. - t) * bvadv(f, k) + t * wnk);
        }
      }
    }
    if (ADV == 2) {  // continuous volumetric flux (src/HO_2D.cpp:2539-2546)
      double c1[K2], c2[K2];
#pragma unroll
      for (int j = 0; j < K; ++j)
#pragma unroll
        for (int i = 0; i < K; ++i) {
          const int v = j * K + i;
          c1[v] = U1[v] + (cvf[0][j] - bu[0][j]) * T.rad[i][0] + (cvf[1][j] - bu[1][j]) * T.rad[i][1];
          c2[v] = U2[v] + (cvf[2][i] - bu[2][i]) * T.rad[j][0] + (cvf[3][i] - bu[3][i]) * T.rad[j][1];
        }
#pragma unroll
      for (int m = 0; m < K2; ++m) { U1[m] = c1[m]; U2[m] = c2[m]; }
    }
    double f1[K2], f2[K2];  // discontinuous fluxes U^1 w, U^2 w at the sps
#pragma unroll
    for (int m = 0; m < K2; ++m) { f1[m] = U1[m] * wadv(m); f2[m] = U2[m] * wadv(m); }
    double bdf[4][K];
#pragma unroll
    for (int ib = 0; ib < 2; ++ib)
#pragma unroll
      for (int j = 0; j < K; ++j) {
        double a = 0., b = 0.;
#pragma unroll
        for (int m = 0; m < K; ++m) { a -= f1[j * K + m] * T.bb[m][ib]; b -= f2[m * K + j] * T.bb[m][ib]; }
        bdf[ib][j] = a + upw[ib][j];
        bdf[2 + ib][j] = b + upw[2 + ib][j];
      }
#pragma unroll
    for (int j = 0; j < K; ++j)
#pragma unroll
      for (int i = 0; i < K; ++i) {
        const int v = j * K + i;
        double r = 0.;
#pragma unroll
        for (int d = 0; d < K; ++d) r -= f1[j * K + d] * T.D[d][i];
#pragma unroll
        for (int d = 0; d < K; ++d) r -= f2[d * K + i] * T.D[d][j];
        r -= bdf[0][j] * T.grad[i][0] + bdf[1][j] * T.grad[i][1];
        r -= bdf[2][i] * T.grad[j][0] + bdf[3][i] * T.grad[j][1];
        const double J = AFF ? ej : G.jac[base + v];
        const double adv = r / J;
        const double ri = I.re_inv ? I.re_inv[G.inst[el]] : re_inv;
        const double dti = I.dt ? I.dt[G.inst[el]] : dt;
        const double diff = dif[v] * (ri / J);
        kk[v] = dti * (adv + diff);
      }
  } else {
#pragma unroll
    for (int m = 0; m < K2; ++m) kk[m] = 0.;
  }
  // ---- outputs: k, and the RK combination
  __syncthreads();
#pragma unroll
  for (int m = 0; m < K2; ++m) lds[threadIdx.x * K2 + m] = kk[m];
  if (so.k) tile_store<K2>(so.k, lds, e0, N);
  if (so.next) {  // next stage input: base + c k
    __syncthreads();
    tile_load<K2>(so.base, lds, e0, N);
#pragma unroll
    for (int m = 0; m < K2; ++m) lds[threadIdx.x * K2 + m] = lds[threadIdx.x * K2 + m] + so.c * kk[m];
    tile_store<K2>(so.next, lds, e0, N);
  }
  if (so.final_mode) {  // vorticity update
    double acc[K2];
    __syncthreads();
    tile_load<K2>(so.base, lds, e0, N);
#pragma unroll
    for (int m = 0; m < K2; ++m) acc[m] = lds[threadIdx.x * K2 + m];
    double k1[K2], k2[K2], k3[K2];
    if (so.final_mode >= 2) {
      __syncthreads();
      tile_load<K2>(so.k1, lds, e0, N);
#pragma unroll
      for (int m = 0; m < K2; ++m) k1[m] = lds[threadIdx.x * K2 + m];
    }
    if (so.final_mode == 4) {
      __syncthreads();
      tile_load<K2>(so.k2, lds, e0, N);
#pragma unroll
      for (int m = 0; m < K2; ++m) k2[m] = lds[threadIdx.x * K2 + m];
      __syncthreads();
      tile_load<K2>(so.k3, lds, e0, N);
#pragma unroll
      for (int m = 0; m < K2; ++m) k3[m] = lds[threadIdx.x * K2 + m];
    }
#pragma unroll
    for (int m = 0; m < K2; ++m) {
      if (so.final_mode == 1) acc[m] += kk[m];
      else if (so.final_mode == 2) acc[m] += 0.5 * (k1[m] + kk[m]);
      else acc[m] += (k1[m] + 2. * k2[m] + 2. * k3[m] + kk[m]) / 6.;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < K2; ++m) lds[threadIdx.x * K2 + m] = acc[m];
    tile_store<K2>(so.base, lds, e0, N);
  }
}

// ==========================================================================================
// Per-stage boundary values at time t (update_BCs, src/HO_2D.cpp:1979-2217).
// ==========================================================================================
__global__ void k_update_bcs(int n, int K, BCDev B, double tc0, InstDev I) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double tc = I.ramp ? I.ramp[I.edge_inst[i / K]] : tc0;
  B.psi[i] = B.psi_r[i] != 0. ? B.psi0[i] * tc : B.psi0[i];
  B.nvel[i] = B.nvel_r[i] != 0. ? B.nvel0[i] * tc : B.nvel0[i];
  B.pvel[i] = B.pvel_r[i] != 0. ? B.pvel0[i] * tc : B.pvel0[i];
  B.diff[i] = B.diff_r[i] != 0. ? B.diff0[i] * tc : B.diff0[i];
}

// ==========================================================================================
// Hybrid boundary values at fraction fac1 of the outer step (update_BCs for problem_type
// 11, src/HO_2D.cpp:2104-2191, then the sign loop :2203-2215) and the blend of the
// vorticity field towards the driver's target (:2179-2186).
// ==========================================================================================
__global__ void k_update_bcs_hybrid(int n, int K, BCDev B, HybridDev H, double fac1, double fac2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int e = i / K;
  const double sg = H.esign[e];
  double nv;
  if (H.kind[e] == 1) {  // open: blended driver values
    const double pv = -(fac1 * H.vp0[i] + fac2 * H.vp1[i]);
    B.pvel[i] = pv;
    B.psi[i] = pv;
    B.diff[i] = fac1 * H.vw0[i] + fac2 * H.vw1[i];
    nv = -(fac1 * H.vn0[i] + fac2 * H.vn1[i]);
  } else {               // set-up values: marched psi, inlet vorticity, zero outlet values
    B.psi[i] = B.psi0[i];
    B.diff[i] = B.diff0[i];
    B.pvel[i] = 0.;
    nv = H.kind[e] == 2 ? -H.inlet_nvel[e] : (H.kind[e] == 0 ? 0. : B.nvel[i]);
  }
  B.nvel[i] = nv * sg;
  if (H.sw_diff[e] == 1) B.diff[i] *= sg;
}

__global__ void k_hybrid_blend(size_t n, double* __restrict__ w, HybridDev H, double fac1, double fac2) {
  for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (size_t)gridDim.x * blockDim.x)
    w[v] = (1.0 - H.wgt[v]) * w[v] + H.wgt[v] * (fac1 * H.w0[v] + fac2 * H.w1[v]);
}

// ==========================================================================================
// Host launchers
// ==========================================================================================
#define HOC_K_SWITCH(KV, BODY)                                   \
  switch (KV) {                                                  \
    case 2: { constexpr int KK = 2; BODY; } break;               \
    case 3: { constexpr int KK = 3; BODY; } break;               \
    case 4: { constexpr int KK = 4; BODY; } break;               \
    case 5: { constexpr int KK = 5; BODY; } break;               \
    case 6: { constexpr int KK = 6; BODY; } break;               \
    default: return hipErrorInvalidValue;                        \
  }

static inline int nblk(int N) { return (N + kWG - 1) / kWG; }

template <class R>
static hipError_t launch_trace_t(int K, int pre, int N, const TabT<R>& T, R* x, const R* a, const R* b, R* d,
                                 double coef, const double* scal, int scal_idx, const int* active, R* tr, hipStream_t s,
                                 const int32_t* ev, const int32_t* einst, const int* iact) {
#define LT(P) hipLaunchKernelGGL((k_trace<KK, P, R>), dim3(nblk(N)), dim3(kWG), 0, s, N, T, x, a, b, d, (R)coef, scal, \
                                 scal_idx, active, tr, ev, einst, iact)
  HOC_K_SWITCH(K, {
    switch (pre) {
      case PRE_COPY: LT(PRE_COPY); break;
      case PRE_CHEB0: LT(PRE_CHEB0); break;
      case PRE_PROL: LT(PRE_PROL); break;
      case PRE_PUPD: LT(PRE_PUPD); break;
      default: return hipErrorInvalidValue;
    }
  });
#undef LT
  return hipGetLastError();
}

template <class R>
static hipError_t launch_lap_t(int K, int epi, int N, const TabT<R>& T, const GeoT<R>& G, R* x, const R* tr, R* y,
                               const R* r, const R* diag, R* d, double c1, double c2, double* part, const int* active,
                               hipStream_t s, R* tr_out, int* ran) {
#define LL(E)                                                                                                        \
  do {                                                                                                               \
    if (G.gel)                                                                                                       \
      hipLaunchKernelGGL((k_lap<KK, E, true, R>), dim3(nblk(N)), dim3(kWG), 0, s, N, T, G, x, tr, y, r, diag, d,      \
                         (R)c1, (R)c2, part, active, tr_out, ran);                                                   \
    else                                                                                                             \
      hipLaunchKernelGGL((k_lap<KK, E, false, R>), dim3(nblk(N)), dim3(kWG), 0, s, N, T, G, x, tr, y, r, diag, d,     \
                         (R)c1, (R)c2, part, active, tr_out, ran);                                                   \
  } while (0)
  HOC_K_SWITCH(K, {
    switch (epi) {
      case EPI_STORE: LL(EPI_STORE); break;
      case EPI_RESID: LL(EPI_RESID); break;
      case EPI_CHEB: LL(EPI_CHEB); break;
      case EPI_CHEBF: LL(EPI_CHEBF); break;
      case EPI_RESTRICT: LL(EPI_RESTRICT); break;
      default: return hipErrorInvalidValue;
    }
  });
#undef LL
  return hipGetLastError();
}

hipError_t launch_trace(int K, int pre, int N, const TabT<double>& T, double* x, const double* a, const double* b,
                        double* d, double coef, const double* scal, int scal_idx, const int* active, double* tr,
                        hipStream_t s, const int32_t* ev, const int32_t* einst, const int* iact) {
  return launch_trace_t<double>(K, pre, N, T, x, a, b, d, coef, scal, scal_idx, active, tr, s, ev, einst, iact);
}
hipError_t launch_trace(int K, int pre, int N, const TabT<float>& T, float* x, const float* a, const float* b,
                        float* d, double coef, const double* scal, int scal_idx, const int* active, float* tr,
                        hipStream_t s, const int32_t* ev, const int32_t* einst, const int* iact) {
  return launch_trace_t<float>(K, pre, N, T, x, a, b, d, coef, scal, scal_idx, active, tr, s, ev, einst, iact);
}
hipError_t launch_trace_pupd_f32z(int K, int N, const TabT<double>& T, double* p, const float* z, const double* scal,
                                  int scal_idx, const int* active, double* tr, hipStream_t s, const int32_t* einst,
                                  const int* iact) {
  HOC_K_SWITCH(K, hipLaunchKernelGGL((k_trace<KK, PRE_PUPD, double, float>), dim3(nblk(N)), dim3(kWG), 0, s, N, T, p,
                                     z, (const double*)nullptr, (double*)nullptr, 0., scal, scal_idx, active, tr,
                                     (const int32_t*)nullptr, einst, iact));
  return hipGetLastError();
}
hipError_t launch_lap(int K, int epi, int N, const TabT<double>& T, const GeoT<double>& G, double* x, const double* tr,
                      double* y, const double* r, const double* diag, double* d, double c1, double c2, double* part,
                      const int* active, hipStream_t s, double* tr_out, int* ran) {
  return launch_lap_t<double>(K, epi, N, T, G, x, tr, y, r, diag, d, c1, c2, part, active, s, tr_out, ran);
}
hipError_t launch_lap(int K, int epi, int N, const TabT<float>& T, const GeoT<float>& G, float* x, const float* tr,
                      float* y, const float* r, const float* diag, float* d, double c1, double c2, double* part,
                      const int* active, hipStream_t s, float* tr_out, int* ran) {
  return launch_lap_t<float>(K, epi, N, T, G, x, tr, y, r, diag, d, c1, c2, part, active, s, tr_out, ran);
}

hipError_t launch_rhs(int K, int N, const Tab& T, const Geo& G, const double* omega, const double* bcpsi, double* bs,
                      double* part, hipStream_t s) {
  HOC_K_SWITCH(K, hipLaunchKernelGGL((k_rhs<KK>), dim3(nblk(N)), dim3(kWG), 0, s, N, T, G, omega, bcpsi, bs, part));
  return hipGetLastError();
}

hipError_t launch_velocity(int K, int N, const Tab& T, const Geo& G, const BCDev& B, const double* psi,
                           const double* trpsi, const double* omega, double* vel, double* U, double* trw,
                           double* tru, double inv_dt_re, const InstDev& I, hipStream_t s) {
  const bool aff = G.gel && G.gface && G.gjm;
  const VelBatch vb{};
  HOC_K_SWITCH(K, {
    if (aff)
      hipLaunchKernelGGL((k_velocity<KK, true>), dim3(nblk(N)), dim3(kWG), 0, s, N, T, G, B, psi, trpsi, omega, vel, U,
                         trw, tru, inv_dt_re, I, vb);
    else
      hipLaunchKernelGGL((k_velocity<KK, false>), dim3(nblk(N)), dim3(kWG), 0, s, N, T, G, B, psi, trpsi, omega, vel, U,
                         trw, tru, inv_dt_re, I, vb);
  });
  return hipGetLastError();
}

hipError_t launch_velocity_batch(int K, int N, const Tab& T, const Geo& G, const BCDev& B, const double* psi,
                                 const double* trpsi, const double* omega, double* vel, double* U, double* trw,
                                 double* tru, double inv_dt_re, const InstDev& I, int nbatch, const VelBatch& vb,
                                 hipStream_t s) {
  if (nbatch < 1 || nbatch > 65535) return hipErrorInvalidValue;
  const bool aff = G.gel && G.gface && G.gjm;
  HOC_K_SWITCH(K, {
    if (aff)
      hipLaunchKernelGGL((k_velocity<KK, true, true>), dim3(nblk(N), nbatch), dim3(kWG), 0, s, N, T, G, B, psi, trpsi,
                         omega, vel, U, trw, tru, inv_dt_re, I, vb);
    else
      hipLaunchKernelGGL((k_velocity<KK, false, true>), dim3(nblk(N), nbatch), dim3(kWG), 0, s, N, T, G, B, psi, trpsi,
                         omega, vel, U, trw, tru, inv_dt_re, I, vb);
  });
  return hipGetLastError();
}

hipError_t launch_stage(int K, int adv, int N, const Tab& T, const Geo& G, const BCDev& B, const double* omega,
                        const double* U, const double* trw, const double* tru, double re_inv, double dt,
                        const InstDev& I, const StageOut& so, hipStream_t s, int mode) {
  const bool aff = G.gel && G.gface && G.gjm;
  if (mode != 0 && (adv != 1 || !so.sig || (mode == 2 && (!so.omega_adv || !so.trw_adv))))
    return hipErrorInvalidValue;
  if (mode == 2 && (so.nbatch < 1 || so.nbatch > 65535)) return hipErrorInvalidValue;
#define LS(A, F, M)                                                                                           \
  hipLaunchKernelGGL((k_stage<KK, A, F, M>), dim3(nblk(N), M == 2 ? so.nbatch : 1), dim3(kWG), 0, s, N, T, G, B, \
                     omega, U, trw, tru, re_inv, dt, I, so)
  HOC_K_SWITCH(K, {
    if (mode == 1) { if (aff) LS(1, true, 1); else LS(1, false, 1); }
    else if (mode == 2) { if (aff) LS(1, true, 2); else LS(1, false, 2); }
    else if (adv == 1) { if (aff) LS(1, true, 0); else LS(1, false, 0); }
    else { if (aff) LS(2, true, 0); else LS(2, false, 0); }
  });
#undef LS
  return hipGetLastError();
}

hipError_t launch_update_bcs(int n, int K, const BCDev& B, double tc, const InstDev& I, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_update_bcs, dim3((n + 255) / 256), dim3(256), 0, s, n, K, B, tc, I);
  return hipGetLastError();
}

hipError_t launch_update_bcs_hybrid(int n, int K, const BCDev& B, const HybridDev& H, double fac1, double fac2,
                                    hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_update_bcs_hybrid, dim3((n + 255) / 256), dim3(256), 0, s, n, K, B, H, fac1, fac2);
  return hipGetLastError();
}

hipError_t launch_hybrid_blend(size_t n, double* w, const HybridDev& H, double fac1, double fac2, hipStream_t s) {
  size_t g = (n + 255) / 256;
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(k_hybrid_blend, dim3((unsigned)g), dim3(256), 0, s, n, w, H, fac1, fac2);
  return hipGetLastError();
}

int lap_blocks(int N) { return nblk(N); }

}  // namespace hoc
