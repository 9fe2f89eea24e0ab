// hoc_out.hip - node-averaged output (save_smooth_vtk, src/HO_2D.cpp:3369-3531) and sample-point
// values (src/HO_2D.cpp:3593-3612).  Output runs once per dump, so one thread per element / node /
// point with direct loads is enough; the arithmetic order follows the reference.
#include "hoc_out.hpp"

namespace hoc {

__global__ void k_smooth_elem(SmoothArgs A) {
  const int el = blockIdx.x * blockDim.x + threadIdx.x;
  if (el >= A.N) return;
  const int K = A.K, L = A.L, K2 = K * K, L2 = L * L;
  double om[16], ps[16], ux[16], uy[16];
  for (int Ly = 0; Ly < L; ++Ly)  // :3413-3428, solution points to geometry nodes
    for (int Lx = 0; Lx < L; ++Lx) {
      const int g = A.slot[Ly * L + Lx];
      double a = 0., b = 0., c = 0., d = 0.;
      for (int ky = 0; ky < K; ++ky)
        for (int kx = 0; kx < K; ++kx) {
          const double cf = A.sgb[ky][Ly] * A.sgb[kx][Lx];
          const size_t v = (size_t)el * K2 + ky * K + kx;
          a += cf * A.w[v];
          b += cf * A.psi[v];
          c += cf * A.vel[2 * v];
          d += cf * A.vel[2 * v + 1];
        }
      om[g] = a; ps[g] = b; ux[g] = c; uy[g] = d;
    }
  // boundary faces, in increasing boundary-edge order (the reference's loop :3485), take the
  // boundary vorticity and the Cartesian boundary velocity (:3431-3478) extrapolated along the face
  int es[4], fs[4], nf = 0;
  for (int f = 0; f < 4; ++f) {
    const int nbv = A.nb[(size_t)el * 4 + f];
    if (nbv >= 0) continue;
    int pos = nf++;
    while (pos > 0 && es[pos - 1] > -1 - nbv) { es[pos] = es[pos - 1]; fs[pos] = fs[pos - 1]; --pos; }
    es[pos] = -1 - nbv;
    fs[pos] = f;
  }
  for (int q = 0; q < nf; ++q) {
    const int e = es[q], ijp = fs[q], d = ijp / 2, t = ijp % 2;
    const int side = ijp == 0 ? 3 : (ijp == 1 ? 1 : (ijp == 2 ? 0 : 2));  // i2f (src/MathHelpers.hpp:23)
    const double parl_sign = -1. + 2. * (side / 2);
    double cx[kMaxKd], cy[kMaxKd];
    for (int k = 0; k < K; ++k) {
      const double* mt = A.bmet + ((size_t)e * K + k) * 4;
      const double dxx = mt[0], dxe = mt[1], dyx = mt[2], dye = mt[3];
      const int idx = (side / 2) * (K - 2 * k - 1) + k;
      const double un = A.bnvel[(size_t)e * K + idx], ut = parl_sign * A.bpvel[(size_t)e * K + idx];
      if (d == 0) {
        const double g2 = sqrt(dxe * dxe + dye * dye);
        cx[k] = (1. / g2) * (un * dye + ut * dxe);
        cy[k] = (1. / g2) * (-un * dxe + ut * dye);
      } else {
        const double g1 = sqrt(dxx * dxx + dyx * dyx);
        cx[k] = (1. / g1) * (-un * dyx + ut * dxx);
        cy[k] = (1. / g1) * (un * dxx + ut * dyx);
      }
    }
    for (int l = 0; l < L; ++l) {
      const int su2pv = l * ((1 - d) * L + d) + t * (L - 1) * (d * (L - 1) + 1);
      const int g = A.slot[su2pv];
      double a = 0., b = 0., c = 0.;
      for (int m = 0; m < K; ++m) {
        const double cf = A.sgb[m][l];
        const int nidx = (side / 2) * (K - 2 * m - 1) + m;
        a += cf * A.bvort[(size_t)e * K + nidx];
        b += cf * cx[m];
        c += cf * cy[m];
      }
      om[g] = a; ux[g] = b; uy[g] = c;
    }
  }
  double* o = A.elemout + (size_t)el * L2 * 4;
  for (int g = 0; g < L2; ++g) { o[4 * g] = om[g]; o[4 * g + 1] = ps[g]; o[4 * g + 2] = ux[g]; o[4 * g + 3] = uy[g]; }
}

__global__ void k_smooth_nodes(int n_nodes, const int32_t* __restrict__ nptr, const int32_t* __restrict__ ninc,
                               const double* __restrict__ elemout, double* __restrict__ out) {
  const int nd = blockIdx.x * blockDim.x + threadIdx.x;
  if (nd >= n_nodes) return;
  double sw = 0., sp = 0., sx = 0., sy = 0.;
  const int q0 = nptr[nd], q1 = nptr[nd + 1];
  for (int q = q0; q < q1; ++q) {
    const double* e = elemout + (size_t)ninc[q] * 4;
    sx += 1. * e[2]; sy += 1. * e[3];
    sw += e[0]; sp += e[1];
  }
  const int rep = q1 - q0;
  const double r = 1. / rep;
  out[4 * (size_t)nd] = sw / rep;
  out[4 * (size_t)nd + 1] = sp / rep;
  out[4 * (size_t)nd + 2] = sx * r;
  out[4 * (size_t)nd + 3] = sy * r;
}

__global__ void k_sample(int n, int L2, const int32_t* __restrict__ sel, const double* __restrict__ shape,
                         const int32_t* __restrict__ elem_nodes, const double* __restrict__ nodes,
                         double* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int el = sel[p];
  double w = 0., u = 0., v = 0.;
  for (int l = 0; l < L2; ++l) {
    const int nd = elem_nodes[(size_t)el * L2 + l];
    const double c = shape[(size_t)p * L2 + l];
    w += c * nodes[4 * (size_t)nd];
    u += c * nodes[4 * (size_t)nd + 2];
    v += c * nodes[4 * (size_t)nd + 3];
  }
  out[3 * (size_t)p] = w;
  out[3 * (size_t)p + 1] = u;
  out[3 * (size_t)p + 2] = v;
}

hipError_t launch_smooth_elem(const SmoothArgs& A, hipStream_t s) {
  if (A.L < 2 || A.L > 4 || A.K < 2 || A.K > kMaxKd) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_smooth_elem, dim3((A.N + 127) / 128), dim3(128), 0, s, A);
  return hipGetLastError();
}
hipError_t launch_smooth_nodes(int n_nodes, const int32_t* nptr, const int32_t* ninc, const double* elemout,
                               double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_smooth_nodes, dim3((n_nodes + 255) / 256), dim3(256), 0, s, n_nodes, nptr, ninc, elemout, out);
  return hipGetLastError();
}
hipError_t launch_sample(int n, int L2, const int32_t* sel, const double* shape, const int32_t* elem_nodes,
                         const double* nodes, double* out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sample, dim3((n + 255) / 256), dim3(256), 0, s, n, L2, sel, shape, elem_nodes, nodes, out);
  return hipGetLastError();
}

}  // namespace hoc
