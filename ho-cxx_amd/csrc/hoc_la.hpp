// hoc_la.hpp - vector/reduction/CSR/probing kernels of the Poisson solver (hoc_la.hip).
//
// Vector and level kernels are instantiated for double (FP64 solve) and float (the
// mixed-precision inner solve); scalars, partial sums and flags are always double/int.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <vector>

#include "hoc_kernels.hpp"

namespace hoc {

// Device scalar slots (double) and flags (int) of the Krylov solvers.
// F_MAXIT: iteration budget read by finalize kernels launched with maxiter < 0.
enum { S_BB = 0, S_RR, S_RZ, S_PQ, S_ALPHA, S_BETA, S_OMEGA, S_RHO, S_TOL2, S_IMAX, S_COUNT = 16 };
enum { F_ACTIVE = 0, F_ITER, F_ZERO, F_HALF, F_MAXIT, F_SEQ, F_COUNT = 8 };
enum { PH_BB = 0, PH_INIT, PH_RZ, PH_PQ, PH_UPD, PH_RHO, PH_ALPHA, PH_SNORM, PH_OMEGA, PH_RNORM };
enum { VOP_PCG_UPD = 0, VOP_BICG_P, VOP_BICG_S, VOP_DOT2, VOP_COPY, VOP_BICG_X };

// K: points per direction of an element (1 for a plain vector); winv: device table of the
// K*K inverse quadrature weights 1/(w_j w_i), indexed j K + i.
struct VecArgs {
  int K;
  const double* winv;
  // ensembles with a step length per instance: VOP_PCG_UPD takes alpha = ialpha[einst[element]]
  // and skips the dofs of an instance whose alpha is 0 (stopped)
  const double* ialpha = nullptr;
  const int32_t* einst = nullptr;
};

int vec_grid(size_t n);
template <class R>
void vec(hipStream_t s, int op, const VecArgs& A, size_t n, R* a, R* b, const R* c, const R* d, R* e, const double* S,
         const int* F, double* part);
template <class R>
void bicg_half(hipStream_t s, size_t n, R* x, const R* ph, const double* S, int* F);
template <class R>
void zero_if_flag(hipStream_t s, size_t n, R* x, const int* F);
// mirror (optional): pinned host int that receives seq * 4 + near * 2 + (active flag) when the
// kernel ends (near: active with the residual within 8x of the tolerance);
// seq < 0: the next number of the device counter F[F_SEQ] (incremented) instead of seq.
// maxiter < 0: the budget is read from F[F_MAXIT].
// irr, ibb (optional, ninst values): per-instance squared residual and right-hand-side norms of
// an ensemble; PH_INIT and PH_UPD then continue while some instance has irr > tol^2 ibb (tol^2 =
// S[S_TOL2]; an instance with ibb = 0 is held to the largest ibb), and S[S_IMAX] keeps the
// largest irr / (tol^2 ibb).
void finalize(hipStream_t s, int phase, const double* part, int np, double* S, int* F, int maxiter,
              int* mirror = nullptr, int seq = 0, const double* irr = nullptr, const double* ibb = nullptr,
              int ninst = 0);
// *p = v, in stream order (a kernel: usable where a memset or copy may not be recorded).
void set_int(hipStream_t s, int* p, int v);
template <class R>
void csr(hipStream_t s, int mode, int n, const int32_t* ptr, const int32_t* col, const R* val, const R* x, const R* f,
         R* y, const R* M, R* xio, const int* F, int lanes_per_row);
int csr_lanes(int n, size_t nnz);  // lanes per row suited to a matrix's average row length
template <class R>
void ell(hipStream_t s, int mode, int n, int w, int lanes, const int32_t* col, const R* val, const R* x, const R* f,
         R* y, const R* M, R* xio, const int* F);
template <class R>
void scale(hipStream_t s, int n, const R* M, const R* f, R* x, const int* F);
template <class R>
void dense(hipStream_t s, int n, const R* Ainv, const R* f, R* x, const int* F);
template <class R>
void q1_gather(hipStream_t s, int nv, const int32_t* vptr, const int32_t* vinc, const R* ec, R* rc, const int* F);

// The FP32 V-cycle inside the FP64 Krylov method: r32 = (float) r64; z64 = (double) z32 with
// partials of r64 . z64 in part (vec_grid(n) blocks; may be null).
void demote(hipStream_t s, size_t n, const double* r64, float* r32, const int* F);
void promote_dot(hipStream_t s, size_t n, const float* z32, double* z64, const double* r64, double* part,
                 const int* F);

// Initial guess of a stage solve from the solutions at the last two distinct stage times:
// extrap = 0: prev = psi (history shift, psi kept);  extrap = 1: psi' = psi + f (psi - prev)
// and prev = psi (extrapolation and history shift in one pass).
void psi_predict(hipStream_t s, size_t n, double f, int extrap, double* psi, double* prev);

// Projection onto a basis of earlier solutions (m <= kProjMax vectors X_j of length n, stored
// contiguously, with AX_j = As X_j; Y = X for the As inner product, Y = AX for the Euclidean one
// of the images; the basis need not be orthogonal: its Gram matrix G = Y^T AX is kept).  Every
// coefficient is per segment of a ProjSeg (an ensemble instance: its own c, e, G), at a stride
// per segment; reductions are deterministic (per-block partials summed in block order).
//   proj_dots2:   out[j] = Y_j . t and out[m + j] = Y_j . b (j < m); part: S.nb * 2m
//   proj_dots3:   proj_dots2 and, from the same pass over Y, gcol[j] = Y_j . ax (j < m; ax the
//                 image of the newest vector Y_{m-1}: its Gram column, as proj_gcol gives it);
//                 part: S.nb * 3m
//   proj_solve:   v[0, m) and v[m, 2m) <- G^-1 v (G in columns of ldg values, see k_proj_solve)
//   proj_combine: out += sum_j c[j] X_j; with v set, also v = -sum_j c[m + j] X_j (the e of a
//                 later append, from the same pass over X)
//   proj_append:  proj_ortho's append when proj_combine has left v = -X e in x: x = psi + x,
//                 with n0 and wcol as proj_ortho writes them
//   proj_ortho:   appends a new solution psi: x = psi - sum_j e[j] X_j (e = G^-1 Y . b from
//                 before the solve: Y . As psi = Y . b up to the solve's residual), psi's
//                 squared norm (mr ? b . b : psi . b) into n0 (the scale of the in-span test),
//                 and wcol[0, ldw): psi's coefficients in the basis with x appended; part: S.nb
//   proj_gcol:    the Gram column of the appended x from its image ax = As x (applied by the
//                 operator, not inferred from the solve's recursive residual): Y_j . ax (j < m)
//                 and (mr ? ax : x) . ax; part: S.nb * (m + 1)
//   proj_gram:    out[i m + j] = Y_i . AX_j (part: S.nb * m * m; out stride m * m)
//   proj_reseed:  slots k < kk of a basis of m vectors become sum_j C[k m + j] X_j (and AX alike)
constexpr int kProjMax = 24;
struct ProjSeg {
  int nseg = 0, nb = 0;              // segments; blocks of 256 threads over all of them
  const int32_t* blk_seg = nullptr;  // [nb] the segment of each block (device)
  const int32_t* bbeg = nullptr;     // [nseg + 1] first block of each segment (device)
  const int64_t* sbeg = nullptr;     // [nseg + 1] first dof of each segment (device)
};
// Host side of a ProjSeg: blocks per segment (about 8 dofs per thread, at least one) from the
// segments' first dofs.
void proj_seg_layout(const std::vector<int64_t>& sbeg, std::vector<int32_t>& bbeg, std::vector<int32_t>& blk_seg);
void proj_dots2(hipStream_t s, size_t n, const ProjSeg& S, int m, const double* Y, const double* t, const double* b,
                double* part, double* out, int ostride);
void proj_dots3(hipStream_t s, size_t n, const ProjSeg& S, int m, const double* Y, const double* ax, const double* t,
                const double* b, double* part, double* out, int ostride, double* gcol, int gstride);
void proj_solve(hipStream_t s, const ProjSeg& S, int m, const double* G, int ldg, int gstride, double* v,
                int vstride);
void proj_combine(hipStream_t s, size_t n, const ProjSeg& S, int m, const double* X, const double* c, int cstride,
                  double* out, double* v = nullptr);
void proj_append(hipStream_t s, size_t n, const ProjSeg& S, int m, int mr, const double* e, int estride,
                 const double* psi, const double* b, double* x, double* part, double* n0, int n0stride,
                 double* wcol, int wstride, int ldw);
void proj_ortho(hipStream_t s, size_t n, const ProjSeg& S, int m, int mr, const double* X, const double* e,
                int estride, const double* psi, const double* b, double* x, double* part, double* n0, int n0stride,
                double* wcol, int wstride, int ldw);
void proj_gram(hipStream_t s, size_t n, const ProjSeg& S, int m, const double* Y, const double* AX, double* part,
               double* out);
void proj_reseed(hipStream_t s, size_t n, const ProjSeg& S, int m, int kk, const double* C, int cstride, double* X,
                 double* AX);
void proj_gcol(hipStream_t s, size_t n, const ProjSeg& S, int m, int mr, const double* Y, const double* x,
               const double* ax, double* part, double* gcol, int gstride);
// PCG over an ensemble with a step length per instance: every instance has its own alpha, beta
// and active flag, so each converges as if solved alone, to its own tolerance; the shared
// kernels run while any instance is active, and an instance that has stopped takes alpha =
// beta = 0, so its x and r stay as they are (its Poisson-operator blocks return at once).
struct EnsKry {
  double* in;        // the per-instance sums of the phase (r.z, p.q or the weighted ||r||^2)
  double* rz;        // r.z of the last direction update
  double* alpha;     // step lengths
  double* beta;      // direction-update coefficients
  const double* bb;  // weighted ||b_i||^2
  int* act;          // per-instance active flags
  int n;
};
enum { ENS_INIT = 0, ENS_RZ, ENS_PQ, ENS_UPD };
// ENS_INIT after the joint PH_INIT finalize (E.in = ||r_i||^2): each instance active when the
// solve is and its residual is above tol^2 bb_i.  ENS_RZ: beta_i (a breakdown stops the
// instance); ENS_PQ: alpha_i; ENS_UPD: the instances' tests, the iteration count, F_ACTIVE =
// any instance active within maxiter, S[S_RR] = sum ||r_i||^2, S[S_IMAX], and the status word as
// finalize publishes it.  An instance with bb_i = 0 is held to the largest bb.
void finalize_ens(hipStream_t s, int phase, const EnsKry& E, double* S, int* F, int maxiter, int* mirror = nullptr,
                  int seq = 0);
// out[s] = sum over segment s of a . b (a FP64, b FP64 or FP32); skipped when F's active flag is
// clear (F may be null), and out[s] left as it is for a segment whose flag in iact (may be null)
// is 0.  part: S.nb values.
template <class RB>
void seg_dot(hipStream_t s, const ProjSeg& S, const double* a, const RB* b, const int* F, double* part, double* out,
             const int* iact = nullptr);
// out[s] = sum over segment s of (r winv)^2 (weighted squared norms per ensemble instance);
// skipped when F (may be null) has its active flag clear, and out[s] left as it is for a segment
// whose flag in iact (may be null) is 0.  part: S.nb values.
void seg_norm2(hipStream_t s, const ProjSeg& S, const VecArgs& A, const double* r, const int* F, double* part,
               double* out, const int* iact = nullptr);
// The RK4 guess and stage-difference history at the first stage of a step, in one pass: psi (the
// new step's first stage solution T1 on entry) becomes the second stage's guess, Q[0..3] the
// completed step's stage differences and S[0] = T1 (see k_rk4_history).
void rk4_history(hipStream_t s, size_t n, int qok, double* psi, double* const* S, double* const* Q);
// psi' = sum_j a[j] x_j (x_j may be psi itself, read before the write), and save = psi (the old
// value) when save is not null; nterm <= 8.
void psi_combine(hipStream_t s, size_t n, int nterm, const double* a, const double* const* x, double* psi,
                 double* save);

// Ensembles: out[i] = ||b_i / W||^2 over instance i's elements [off[i], off[i+1]) (K2 values
// per element), one workgroup per instance (deterministic order).
void inst_norms(hipStream_t s, const VecArgs& A, int n_inst, const int32_t* off, int K2, const double* b, double* out);

// y = -y / (w[j] w[i]) at every solution point (v = el K^2 + j K + i); w: K host values (K <= 6).
void unweight(hipStream_t s, size_t n, int K, const double* w, double* y);

// Setup-time probing (FP64).
void probe_set(hipStream_t s, int N, int K2, const int32_t* color, int c, int rs, double* x);
void probe_diag(hipStream_t s, int N, int K2, const int32_t* color, int c, int rs, const double* y, double* diag);
void q1_probe_accum(hipStream_t s, int N, int K, const Tab& T, int rs, const double* y, double* M);

}  // namespace hoc
