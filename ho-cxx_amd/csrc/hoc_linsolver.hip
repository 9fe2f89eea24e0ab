// hoc_linsolver.hip - GPU plug-in with the shape of the reference's linear-solver
// interface BaseLinearSolver (old/cxx/linsolvers/linsolver.hpp:42-109): build(num_rows,
// rowptr, colidx, values), then solve(f, u) with u as initial guess and result.
//
// Algorithm: the one the reference configures at its Poisson call site
// (src/HO_2D.cpp:3011-3086, AMGCL 1.4.0 make_solver<amg<smoothed_aggregation, spai0>,
// bicgstab>): smoothed-aggregation AMG built on the host (plain aggregates with strength
// threshold eps = 0.08 halved per level, prolongator smoothing with omega = 2/3, spai0
// relaxation, Galerkin coarse operators) used as a right preconditioner (one V(1,1) cycle)
// of BiCGSTAB, which runs on the GPU with its scalars kept on the device.  Stopping test
// as AMGCL's: ||f - A u|| <= max(reltol ||f||, abstol).  The coarsest level (<= 400 rows)
// is solved with a dense inverse instead of AMGCL's skyline LU (same result to rounding).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hoc.h"
#include "hoc_dev.hpp"

namespace hoc {

struct LinSolver {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // BaseLinearSolver defaults (old/cxx/linsolvers/linsolver.hpp:78-83)
  int maxiters = 200;
  double reltol = 0., abstol = 1e-10;
  int n = 0;
  bool built = false;
  std::vector<std::unique_ptr<DevLevel<double>>> lv;
  DBuf<double> coarse_inv, d_winv;
  int coarse_n = 0;
  DBuf<double> x, f, r, rh, p, v, ph, sh, t, part, S;
  DBuf<int> F;
  int* pinned = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  VecArgs VA{};
  int iters = 0;
  double resid = 0.;

  ~LinSolver() {
    if (pinned) (void)hipHostFree(pinned);
    for (auto& e : ev) if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }

  void build(int nrows, const int32_t* rowptr, const int32_t* colidx, const double* values) {
    HIPCHK(hipSetDevice(device));
    if (!stream) HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    if (nrows <= 0 || !rowptr || !colidx || !values) throw StatusError(HOC_EINVAL, "empty matrix");
    if (rowptr[0] != 0) throw StatusError(HOC_EINVAL, "rowptr[0] must be 0");
    for (int i = 0; i < nrows; ++i)
      if (rowptr[i + 1] < rowptr[i]) throw StatusError(HOC_EINVAL, "rowptr must be non-decreasing");
    CSR A;
    A.n = A.m = nrows;
    A.ptr.assign(rowptr, rowptr + nrows + 1);
    A.col.assign(colidx, colidx + rowptr[nrows]);
    A.val.assign(values, values + rowptr[nrows]);
    for (int32_t c : A.col) if (c < 0 || c >= nrows) throw StatusError(HOC_EINVAL, "column index out of range");
    for (double v0 : A.val) if (!std::isfinite(v0)) throw StatusError(HOC_ENAN, "non-finite matrix value");
    AMGHost H;
    build_sa_amg(A, 400, H);
    lv.clear();
    // level 0 is the system matrix itself; when it is small enough it is also the coarsest
    for (size_t l = 0; l < H.levels.size(); ++l) {
      auto L = std::make_unique<DevLevel<double>>();
      L->A.upload(H.levels[l].A);
      if (l + 1 < H.levels.size()) {
        L->P.upload(H.levels[l].P);
        L->Rs.upload(H.levels[l].R);
        L->M.upload_as(H.levels[l].Minv);
      }
      const int m = H.levels[l].A.n;
      L->f.alloc(m); L->u.alloc(m); L->t.alloc(m);
      lv.push_back(std::move(L));
    }
    coarse_n = H.levels.back().A.n;
    coarse_inv.upload_as(H.coarse_inv);
    n = nrows;
    for (auto* b : {&x, &f, &r, &rh, &p, &v, &ph, &sh, &t}) b->alloc(n);
    part.alloc(2 * (size_t)vec_grid(n));
    S.alloc(S_COUNT); F.alloc(F_COUNT);
    if (!pinned) HIPCHK(hipHostMalloc((void**)&pinned, 2 * F_COUNT * sizeof(int), hipHostMallocDefault));
    for (auto& e : ev) if (!e) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    d_winv.upload(std::vector<double>{1.});  // unit weights: plain Euclidean norms
    VA.K = 1;
    VA.winv = d_winv.p;
    built = true;
  }

  int poll_k = 0;
  bool poll(bool final_wait) {
    const int slot = poll_k & 1;
    HIPCHK(hipMemcpyAsync(pinned + slot * F_COUNT, F.p, F_COUNT * sizeof(int), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipEventRecord(ev[slot], stream));
    ++poll_k;
    if (final_wait) { HIPCHK(hipEventSynchronize(ev[slot])); return pinned[slot * F_COUNT + F_ACTIVE] != 0; }
    if (poll_k < 2) return true;
    const int prev = (poll_k - 2) & 1;
    HIPCHK(hipEventSynchronize(ev[prev]));
    return pinned[prev * F_COUNT + F_ACTIVE] != 0;
  }

  void vcycle(const double* in, double* out) {
    csr_vcycle<double>(stream, lv, coarse_inv, coarse_n, 0, in, out, F.p);
  }
  void spmv(const double* in, double* out) {
    lv[0]->A.apply(stream, 0, in, nullptr, out, nullptr, nullptr, F.p);
  }

  // Returns true when the stopping test was met.
  bool solve(const double* fh, double* uh) {
    if (!built) throw StatusError(HOC_ESTATE, "hoc_linsolver_solve before hoc_linsolver_build");
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipMemcpyAsync(f.p, fh, n * sizeof(double), hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(x.p, uh, n * sizeof(double), hipMemcpyHostToDevice, stream));
    const int vg = vec_grid(n);
    const double* none = nullptr;
    vec<double>(stream, VOP_DOT2, VA, n, f, f, f, f, nullptr, nullptr, nullptr, part);
    finalize(stream, PH_BB, part, vg, S, F, maxiters);
    double bb = 0.;
    HIPCHK(hipMemcpyAsync(&bb, S.p + S_BB, sizeof(double), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    if (!std::isfinite(bb)) throw StatusError(HOC_ENAN, "non-finite right-hand side");
    // ||r||^2 <= tol2 * ||f||^2 with tol2 = max(reltol^2, abstol^2 / ||f||^2)
    const double tol2 = bb > 0. ? std::max(reltol * reltol, abstol * abstol / bb) : 0.;
    HIPCHK(hipMemcpyAsync(S.p + S_TOL2, &tol2, sizeof(double), hipMemcpyHostToDevice, stream));
    // r = f - A x (the spmv's gating flag is not set yet: run it unpredicated)
    lv[0]->A.apply(stream, 1, x, f, r, nullptr, nullptr, nullptr);
    vec<double>(stream, VOP_DOT2, VA, n, r, r, r, r, nullptr, nullptr, nullptr, part);
    finalize(stream, PH_INIT, part, vg, S, F, maxiters);
    zero_if_flag<double>(stream, n, x, F);
    vec<double>(stream, VOP_COPY, VA, n, rh, nullptr, r, none, nullptr, nullptr, nullptr, nullptr);
    HIPCHK(hipMemsetAsync(p.p, 0, n * sizeof(double), stream));
    HIPCHK(hipMemsetAsync(v.p, 0, n * sizeof(double), stream));
    poll_k = 0;
    for (int it = 0;; ++it) {
      vec<double>(stream, VOP_DOT2, VA, n, rh, rh, r, r, nullptr, S, F, part);
      finalize(stream, PH_RHO, part, vg, S, F, maxiters);
      vec<double>(stream, VOP_BICG_P, VA, n, p, nullptr, r, v, nullptr, S, F, nullptr);
      vcycle(p, ph);
      spmv(ph, v);
      vec<double>(stream, VOP_DOT2, VA, n, rh, rh, v, v, nullptr, S, F, part);
      finalize(stream, PH_ALPHA, part, vg, S, F, maxiters);
      vec<double>(stream, VOP_BICG_S, VA, n, nullptr, r, none, v, nullptr, S, F, part);
      finalize(stream, PH_SNORM, part, vg, S, F, maxiters);
      bicg_half<double>(stream, n, x, ph, S, F);
      vcycle(r, sh);
      spmv(sh, t);
      vec<double>(stream, VOP_DOT2, VA, n, t, t, r, t, nullptr, S, F, part);
      finalize(stream, PH_OMEGA, part, vg, S, F, maxiters);
      vec<double>(stream, VOP_BICG_X, VA, n, x, r, ph, sh, t, S, F, part);
      finalize(stream, PH_RNORM, part, vg, S, F, maxiters);
      if (!poll(it > maxiters + 2) || it > maxiters + 4) break;
    }
    double hs[S_COUNT];
    int hf[F_COUNT];
    HIPCHK(hipMemcpyAsync(hs, S.p, sizeof(hs), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(hf, F.p, sizeof(hf), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(uh, x.p, n * sizeof(double), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    iters = hf[F_ITER];
    resid = hs[S_BB] > 0. ? std::sqrt(hs[S_RR] / hs[S_BB]) : 0.;
    if (hf[F_ZERO]) resid = 0.;
    if (!std::isfinite(resid)) throw StatusError(HOC_ENAN, "NaN/Inf in the linear solve");
    return hf[F_ZERO] || hs[S_RR] <= tol2 * hs[S_BB];
  }
};

}  // namespace hoc

struct hoc_linsolver {
  hoc::LinSolver s;
};

static std::string g_ls_err;

static int ls_fail(hoc_ls_t h, int code, const std::string& msg) {
  if (h) h->s.err = msg;
  g_ls_err = msg;
  return code;
}

template <class Fn>
static int ls_guarded(hoc_ls_t h, Fn fn) {
  if (!h) return ls_fail(nullptr, HOC_EINVAL, "null handle");
  try {
    fn();
  } catch (const hoc::StatusError& e) {
    return ls_fail(h, e.code, e.what());
  } catch (const hoc::HipError& e) {
    return ls_fail(h, HOC_EDEVICE, e.what());
  } catch (const std::bad_alloc&) {
    return ls_fail(h, HOC_EDEVICE, "host allocation failed");
  } catch (const std::exception& e) {
    return ls_fail(h, HOC_EINVAL, e.what());
  }
  return HOC_OK;
}

extern "C" {

int hoc_linsolver_create(hoc_ls_t* out, int32_t device) {
  if (!out) return ls_fail(nullptr, HOC_EINVAL, "null argument");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return ls_fail(nullptr, HOC_EDEVICE, "no HIP device");
  if (device < 0 || device >= ndev) return ls_fail(nullptr, HOC_EDEVICE, "device ordinal out of range");
  hoc_ls_t h = new (std::nothrow) hoc_linsolver();
  if (!h) return ls_fail(nullptr, HOC_EDEVICE, "allocation failed");
  h->s.device = device;
  *out = h;
  return HOC_OK;
}

int hoc_linsolver_set_tolerances(hoc_ls_t h, int32_t maxiters, double reltol, double abstol) {
  return ls_guarded(h, [&] {
    if (maxiters < 0 || !(reltol >= 0.) || !(abstol >= 0.)) throw hoc::StatusError(HOC_EINVAL, "bad tolerances");
    h->s.maxiters = maxiters; h->s.reltol = reltol; h->s.abstol = abstol;
  });
}

int hoc_linsolver_build(hoc_ls_t h, int32_t num_rows, const int32_t* rowptr, const int32_t* colidx,
                        const double* values) {
  return ls_guarded(h, [&] { h->s.build(num_rows, rowptr, colidx, values); });
}

int hoc_linsolver_solve(hoc_ls_t h, const double* f, double* u) {
  return ls_guarded(h, [&] {
    if (!f || !u) throw hoc::StatusError(HOC_EINVAL, "null vector");
    if (!h->s.solve(f, u))
      throw hoc::StatusError(HOC_ENOCONV, "linear solve did not reach the tolerance: rel residual " +
                                              std::to_string(h->s.resid));
  });
}

int32_t hoc_linsolver_num_iterations(hoc_ls_t h) { return h ? h->s.iters : 0; }
double hoc_linsolver_residual(hoc_ls_t h) { return h ? h->s.resid : 0.; }
int32_t hoc_linsolver_levels(hoc_ls_t h) { return h ? (int32_t)h->s.lv.size() : 0; }
const char* hoc_linsolver_last_error(hoc_ls_t h) { return h ? h->s.err.c_str() : g_ls_err.c_str(); }

void hoc_linsolver_destroy(hoc_ls_t h) {
  if (!h) return;
  (void)hipSetDevice(h->s.device);
  if (h->s.stream) (void)hipStreamSynchronize(h->s.stream);
  delete h;
}

}  // extern "C"
