// hoc_solver.hip - device state, Poisson solver, RK driver and the C ABI (include/hoc.h).
//
// Per RK stage (src/HO_2D.cpp:1806-1839):
//   update_BCs(t)            k_update_bcs
//   solve_Poisson            k_rhs, then PCG or BiCGSTAB on As = -W A (W = Gauss weights
//                            w_i w_j, which makes As symmetric on affine meshes) with a
//                            V-cycle: matrix-free fine level + Chebyshev(3) Jacobi
//                            smoothing, continuous bilinear coarse space on the element
//                            corners (Ac = Pc^T As Pc assembled from probes of As), then
//                            smoothed aggregation below
//   velocity, wall vorticity k_trace + k_velocity
//   diffusion + advection    k_stage (fused with k = dt (adv + diff) and the RK update)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/hoc.h"
#include "hoc_adj.hpp"
#include "hoc_dev.hpp"
#include "hoc_fr.hpp"
#include "hoc_host.hpp"
#include "hoc_la.hpp"
#include "hoc_out.hpp"

namespace hoc {

// ---- scratch of the no-allocation operator apply (hoc_apply_laplacian_device) ---------------
// The face traces of that entry point live in this library-wide array instead of a buffer of
// the handle: every GPU the library runs on holds one copy of kNoAllocScratchBytes until the
// process ends, shared by every handle on that GPU (calls that overlap on one GPU race on it).
// Host code never passes the array's name to a kernel: it uses the per-device address that
// noalloc_scratch_init() takes from hipGetSymbolAddress.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices] = {nullptr};

// Once per device, with that device current: the device address of `scratch`.
static void noalloc_scratch_init(int dev) {
  if (dev < 0 || dev >= kMaxDevices) throw StatusError(HOC_EINVAL, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev]) return;
  void* p = nullptr;
  HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) throw HipError("hipGetSymbolAddress returned a null address for the scratch array");
  g_scratch_ptr[dev] = static_cast<double*>(p);
}
static double* noalloc_scratch(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[dev];
}

// Preconditioned-Krylov state in working precision R: the fine operator's tables and
// metrics, the Krylov and smoother vectors, face traces and the V-cycle hierarchy.
// FP64 (R = double) serves the plain solve, the outer residual of the mixed solve and
// setup probing; FP32 (R = float) is the mixed solve's inner Krylov + V-cycle.
template <class R>
struct Prec {
  TabT<R> T{};
  GeoT<R> G{};
  DBuf<R> Ga, Gb, Gc, fa, fb, fA, fB, fjac, fAn, gel, gface;  // FP32 copies (unused for FP64)
  // dinv: 1 / diag(As); dch, dch2: Chebyshev directions (ping-pong)
  DBuf<R> r, z, p, q, rh, v, ph, sh, t, x, dch, dch2, dinv, tr, tr2, ec, rc, xc;
  std::vector<std::unique_ptr<DevLevel<R>>> lv;  // Q1 vertex level and the SA levels below
  DBuf<R> coarse_inv;
  int coarse_n = 0;
  int prof_off = 0;  // offset of this precision's entries in the live kernel profile
  // set only while the iterations of a batched ensemble PCG are queued: element -> instance and
  // the instances' live flags; fine-level blocks of a stopped instance then return at once
  const int32_t* einst = nullptr;
  const int* iact = nullptr;
  void alloc_vectors(size_t D, int N, int K) {
    for (auto* f : {&r, &z, &p, &q, &rh, &v, &ph, &sh, &t, &x, &dch, &dch2, &dinv}) f->alloc(D);
    tr.alloc((size_t)N * 8 * K); tr2.alloc((size_t)N * 8 * K);
    ec.alloc((size_t)N * 4);
  }
  void upload_levels(const AMGHost& H) {
    lv.clear();
    for (size_t l = 0; l < H.levels.size(); ++l) {
      auto L = std::make_unique<DevLevel<R>>();
      L->A.upload(H.levels[l].A);
      if (l + 1 < H.levels.size()) {
        L->P.upload(H.levels[l].P);
        L->Rs.upload(H.levels[l].R);
        L->M.upload_as(H.levels[l].Minv);
      }
      const int n = H.levels[l].A.n;
      L->f.alloc(n); L->u.alloc(n); L->t.alloc(n);
      lv.push_back(std::move(L));
    }
    coarse_n = H.levels.back().A.n;
    coarse_inv.upload_as(H.coarse_inv);
    rc.alloc(H.levels[0].A.n); xc.alloc(H.levels[0].A.n);
  }
};

constexpr int kMaxCheb = 8;

struct Solver {
  hoc_config cfg{};
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;

  // host
  Mesh mesh;
  bool mesh_loaded = false, ready = false;
  Tables tab;
  Geometry geo;
  BCs bch;
  int K = 0, K2 = 0, N = 0, E = 0;
  size_t D = 0;
  double time = 0., re_inv = 0.;
  Tab T{};
  VecArgs VA{};
  DBuf<double> d_winv;  // 1/(w_j w_i), the inverse quadrature weights VA.winv points at
  bool mixed = false;

  // device: geometry and connectivity
  DBuf<double> d_jac, d_xxi, d_xeta, d_yxi, d_yeta, d_Ga, d_Gb, d_Gc, d_fa, d_fb, d_fA, d_fB, d_fjac, d_fAn;
  DBuf<int32_t> d_nb;
  DBuf<int8_t> d_nbf, d_swpsi, d_swdiff, d_noslip;
  DBuf<double> d_esign, d_gel, d_gface, d_gjm;
  Geo G{};
  // boundary values
  DBuf<double> b_psi, b_diff, b_nvel, b_pvel, b_vj, b_vort, b_adv;
  DBuf<double> b_psi0, b_psi_r, b_diff0, b_diff_r, b_nvel0, b_nvel_r, b_pvel0, b_pvel_r;
  BCDev B{};
  // fields
  DBuf<double> w, psi, vel, U, k1, k2, k3, k4, vort, kscratch, bs;
  DBuf<double> trw, tru;
  // Poisson solver: FP64 state (the Krylov method, the initial residual, probing) and FP32
  // state (the mixed-precision V-cycle); S/F = scalars and flags of the Krylov solve.
  Prec<double> P64;
  Prec<float> P32;
  DBuf<double> part, S;
  DBuf<int> F;
  // Krylov progress published by the end-of-iteration finalize kernel (seq * 4 + near * 2 + active), in
  // pinned coherent host memory; seq counts iterations over the handle's lifetime.
  int* mirror = nullptr;
  int* mirror_dev = nullptr;
  int seq_next = 0;
  int nparts = 0;
  int nv = 0;                             // Q1 coarse space: element-corner vertices
  DBuf<int32_t> d_ev, d_vptr, d_vinc;     // element -> 4 vertices; vertex -> (element*4+corner)
  // Chebyshev smoother of degree cheb_m: step 0 scales by 1/theta, step k >= 1 uses
  // d = cheb_c1[k] d + cheb_c2[k] D^-1 res
  int cheb_m = 2;
  double cheb_theta = 1., cheb_c1[kMaxCheb] = {0}, cheb_c2[kMaxCheb] = {0};
  double lam_max = 0.;
  int krylov = HOC_KRYLOV_BICGSTAB;
  // live kernel timing (hoc_profile_*)
  struct Prof {
    bool on = false;
    int every = 1;      // record every every-th launch
    int64_t seen = 0;   // launches since profiling was turned on
    std::vector<hipEvent_t> e0, e1;
    std::vector<int> var;
    std::vector<double> nbytes;
    DBuf<int> ran;        // per recorded launch: 1 when it worked, 0 when it was a predicated no-op
    int64_t skipped = 0;  // recorded launches left out as no-ops
    int n = 0;
    double ms[HOC_PROF_VARIANTS] = {0};
    double bytes[HOC_PROF_VARIANTS] = {0};
    int64_t cnt[HOC_PROF_VARIANTS] = {0};
  } prof;
  // stats
  // hybrid operation (problem_type 11): values carried by boundary edges, outer-step window,
  // open-boundary edges in driver order, start/end values and blend targets
  std::vector<double> edge_nvel_in;
  std::vector<int8_t> edge_has_nvel;
  double time_start = 0., time_end = 0.;
  std::vector<int32_t> open_edges;
  std::vector<double> h_vn1, h_vp1, h_vw1;
  DBuf<uint8_t> d_hkind;
  DBuf<double> d_hinlet, d_vn0, d_vn1, d_vp0, d_vp1, d_vw0, d_vw1, d_hw0, d_hw1, d_hwgt;
  HybridDev H{};
  // ensembles: instance i owns elements [inst_off[i], inst_off[i+1]) and has its own Re, dt, clock
  int n_inst = 1;
  std::vector<int32_t> inst_off;
  std::vector<double> inst_re, inst_dt, inst_time;
  DBuf<int32_t> d_inst, d_einst, d_ioff;
  DBuf<double> d_ireinv, d_idt, d_iidr, d_iramp, d_inorm;
  InstDev I{};
  int last_iters = 0, last_outer = 0;
  double last_rel = 0.;
  int64_t total_iters = 0, n_solves = 0;
  double setup_seconds = 0.;

  ~Solver() {
    clear_graphs();
    for (auto e : prof.e0) (void)hipEventDestroy(e);
    for (auto e : prof.e1) (void)hipEventDestroy(e);
    if (mirror) (void)hipHostFree(mirror);
    if (stream) (void)hipStreamDestroy(stream);
  }

  // ---------------------------------------------------------------------------------------
  void setup() {
    auto t0 = std::chrono::steady_clock::now();
    HIPCHK(hipSetDevice(device));
    if (!stream) HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    clear_graphs();
    noalloc_scratch_init(device);
    connect(mesh);
    build_tables(cfg.K, mesh.lnod, tab);
    build_geometry(mesh, tab, geo);
    std::vector<int32_t> edge_inst;
    if (n_inst > 1 && (int)inst_off.size() == n_inst + 1) {
      edge_inst.resize(mesh.n_bedges);
      for (int e = 0; e < mesh.n_bedges; ++e)
        edge_inst[e] = (int32_t)(std::upper_bound(inst_off.begin(), inst_off.end(), mesh.be_elem[e]) - inst_off.begin() - 1);
    }
    build_bcs(mesh, tab, cfg.problem_type, bch, edge_inst, edge_nvel_in, edge_has_nvel);
    K = cfg.K; K2 = K * K; N = mesh.n_elem; E = mesh.n_bedges; D = (size_t)N * K2;
    re_inv = cfg.Re > 1.e-3 ? 1. / cfg.Re : -1.;  // src/HO_2D.cpp:255
    // tables
    std::memset(&T, 0, sizeof(T));
    for (int m = 0; m < K; ++m) {
      for (int s = 0; s < 2; ++s) {
        T.bb[m][s] = tab.bb[m][s]; T.bgb[m][s] = tab.bgb[m][s];
        T.rad[m][s] = tab.rad[m][s]; T.grad[m][s] = tab.grad[m][s];
      }
      for (int k = 0; k < K; ++k) T.D[m][k] = tab.D[m][k];
      T.w[m] = tab.wgt[m];
      T.x[m] = tab.sps[m];
    }
    T.gp[0] = tab.gprime[0]; T.gp[1] = tab.gprime[1];
    {
      std::vector<double> wi((size_t)K * K);
      for (int j = 0; j < K; ++j)
        for (int i = 0; i < K; ++i) wi[(size_t)j * K + i] = 1. / (tab.wgt[j] * tab.wgt[i]);
      d_winv.upload(wi);
      VA.K = K;
      VA.winv = d_winv.p;
    }
    for (size_t v = 0; v < D; ++v)
      if (!(geo.jac[v] > 0.)) throw StatusError(HOC_EMESH, "non-positive Jacobian at a solution point (element orientation must be CCW)");
    // geometry upload
    std::vector<double> Ga(D), Gb(D), Gc(D);
    for (size_t v = 0; v < D; ++v) {
      const double J = geo.jac[v];
      Ga[v] = (geo.xeta[v] * geo.xeta[v] + geo.yeta[v] * geo.yeta[v]) / J;
      Gb[v] = (geo.xxi[v] * geo.xeta[v] + geo.yxi[v] * geo.yeta[v]) / J;
      Gc[v] = (geo.xxi[v] * geo.xxi[v] + geo.yxi[v] * geo.yxi[v]) / J;
    }
    std::vector<double> fa(geo.fA.size()), fb(geo.fA.size());
    for (size_t i = 0; i < fa.size(); ++i) { fa[i] = geo.fA[i] / geo.fjac[i]; fb[i] = geo.fB[i] / geo.fjac[i]; }
    d_jac.upload(geo.jac); d_xxi.upload(geo.xxi); d_xeta.upload(geo.xeta); d_yxi.upload(geo.yxi); d_yeta.upload(geo.yeta);
    d_Ga.upload(Ga); d_Gb.upload(Gb); d_Gc.upload(Gc);
    d_fa.upload(fa); d_fb.upload(fb); d_fA.upload(geo.fA); d_fB.upload(geo.fB); d_fjac.upload(geo.fjac); d_fAn.upload(geo.fAn);
    d_nb.upload(mesh.nb);
    d_nbf.upload(mesh.nb_face);
    std::vector<int8_t> swp(E), swd(E), ns(E);
    for (int e = 0; e < E; ++e) { swp[e] = (int8_t)bch.sw_psi[e]; swd[e] = (int8_t)bch.sw_diff[e]; ns[e] = (int8_t)bch.no_slip_edge[e]; }
    d_swpsi.upload(swp); d_swdiff.upload(swd); d_noslip.upload(ns);
    d_esign.upload(bch.edge_sign);
    // affine fast path: every element's metrics constant over its sps and faces
    // (tolerances relative to each element's metric scale: rounding leaves ~1e-16 noise
    // in components that vanish on rectangles)
    bool affine = true;
    std::vector<double> gel((size_t)N * 3), gface((size_t)N * 8), gjm((size_t)N * 5);
    for (int e = 0; e < N && affine; ++e) {
      const size_t b0 = (size_t)e * K2;
      const double sc = std::max(std::fabs(Ga[b0]), std::fabs(Gc[b0]));
      gel[(size_t)e * 3] = Ga[b0]; gel[(size_t)e * 3 + 1] = Gb[b0]; gel[(size_t)e * 3 + 2] = Gc[b0];
      const double* jm[5] = {geo.xxi.data(), geo.xeta.data(), geo.yxi.data(), geo.yeta.data(), geo.jac.data()};
      const double sj = std::max(std::max(std::fabs(jm[0][b0]), std::fabs(jm[1][b0])),
                                 std::max(std::fabs(jm[2][b0]), std::fabs(jm[3][b0])));
      for (int c = 0; c < 5; ++c) gjm[(size_t)e * 5 + c] = jm[c][b0];
      for (int m = 1; m < K2 && affine; ++m) {
        affine = std::fabs(Ga[b0 + m] - Ga[b0]) <= 1e-12 * sc && std::fabs(Gb[b0 + m] - Gb[b0]) <= 1e-12 * sc &&
                 std::fabs(Gc[b0 + m] - Gc[b0]) <= 1e-12 * sc;
        for (int c = 0; c < 4 && affine; ++c) affine = std::fabs(jm[c][b0 + m] - jm[c][b0]) <= 1e-12 * sj;
        affine = affine && std::fabs(jm[4][b0 + m] - jm[4][b0]) <= 1e-12 * std::fabs(jm[4][b0]);
      }
      for (int f = 0; f < 4 && affine; ++f) {
        const size_t q0 = ((size_t)e * 4 + f) * K;
        const double fs = std::max(std::fabs(fa[q0]), 1e-300);
        gface[((size_t)e * 4 + f) * 2] = fa[q0]; gface[((size_t)e * 4 + f) * 2 + 1] = fb[q0];
        for (int k = 1; k < K && affine; ++k)
          affine = std::fabs(fa[q0 + k] - fa[q0]) <= 1e-12 * fs && std::fabs(fb[q0 + k] - fb[q0]) <= 1e-12 * fs;
      }
    }
    if (affine) { d_gel.upload(gel); d_gface.upload(gface); d_gjm.upload(gjm); }
    G = Geo{N, E, d_jac, d_xxi, d_xeta, d_yxi, d_yeta, d_Ga, d_Gb, d_Gc, d_fa, d_fb, d_fA, d_fB, d_fjac, d_fAn,
            affine ? d_gel.p : nullptr, affine ? d_gface.p : nullptr, affine ? d_gjm.p : nullptr,
            d_nb, d_nbf, d_swpsi, d_swdiff, d_noslip, d_esign, nullptr};
    setup_instances();
    setup_hybrid();
    const size_t EK = std::max<size_t>((size_t)E * K, 1);
    for (auto* b : {&b_psi, &b_diff, &b_nvel, &b_pvel, &b_vj, &b_vort, &b_adv}) b->alloc(EK);
    auto upl = [&](DBuf<double>& d, const std::vector<double>& h) { if (h.empty()) d.alloc(1); else d.upload(h); };
    upl(b_psi0, bch.psi); upl(b_psi_r, bch.psi_ramp); upl(b_diff0, bch.diff); upl(b_diff_r, bch.diff_ramp);
    upl(b_nvel0, bch.nvel); upl(b_nvel_r, bch.nvel_ramp); upl(b_pvel0, bch.pvel); upl(b_pvel_r, bch.pvel_ramp);
    B = BCDev{b_psi, b_diff, b_nvel, b_pvel, b_vj, b_vort, b_adv, b_psi0, b_psi_r, b_diff0, b_diff_r, b_nvel0, b_nvel_r, b_pvel0, b_pvel_r};
    // fields
    for (auto* f : {&w, &psi, &k1, &k2, &k3, &k4, &vort, &kscratch, &bs}) f->alloc(D);
    vel.alloc(2 * D); U.alloc(2 * D);
    trw.alloc((size_t)N * 8 * K); tru.alloc((size_t)N * 4 * K);
    mixed = cfg.precision == HOC_MIXED;
    P64.T = T; P64.G = G; P64.prof_off = 0;
    P64.alloc_vectors(D, N, K);
    if (mixed) {  // FP32 tables and metrics of the inner solve's fine operator
      Prec<float>& P = P32;
      static_assert(2 * sizeof(TabT<float>) == sizeof(TabT<double>), "table layout");
      float* Tf = reinterpret_cast<float*>(&P.T);
      const double* Td = reinterpret_cast<const double*>(&T);
      for (size_t i = 0; i < sizeof(T) / sizeof(double); ++i) Tf[i] = (float)Td[i];
      P.Ga.upload_as(Ga); P.Gb.upload_as(Gb); P.Gc.upload_as(Gc);
      P.fa.upload_as(fa); P.fb.upload_as(fb); P.fA.upload_as(geo.fA); P.fB.upload_as(geo.fB);
      P.fjac.upload_as(geo.fjac); P.fAn.upload_as(geo.fAn);
      if (affine) { P.gel.upload_as(gel); P.gface.upload_as(gface); }
      P.G = GeoT<float>{N, E, nullptr, nullptr, nullptr, nullptr, nullptr, P.Ga, P.Gb, P.Gc, P.fa, P.fb, P.fA, P.fB,
                        P.fjac, P.fAn, affine ? P.gel.p : nullptr, affine ? P.gface.p : nullptr, nullptr,
                        d_nb, d_nbf, d_swpsi, d_swdiff, d_noslip, nullptr, nullptr};
      P.prof_off = HOC_PROF_VARIANTS / 2;
      P.alloc_vectors(D, N, K);
    }
    nparts = std::max(lap_blocks(N), vec_grid(D));
    part.alloc(2 * (size_t)nparts);
    S.alloc(S_COUNT); F.alloc(F_COUNT);
    if (!mirror) {
      HIPCHK(hipHostMalloc((void**)&mirror, 64, hipHostMallocCoherent | hipHostMallocMapped));
      __atomic_store_n(mirror, -2, __ATOMIC_RELEASE);
      HIPCHK(hipHostGetDevicePointer((void**)&mirror_dev, mirror, 0));
    }
    build_hierarchy();
    // initial-guess history (its mode depends on the Krylov method chosen above)
    psi_prev.alloc(guess_mode() >= 2 ? D : 0);
    for (auto& b : rkS) b.alloc(guess_mode() >= 3 ? D : 0);
    for (auto& b : rkQ) b.alloc(guess_mode() >= 4 ? D : 0);
    proj_fwd.X.alloc(0); proj_fwd.AX.alloc(0); proj_fwd.m = 0;
    proj_adj.X.alloc(0); proj_adj.AX.alloc(0); proj_adj.m = 0;
    reset_adjoint();
    ensure_segments();  // (the timed iterations of an ensemble run its per-instance test)
    tune_launch_mode();
    invalidate_guess();
    time = 0.;
    smooth_ready = false;
    n_sp = 0;
    ready = true;
    HIPCHK(hipStreamSynchronize(stream));
    setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }

  // Fine-level operator launch in precision R, bracketed by events when profiling.
  template <class R>
  void lap(Prec<R>& P, int epi, R* x, R* y, const R* rr, const R* dg, R* d, double c1, double c2, double* partp,
           const int* active, const R* tin = nullptr, R* tout = nullptr) {
    const bool rec = prof.on && (prof.seen++ % prof.every) == 0;
    if (rec) HIPCHK(hipEventRecord(prof.e0[prof.n], stream));
    GeoT<R> G = P.G;
    if (P.iact) { G.inst = P.einst; G.iact = P.iact; }
    HIPCHK(launch_lap(K, epi, N, P.T, G, x, tin ? tin : P.tr.p, y, rr, dg, d, c1, c2, partp, active, stream, tout,
                      rec ? prof.ran.p + prof.n : nullptr));
    if (rec) {
      HIPCHK(hipEventRecord(prof.e1[prof.n], stream));
      prof.var[prof.n] = P.prof_off + epi;
      prof.nbytes[prof.n] = lap_bytes(epi, (double)sizeof(R), tout != nullptr, y != nullptr);
      if (++prof.n == (int)prof.e0.size()) prof_flush();
    }
  }
  // Only launches that worked count: a launch predicated off after convergence returns at once
  // and would book its algorithmic bytes against a near-zero duration.
  void prof_flush() {
    if (prof.n == 0) return;
    HIPCHK(hipEventSynchronize(prof.e1[prof.n - 1]));
    std::vector<int> ran(prof.n);
    HIPCHK(hipMemcpy(ran.data(), prof.ran.p, prof.n * sizeof(int), hipMemcpyDeviceToHost));
    for (int i = 0; i < prof.n; ++i) {
      if (!ran[i]) { ++prof.skipped; continue; }
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, prof.e0[i], prof.e1[i]));
      prof.ms[prof.var[i]] += ms;
      prof.cnt[prof.var[i]] += 1;
      prof.bytes[prof.var[i]] += prof.nbytes[i];
    }
    prof.n = 0;
  }
  // Algorithmic HBM bytes of one fine-operator launch per epilogue, for values of vb bytes:
  // the field, the neighbour face traces (value + derivative), metrics, connectivity, the
  // epilogue's own vectors and (traces_out) the updated iterate's face traces, each read or
  // written once; dir_out: a Chebyshev step stores its direction.
  double lap_bytes(int epi, double vb = 8., bool traces_out = false, bool dir_out = true) const {
    const double vec = (double)N * K2 * vb;
    double b = vec                                   // x
             + (double)N * 4 * 2 * K * vb            // neighbour traces
             + (double)N * 4 * (4 + 1);              // nb, nbf
    if (G.gel) b += (double)N * (3 + 8 + 8) * vb;    // affine: element + own/neighbour face metrics
    else b += 3 * vec + (double)N * 4 * K * 2 * vb * 2;
    switch (epi) {
      case EPI_STORE: b += vec; break;                        // y
      case EPI_RESID: b += 2 * vec; break;                    // r, y
      case EPI_CHEB: b += (dir_out ? 5 : 4) * vec; break;     // r, dinv, d, x out (+ direction out)
      case EPI_CHEBF: b += (dir_out ? 4 : 3) * vec; break;    // r, dinv, x out (+ direction out)
      case EPI_RESTRICT: b += vec + (double)N * 4 * vb; break;  // r, corner sums
    }
    if (traces_out) b += (double)N * 4 * 2 * K * vb;
    return b;
  }

  // As x into y (FP64 trace + apply); x is not modified.
  void apply_As(double* x, double* y) {
    HIPCHK(launch_trace(K, PRE_COPY, N, T, x, nullptr, nullptr, nullptr, 0., nullptr, 0, nullptr, P64.tr.p, stream));
    lap<double>(P64, EPI_STORE, x, y, nullptr, nullptr, nullptr, 0., 0., nullptr, nullptr);
  }

  // Stream-ordered copies that return when the data has arrived.
  void d2h(void* dst, const void* src, size_t bytes) {
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
  }
  void h2d(void* dst, const void* src, size_t bytes) {
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
  }

  // Fine-level diagonal and element-level operator by probing, then SA below it.
  void build_hierarchy() {
    std::vector<int32_t> col1, col2;
    const int nc1 = color_elements(mesh, 1, col1);
    const int nc2 = color_elements(mesh, 2, col2);
    DBuf<int32_t> dcol1, dcol2;
    dcol1.upload(col1); dcol2.upload(col2);
    for (int c = 0; c < nc1; ++c)
      for (int rs = 0; rs < K2; ++rs) {
        probe_set(stream, N, K2, dcol1, c, rs, P64.z);
        apply_As(P64.z, P64.q);
        probe_diag(stream, N, K2, dcol1, c, rs, P64.q, P64.dinv);
      }
    // Continuous bilinear coarse space on the element corners: Pc interpolates vertex values
    // to every element's solution points; Ac = Pc^T As Pc is assembled from probes of As.
    std::vector<int32_t> vid((size_t)mesh.n_nodes, -1), ev((size_t)N * 4);
    nv = 0;
    for (int e = 0; e < N; ++e)
      for (int a = 0; a < 4; ++a) {
        const int nd = mesh.elem[(size_t)e * mesh.lnod * mesh.lnod + a];
        if (vid[nd] < 0) vid[nd] = nv++;
        ev[(size_t)e * 4 + a] = vid[nd];
      }
    std::vector<int32_t> vptr(nv + 1, 0), vinc((size_t)N * 4);
    for (int32_t v : ev) vptr[v + 1]++;
    for (int v = 0; v < nv; ++v) vptr[v + 1] += vptr[v];
    {
      std::vector<int32_t> pos(vptr.begin(), vptr.end() - 1);
      for (int e = 0; e < N; ++e)
        for (int a = 0; a < 4; ++a) vinc[pos[ev[(size_t)e * 4 + a]]++] = e * 4 + a;
    }
    d_ev.upload(ev); d_vptr.upload(vptr); d_vinc.upload(vinc);
    inst_coupled = false;  // an ensemble whose instances share a corner vertex
    if (n_inst > 1 && (int)inst_off.size() == n_inst + 1) {
      auto inst_of = [&](int e) { return std::upper_bound(inst_off.begin(), inst_off.end(), e) - inst_off.begin() - 1; };
      for (int v = 0; v < nv && !inst_coupled; ++v)
        for (int k = vptr[v] + 1; k < vptr[v + 1]; ++k)
          if (inst_of(vinc[k] / 4) != inst_of(vinc[vptr[v]] / 4)) { inst_coupled = true; break; }
    }
    std::vector<std::vector<std::pair<int, double>>> rows(nv);
    auto add = [&](int r, int c, double val) {
      for (auto& pr : rows[r]) if (pr.first == c) { pr.second += val; return; }
      rows[r].push_back({c, val});
    };
    DBuf<double> dM;
    dM.alloc((size_t)N * 16);
    std::vector<double> hM((size_t)N * 16);
    for (int c = 0; c < nc2; ++c) {
      HIPCHK(hipMemsetAsync(dM.p, 0, (size_t)N * 16 * sizeof(double), stream));
      for (int rs = 0; rs < K2; ++rs) {
        probe_set(stream, N, K2, dcol2, c, rs, P64.z);
        apply_As(P64.z, P64.q);
        q1_probe_accum(stream, N, K, T, rs, P64.q, dM);
      }
      d2h(hM.data(), dM.p, hM.size() * sizeof(double));
      for (int f = 0; f < N; ++f) {
        int e = -1;
        if (col2[f] == c) e = f;
        else
          for (int k = 0; k < 4; ++k) {
            const int nb = mesh.nb[(size_t)f * 4 + k];
            if (nb >= 0 && col2[nb] == c) { e = nb; break; }
          }
        if (e < 0) continue;
        for (int a = 0; a < 4; ++a)
          for (int b = 0; b < 4; ++b) {
            const double val = hM[(size_t)f * 16 + a * 4 + b];
            if (val != 0.) add(ev[(size_t)f * 4 + a], ev[(size_t)e * 4 + b], val);
          }
      }
    }
    CSR A1;
    A1.n = A1.m = nv;
    A1.ptr.assign(nv + 1, 0);
    for (int r0 = 0; r0 < nv; ++r0) {
      std::sort(rows[r0].begin(), rows[r0].end());
      for (auto& pr : rows[r0]) { A1.col.push_back(pr.first); A1.val.push_back(pr.second); }
      A1.ptr[r0 + 1] = (int)A1.col.size();
    }
    // Aggregation stops at <= 768 unknowns, solved by a dense inverse: one launch reading
    // at most 4.7 MB replaces a level's four sparse launches (each latency-bound at ~5 us).
    AMGHost H;
    build_sa_amg(A1, 768, H);
    if (mixed) P32.upload_levels(H);
    else P64.upload_levels(H);
    // Chebyshev interval from lambda_max(D^-1 As) by power iteration
    std::vector<double> hx(D);
    uint64_t st = 88172645463325252ull;
    for (size_t i = 0; i < D; ++i) { st ^= st << 13; st ^= st >> 7; st ^= st << 17; hx[i] = (double)(st % 2000001) / 1e6 - 1.; }
    h2d(P64.z.p, hx.data(), D * sizeof(double));
    std::vector<double> hd(D), hy(D);
    d2h(hd.data(), P64.dinv.p, D * sizeof(double));  // the diagonal, inverted below
    for (size_t i = 0; i < D; ++i)
      if (!(hd[i] > 0.)) throw StatusError(HOC_EINVAL, "Poisson operator has a non-positive diagonal entry");
    double lam = 0.;
    for (int it = 0; it < 30; ++it) {
      apply_As(P64.z, P64.q);
      d2h(hy.data(), P64.q.p, D * sizeof(double));
      double nrm = 0., ray = 0., xx = 0.;
      for (size_t i = 0; i < D; ++i) { hy[i] /= hd[i]; nrm += hy[i] * hy[i]; ray += hy[i] * hx[i]; xx += hx[i] * hx[i]; }
      lam = ray / xx;
      nrm = std::sqrt(nrm);
      for (size_t i = 0; i < D; ++i) hx[i] = hy[i] / nrm;
      h2d(P64.z.p, hx.data(), D * sizeof(double));
    }
    lam_max = lam;
    // symmetry of As decides the Krylov method
    krylov = cfg.krylov;
    if (krylov == HOC_KRYLOV_AUTO) {
      std::vector<double> hx2(D), hAx(D), hAy(D);
      for (size_t i = 0; i < D; ++i) { st ^= st << 13; st ^= st >> 7; st ^= st << 17; hx2[i] = (double)(st % 2000001) / 1e6 - 1.; }
      for (size_t i = 0; i < D; ++i) { st ^= st << 13; st ^= st >> 7; st ^= st << 17; hx[i] = (double)(st % 2000001) / 1e6 - 1.; }
      h2d(P64.z.p, hx.data(), D * sizeof(double));
      apply_As(P64.z, P64.q);
      d2h(hAx.data(), P64.q.p, D * sizeof(double));
      h2d(P64.z.p, hx2.data(), D * sizeof(double));
      apply_As(P64.z, P64.q);
      d2h(hAy.data(), P64.q.p, D * sizeof(double));
      double a = 0., b = 0., na = 0., nb = 0.;
      for (size_t i = 0; i < D; ++i) { a += hx2[i] * hAx[i]; b += hx[i] * hAy[i]; na += std::fabs(hx2[i] * hAx[i]); nb += std::fabs(hx[i] * hAy[i]); }
      krylov = std::fabs(a - b) <= 1e-12 * std::max(na, nb) ? HOC_KRYLOV_PCG : HOC_KRYLOV_BICGSTAB;
    }
    const double lmax = 1.1 * lam, lmin = lmax * (cfg.cheb_lower > 0. ? cfg.cheb_lower : 1. / 60.);
    const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
    const double sigma = theta / delta;
    // default degree 3; 4 for the FP32 cycle of a mixed-precision PCG, where the two extra operator
    // applies cost less than the residual reduction they add.  BiCGSTAB (non-affine meshes) keeps 3
    // in both precisions: there degree 4 did not pay at the cylinder, and at the step it made the
    // two precisions stop after different iteration counts, moving the mixed fields away from the
    // FP64 ones by more than the 1e-8 budget at a Krylov tolerance of 1e-12.
    cheb_m = cfg.cheb_degree > 0 ? cfg.cheb_degree : (mixed && krylov == HOC_KRYLOV_PCG) ? 4 : 3;
    cheb_theta = theta;
    double rho = 1. / sigma;
    for (int k = 1; k < cheb_m; ++k) {
      const double rho_new = 1. / (2. * sigma - rho);
      cheb_c1[k] = rho_new * rho;
      cheb_c2[k] = 2. * rho_new / delta;
      rho = rho_new;
    }
    for (size_t i = 0; i < D; ++i) hd[i] = 1. / hd[i];
    h2d(P64.dinv.p, hd.data(), D * sizeof(double));
    if (mixed) P32.dinv.upload_as(hd);
    HIPCHK(hipMemsetAsync(P64.z.p, 0, D * sizeof(double), stream));
    HIPCHK(hipStreamSynchronize(stream));
  }

  // ---------------------------------------------------------------------------------------
  // V-cycle on the Q1 vertex level and below: u_l = V_l(f_l)
  template <class R>
  void coarse_cycle(Prec<R>& P, size_t l, const R* f, R* u, const int* Fp) {
    csr_vcycle<R>(stream, P.lv, P.coarse_inv, P.coarse_n, l, f, u, Fp);
  }

  // zout = V(rin); if part != nullptr, partials of rin.zout are left in part.
  template <class R>
  void vcycle(Prec<R>& P, const R* rin, R* zout, double* partp, const int* Fp) {
    R* rr = const_cast<R*>(rin);
    R* cur = P.tr.p;   // traces of the current iterate
    R* nxt = P.tr2.p;
    // pre-smoothing: Chebyshev(m) from zero
    // directions ping-pong between dch and dch2; the last step of a sweep stores none
    R* din = P.dch.p;
    R* dout = P.dch2.p;
    HIPCHK(launch_trace(K, PRE_CHEB0, N, P.T, zout, rr, P.dinv.p, din, 1. / cheb_theta, nullptr, 0, Fp, cur, stream,
                        nullptr, P.einst, P.iact));
    for (int k = 1; k < cheb_m; ++k) {
      lap(P, EPI_CHEB, zout, k + 1 < cheb_m ? dout : (R*)nullptr, rr, P.dinv.p, din, cheb_c1[k], cheb_c2[k], nullptr,
          Fp, cur, nxt);
      std::swap(cur, nxt);
      std::swap(din, dout);
    }
    // restrict the residual to the corner vertices
    lap(P, EPI_RESTRICT, zout, P.ec.p, rr, (const R*)nullptr, (R*)nullptr, 0., 0., nullptr, Fp, cur);
    q1_gather<R>(stream, nv, d_vptr, d_vinc, P.ec, P.rc, Fp);
    coarse_cycle(P, 0, P.rc.p, P.xc.p, Fp);
    // prolongate and post-smooth: Chebyshev(m) from the corrected iterate
    cur = P.tr.p; nxt = P.tr2.p;
    HIPCHK(launch_trace(K, PRE_PROL, N, P.T, zout, P.xc.p, (const R*)nullptr, (R*)nullptr, 0., nullptr, 0, Fp, cur,
                        stream, d_ev.p, P.einst, P.iact));
    din = P.dch.p;
    dout = P.dch2.p;
    lap(P, EPI_CHEBF, zout, cheb_m > 1 ? din : (R*)nullptr, rr, P.dinv.p, (R*)nullptr, 0., 1. / cheb_theta,
        cheb_m == 1 ? partp : nullptr, Fp, cur, cheb_m == 1 ? (R*)nullptr : nxt);
    std::swap(cur, nxt);
    for (int k = 1; k < cheb_m; ++k) {
      const bool last = k == cheb_m - 1;
      lap(P, EPI_CHEB, zout, last ? (R*)nullptr : dout, rr, P.dinv.p, din, cheb_c1[k], cheb_c2[k],
          last ? partp : nullptr, Fp, cur, last ? (R*)nullptr : nxt);
      std::swap(cur, nxt);
      std::swap(din, dout);
    }
  }

  // z = M r: the V-cycle in the precision of P; for the FP64 Krylov method of a mixed-precision
  // handle, the FP32 V-cycle between a demotion of r and a promotion of z.  With partp set, the
  // partials of r . z are left there; returns their count (0 when partp is null).
  template <class R>
  int precond(Prec<R>& P, const R* rin, R* zout, double* partp, const int* Fp) {
    if constexpr (std::is_same<R, double>::value) {
      if (mixed) {
        demote(stream, D, rin, P32.r.p, Fp);
        vcycle(P32, P32.r.p, P32.z.p, nullptr, Fp);
        promote_dot(stream, D, P32.z.p, zout, rin, partp, Fp);
        return partp ? vec_grid(D) : 0;
      }
    }
    vcycle(P, rin, zout, partp, Fp);
    return partp ? lap_blocks(N) : 0;
  }

  // Wait, lag iterations behind (none when final_wait), until the finalize kernel of the
  // iteration numbered seq has published its status; returns false when the solve ended.
  // seq_first: the first iteration of this solve (no wait before it has run).  The wait spins
  // on the pinned word and checks the stream every 4096 reads, so an idle stream without the
  // expected status (or a device error) ends it with an error instead of a hang.
  bool poll(int seq, int seq_first, bool final_wait, int lag = 1, bool* near = nullptr) {
    const int target = final_wait ? seq : seq - lag;
    if (near) *near = false;
    if (target < seq_first) return true;
    for (unsigned spin = 1;; ++spin) {
      const int v = __atomic_load_n(mirror, __ATOMIC_ACQUIRE);
      if ((v >> 2) >= target) {
        if (near) *near = (v & 2) != 0;
        return (v & 1) != 0;
      }
      if ((spin & 4095) == 0) {
        const hipError_t q = hipStreamQuery(stream);
        if (q == hipSuccess) {
          const int v2 = __atomic_load_n(mirror, __ATOMIC_ACQUIRE);
          if ((v2 >> 2) >= target) {
            if (near) *near = (v2 & 2) != 0;
            return (v2 & 1) != 0;
          }
          throw HipError("Krylov status of an iteration was not published");
        }
        if (q != hipErrorNotReady) HIPCHK(q);
      }
    }
  }

  // Preconditioned Krylov iterations on As x = r in precision R, from the state left by
  // PH_INIT in (Sp, Fp): r holds the residual of x.  PCG when As is symmetric, else
  // right-preconditioned BiCGSTAB.  Every kernel is predicated on Fp's active flag.
  // maxit bounds the host loop; maxit_dev is the device budget (< 0: Fp[F_MAXIT]).
  //
  // Launch modes.  Direct: the host queues each iteration's kernels and waits for the status of
  // the iteration before the last one it queued.  Graph: a hipGraph holding kGraphChunk
  // iterations, captured once per set of buffers, is replayed; its end-of-iteration finalize
  // kernels number themselves from the device counter F[F_SEQ], and the host waits one chunk
  // behind.  The kernels and their order are the same, so both modes give the same iterates;
  // graphs pay where the host's launch cost exceeds the kernels' run time (small meshes), and
  // setup chooses by timing both (launch_mode 0).
  static constexpr int kGraphChunk = 4;
  struct KGraph {
    hipGraph_t g = nullptr;
    hipGraphExec_t e = nullptr;
    const void* key[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    int maxit = 0;
  };
  std::vector<KGraph> kgraphs;
  bool use_graphs = false;
  double launch_tune_ratio = 0.;  // graph / direct time of the same iterations at setup (0: not timed)
  void clear_graphs() {
    for (auto& k : kgraphs) {
      if (k.e) (void)hipGraphExecDestroy(k.e);
      if (k.g) (void)hipGraphDestroy(k.g);
    }
    kgraphs.clear();
  }
  // Before the first iteration.
  template <class R>
  void krylov_begin(Prec<R>& P, R* r) {
    if (krylov == HOC_KRYLOV_PCG) {
      HIPCHK(hipMemsetAsync(P.p.p, 0, D * sizeof(R), stream));
    } else {
      vec<R>(stream, VOP_COPY, VA, D, P.rh, nullptr, r, nullptr, nullptr, nullptr, nullptr, nullptr);
      HIPCHK(hipMemsetAsync(P.p.p, 0, D * sizeof(R), stream));
      HIPCHK(hipMemsetAsync(P.v.p, 0, D * sizeof(R), stream));
    }
  }
  // One iteration's launches.  seq >= 0: the end-of-iteration finalize publishes seq; seq < 0:
  // it takes the next number from F[F_SEQ].
  template <class R>
  void krylov_iter(Prec<R>& P, R* x, R* r, double* Sp, int* Fp, int maxit, int seq) {
    const int lb = lap_blocks(N);
    const int vg = vec_grid(D);
    if constexpr (std::is_same<R, double>::value) {
      if (ens_batched()) {
        ens_iter(x, r, Sp, Fp, maxit, seq);
        return;
      }
    }
    if (krylov == HOC_KRYLOV_PCG) {
      bool zf = false;  // z left in FP32 by the mixed-precision V-cycle
      if constexpr (std::is_same<R, double>::value) {
        if (mixed) {  // r.z from the FP32 V-cycle's last sweep; p = z + beta p reads z in FP32
          demote(stream, D, r, P32.r.p, Fp);
          vcycle(P32, P32.r.p, P32.z.p, part.p, Fp);
          finalize(stream, PH_RZ, part, lb, Sp, Fp, maxit);
          HIPCHK(launch_trace_pupd_f32z(K, N, P.T, P.p.p, P32.z.p, Sp, S_BETA, Fp, P.tr.p, stream));
          zf = true;
        }
      }
      if (!zf) {
        const int np = precond(P, r, P.z.p, part.p, Fp);
        finalize(stream, PH_RZ, part, np, Sp, Fp, maxit);
        HIPCHK(launch_trace(K, PRE_PUPD, N, P.T, P.p.p, P.z.p, (const R*)nullptr, (R*)nullptr, 0., Sp, S_BETA, Fp,
                            P.tr.p, stream));
      }
      lap(P, EPI_STORE, P.p.p, P.q.p, (const R*)nullptr, (const R*)nullptr, (R*)nullptr, 0., 0., part.p, Fp);
      finalize(stream, PH_PQ, part, lb, Sp, Fp, maxit);
      vec<R>(stream, VOP_PCG_UPD, VA, D, x, r, P.p, P.q, nullptr, Sp, Fp, part);
      if constexpr (std::is_same<R, double>::value) {
        if (ens_pcg()) {  // every instance against its own right-hand side
          seg_norm2(stream, pseg, VA, r, Fp, ens_part.p, ens_rr.p);
          finalize(stream, PH_UPD, part, vg, Sp, Fp, maxit, mirror_dev, seq, ens_rr.p, ens_bb.p, pseg.nseg);
          return;
        }
      }
      finalize(stream, PH_UPD, part, vg, Sp, Fp, maxit, mirror_dev, seq);
    } else {
      const R* none = nullptr;
      vec<R>(stream, VOP_DOT2, VA, D, P.rh, P.rh, r, r, nullptr, Sp, Fp, part);
      finalize(stream, PH_RHO, part, vg, Sp, Fp, maxit);
      vec<R>(stream, VOP_BICG_P, VA, D, P.p, nullptr, r, P.v, nullptr, Sp, Fp, nullptr);
      precond(P, P.p.p, P.ph.p, nullptr, Fp);
      HIPCHK(launch_trace(K, PRE_COPY, N, P.T, P.ph.p, none, none, (R*)nullptr, 0., nullptr, 0, Fp, P.tr.p, stream));
      lap(P, EPI_STORE, P.ph.p, P.v.p, none, none, (R*)nullptr, 0., 0., nullptr, Fp);
      vec<R>(stream, VOP_DOT2, VA, D, P.rh, P.rh, P.v, P.v, nullptr, Sp, Fp, part);
      finalize(stream, PH_ALPHA, part, vg, Sp, Fp, maxit);
      vec<R>(stream, VOP_BICG_S, VA, D, nullptr, r, nullptr, P.v, nullptr, Sp, Fp, part);
      finalize(stream, PH_SNORM, part, vg, Sp, Fp, maxit);
      bicg_half<R>(stream, D, x, P.ph, Sp, Fp);
      precond(P, r, P.sh.p, nullptr, Fp);  // gated by F_ACTIVE only; harmless when converged at the half step
      HIPCHK(launch_trace(K, PRE_COPY, N, P.T, P.sh.p, none, none, (R*)nullptr, 0., nullptr, 0, Fp, P.tr.p, stream));
      lap(P, EPI_STORE, P.sh.p, P.t.p, none, none, (R*)nullptr, 0., 0., nullptr, Fp);
      vec<R>(stream, VOP_DOT2, VA, D, P.t, P.t, r, P.t, nullptr, Sp, Fp, part);
      finalize(stream, PH_OMEGA, part, vg, Sp, Fp, maxit);
      vec<R>(stream, VOP_BICG_X, VA, D, x, r, P.ph, P.sh, P.t, Sp, Fp, part);
      finalize(stream, PH_RNORM, part, vg, Sp, Fp, maxit, mirror_dev, seq);
    }
  }
  // One iteration of the batched ensemble PCG (ens_batched): PCG with a step length and a
  // direction coefficient per instance (r_i.z_i, p_i.q_i and ||r_i|| by segment), each instance
  // stopping on its own test; while an instance is stopped, the fine-level kernels skip the
  // blocks that hold only its elements, and its x and r stay as they are.  The coarse levels run
  // over every instance: with no Q1 vertex shared between instances they stay block-diagonal.
  void ens_iter(double* x, double* r, double* Sp, int* Fp, int maxit, int seq) {
    struct Mask {  // the masks hold for this iteration's launches only
      Solver& s;
      explicit Mask(Solver& so) : s(so) {
        s.P64.einst = s.P32.einst = s.d_inst.p;
        s.P64.iact = s.P32.iact = s.ens_act.p;
      }
      ~Mask() {
        s.P64.einst = s.P32.einst = nullptr;
        s.P64.iact = s.P32.iact = nullptr;
      }
    } mask(*this);
    Prec<double>& P = P64;
    const EnsKry E = ensk();
    VecArgs va = VA;  // x += alpha_i p, r -= alpha_i q
    va.ialpha = ens_alpha.p;
    va.einst = d_inst.p;
    if (mixed) {
      demote(stream, D, r, P32.r.p, Fp);
      vcycle(P32, P32.r.p, P32.z.p, nullptr, Fp);
      seg_dot<float>(stream, pseg, r, P32.z.p, Fp, ens_part.p, ens_in.p, ens_act.p);
      finalize_ens(stream, ENS_RZ, E, Sp, Fp, maxit);
      HIPCHK(launch_trace_pupd_f32z(K, N, P.T, P.p.p, P32.z.p, ens_beta.p, -1, Fp, P.tr.p, stream, d_inst.p,
                                    ens_act.p));
    } else {
      vcycle(P, r, P.z.p, nullptr, Fp);
      seg_dot<double>(stream, pseg, r, P.z.p, Fp, ens_part.p, ens_in.p, ens_act.p);
      finalize_ens(stream, ENS_RZ, E, Sp, Fp, maxit);
      HIPCHK(launch_trace(K, PRE_PUPD, N, P.T, P.p.p, P.z.p, nullptr, nullptr, 0., ens_beta.p, -1, Fp, P.tr.p, stream,
                          nullptr, d_inst.p, ens_act.p));
    }
    lap(P, EPI_STORE, P.p.p, P.q.p, (const double*)nullptr, (const double*)nullptr, (double*)nullptr, 0., 0., nullptr,
        Fp);
    seg_dot<double>(stream, pseg, P.p.p, P.q.p, Fp, ens_part.p, ens_in.p, ens_act.p);
    finalize_ens(stream, ENS_PQ, E, Sp, Fp, maxit);
    vec<double>(stream, VOP_PCG_UPD, va, D, x, r, P.p, P.q, nullptr, Sp, Fp, nullptr);
    // ||r_i||^2 into ens_rr, where a stopped instance keeps its last value (its r no longer changes)
    seg_norm2(stream, pseg, VA, r, Fp, ens_part.p, ens_rr.p, ens_act.p);
    EnsKry Eu = E;
    Eu.in = ens_rr.p;
    finalize_ens(stream, ENS_UPD, Eu, Sp, Fp, maxit, mirror_dev, seq);
  }
  // The instantiated graph of kGraphChunk iterations on these buffers (captured on first use).
  template <class R>
  hipGraphExec_t krylov_graph(Prec<R>& P, R* x, R* r, double* Sp, int* Fp, int maxit) {
    const void* key[6] = {&P, x, r, Sp, Fp, mixed ? (const void*)&P32 : nullptr};
    for (auto& k : kgraphs)
      if (k.maxit == maxit && std::equal(key, key + 6, k.key)) return k.e;
    KGraph kg;
    std::copy(key, key + 6, kg.key);
    kg.maxit = maxit;
    HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    try {
      for (int i = 0; i < kGraphChunk; ++i) krylov_iter(P, x, r, Sp, Fp, maxit, -1);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(stream, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIPCHK(hipStreamEndCapture(stream, &kg.g));
    HIPCHK(hipGraphInstantiate(&kg.e, kg.g, nullptr, nullptr, 0));
    kgraphs.push_back(kg);
    return kg.e;
  }
  template <class R>
  void krylov_solve(Prec<R>& P, R* x, R* r, double* Sp, int* Fp, int maxit_host, int maxit_dev) {
    const int maxit = maxit_dev;
    const int seq0 = seq_next;
    krylov_begin(P, r);
    if (use_graphs && !prof.on) {
      hipGraphExec_t ge = krylov_graph(P, x, r, Sp, Fp, maxit);
      set_int(stream, Fp + F_SEQ, seq_next - 1);
      for (int done = 0;;) {
        HIPCHK(hipGraphLaunch(ge, stream));
        const int last = seq_next + kGraphChunk - 1;
        seq_next += kGraphChunk;
        done += kGraphChunk;
        const bool fin = done > maxit_host + 2;
        if (!poll(last, seq0, fin, kGraphChunk) || done > maxit_host + 4) break;
      }
      return;
    }
    for (int it = 0;; ++it) {
      const int seq = seq_next++;
      krylov_iter(P, x, r, Sp, Fp, maxit, seq);
      bool near = false;
      if (!poll(seq, seq0, it > maxit_host + 2, 1, &near) || it > maxit_host + 4) break;
      // the previous iteration ended within 8x of the tolerance: the one just queued is likely
      // the last, so wait for it instead of queuing one that would run predicated off
      if (near && !poll(seq, seq0, true)) break;
    }
  }

  // Launch mode (launch_mode 0): the same 2 kGraphChunk iterations of a solve on a random
  // right-hand side, queued directly and replayed from the captured graph; graphs are used when
  // they take under 85% of the direct time.  The solve's state is reset afterwards.
  void tune_launch_mode() {
    clear_graphs();
    use_graphs = cfg.launch_mode == 2;
    launch_tune_ratio = 0.;
    if (cfg.launch_mode != 0) return;
    std::vector<double> hb(D);
    uint64_t st = 0x9E3779B97F4A7C15ull;
    for (size_t i = 0; i < D; ++i) { st ^= st << 13; st ^= st >> 7; st ^= st << 17; hb[i] = (double)(st % 2000001) / 1e6 - 1.; }
    h2d(bs.p, hb.data(), D * sizeof(double));
    const double sv[2] = {1., 0.};  // S_BB = 1, S_TOL2 = 0: every tuning iteration stays active
    h2d(S.p + S_BB, &sv[0], sizeof(double));
    h2d(S.p + S_TOL2, &sv[1], sizeof(double));
    const int maxit = std::max(cfg.maxiter, 4 * kGraphChunk);
    auto run = [&](bool graph) {
      HIPCHK(hipMemsetAsync(psi.p, 0, D * sizeof(double), stream));
      outer_residual();
      krylov_begin(P64, P64.r.p);
      hipGraphExec_t ge = graph ? krylov_graph(P64, psi.p, P64.r.p, S.p, F.p, maxit) : nullptr;
      HIPCHK(hipStreamSynchronize(stream));
      const auto t0 = std::chrono::steady_clock::now();
      if (graph) {
        set_int(stream, F.p + F_SEQ, seq_next - 1);
        for (int c = 0; c < 2; ++c) HIPCHK(hipGraphLaunch(ge, stream));
        seq_next += 2 * kGraphChunk;
      } else {
        for (int i = 0; i < 2 * kGraphChunk; ++i) krylov_iter(P64, psi.p, P64.r.p, S.p, F.p, maxit, seq_next++);
      }
      HIPCHK(hipStreamSynchronize(stream));
      return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    };
    run(false);  // warm-up
    const double td = run(false);
    run(true);   // capture and warm-up
    const double tg = run(true);
    launch_tune_ratio = td > 0. ? tg / td : 0.;
    use_graphs = tg < 0.85 * td;
    HIPCHK(hipMemsetAsync(psi.p, 0, D * sizeof(double), stream));
    HIPCHK(hipMemsetAsync(bs.p, 0, D * sizeof(double), stream));
    HIPCHK(hipStreamSynchronize(stream));
  }

  // FP64 residual r = bs - As psi with the reference's unscaled norm, and the solve's
  // convergence state in (S, F).
  void outer_residual() {
    const int lb = lap_blocks(N);
    HIPCHK(launch_trace(K, PRE_COPY, N, T, psi.p, nullptr, nullptr, nullptr, 0., nullptr, 0, nullptr, P64.tr.p, stream));
    lap<double>(P64, EPI_RESID, psi.p, P64.r.p, bs.p, nullptr, nullptr, 0., 0., part.p, nullptr);
    if (ens_pcg()) {
      seg_norm2(stream, pseg, VA, P64.r.p, nullptr, ens_part.p, ens_rr.p);
      finalize(stream, PH_INIT, part, lb, S, F, cfg.maxiter, nullptr, 0, ens_rr.p, ens_bb.p, pseg.nseg);
      if (ens_batched()) {
        EnsKry E = ensk();
        E.in = ens_rr.p;
        finalize_ens(stream, ENS_INIT, E, S, F, cfg.maxiter);
      }
    } else {
      finalize(stream, PH_INIT, part, lb, S, F, cfg.maxiter);
    }
  }

  // solve_Poisson (src/HO_2D.cpp:1841-1977) for omega at the current boundary values.
  //  FP64:  PCG/BiCGSTAB with the FP64 V-cycle on the scaled system, stopping on the
  //         reference's ||b - A psi|| <= tol ||b||.
  //  mixed: the same FP64 Krylov method, the same operator and the same stopping test, with
  //         the V-cycle in FP32 as its preconditioner: the residual is demoted at the cycle's
  //         entry and the cycle's output enters the FP64 direction update (precond, and the
  //         PCG loop's FP32-input direction update).
  const double* rhs_bc = nullptr;  // boundary stream-function values of the next solve (null: b_psi)

  // Initial guess of the stage solves.  Mode 1: the previous solution, as the reference
  // (src/HO_2D.cpp:1949-1953).  Mode 2: psi is affine in omega and the RK stage inputs advance
  // smoothly in the stage time tau = steps + c, so the guess at tau is the linear extrapolation
  // psi_a + (tau - tau_a) / (tau_a - tau_b) (psi_a - psi_b) from the solutions at the last two
  // distinct stage times (the previous solution where tau = tau_a).  Only the iteration count
  // changes: every solve still stops on the same residual test.
  DBuf<double> psi_prev;        // psi_b
  double tau_a = 0., tau_b = 0.;
  int n_hist = 0;               // solutions in the history (0: psi is not a stage solution)
  int64_t step_count = 0;
  double stage_tau = std::nan("");  // stage time of the solve in progress (NaN: not a stage solve)
  // default: the projection (mode 5)
  int guess_mode() const { return cfg.psi_guess != 0 ? cfg.psi_guess : 5; }
  // Mode 3 (RK4): psi(omega_n + c k) = psi(omega_n) + c Psi(k) with Psi linear, so the stage
  // solutions S1..S4 of a step give Psi(k1) = 2 (S2 - S1), Psi(k2) = 2 (S3 - S1), Psi(k3) = S4 - S1,
  // and the next step's first solve is S1 + (Psi(k1) + 2 Psi(k2) + 2 Psi(k3) + Psi(k4)) / 6 with
  // Psi(k4) ~ Psi(k3): (-3 S1 + 2 S2 + 4 S3 + 3 S4) / 6.  Its solution T1 then gives Psi(k4) exactly,
  // and the second stage, T1 + Psi(k1') / 2 with Psi(k1') ~ Psi(k4): 4 T1 + S1 - S2 - 2 S3 - S4.
  // Stages 3 and 4 as in mode 2 (S2; 2 S3 - S1).  Other integrators use mode 2 under modes 3 and 4.
  DBuf<double> rkS[4];
  DBuf<double> rkQ[4];            // mode 4: Psi(k1..k4) of the step before S1..S4's
  int rk_stage = -1, rk_cur = 0;  // stage index of the RK4 solve in progress; stages of this step solved
  bool rk_full = false;           // S1..S4 hold a complete step
  bool rk_qok = false;            // rkQ holds a complete step (mode 4)
  // Mode 4: as mode 3, with each predicted Psi(k) corrected by the same stage difference one step
  // earlier (Q = Psi(k) of that step): stage 1 Psi(k4) ~ Psi(k3) + Q4 - Q3; stage 2
  // Psi(k1') ~ Psi(k4) + Q1 - Q4; stage 3 Psi(k2') ~ Psi(k1') + Q2 - Q1; stage 4 Psi(k3') ~ Psi(k2') + Q3 - Q2.
  void rk4_guess_before() {
    const int st = rk_stage;
    double* S[4] = {rkS[0], rkS[1], rkS[2], rkS[3]};
    double* Q[4] = {rkQ[0], rkQ[1], rkQ[2], rkQ[3]};
    const double q = rk_qok ? 1. : 0.;
    if (st == 0 && rk_full) {
      const double a[6] = {-0.5, 1. / 3., 2. / 3., 0.5, q / 6., -q / 6.};
      const double* x[6] = {S[0], S[1], S[2], S[3], Q[3], Q[2]};
      psi_combine(stream, D, rk_qok ? 6 : 4, a, x, psi, nullptr);
    } else if (st == 1 && rk_cur == 1) {
      if (rk_full) {
        // T1 = psi; guess 4 T1 - 2 S3 - S4 - Q4/2 (Q4 of the older step), then the Q history of the
        // completed step: Q1 = 2 (S2 - S1), Q2 = 2 (S3 - S1), Q3 = S4 - S1, Q4 = 6 T1 + 2 S1 - 2 S2 - 4 S3 - 2 S4
        // (mode 3's guess 4 T1 + S1 - S2 - 2 S3 - S4 while Q holds no complete step), and S1 = T1, in one pass
        rk4_history(stream, D, rk_qok ? 1 : 0, psi, S, Q);
        rk_qok = true;
      } else {
        dcopy(rkS[0], psi, D);
      }
    } else if (st == 2 && rk_cur == 2 && rk_qok) {
      const double a[3] = {1., 0.5, -0.5};
      const double* x[3] = {psi, Q[1], Q[0]};
      psi_combine(stream, D, 3, a, x, psi, nullptr);
    } else if (st == 3 && rk_cur == 3) {
      const double a[4] = {2., -1., q, -q};
      const double* x[4] = {psi, rkS[0], Q[2], Q[1]};
      psi_combine(stream, D, rk_qok ? 4 : 2, a, x, psi, nullptr);
    }
  }
  void rk_guess_before() {
    if (guess_mode() >= 4) { rk4_guess_before(); return; }
    const int st = rk_stage;
    if (st == 0 && rk_full) {
      const double a[4] = {-0.5, 1. / 3., 2. / 3., 0.5};
      const double* x[4] = {rkS[0], rkS[1], rkS[2], rkS[3]};
      psi_combine(stream, D, 4, a, x, psi, nullptr);
    } else if (st == 1 && rk_cur == 1) {
      if (rk_full) {
        const double a[5] = {4., 1., -1., -2., -1.};
        const double* x[5] = {psi, rkS[0], rkS[1], rkS[2], rkS[3]};
        psi_combine(stream, D, 5, a, x, psi, rkS[0]);
      } else {
        dcopy(rkS[0], psi, D);
      }
    } else if (st == 3 && rk_cur == 3) {
      const double a[2] = {2., -1.};
      const double* x[2] = {psi, rkS[0]};
      psi_combine(stream, D, 2, a, x, psi, nullptr);
    }
  }
  void rk_guess_after() {
    const int st = rk_stage;
    if (st != rk_cur % 4 || (st > 0 && rk_cur != st)) { rk_full = false; rk_qok = false; rk_cur = 0; return; }
    if (st >= 1) dcopy(rkS[st], psi, D);
    rk_cur = st + 1;
    if (st == 3) { rk_full = true; rk_cur = 0; }
  }
  // Mode 5: projection for successive right-hand sides (P. F. Fischer, Comput. Methods Appl. Mech.
  // Engrg. 163 (1998) 193-204), applied to the correction of mode 4's guess x0 (with Euler and RK2,
  // whose stiff modes alternate in sign from step to step: of the previous solution).  A basis X of
  // up to M = kProjVectors vectors spanning earlier stage solutions, with their images AX = As X and Gram
  // matrix G = Y^T AX, gives the start x0 + X c with c = G^-1 Y^T (b - As x0):
  //   symmetric As (PCG): Y = X, the As-norm best correction in the span;
  //   otherwise (BiCGSTAB): Y = AX, the least-squares residual.
  // Either is never worse than x0 in its norm.  A new solution x is appended after its solve as
  // x - X e with e = G^-1 Y^T As x = G^-1 (Y^T b - Y^T r); the Y^T b part is formed before the
  // solve in the pass that forms c, and the Y^T r part (within the solve's tolerance) is left in
  // place.  The appended vector's image is applied by the operator (one apply per solve): images
  // inferred from the solve's recursive residual, b - r - AX e, carried the residual's drift, and
  // the projection built on them degraded as the basis aged.  Its Gram column is then measured
  // exactly, so the basis is not kept orthogonal and needs no second pass or normalisation.
  // Each appended solution's coefficients in the basis are kept (W).  A full basis restarts from
  // the span of its newest kProjKeep solutions: their coefficient columns, orthonormalised on the
  // host in the inner product of G (every column of it measured: at an append, or after the last
  // restart), recombine the basis in one pass; a pass over the recombined vectors then measures
  // their Gram matrix (about I), so the start of every solve is projected on at least kProjKeep
  // earlier solutions.  24 kept 12 took fewer iterations than 16 kept 8 for less than the extra
  // upkeep (at 512^2, K = 5: 2.05 against 2.44 per solve, 463 against 438 FP64 Mdof-steps/s);
  // 32 kept 16 took 1.89 for the same throughput.
  static constexpr int kProjVectors = 24;
  static constexpr int kProjKeep = 12;
  static constexpr int kProjLdg = kProjVectors + 1;  // Gram column: M entries and the solution's norm
  static constexpr int kProjCoef = 2 * kProjMax;      // c and e of one segment
  static_assert(2 * kProjKeep <= kProjVectors && kProjVectors <= kProjMax, "basis restart sizes");
  struct ProjBasis {
    DBuf<double> X, AX;
    // per segment: Gram columns (M x kProjLdg), solution coefficients (M x M), restart combination
    // or measured Gram matrix (M x M)
    DBuf<double> G, W, C;
    int m = 0;       // vectors held
    int e_m = -1;    // basis size the coefficients e of the solve in progress were formed for
    int v_slot = -1; // the slot holding -X e for the append after the solve in progress (or -1)
    int gcol_pending = -1;  // the newest slot, whose Gram column the next projected solve forms
    int nseg = 0;    // segments the per-segment arrays were sized for
    int restarts = 0;
  };
  ProjBasis proj_fwd, proj_adj;     // stage solves; adjoint (transposed) solves
  ProjBasis* proj_cur = &proj_fwd;  // the basis of the solve in progress
  bool proj_force = false;          // project a solve that is not a stage solve (adjoint solves)
  DBuf<double> pcoef, ppart, ptmp;  // pcoef: per segment c (m) and e (m)
  // The projection's dof segments: one per ensemble instance (each projects on its own
  // coefficients), one for a single problem.
  ProjSeg pseg;
  DBuf<int32_t> pseg_blk, pseg_bbeg;
  DBuf<int64_t> pseg_sbeg;
  std::vector<int64_t> pseg_key;  // the segments' first dofs the map was built for
  bool proj_on() const { return guess_mode() >= 5; }
  bool proj_mr() const { return krylov != HOC_KRYLOV_PCG; }  // minimal-residual form
  // The segment map (rebuilt when the mesh or the instances change) and the per-instance norms
  // of an ensemble's convergence test.
  DBuf<double> ens_part, ens_rr, ens_bb;
  bool ens_pcg() const { return n_inst > 1 && krylov == HOC_KRYLOV_PCG; }
  // The batched ensemble PCG (ens_iter) when no Q1 vertex is shared between instances (else the
  // joint PCG with the per-instance stopping test).
  bool inst_coupled = false;
  bool ens_batched() const { return ens_pcg() && !inst_coupled; }
  DBuf<double> ens_in, ens_rz, ens_alpha, ens_beta;
  DBuf<int> ens_act;
  EnsKry ensk() {
    return EnsKry{ens_in.p, ens_rz.p, ens_alpha.p, ens_beta.p, ens_bb.p, ens_act.p, pseg.nseg};
  }
  void ensure_segments() {
    std::vector<int64_t> sb;
    if (n_inst > 1 && (int)inst_off.size() == n_inst + 1)
      for (int32_t e : inst_off) sb.push_back((int64_t)e * K2);
    else
      sb = {0, (int64_t)D};
    if (sb == pseg_key) return;
    std::vector<int32_t> bbeg, blk;
    proj_seg_layout(sb, bbeg, blk);
    pseg_sbeg.upload(sb); pseg_bbeg.upload(bbeg); pseg_blk.upload(blk);
    pseg.nseg = (int)sb.size() - 1;
    pseg.nb = (int)blk.size();
    pseg.blk_seg = pseg_blk.p; pseg.bbeg = pseg_bbeg.p; pseg.sbeg = pseg_sbeg.p;
    pseg_key = sb;
    ptmp.alloc(0);
    ens_part.alloc(pseg.nb); ens_rr.alloc(pseg.nseg); ens_bb.alloc(pseg.nseg);
    for (auto* b : {&ens_in, &ens_rz, &ens_alpha, &ens_beta}) b->alloc(pseg.nseg);
    ens_act.alloc(pseg.nseg);
    HIPCHK(hipMemsetAsync(ens_alpha.p, 0, pseg.nseg * sizeof(double), stream));
    HIPCHK(hipMemsetAsync(ens_act.p, 0, pseg.nseg * sizeof(int), stream));
  }
  void proj_alloc() {
    ensure_segments();
    if (ptmp.n < D) {  // t of proj_before; coefficients; partial sums
      ptmp.alloc(D);
      pcoef.alloc((size_t)pseg.nseg * kProjCoef);
      ppart.alloc((size_t)pseg.nb * std::max(2 * kProjMax + 2, kProjVectors * kProjVectors));
    }
  }
  void proj_before(const double* b) {
    ProjBasis& P = *proj_cur;
    P.e_m = -1;
    P.v_slot = -1;
    if (P.m == 0) return;
    proj_alloc();
    if (P.nseg != pseg.nseg) { P.m = 0; return; }
    double* t = ptmp.p;
    HIPCHK(launch_trace(K, PRE_COPY, N, T, psi.p, nullptr, nullptr, nullptr, 0., nullptr, 0, nullptr, P64.tr.p, stream));
    lap<double>(P64, EPI_RESID, psi.p, t, b, nullptr, nullptr, 0., 0., nullptr, nullptr);  // t = b - As x0
    const double* Y = proj_mr() ? P.AX.p : P.X.p;
    if (P.gcol_pending >= 0 && P.gcol_pending == P.m - 1) {  // Y^T t, Y^T b and the newest Gram column
      const int j = P.m - 1;
      proj_dots3(stream, D, pseg, P.m, Y, P.AX.p + (size_t)j * D, t, b, ppart, pcoef, kProjCoef,
                 P.G.p + (size_t)j * kProjLdg, kProjVectors * kProjLdg);
    } else {
      proj_dots2(stream, D, pseg, P.m, Y, t, b, ppart, pcoef, kProjCoef);  // Y^T t, Y^T b
    }
    P.gcol_pending = -1;
    proj_solve(stream, pseg, P.m, P.G.p, kProjLdg, kProjVectors * kProjLdg, pcoef, kProjCoef);    // c, e
    // with a free slot, the same pass leaves -X e there for the append after this solve
    const bool pre = P.m < kProjVectors;
    proj_combine(stream, D, pseg, P.m, P.X.p, pcoef, kProjCoef, psi, pre ? P.X.p + (size_t)P.m * D : nullptr);
    P.e_m = P.m;
    P.v_slot = pre ? P.m : -1;
  }
  // A full basis: keep, per segment, the span of the newest kProjKeep appended solutions
  // (columns M - kProjKeep .. M - 1 of W: with kProjKeep <= M / 2 these were all appended after
  // the last restart).  The new basis is orthonormalised in the inner product of the Gram matrix
  // held in G's columns, and the new vectors' Gram matrix, measured after the recombination,
  // replaces it; a segment whose newest solutions span fewer than the most vectors kept by any
  // segment gets zero vectors in the remaining slots (left out of its solves).
  void proj_restart(ProjBasis& P) {
    constexpr int M = kProjVectors, KP = kProjKeep, LG = kProjLdg;
    const int ns = pseg.nseg;
    P.m = 0;
    P.e_m = -1;
    P.v_slot = -1;
    P.gcol_pending = -1;
    ++P.restarts;
    if (KP == 0) return;
    std::vector<double> w((size_t)ns * M * M), gc((size_t)ns * M * LG), gm((size_t)M * M);
    HIPCHK(hipStreamSynchronize(stream));
    d2h(w.data(), P.W.p, w.size() * sizeof(double));
    d2h(gc.data(), P.G.p, gc.size() * sizeof(double));
    std::vector<std::vector<double>> cs(ns);
    int kk = 0;
    for (int sg = 0; sg < ns; ++sg) {
      const double* W = w.data() + (size_t)sg * M * M;
      const double* Gc = gc.data() + (size_t)sg * M * LG;
      for (int j = 0; j < M; ++j)  // the symmetric Gram matrix from its columns (upper triangle)
        for (int i = 0; i <= j; ++i) gm[(size_t)i * M + j] = gm[(size_t)j * M + i] = Gc[(size_t)j * LG + i];
      const int k = proj_restart_basis(M, KP, Gc, LG, gm.data(), W, cs[sg]);
      kk = std::max(kk, k);
    }
    if (kk == 0) return;
    // combinations (zero rows pad a segment to kk), then the recombined vectors' Gram matrix
    std::vector<double> ca((size_t)ns * M * M, 0.), gn((size_t)ns * kk * kk), gi((size_t)ns * M * LG, 0.);
    for (int sg = 0; sg < ns; ++sg) std::copy(cs[sg].begin(), cs[sg].end(), ca.begin() + (size_t)sg * M * M);
    HIPCHK(hipMemcpyAsync(P.C.p, ca.data(), ca.size() * sizeof(double), hipMemcpyHostToDevice, stream));
    proj_reseed(stream, D, pseg, M, kk, P.C.p, M * M, P.X.p, P.AX.p);
    proj_gram(stream, D, pseg, kk, proj_mr() ? P.AX.p : P.X.p, P.AX.p, ppart, P.C.p);  // kk x kk per segment
    HIPCHK(hipStreamSynchronize(stream));  // (ca is a local)
    d2h(gn.data(), P.C.p, gn.size() * sizeof(double));
    // the Gram columns of the new vectors, each with a unit solution norm (zero vectors padding a
    // segment have G_jj = 0 and are left out)
    for (int sg = 0; sg < ns; ++sg)
      for (int j = 0; j < kk; ++j) {
        double* col = gi.data() + (size_t)sg * M * LG + (size_t)j * LG;
        for (int i = 0; i <= j; ++i) col[i] = gn[(size_t)sg * kk * kk + (size_t)i * kk + j];
        col[j + 1] = 1.;
      }
    HIPCHK(hipMemcpyAsync(P.G.p, gi.data(), gi.size() * sizeof(double), hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));  // (gi is a local)
    P.m = kk;
  }
  void proj_after(const double* b) {
    constexpr int M = kProjVectors;
    ProjBasis& P = *proj_cur;
    proj_alloc();
    if (P.X.n < (size_t)M * D || P.nseg != pseg.nseg) {
      P.X.alloc((size_t)M * D); P.AX.alloc((size_t)M * D);
      const size_t ns = pseg.nseg;
      P.G.alloc(ns * M * kProjLdg); P.W.alloc(ns * M * M); P.C.alloc(ns * M * M);
      P.nseg = pseg.nseg;
      P.m = 0;
    }
    proj_flush_gcol(P);
    if (P.m >= M) proj_restart(P);
    const int j = P.m;
    if (j > 0 && P.e_m != j) {  // e = G^-1 Y^T b
      proj_dots2(stream, D, pseg, j, proj_mr() ? P.AX.p : P.X.p, b, b, ppart, pcoef, kProjCoef);
      proj_solve(stream, pseg, j, P.G.p, kProjLdg, M * kProjLdg, pcoef, kProjCoef);
    }
    // x = psi - X e into slot j, its exact image As x, and its Gram column
    double* xj = P.X.p + (size_t)j * D;
    double* axj = P.AX.p + (size_t)j * D;
    double* gcol = P.G.p + (size_t)j * kProjLdg;
    if (P.v_slot == j && P.e_m == j)  // slot j holds -X e from proj_before's pass
      proj_append(stream, D, pseg, j, proj_mr() ? 1 : 0, pcoef.p + j, kProjCoef, psi, b, xj, ppart, gcol + j + 1,
                  M * kProjLdg, P.W.p + (size_t)j * M, M * M, M);
    else
      proj_ortho(stream, D, pseg, j, proj_mr() ? 1 : 0, P.X.p, pcoef.p + j, kProjCoef, psi, b, xj, ppart,
                 gcol + j + 1, M * kProjLdg, P.W.p + (size_t)j * M, M * M, M);
    HIPCHK(launch_trace(K, PRE_COPY, N, T, xj, nullptr, nullptr, nullptr, 0., nullptr, 0, nullptr, P64.tr.p, stream));
    lap<double>(P64, EPI_STORE, xj, axj, nullptr, nullptr, nullptr, 0., 0., nullptr, nullptr);
    // its Gram column (Y_i . axj, i <= j) comes from the next projected solve's pass over Y
    // (proj_dots3), or from proj_flush_gcol when the basis is used otherwise first
    P.m = j + 1;
    P.e_m = -1;
    P.v_slot = -1;
    P.gcol_pending = j;
  }
  // The Gram column of the newest vector when it is still pending (its vector and image are in
  // their slots); a pending column of a basis since reset is dropped.
  void proj_flush_gcol(ProjBasis& P) {
    const int j = P.gcol_pending;
    P.gcol_pending = -1;
    if (j < 0 || j != P.m - 1) return;
    constexpr int M = kProjVectors;
    proj_gcol(stream, D, pseg, j, proj_mr() ? 1 : 0, proj_mr() ? P.AX.p : P.X.p, P.X.p + (size_t)j * D,
              P.AX.p + (size_t)j * D, ppart, P.G.p + (size_t)j * kProjLdg, M * kProjLdg);
  }
  // The adjoint's colouring, velocity map, buffers and recorded trajectory belong to one mesh and
  // one set-up: a new mesh or set-up rebuilds them on the next adjoint call.
  void reset_adjoint() {
    adj.init = false;
    adj.have_forward = false;
    adj.zn = 0;
    proj_adj.m = 0;
  }
  // Called when the state is set from outside: the histories restart, so that the same state gives
  // the same iterations (the projection basis would stay valid, holding solutions of As x = b only).
  void invalidate_guess() {
    proj_fwd.m = 0;
    n_hist = 0;
    rk_full = false;
    rk_qok = false;
    rk_cur = 0;
  }
  void guess_before_solve() {
    if (guess_mode() >= 3 && rk_stage >= 0 && !std::isnan(stage_tau)) { rk_guess_before(); return; }
    if (guess_mode() < 2 || guess_mode() == 5 || std::isnan(stage_tau)) return;
    const double tau = stage_tau;
    if (n_hist == 0 || tau == tau_a) return;  // no history, or the same stage time: the previous solution
    const bool ok = n_hist == 2 && tau > tau_a && tau - tau_a <= 1. && tau_a > tau_b && tau_a - tau_b <= 1.;
    psi_predict(stream, D, ok ? (tau - tau_a) / (tau_a - tau_b) : 0., ok ? 1 : 0, psi, psi_prev);
  }
  void guess_after_solve() {
    if (guess_mode() >= 3 && rk_stage >= 0 && !std::isnan(stage_tau)) { rk_guess_after(); return; }
    rk_full = false;
    rk_qok = false;
    rk_cur = 0;
    if (guess_mode() < 2 || std::isnan(stage_tau)) { n_hist = 0; return; }
    if (n_hist >= 1 && stage_tau == tau_a) return;
    if (n_hist == 0) { tau_a = stage_tau; n_hist = 1; return; }
    tau_b = tau_a;
    tau_a = stage_tau;
    n_hist = 2;
  }

  void poisson(const double* omega) {
    guess_before_solve();
    const int lb = lap_blocks(N);
    const int maxit = cfg.maxiter;
    const double tol2 = cfg.tol * cfg.tol;
    HIPCHK(hipMemcpyAsync(S.p + S_TOL2, &tol2, sizeof(double), hipMemcpyHostToDevice, stream));
    HIPCHK(launch_rhs(K, N, T, G, omega, rhs_bc ? rhs_bc : b_psi.p, bs, part, stream));
    const bool proj = proj_on() && (!std::isnan(stage_tau) || proj_force);
    if (proj) proj_before(bs);
    finalize(stream, PH_BB, part, lb, S, F, maxit);
    if (ens_pcg()) {
      // PCG: the iterations continue while some instance has ||r_i|| > tol ||b_i||
      ensure_segments();
      seg_norm2(stream, pseg, VA, bs, nullptr, ens_part.p, ens_bb.p);
    } else if (n_inst > 1) {
      // BiCGSTAB: ||b - A psi|| <= tol min_i ||b_i|| bounds every instance's own relative residual
      // (instances with a zero right-hand side are left out of the minimum)
      inst_norms(stream, VA, n_inst, d_ioff, K2, bs, d_inorm);
      std::vector<double> bi(n_inst);
      double bb = 0.;
      HIPCHK(hipMemcpyAsync(bi.data(), d_inorm.p, n_inst * sizeof(double), hipMemcpyDeviceToHost, stream));
      HIPCHK(hipMemcpyAsync(&bb, S.p + S_BB, sizeof(double), hipMemcpyDeviceToHost, stream));
      HIPCHK(hipStreamSynchronize(stream));
      double mx = 0., mn = HUGE_VAL;
      for (double v : bi) mx = std::max(mx, v);
      for (double v : bi) if (v > 1e-30 * mx) mn = std::min(mn, v);
      const double t2 = (bb > 0. && mn < HUGE_VAL) ? tol2 * mn / bb : tol2;
      HIPCHK(hipMemcpyAsync(S.p + S_TOL2, &t2, sizeof(double), hipMemcpyHostToDevice, stream));
      HIPCHK(hipStreamSynchronize(stream));
    }
    outer_residual();
    zero_if_flag<double>(stream, D, psi, F);
    double hs[S_COUNT];
    int hf[F_COUNT];
    auto read_state = [&](const double* Sp, const int* Fp) {
      HIPCHK(hipMemcpyAsync(hs, Sp, sizeof(hs), hipMemcpyDeviceToHost, stream));
      HIPCHK(hipMemcpyAsync(hf, Fp, sizeof(hf), hipMemcpyDeviceToHost, stream));
      HIPCHK(hipStreamSynchronize(stream));
    };
    // FP64: the FP64 V-cycle; mixed: the FP32 V-cycle (precond), the Krylov vectors, the operator
    // application and the stopping test in FP64 either way.
    krylov_solve(P64, psi.p, P64.r.p, S.p, F.p, maxit, maxit);
    read_state(S, F);
    const int iters = hf[F_ITER], outer = 0;
    // statistics
    last_iters = iters;
    last_outer = outer;
    if (!std::isfinite(hs[S_BB]) || !std::isfinite(hs[S_RR])) {
      last_rel = std::nan("");
      throw StatusError(HOC_ENAN, "NaN/Inf in the Poisson right-hand side or residual");
    }
    last_rel = hs[S_BB] > 0. ? std::sqrt(hs[S_RR] / hs[S_BB]) : 0.;
    if (hf[F_ZERO]) last_rel = 0.;
    total_iters += last_iters;
    ++n_solves;
    if (proj && !hf[F_ZERO]) proj_after(bs);
    guess_after_solve();
    if (!std::isfinite(last_rel)) throw StatusError(HOC_ENAN, "NaN/Inf in the Poisson solve");
    if (cfg.fail_on_noconv && last_rel > cfg.tol)
      throw StatusError(HOC_ENOCONV, "Poisson solve did not converge: rel residual " + std::to_string(last_rel));
  }

  // ---- node-averaged output and sample points ---------------------------------------------
  DBuf<int32_t> d_elnodes, d_nptr, d_ninc, d_sel;
  DBuf<double> d_bmet, d_elemout, d_nodeout, d_sshape, d_sval;
  bool smooth_ready = false;
  int n_sp = 0;
  std::vector<double> sp_xy, sp_loc;
  std::vector<int32_t> sp_elem;

  // Element node lists, node incidences in element order and the boundary faces' metrics,
  // uploaded on the first output call.
  void smooth_setup() {
    if (smooth_ready) return;
    const int L2 = mesh.lnod * mesh.lnod;
    d_elnodes.upload(mesh.elem);
    std::vector<int32_t> ptr((size_t)mesh.n_nodes + 1, 0), inc((size_t)N * L2);
    for (int32_t nd : mesh.elem) ptr[(size_t)nd + 1]++;
    for (int i = 0; i < mesh.n_nodes; ++i) ptr[i + 1] += ptr[i];
    std::vector<int32_t> pos(ptr.begin(), ptr.end() - 1);
    for (int el = 0; el < N; ++el)
      for (int l = 0; l < L2; ++l) inc[pos[mesh.elem[(size_t)el * L2 + l]]++] = el * L2 + l;
    d_nptr.upload(ptr);
    d_ninc.upload(inc);
    std::vector<double> bm((size_t)std::max(E, 1) * K * 4, 0.);
    for (int e = 0; e < E; ++e)
      for (int k = 0; k < K; ++k) {
        const size_t q = ((size_t)mesh.be_elem[e] * 4 + mesh.be_ijp[e]) * K + k;
        double* o = &bm[((size_t)e * K + k) * 4];
        o[0] = geo.fxxi[q]; o[1] = geo.fxeta[q]; o[2] = geo.fyxi[q]; o[3] = geo.fyeta[q];
      }
    d_bmet.upload(bm);
    d_elemout.alloc((size_t)N * L2 * 4);
    d_nodeout.alloc((size_t)mesh.n_nodes * 4);
    smooth_ready = true;
  }
  // What the reference's main loop runs before each dump (src/HO_2D.cpp:1411-1414): boundary
  // values at the current time, the stream function and the velocity of the current vorticity.
  void refresh_fields() {
    update_bcs(0.);
    poisson(w);
    HIPCHK(launch_trace(K, PRE_COPY, N, T, psi, nullptr, nullptr, nullptr, 0., nullptr, 0, nullptr, P64.tr.p, stream));
    HIPCHK(launch_velocity(K, N, T, G, B, psi, P64.tr, w, vel, U, trw, tru, 1. / (cfg.dt * re_inv), I, stream));
  }
  // Node-averaged vorticity, psi and velocity into d_nodeout (save_smooth_vtk,
  // src/HO_2D.cpp:3369-3531).
  void node_fields(bool refresh) {
    smooth_setup();
    if (refresh) refresh_fields();
    SmoothArgs A{};
    A.N = N; A.K = K; A.L = mesh.lnod;
    for (int k = 0; k < K; ++k)  // sps_gps_basis (src/HO_2D.cpp:571-583)
      for (int l = 0; l < A.L; ++l) {
        double num = 1., den = 1.;
        for (int j = 0; j < K; ++j) {
          if (j == k) continue;
          den *= tab.sps[k] - tab.sps[j];
          num *= tab.gps[l] - tab.sps[j];
        }
        A.sgb[k][l] = num / den;
      }
    for (int j = 0; j < A.L; ++j)
      for (int i = 0; i < A.L; ++i) A.slot[j * A.L + i] = gmsh_slot(A.L, i, j);
    A.w = w; A.psi = psi; A.vel = vel; A.nb = d_nb;
    A.bvort = b_vort; A.bnvel = b_nvel; A.bpvel = b_pvel; A.bmet = d_bmet; A.elemout = d_elemout;
    HIPCHK(launch_smooth_elem(A, stream));
    HIPCHK(launch_smooth_nodes(mesh.n_nodes, d_nptr, d_ninc, d_elemout, d_nodeout, stream));
  }
  void set_sample_points(int n, const double* xy) {
    std::vector<double> shape;
    try {
      locate_sample_points(mesh, tab, n, xy, sp_elem, sp_loc, shape);
    } catch (const std::runtime_error& e) {
      throw StatusError(HOC_EINVAL, e.what());
    }
    sp_xy.assign(xy, xy + 2 * (size_t)n);
    n_sp = n;
    if (n > 0) { d_sel.upload(sp_elem); d_sshape.upload(shape); d_sval.alloc(3 * (size_t)n); }
  }
  void sample_values(bool refresh, double* out) {
    node_fields(refresh);
    HIPCHK(launch_sample(n_sp, mesh.lnod * mesh.lnod, d_sel, d_sshape, d_elnodes, d_nodeout, d_sval, stream));
    if (n_sp > 0) d2h(out, d_sval.p, 3 * (size_t)n_sp * sizeof(double));
    else HIPCHK(hipStreamSynchronize(stream));
  }

  // Validates a mesh given as arrays or read from a file and makes it the handle's mesh.
  void accept_mesh(Mesh&& m) {
    for (int32_t v : m.elem) if (v < 0 || v >= m.n_nodes) throw StatusError(HOC_EINVAL, "element node id out of range");
    for (auto& e : m.edges)
      for (int32_t v : e) if (v < 0 || v >= m.n_nodes) throw StatusError(HOC_EINVAL, "edge node id out of range");
    for (int32_t b : m.bedge_bdry)
      if (b < 0 || b >= (int)m.bnames.size()) throw StatusError(HOC_EINVAL, "boundary id out of range");
    mesh = std::move(m);
    mesh_loaded = true;
    ready = false;
    reset_adjoint();
    n_inst = 1;
    inst_off.clear(); inst_re.clear(); inst_dt.clear(); inst_time.clear();
    edge_nvel_in.clear(); edge_has_nvel.clear();
  }

  double ramp(double t) const { return bch.time_dependent ? 2. * std::fabs(std::sin(M_PI / 40. * t)) : 1.; }

  // Boundary values at stage fraction c of the step (each instance at its own clock), or at
  // the absolute time t_abs for all instances when t_abs is given.
  void update_bcs(double c, double t_abs = std::nan("")) {
    if (cfg.problem_type == 11) {
      const double t = std::isnan(t_abs) ? time + c * cfg.dt : t_abs;
      const double span = time_end - time_start;
      const double fac1 = span > 0. ? (t - time_start) / span : 0.;
      HIPCHK(launch_update_bcs_hybrid(E * K, K, B, H, fac1, 1. - fac1, stream));
      HIPCHK(launch_hybrid_blend(D, w, H, fac1, 1. - fac1, stream));
      return;
    }
    if (n_inst > 1 && std::isnan(t_abs)) {
      std::vector<double> r(n_inst);
      for (int i = 0; i < n_inst; ++i) r[i] = ramp(inst_time[i] + c * inst_dt[i]);
      HIPCHK(hipMemcpyAsync(d_iramp.p, r.data(), n_inst * sizeof(double), hipMemcpyHostToDevice, stream));
      HIPCHK(launch_update_bcs(E * K, K, B, 1., I, stream));
    } else {
      InstDev I0 = I;
      I0.ramp = nullptr;
      HIPCHK(launch_update_bcs(E * K, K, B, ramp(std::isnan(t_abs) ? time + c * cfg.dt : t_abs), I0, stream));
    }
  }

  // Euler_time_integrate (src/HO_2D.cpp:1806-1839) on input win; boundary values already set.
  double* capture_psi = nullptr;  // when set, RK4 stage s copies its solution to capture_psi + s D
  void stage_body(const double* win, const StageOut& so) {
    poisson(win);
    if (capture_psi && rk_stage >= 0) dcopy(capture_psi + (size_t)rk_stage * D, psi, D);
    HIPCHK(launch_trace(K, PRE_COPY, N, T, psi, nullptr, nullptr, nullptr, 0., nullptr, 0, nullptr, P64.tr.p, stream));
    // the Cartesian velocity is an output of the step's last stage only (what the getters and
    // writers read); intermediate RK stages need the contravariant fluxes U alone
    double* v_out = (so.next && !so.final_mode) ? nullptr : vel.p;
    HIPCHK(launch_velocity(K, N, T, G, B, psi, P64.tr, win, v_out, U, trw, tru, 1. / (cfg.dt * re_inv), I, stream));
    HIPCHK(launch_stage(K, cfg.advection_type, N, T, G, B, win, U, trw, tru, re_inv, cfg.dt, I, so, stream));
  }
  void stage(const double* win, double c, const StageOut& so) {
    update_bcs(c);
    stage_tau = (double)step_count + c;
    try {
      stage_body(win, so);
    } catch (...) {
      stage_tau = std::nan("");
      rk_stage = -1;
      n_hist = 0;
      rk_full = false;
      rk_qok = false;
      rk_cur = 0;
      throw;
    }
    stage_tau = std::nan("");
  }

  // solve_advection_diffusion (src/HO_2D.cpp:1465-1531)
  void rk_step() {
    StageOut so;
    if (cfg.time_integration_type == 1) {
      so.base = w; so.final_mode = 1;
      stage(w, 0., so);
    } else if (cfg.time_integration_type == 2) {
      so.k = k1; so.next = vort; so.c = 1.; so.base = w;
      stage(w, 0., so);
      StageOut s2; s2.base = w; s2.final_mode = 2; s2.k1 = k1;
      stage(vort, 1., s2);
    } else if (cfg.time_integration_type == 4) {
      const double c[4] = {0., 0.5, 0.5, 1.};
      double* ks[3] = {k1, k2, k3};
      rk_stage = 0;
      stage_rk4(w, c[0], ks[0], 0.5);
      rk_stage = 1;
      stage_rk4(vort, c[1], ks[1], 0.5);
      rk_stage = 2;
      stage_rk4(vort, c[2], ks[2], 1.0);
      StageOut s4; s4.base = w; s4.final_mode = 4; s4.k1 = k1; s4.k2 = k2; s4.k3 = k3;
      rk_stage = 3;
      stage(vort, c[3], s4);
      rk_stage = -1;
    } else {
      throw StatusError(HOC_EINVAL, "time_integration_type must be 1, 2 or 4");
    }
    time += cfg.dt;
    ++step_count;
    for (int i = 0; i < n_inst && n_inst > 1; ++i) inst_time[i] += inst_dt[i];
  }
  void stage_rk4(const double* win, double c, double* kout, double cnext) {
    StageOut so; so.k = kout; so.next = vort; so.c = cnext; so.base = w;
    stage(win, c, so);
  }

  // Hybrid tables: edge kinds, open-boundary order, zero start/end values and targets.
  void setup_hybrid() {
    H = HybridDev{};
    open_edges.clear();
    time_start = time_end = 0.;
    if (cfg.problem_type != 11) return;
    for (int b = 0; b < (int)mesh.bnames.size(); ++b)
      if (mesh.bnames[b] == "open")
        for (int e = 0; e < E; ++e) if (mesh.bedge_bdry[e] == b) open_edges.push_back(e);
    const size_t EK = std::max<size_t>((size_t)E * K, 1);
    d_hkind.upload(bch.hybrid_kind);
    d_hinlet.upload(bch.inlet_nvel);
    for (auto* b : {&d_vn0, &d_vn1, &d_vp0, &d_vp1, &d_vw0, &d_vw1}) b->alloc(EK);
    for (auto* b : {&d_hw0, &d_hw1, &d_hwgt}) b->alloc(D);
    h_vn1.assign(EK, 0.); h_vp1.assign(EK, 0.); h_vw1.assign(EK, 0.);
    H = HybridDev{d_hkind, d_hinlet, d_esign, d_swdiff, d_vn0, d_vn1, d_vp0, d_vp1, d_vw0, d_vw1, d_hw0, d_hw1, d_hwgt};
  }
  void need_hybrid() const {
    if (cfg.problem_type != 11) throw StatusError(HOC_ESTATE, "hybrid driver calls need problem_type 11");
  }
  // getopenpts_d (src/HO_2D.cpp:4007-4046): K points along each open edge, edge node order
  void open_points(double* xy) const {
    const int L = mesh.lnod;
    for (size_t i = 0; i < open_edges.size(); ++i) {
      const auto& nd = mesh.edges[open_edges[i]];  // tensor order after connect()
      for (int j = 0; j < K; ++j) {
        double x = 0., y = 0.;
        for (int l = 0; l < L; ++l) {
          x += tab.gsb[l][j] * mesh.xy[2 * (size_t)nd[l]];
          y += tab.gsb[l][j] * mesh.xy[2 * (size_t)nd[l] + 1];
        }
        xy[2 * (i * K + j)] = x; xy[2 * (i * K + j) + 1] = y;
      }
    }
  }
  // setopenvels_d (src/HO_2D.cpp:4050-4117): start = end, then the new Cartesian velocities
  // as outward-normal and tangential components with the owner face's metrics
  void set_open_velocities(const double* uv) {
    const size_t EK = std::max<size_t>((size_t)E * K, 1);
    HIPCHK(hipMemcpyAsync(d_vn0.p, d_vn1.p, EK * sizeof(double), hipMemcpyDeviceToDevice, stream));
    HIPCHK(hipMemcpyAsync(d_vp0.p, d_vp1.p, EK * sizeof(double), hipMemcpyDeviceToDevice, stream));
    for (size_t i = 0; i < open_edges.size(); ++i) {
      const int e = open_edges[i];
      const int el = mesh.be_elem[e], ijp = mesh.be_ijp[e];
      const double coeff = 2. * (ijp % 2) - 1.;
      for (int j = 0; j < K; ++j) {
        const size_t q = ((size_t)el * 4 + ijp) * K + j;
        const double u = uv[2 * (i * K + j)], v = uv[2 * (i * K + j) + 1];
        double nx, ny;
        if (ijp / 2 == 1) { nx = -geo.fyxi[q]; ny = geo.fxxi[q]; }
        else { nx = geo.fyeta[q]; ny = -geo.fxeta[q]; }
        nx *= coeff / geo.fAn[q];
        ny *= coeff / geo.fAn[q];
        h_vn1[(size_t)e * K + j] = nx * u + ny * v;
        h_vp1[(size_t)e * K + j] = nx * v - ny * u;
      }
    }
    h2d(d_vn1.p, h_vn1.data(), EK * sizeof(double));
    h2d(d_vp1.p, h_vp1.data(), EK * sizeof(double));
  }
  // setopenvort_d (src/HO_2D.cpp:4120-4141)
  void set_open_vorticity(const double* wv) {
    const size_t EK = std::max<size_t>((size_t)E * K, 1);
    HIPCHK(hipMemcpyAsync(d_vw0.p, d_vw1.p, EK * sizeof(double), hipMemcpyDeviceToDevice, stream));
    for (size_t i = 0; i < open_edges.size(); ++i)
      for (int j = 0; j < K; ++j) h_vw1[(size_t)open_edges[i] * K + j] = wv[K * i + j];
    h2d(d_vw1.p, h_vw1.data(), EK * sizeof(double));
  }
  // setsolnvort_d (src/HO_2D.cpp:4144-4166)
  void set_vorticity_target(const double* wv) {
    HIPCHK(hipMemcpyAsync(d_hw0.p, d_hw1.p, D * sizeof(double), hipMemcpyDeviceToDevice, stream));
    h2d(d_hw1.p, wv, D * sizeof(double));
  }

  // ---- discrete adjoint of the RK4 trajectory -------------------------------------------
  // One stage maps omega_s to k_s = dt R(omega_s), with psi_s = A^-1 b(omega_s) and
  // b = -omega J - (boundary source).  Its transpose is applied as
  //   omega_bar  = (dk/domega)^T lam                 (probes: omega and its face traces, 1-ring)
  //   U_bar      = (dk/dU)^T lam,  diff_bar = (dk/ddiff)^T lam   (probes: velocity, wall flux)
  //   pvel_bar  += sign idr diff_bar                 (no-slip walls: diff = (vjump + sign pvel) idr)
  //   psi_bar    = (dU/dpsi)^T U_bar + (ddiff/dpsi)^T diff_bar   (assembled once: the map is linear
  //                                                              and state-independent)
  //   omega_bar += J W z,   As z = psi_bar           (A^-T = -W As^-1, As symmetric)
  // the probes as coloured, batched launches of the frozen-switch stage (launch_stage mode 2).
  struct Adj {
    bool init = false;
    int nc1 = 0;
    DBuf<int32_t> col1, r1p, r1i, own_el;
    DBuf<int32_t> members;            // elements sorted by colour, colour c at [moff[c], moff[c + 1])
    std::vector<int> moff;
    DBuf<int8_t> own_f;
    DBuf<double> ws[4], Us[4], trws[4], trus[4];  // recorded stages of one step
    DBuf<int8_t> sig[4];
    // probe batches (nb = max(K^2, 4K) probes): seeds and their traces, products, velocities
    DBuf<double> seed, seed_tr, y, p, dU, dvel, dtrw, dtru, lin_diff, lin_vj, pvel_seed;
    DBuf<double> zero_d, zero_tr, zero_e, lin_nvel, lin_vort, lin_adv;
    DBuf<double> lam, xbar, psibar, z, fake, wbar, kbar[4], pvbar;
    DBuf<double> Mv, Md, Ubar, diffbar, zero_u, zero_tu;  // assembled velocity map and its adjoints
    // adjoint solves: the previous solution z (in z) and the one before (zprev) with their stage
    // times, for a guess extrapolated in the reversed stage time
    DBuf<double> zprev;
    double ztau_a = 0., ztau_b = 0.;
    int zn = 0;
    DBuf<double> ckbuf, segbuf, segpsi, save_w, save_psi, save_vel;
    // schedule: 0 uniform segments, 1 every step recorded by the forward run, 2 binomial
    int sched = 0;
    DBuf<double> trajw, trajpsi;   // sched 1: omega_n and the four stage solutions of every step
    DBuf<double> snap;             // sched 2: omega snapshots
    int nsnap = 0;
    std::vector<int> snap_pos;     // sched 2: step held by each snapshot slot (-1: free)
    std::vector<int32_t> bsplit;   // sched 2: optimal first advance for (range, free slots)
    int bsplit_L = 0, bsplit_S = 0;
    int nsteps = 0, every = 0;
    double t0 = 0.;
    bool have_forward = false;
    BCDev Bz{}, Bl{};
    BCDev Bl_seed() const {  // homogeneous boundary values with the wall flux from the edge seeds
      BCDev b = Bl;
      b.diff = pvel_seed;
      return b;
    }
  } adj;

  void adjoint_check() const {
    if (!ready) throw StatusError(HOC_ESTATE, "call hoc_setup first");
    if (mixed) throw StatusError(HOC_EINVAL, "the adjoint needs precision HOC_FP64");
    if (n_inst > 1) throw StatusError(HOC_EINVAL, "the adjoint runs on one instance");
    if (cfg.problem_type != 1 || bch.time_dependent)
      throw StatusError(HOC_EINVAL, "the adjoint needs problem_type 1 (fixed walls, time-independent boundary values)");
    if (cfg.advection_type != 1) throw StatusError(HOC_EINVAL, "the adjoint needs advection_type 1");
    if (cfg.time_integration_type != 4) throw StatusError(HOC_EINVAL, "the adjoint needs time_integration_type 4");
    if (krylov != HOC_KRYLOV_PCG)
      throw StatusError(HOC_EINVAL, "the adjoint needs a symmetric Poisson operator (affine elements)");
    for (int e = 0; e < E; ++e)
      if (bch.sw_psi[e] != 0) throw StatusError(HOC_EINVAL, "the adjoint needs Dirichlet stream-function walls");
  }

  // Greedy colouring in which elements within face distance dist of each other differ.
  int color_within(int dist, std::vector<int32_t>& color) const {
    color.assign(N, -1);
    std::vector<int> mark(N, -1), used;
    std::vector<int> front, next, seen;
    int nc = 0;
    for (int e = 0; e < N; ++e) {
      seen.clear();
      front.assign(1, e);
      mark[e] = e;
      for (int d = 0; d < dist; ++d) {
        next.clear();
        for (int u : front)
          for (int f = 0; f < 4; ++f) {
            const int v = mesh.nb[(size_t)u * 4 + f];
            if (v >= 0 && mark[v] != e) { mark[v] = e; next.push_back(v); seen.push_back(v); }
          }
        front.swap(next);
      }
      for (int v : seen)
        if (color[v] >= 0) used[color[v]] = e;
      int c = 0;
      while (c < (int)used.size() && used[c] == e) ++c;
      if (c == (int)used.size()) used.push_back(-1);
      color[e] = c;
      nc = std::max(nc, c + 1);
    }
    return nc;
  }
  // Elements within face distance dist of each element (itself included), as CSR.
  void rings(int dist, std::vector<int32_t>& ptr, std::vector<int32_t>& idx) const {
    ptr.assign(1, 0);
    idx.clear();
    std::vector<int> mark(N, -1), front, next;
    for (int e = 0; e < N; ++e) {
      mark[e] = e;
      idx.push_back(e);
      front.assign(1, e);
      for (int d = 0; d < dist; ++d) {
        next.clear();
        for (int u : front)
          for (int f = 0; f < 4; ++f) {
            const int v = mesh.nb[(size_t)u * 4 + f];
            if (v >= 0 && mark[v] != e) { mark[v] = e; next.push_back(v); idx.push_back(v); }
          }
        front.swap(next);
      }
      ptr.push_back((int32_t)idx.size());
    }
  }

  void adjoint_setup() {
    adjoint_check();
    if (adj.init) return;
    std::vector<int32_t> c1, p1, i1;
    adj.nc1 = color_within(2, c1);
    rings(1, p1, i1);
    adj.col1.upload(c1);
    adj.r1p.upload(p1); adj.r1i.upload(i1);
    {
      std::vector<int32_t> mem;
      adj.moff.assign(1, 0);
      for (int c = 0; c < adj.nc1; ++c) {
        for (int e = 0; e < N; ++e)
          if (c1[e] == c) mem.push_back(e);
        adj.moff.push_back((int)mem.size());
      }
      adj.members.upload(mem);
    }
    std::vector<int32_t> oe(std::max(E, 1), 0);
    std::vector<int8_t> of(std::max(E, 1), -1);
    for (int e = 0; e < E; ++e) { oe[e] = mesh.be_elem[e]; of[e] = mesh.be_ijp[e]; }
    adj.own_el.upload(oe); adj.own_f.upload(of);
    const size_t TR = (size_t)4 * N * 2 * K, TU = (size_t)4 * N * K, EK = std::max<size_t>((size_t)E * K, 1);
    const size_t NB = std::max(K2, 4 * K);
    for (int s4 = 0; s4 < 4; ++s4) {
      adj.ws[s4].alloc(D); adj.Us[s4].alloc(2 * D); adj.trws[s4].alloc(TR); adj.trus[s4].alloc(TU);
      adj.sig[s4].alloc((size_t)4 * N * K); adj.kbar[s4].alloc(D);
    }
    for (DBuf<double>* b : {&adj.zero_d, &adj.lam, &adj.xbar, &adj.psibar, &adj.z, &adj.fake, &adj.wbar, &adj.save_w,
                            &adj.save_psi})
      b->alloc(D);
    adj.save_vel.alloc(2 * D);
    adj.zero_tr.alloc(TR);
    adj.seed.alloc(NB * D); adj.y.alloc(NB * D); adj.seed_tr.alloc(NB * TR); adj.dtrw.alloc(NB * TR);
    adj.dU.alloc(NB * 2 * D); adj.dvel.alloc(NB * 2 * D); adj.dtru.alloc(NB * TU);
    adj.p.alloc(NB * std::max(N, 1));
    adj.lin_diff.alloc(NB * EK); adj.lin_vj.alloc(NB * EK); adj.pvel_seed.alloc((size_t)4 * K * EK);
    // seeds of the wall-flux probes are edge arrays like the wall velocity's
    for (DBuf<double>* b : {&adj.zero_e, &adj.lin_vort, &adj.lin_adv, &adj.pvbar}) b->alloc(EK);
    adj.lin_nvel.alloc(EK);
    HIPCHK(hipMemcpy(adj.lin_nvel.p, b_nvel.p, EK * sizeof(double), hipMemcpyDeviceToDevice));
    // homogeneous boundary values: zero psi (null), zero wall velocity and flux; the walls' normal
    // velocity (a constant) multiplies the boundary vorticity's variation
    adj.Bl = B;
    adj.Bl.psi = nullptr;
    adj.Bl.diff = adj.lin_diff; adj.Bl.vjump = adj.lin_vj; adj.Bl.pvel = adj.zero_e;
    adj.Bl.nvel = adj.lin_nvel; adj.Bl.vort = adj.lin_vort; adj.Bl.adv = adj.lin_adv;
    adj.Bz = adj.Bl;
    adj.Bz.diff = adj.zero_e;
    adj.Mv.alloc((size_t)N * 5 * K2 * 2 * K2);
    adj.Md.alloc(EK * K2);
    adj.Ubar.alloc(2 * D); adj.zero_u.alloc(2 * D); adj.zero_tu.alloc(TU); adj.diffbar.alloc(EK);
    adj.zprev.alloc(D);
    // the velocity map psi -> (U, wall flux): K^2 probes per colour of the distance-2 colouring
    // (U of an element depends on psi of its face ring), velocity with homogeneous boundary values
    VelBatch vb;
    vb.psi = D; vb.trpsi = TR; vb.vel = 2 * D; vb.U = 2 * D; vb.trw = TR; vb.tru = TU; vb.vjump = EK; vb.diff = EK;
    for (int c = 0; c < adj.nc1; ++c) {
      HIPCHK(launch_adj_seed_batch(N, K, adj.col1, c, T, adj.seed, adj.seed_tr, stream));
      HIPCHK(launch_velocity_batch(K, N, T, G, adj.Bl, adj.seed, adj.seed_tr, adj.zero_d, adj.dvel, adj.dU, adj.dtrw,
                                   adj.dtru, 1. / (cfg.dt * re_inv), I, K2, vb, stream));
      HIPCHK(launch_adj_assemble_vel(N, E, K, adj.col1, c, d_nb, adj.own_el, adj.dU, adj.lin_diff, adj.Mv, adj.Md,
                                     stream));
    }
    HIPCHK(hipStreamSynchronize(stream));
    adj.init = true;
  }

  void dcopy(double* dst, const double* src, size_t n) {
    HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, stream));
  }

  // One RK4 step from w = omega_n recording every stage's input, velocity, traces and upwind
  // switches (w itself is not advanced).  stage_psi: the stage solutions of the same step kept by
  // the segment's recompute (their inputs are recomputed here by the same kernels), or null to
  // solve.
  void adj_record_step(const double* stage_psi) {
    const double c[4] = {0., 0.5, 0.5, 1.};
    const double cn[3] = {0.5, 0.5, 1.};
    dcopy(adj.ws[0], w, D);
    for (int st = 0; st < 4; ++st) {
      update_bcs(c[st]);
      if (stage_psi) dcopy(psi, stage_psi + (size_t)st * D, D);
      else poisson(adj.ws[st]);
      HIPCHK(launch_trace(K, PRE_COPY, N, T, psi, nullptr, nullptr, nullptr, 0., nullptr, 0, nullptr, P64.tr.p, stream));
      HIPCHK(launch_velocity(K, N, T, G, B, psi, P64.tr, adj.ws[st], vel, adj.Us[st], adj.trws[st], adj.trus[st],
                             1. / (cfg.dt * re_inv), I, stream));
      StageOut so;
      so.k = kscratch;
      so.sig = adj.sig[st];
      if (st < 3) { so.next = adj.ws[st + 1]; so.c = cn[st]; so.base = w; }
      HIPCHK(launch_stage(K, 1, N, T, G, B, adj.ws[st], adj.Us[st], adj.trws[st], adj.trus[st], re_inv, cfg.dt, I, so,
                          stream, 1));
    }
  }

  // xbar = (dk_s/domega_s)^T lam for recorded stage st, adding (dk_s/dpvel)^T lam to adj.pvbar.
  // The K^2 probes of a colour (and the 4K wall-velocity probes) run as one batched launch each.
  void adj_stage(int st, const double* lam, double* xbar, double tau = std::nan("")) {
    const double idr = 1. / (cfg.dt * re_inv);
    const size_t TR = (size_t)4 * N * 2 * K, TU = (size_t)4 * N * K, EK = std::max<size_t>((size_t)E * K, 1);
    StageOut so;
    so.k = adj.y;
    so.bs_k = D;
    so.sig = adj.sig[st];
    // (1) omega and its face traces, with psi (so U and the wall flux) fixed
    so.nbatch = K2;
    so.omega_adv = adj.seed; so.bs_wa = D;
    so.trw_adv = adj.seed_tr; so.bs_trwa = TR;
    so.bs_w = D; so.bs_trw = TR;
    for (int c = 0; c < adj.nc1; ++c) {
      HIPCHK(launch_adj_seed_batch(N, K, adj.col1, c, T, adj.seed, adj.seed_tr, stream));
      HIPCHK(launch_stage(K, 1, N, T, G, adj.Bz, adj.seed, adj.Us[st], adj.seed_tr, adj.trus[st], re_inv, cfg.dt, I,
                          so, stream, 2));
      HIPCHK(launch_adj_dot_batch(N, K2, K2, lam, adj.y, adj.p, stream));
      HIPCHK(launch_adj_gather_batch(N, K2, adj.moff[c + 1] - adj.moff[c], adj.members.p + adj.moff[c], adj.r1p,
                                     adj.r1i, adj.p, xbar, K2, 0, stream));
    }
    // (2) velocity: K^2 probes per colour and component (the stage depends on U over the face ring),
    //     the stage advecting the recorded vorticity with the seeded velocity
    so.omega_adv = adj.ws[st]; so.bs_wa = 0;
    so.trw_adv = adj.trws[st]; so.bs_trwa = 0;
    so.bs_w = 0; so.bs_trw = 0; so.bs_U = 2 * D; so.bs_tru = TU; so.bs_diff = 0;
    so.nbatch = K2;
    so.adv_only = 1;  // zero vorticity and wall flux for the diffusion term
    for (int c = 0; c < adj.nc1; ++c)
      for (int comp = 0; comp < 2; ++comp) {
        HIPCHK(launch_adj_seed_u_batch(N, K, adj.col1, c, comp, T, adj.dU, adj.dtru, stream));
        HIPCHK(launch_stage(K, 1, N, T, G, adj.Bz, adj.zero_d, adj.dU, adj.zero_tr, adj.dtru, re_inv, cfg.dt, I, so,
                            stream, 2));
        HIPCHK(launch_adj_dot_batch(N, K2, K2, lam, adj.y, adj.p, stream));
        HIPCHK(launch_adj_gather_batch(N, K2, adj.moff[c + 1] - adj.moff[c], adj.members.p + adj.moff[c], adj.r1p,
                                       adj.r1i, adj.p, adj.Ubar, 2 * K2, comp * K2,
                                       stream));
      }
    // (3) wall flux: probe f K + j seeds edge flux point j of the edges owned through face f, which
    //     reaches only the owner elements; then the wall velocity's share of it
    HIPCHK(hipMemsetAsync(adj.diffbar.p, 0, EK * sizeof(double), stream));
    if (E > 0) {
      HIPCHK(launch_adj_edge_seed_batch(E, K, adj.own_f, adj.pvel_seed, stream));
      so.omega_adv = adj.zero_d; so.trw_adv = adj.zero_tr;
      so.bs_U = 0; so.bs_tru = 0; so.bs_diff = EK;
      so.adv_only = 0;
      so.nbatch = 4 * K;
      HIPCHK(launch_stage(K, 1, N, T, G, adj.Bl_seed(), adj.zero_d, adj.zero_u, adj.zero_tr, adj.zero_tu, re_inv,
                          cfg.dt, I, so, stream, 2));
      HIPCHK(launch_adj_dot_batch(N, K2, 4 * K, lam, adj.y, adj.p, stream));
      HIPCHK(launch_adj_edge_gather_batch(N, E, K, adj.own_el, adj.own_f, adj.p, adj.diffbar, stream));
      HIPCHK(launch_adj_wall(E, K, d_noslip, d_esign, idr, adj.diffbar, adj.pvbar, stream));
    }
    // psi_bar through the assembled velocity map
    HIPCHK(launch_adj_vel_transpose(N, K, d_nb, d_nbf, adj.Mv, adj.Md, adj.Ubar, adj.diffbar, adj.psibar, stream));
    // (4) the Poisson solve: As z = psi_bar (a right-hand side -W b = psi_bar comes from the
    //     vorticity psi_bar / (W J) with zero boundary values), then omega_bar += W J z
    HIPCHK(launch_adj_wj(N, K, tab.wgt, d_jac, adj.psibar, adj.fake, 0, stream));
    if (std::isnan(tau)) {
      adj.zn = 0;
    } else if (adj.zn >= 1 && tau != adj.ztau_a) {
      const double d1 = tau - adj.ztau_a, d2 = adj.ztau_a - adj.ztau_b;
      const bool ok = adj.zn == 2 && d1 * d2 > 0. && std::fabs(d1) <= 1. && std::fabs(d2) <= 1.;
      psi_predict(stream, D, ok ? d1 / d2 : 0., ok ? 1 : 0, adj.z, adj.zprev);
    }
    rhs_bc = adj.zero_e;
    std::swap(psi.p, adj.z.p);
    proj_cur = &proj_adj;
    proj_force = true;
    try {
      poisson(adj.fake);
    } catch (...) {
      std::swap(psi.p, adj.z.p);
      rhs_bc = nullptr;
      proj_cur = &proj_fwd;
      proj_force = false;
      throw;
    }
    std::swap(psi.p, adj.z.p);
    rhs_bc = nullptr;
    proj_cur = &proj_fwd;
    proj_force = false;
    if (!std::isnan(tau) && !(adj.zn >= 1 && tau == adj.ztau_a)) {
      adj.ztau_b = adj.ztau_a;
      adj.ztau_a = tau;
      adj.zn = std::min(adj.zn + 1, 2);
    }
    HIPCHK(launch_adj_wj(N, K, tab.wgt, d_jac, adj.z, xbar, 1, stream));
  }

  // Binomial checkpointing: F(L, s) = the fewest step advances that reverse L steps from a
  // snapshot at the range's start with s free snapshot slots, each step then recorded once:
  // F(1, s) = 0, F(L, 0) = L (L - 1) / 2, F(L, s) = min_d [d + F(L - d, s - 1) + F(d, s)]
  // (advance d steps and snapshot, reverse the tail with one slot less, then the head).  The
  // table of optimal first advances d is built once per (L, s); above kBinoMaxL steps the range
  // is halved instead.
  static constexpr int kBinoMaxL = 2048;
  void binomial_table(int L, int S) {
    if (L <= adj.bsplit_L && S <= adj.bsplit_S) return;
    L = std::min(std::max(L, adj.bsplit_L), kBinoMaxL);
    S = std::max(S, adj.bsplit_S);
    std::vector<int64_t> F((size_t)(L + 1) * (S + 1), 0);
    adj.bsplit.assign((size_t)(L + 1) * (S + 1), 1);
    auto at = [&](int l, int sl) -> int64_t& { return F[(size_t)l * (S + 1) + sl]; };
    for (int l = 1; l <= L; ++l) at(l, 0) = (int64_t)l * (l - 1) / 2;
    for (int sl = 1; sl <= S; ++sl)
      for (int l = 2; l <= L; ++l) {
        int64_t best = INT64_MAX;
        int bd = 1;
        for (int d = 1; d < l; ++d) {
          const int64_t c = d + at(l - d, sl - 1) + at(d, sl);
          if (c < best) { best = c; bd = d; }
        }
        at(l, sl) = best;
        adj.bsplit[(size_t)l * (S + 1) + sl] = bd;
      }
    adj.bsplit_L = L;
    adj.bsplit_S = S;
  }
  int binomial_split(int L, int S) const {
    if (L > adj.bsplit_L || S > adj.bsplit_S) return std::max(1, L / 2);
    return adj.bsplit[(size_t)L * (adj.bsplit_S + 1) + S];
  }
  int snap_take(int pos) {
    for (int i = 0; i < adj.nsnap; ++i)
      if (adj.snap_pos[i] < 0) {
        adj.snap_pos[i] = pos;
        dcopy(adj.snap.p + (size_t)i * D, w, D);
        return i;
      }
    throw StatusError(HOC_EINVAL, "binomial schedule ran out of snapshot slots");
  }
  int snap_find(int pos) const {
    for (int i = 0; i < adj.nsnap; ++i)
      if (adj.snap_pos[i] == pos) return i;
    return -1;
  }

  // Forward run of nsteps RK4 steps recording what the reverse sweep needs (hoc_adjoint_forward):
  //  every >= 1: the vorticity every `every` steps (uniform segments, each recomputed once);
  //  every == 0: the vorticity and the four stage solutions of every step when they fit in half
  //              the free device memory (no recompute), otherwise binomial checkpointing with the
  //              snapshots that fit;
  //  every < 0:  binomial checkpointing with -every vorticity snapshots.
  void adjoint_forward(int nsteps, int every) {
    adjoint_setup();
    if (nsteps < 0) throw StatusError(HOC_EINVAL, "nsteps must be >= 0");
    adj.have_forward = false;
    adj.trajw.alloc(0); adj.trajpsi.alloc(0); adj.snap.alloc(0);
    adj.t0 = time;
    int sched = every >= 1 ? 0 : 2;
    int nsnap = every < 0 ? -every : 0;
    if (every == 0) {
      size_t fr = 0, tot = 0;
      HIPCHK(hipMemGetInfo(&fr, &tot));
      const double budget = 0.5 * (double)fr;
      const double per = (double)D * sizeof(double);
      if (5. * per * std::max(nsteps, 1) <= budget) sched = 1;
      else nsnap = (int)std::min<double>(budget / per, (double)std::max(nsteps, 1));
      if (sched == 2 && nsnap < 1) throw StatusError(HOC_EDEVICE, "no device memory for adjoint snapshots");
    }
    adj.sched = sched;
    if (sched == 0) {
      const int nck = std::max(1, (nsteps + every - 1) / every);
      adj.ckbuf.alloc((size_t)nck * D);
      adj.segbuf.alloc((size_t)std::min(every, std::max(nsteps, 1)) * D);
      adj.segpsi.alloc((size_t)std::min(every, std::max(nsteps, 1)) * 4 * D);
      for (int n = 0; n < nsteps; ++n) {
        if (n % every == 0) dcopy(adj.ckbuf.p + (size_t)(n / every) * D, w, D);
        rk_step();
      }
    } else if (sched == 1) {
      adj.trajw.alloc((size_t)std::max(nsteps, 1) * D);
      adj.trajpsi.alloc((size_t)std::max(nsteps, 1) * 4 * D);
      for (int n = 0; n < nsteps; ++n) {
        dcopy(adj.trajw.p + (size_t)n * D, w, D);
        capture_psi = adj.trajpsi.p + (size_t)n * 4 * D;
        rk_step();
        capture_psi = nullptr;
      }
    } else {
      // snapshot slots: the start and the first descent of the schedule, taken on the way
      nsnap = std::max(1, std::min(nsnap, std::max(nsteps, 1)));
      adj.nsnap = nsnap;
      adj.snap.alloc((size_t)nsnap * D);
      adj.snap_pos.assign(nsnap, -1);
      binomial_table(std::max(nsteps, 1), nsnap - 1);
      snap_take(0);
      int a = 0, free_slots = nsnap - 1;
      int next = (nsteps - a > 1 && free_slots > 0) ? a + binomial_split(nsteps - a, free_slots) : -1;
      for (int n = 0; n < nsteps; ++n) {
        rk_step();
        if (n + 1 == next) {
          snap_take(next);
          a = next;
          --free_slots;
          next = (nsteps - a > 1 && free_slots > 0) ? a + binomial_split(nsteps - a, free_slots) : -1;
        }
      }
    }
    HIPCHK(hipStreamSynchronize(stream));
    adj.nsteps = nsteps;
    adj.every = every;
    adj.have_forward = true;
  }

  // Reverse sweep over the recorded trajectory.  Each step is re-run from its vorticity recording
  // its stages (with the stage solutions kept by the forward run or a recompute, or solving them)
  // and transposed stage by stage:
  //   k4_bar = w_bar/6, k3_bar = k2_bar = w_bar/3, k1_bar = w_bar/6,
  //   for s = 4..1: x = (dk_s/domega_s)^T k_s_bar;  w_bar += x;  k_{s-1}_bar += c_s x
  // with omega_s = omega_n + c_s k_{s-1}, c = (1/2, 1/2, 1).  Uniform segments are recomputed from
  // their checkpoint keeping every step's vorticity and stage solutions; recorded trajectories
  // need no recompute; the binomial schedule advances from its snapshots as it goes.  The
  // handle's state (vorticity, stream function, velocity, time) is restored afterwards.
  void adjoint_backward(const double* wbar_h, double* w0bar_h, double* pvbar_h) {
    if (!adj.have_forward) throw StatusError(HOC_ESTATE, "call hoc_adjoint_forward first");
    adjoint_check();
    const double time_end_ = time;
    dcopy(adj.save_w, w, D);
    dcopy(adj.save_psi, psi, D);
    dcopy(adj.save_vel, vel, 2 * D);
    h2d(adj.wbar.p, wbar_h, D * sizeof(double));
    const size_t EK = std::max<size_t>((size_t)E * K, 1);
    HIPCHK(hipMemsetAsync(adj.pvbar.p, 0, EK * sizeof(double), stream));
    const double cn[3] = {0.5, 0.5, 1.};
    const double cst[4] = {0., 0.5, 0.5, 1.};
    const int every = adj.every, nsteps = adj.nsteps;
    adj.zn = 0;
    proj_adj.m = 0;
    auto restore = [&] {
      dcopy(w, adj.save_w, D);
      dcopy(psi, adj.save_psi, D);
      dcopy(vel, adj.save_vel, 2 * D);
      time = time_end_;
      invalidate_guess();   // the recompute left the histories of the trajectory's early steps
    };
    // w = omega_n on entry
    auto reverse_step = [&](int n, const double* stage_psi) {
      time = adj.t0 + n * cfg.dt;
      adj_record_step(stage_psi);
      HIPCHK(launch_adj_axpby(D, 1. / 6., adj.wbar, 0., adj.kbar[3], stream));
      HIPCHK(launch_adj_axpby(D, 1. / 3., adj.wbar, 0., adj.kbar[2], stream));
      HIPCHK(launch_adj_axpby(D, 1. / 3., adj.wbar, 0., adj.kbar[1], stream));
      HIPCHK(launch_adj_axpby(D, 1. / 6., adj.wbar, 0., adj.kbar[0], stream));
      for (int st = 3; st >= 0; --st) {
        adj_stage(st, adj.kbar[st], adj.xbar, n + cst[st]);
        HIPCHK(launch_adj_axpby(D, 1., adj.xbar, 1., adj.wbar, stream));
        if (st > 0) HIPCHK(launch_adj_axpby(D, cn[st - 1], adj.xbar, 1., adj.kbar[st - 1], stream));
      }
    };
    // binomial schedule: w = omega_{a+d}, advanced from the snapshot at a
    auto advance_from = [&](int a, int d) {
      const int sl = snap_find(a);
      if (sl < 0) throw StatusError(HOC_ESTATE, "binomial schedule lost a snapshot");
      dcopy(w, adj.snap.p + (size_t)sl * D, D);
      time = adj.t0 + a * cfg.dt;
      invalidate_guess();
      for (int i = 0; i < d; ++i) rk_step();
    };
    std::function<void(int, int, int)> rev = [&](int a, int b, int free_slots) {
      if (b <= a) return;
      if (b - a == 1) { advance_from(a, 0); reverse_step(a, nullptr); return; }
      if (free_slots == 0) {
        for (int n = b - 1; n >= a; --n) { advance_from(a, n - a); reverse_step(n, nullptr); }
        return;
      }
      const int d = binomial_split(b - a, free_slots);
      int sl = snap_find(a + d);
      if (sl < 0) { advance_from(a, d); sl = snap_take(a + d); }
      rev(a + d, b, free_slots - 1);
      adj.snap_pos[sl] = -1;
      rev(a, a + d, free_slots);
    };
    try {
      if (adj.sched == 0) {
        const int nseg = (nsteps + every - 1) / every;
        for (int sg = nseg - 1; sg >= 0; --sg) {
          const int n0 = sg * every, n1 = std::min(n0 + every, nsteps);
          dcopy(w, adj.ckbuf.p + (size_t)sg * D, D);
          time = adj.t0 + n0 * cfg.dt;
          invalidate_guess();
          for (int n = n0; n < n1; ++n) {
            dcopy(adj.segbuf.p + (size_t)(n - n0) * D, w, D);
            capture_psi = adj.segpsi.p + (size_t)(n - n0) * 4 * D;
            rk_step();
            capture_psi = nullptr;
          }
          for (int n = n1 - 1; n >= n0; --n) {
            dcopy(w, adj.segbuf.p + (size_t)(n - n0) * D, D);
            reverse_step(n, adj.segpsi.p + (size_t)(n - n0) * 4 * D);
          }
        }
      } else if (adj.sched == 1) {
        for (int n = nsteps - 1; n >= 0; --n) {
          dcopy(w, adj.trajw.p + (size_t)n * D, D);
          reverse_step(n, adj.trajpsi.p + (size_t)n * 4 * D);
        }
      } else {
        rev(0, nsteps, adj.nsnap - 1);
      }
      d2h(w0bar_h, adj.wbar.p, D * sizeof(double));
      if (pvbar_h && E > 0) d2h(pvbar_h, adj.pvbar.p, (size_t)E * K * sizeof(double));
    } catch (...) {
      capture_psi = nullptr;
      restore();
      throw;
    }
    restore();
    HIPCHK(hipStreamSynchronize(stream));
  }

  // (dk/domega)^T lam and (dk/dpvel)^T lam of one stage at time t (hoc_stage_residual's map).
  void stage_vjp(const double* omega_h, double t, const double* lam_h, double* out_h, double* pvbar_h) {
    adjoint_setup();
    proj_adj.m = 0;
    const size_t EK = std::max<size_t>((size_t)E * K, 1);
    h2d(adj.ws[0].p, omega_h, D * sizeof(double));
    h2d(adj.lam.p, lam_h, D * sizeof(double));
    update_bcs(0., t);
    poisson(adj.ws[0]);
    HIPCHK(launch_trace(K, PRE_COPY, N, T, psi, nullptr, nullptr, nullptr, 0., nullptr, 0, nullptr, P64.tr.p, stream));
    HIPCHK(launch_velocity(K, N, T, G, B, psi, P64.tr, adj.ws[0], vel, adj.Us[0], adj.trws[0], adj.trus[0],
                           1. / (cfg.dt * re_inv), I, stream));
    StageOut so;
    so.k = kscratch;
    so.sig = adj.sig[0];
    HIPCHK(launch_stage(K, 1, N, T, G, B, adj.ws[0], adj.Us[0], adj.trws[0], adj.trus[0], re_inv, cfg.dt, I, so,
                        stream, 1));
    HIPCHK(hipMemsetAsync(adj.pvbar.p, 0, EK * sizeof(double), stream));
    adj_stage(0, adj.lam, adj.xbar);
    d2h(out_h, adj.xbar.p, D * sizeof(double));
    if (pvbar_h && E > 0) d2h(pvbar_h, adj.pvbar.p, (size_t)E * K * sizeof(double));
  }

  // Boundary flux points [E][K] (edge flux-point order of the boundary arrays), x and y.
  void boundary_points(double* xy) const {
    const int L = mesh.lnod;
    for (int e = 0; e < E; ++e) {
      const auto& nd = mesh.edges[e];
      for (int j = 0; j < K; ++j) {
        double x = 0., y = 0.;
        for (int l = 0; l < L; ++l) {
          x += tab.gsb[l][j] * mesh.xy[2 * (size_t)nd[l]];
          y += tab.gsb[l][j] * mesh.xy[2 * (size_t)nd[l] + 1];
        }
        xy[2 * ((size_t)e * K + j)] = x;
        xy[2 * ((size_t)e * K + j) + 1] = y;
      }
    }
  }

  // Ensemble tables (instance of every element and boundary edge, per-instance Re and dt).
  void setup_instances() {
    I = InstDev{};
    G.inst = nullptr;
    if (n_inst <= 1) return;
    if ((int)inst_off.size() != n_inst + 1 || inst_off[0] != 0 || inst_off[n_inst] != N)
      throw StatusError(HOC_EINVAL, "instance element offsets do not cover the mesh");
    std::vector<int32_t> ei(N), bi(std::max(E, 1), 0);
    for (int i = 0; i < n_inst; ++i) {
      if (inst_off[i + 1] <= inst_off[i]) throw StatusError(HOC_EINVAL, "instance element offsets must increase");
      for (int e = inst_off[i]; e < inst_off[i + 1]; ++e) ei[e] = i;
    }
    // instances must not share faces
    for (int e = 0; e < N; ++e)
      for (int f = 0; f < 4; ++f) {
        const int nb = mesh.nb[(size_t)e * 4 + f];
        if (nb >= 0 && ei[nb] != ei[e]) throw StatusError(HOC_EINVAL, "ensemble instances share a face");
      }
    for (int e = 0; e < E; ++e) bi[e] = ei[mesh.be_elem[e]];
    std::vector<double> ri(n_inst), idr(n_inst);
    for (int i = 0; i < n_inst; ++i) { ri[i] = 1. / inst_re[i]; idr[i] = 1. / (inst_dt[i] * ri[i]); }
    d_inst.upload(ei); d_einst.upload(bi); d_ioff.upload(inst_off);
    d_ireinv.upload(ri); d_idt.upload(inst_dt); d_iidr.upload(idr);
    d_iramp.alloc(n_inst); d_inorm.alloc(n_inst);
    inst_time.assign(n_inst, 0.);
    G.inst = d_inst;
    I = InstDev{d_ireinv, d_idt, d_iidr, d_iramp, d_einst};
  }
};

}  // namespace hoc

using hoc::Solver;

struct hoc_solver {
  Solver s;
};

static std::string g_err;

static int hoc_fail(hoc_t h, int code, const std::string& msg) {
  if (h) h->s.err = msg;
  g_err = msg;
  return code;
}

template <class Fn>
static int guarded(hoc_t h, Fn fn) {
  if (!h) return hoc_fail(nullptr, HOC_EINVAL, "null handle");
  try {
    fn();
  } catch (const hoc::StatusError& e) {
    return hoc_fail(h, e.code, e.what());
  } catch (const hoc::HipError& e) {
    return hoc_fail(h, HOC_EDEVICE, e.what());
  } catch (const std::bad_alloc&) {
    return hoc_fail(h, HOC_EDEVICE, "host allocation failed");
  } catch (const std::exception& e) {
    return hoc_fail(h, HOC_EMESH, e.what());
  }
  return HOC_OK;
}

extern "C" {

void hoc_default_config(hoc_config* c) {
  std::memset(c, 0, sizeof(*c));
  c->K = 3;                      // set_defaults (src/HO_2D.cpp:3699)
  c->problem_type = 11;             // :3702 (the hybrid driver's default)
  c->Re = 1.0;
  c->dt = 1.0;
  c->time_integration_type = 2;  // :3703
  c->advection_type = 2;         // :3704
  c->tol = 1e-15;                // :3067
  c->maxiter = 150;              // :3069
  c->precision = HOC_FP64;
  c->krylov = HOC_KRYLOV_AUTO;
  c->device = 0;
  c->fail_on_noconv = 0;
  c->mixed_inner_tol = 0.;
  c->cheb_degree = 0;
  c->cheb_lower = 0.;
  c->launch_mode = 0;
}

int hoc_create(hoc_t* out, const hoc_config* cfg) {
  if (!out || !cfg) return hoc_fail(nullptr, HOC_EINVAL, "null argument");
  *out = nullptr;
  if (cfg->K < 2 || cfg->K > 6) return hoc_fail(nullptr, HOC_EINVAL, "K must be in 2..6");
  if (cfg->time_integration_type != 1 && cfg->time_integration_type != 2 && cfg->time_integration_type != 4)
    return hoc_fail(nullptr, HOC_EINVAL, "time_integration_type must be 1, 2 or 4");
  if (cfg->advection_type != 1 && cfg->advection_type != 2) return hoc_fail(nullptr, HOC_EINVAL, "advection_type must be 1 or 2");
  if ((cfg->problem_type < 1 || cfg->problem_type > 3) && cfg->problem_type != 11)
    return hoc_fail(nullptr, HOC_EINVAL, "problem_type must be 1, 2, 3 or 11");
  if (cfg->precision != HOC_FP64 && cfg->precision != HOC_MIXED) return hoc_fail(nullptr, HOC_EINVAL, "bad precision");
  if (!(cfg->dt > 0.) || cfg->maxiter < 0 || !(cfg->tol >= 0.)) return hoc_fail(nullptr, HOC_EINVAL, "bad dt/tol/maxiter");
  if (!(cfg->mixed_inner_tol >= 0. && cfg->mixed_inner_tol < 1.)) return hoc_fail(nullptr, HOC_EINVAL, "mixed_inner_tol must be in [0, 1)");
  if (cfg->cheb_degree < 0 || cfg->cheb_degree > hoc::kMaxCheb) return hoc_fail(nullptr, HOC_EINVAL, "cheb_degree must be in 0..8");
  if (!(cfg->cheb_lower >= 0. && cfg->cheb_lower < 1.)) return hoc_fail(nullptr, HOC_EINVAL, "cheb_lower must be in [0, 1)");
  if (cfg->launch_mode < 0 || cfg->launch_mode > 2) return hoc_fail(nullptr, HOC_EINVAL, "launch_mode must be 0, 1 or 2");
  if (cfg->psi_guess < 0 || cfg->psi_guess > 5) return hoc_fail(nullptr, HOC_EINVAL, "psi_guess must be 0..5");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return hoc_fail(nullptr, HOC_EDEVICE, "no HIP device");
  if (cfg->device < 0 || cfg->device >= ndev) return hoc_fail(nullptr, HOC_EDEVICE, "device ordinal out of range");
  hoc_t h = new (std::nothrow) hoc_solver();
  if (!h) return hoc_fail(nullptr, HOC_EDEVICE, "allocation failed");
  h->s.cfg = *cfg;
  h->s.device = cfg->device;
  *out = h;
  return HOC_OK;
}

int hoc_load_mesh(hoc_t h, int32_t lnod, int32_t n_nodes, const double* xy, int32_t n_elem, const int32_t* elem_nodes,
                  int32_t n_bedges, const int32_t* bedge_nodes, const int32_t* bedge_bdry, int32_t n_bdry,
                  const char* bdry_names) {
  return guarded(h, [&] {
    Solver& s = h->s;
    if (lnod < 2 || lnod > 4) throw hoc::StatusError(HOC_EINVAL, "lnod must be 2, 3 or 4");
    if (n_nodes <= 0 || n_elem <= 0 || n_bedges < 0 || n_bdry < 0 || !xy || !elem_nodes)
      throw hoc::StatusError(HOC_EINVAL, "empty mesh");
    if (n_bedges > 0 && (!bedge_nodes || !bedge_bdry)) throw hoc::StatusError(HOC_EINVAL, "missing boundary arrays");
    hoc::Mesh m;
    m.lnod = lnod; m.n_nodes = n_nodes; m.n_elem = n_elem; m.n_bedges = n_bedges;
    m.xy.assign(xy, xy + 2 * (size_t)n_nodes);
    m.elem.assign(elem_nodes, elem_nodes + (size_t)n_elem * lnod * lnod);
    m.edges.resize(n_bedges);
    for (int e = 0; e < n_bedges; ++e) m.edges[e].assign(bedge_nodes + (size_t)e * lnod, bedge_nodes + (size_t)(e + 1) * lnod);
    m.bedge_bdry.assign(bedge_bdry, bedge_bdry + n_bedges);
    std::string names(bdry_names ? bdry_names : "");
    size_t p0 = 0;
    for (int b = 0; b < n_bdry; ++b) {
      size_t p1 = names.find('\n', p0);
      m.bnames.push_back(names.substr(p0, p1 == std::string::npos ? std::string::npos : p1 - p0));
      p0 = p1 == std::string::npos ? names.size() : p1 + 1;
    }
    s.accept_mesh(std::move(m));
  });
}

int hoc_load_gmsh(hoc_t h, const char* path, int32_t occ) {
  return guarded(h, [&] {
    if (!path) throw hoc::StatusError(HOC_EINVAL, "null path");
    hoc::Mesh m;
    try {
      hoc::read_gmsh41(path, occ, m);
    } catch (const std::runtime_error& e) {
      throw hoc::StatusError(HOC_EMESH, e.what());
    }
    h->s.accept_mesh(std::move(m));
  });
}

int hoc_set_instances(hoc_t h, int32_t n_inst, const int32_t* elem_offsets, const double* Re, const double* dt) {
  return guarded(h, [&] {
    Solver& s = h->s;
    if (!s.mesh_loaded) throw hoc::StatusError(HOC_ESTATE, "hoc_set_instances before hoc_load_mesh");
    if (n_inst < 1 || !elem_offsets || !Re || !dt) throw hoc::StatusError(HOC_EINVAL, "bad instance arrays");
    if (elem_offsets[0] != 0 || elem_offsets[n_inst] != s.mesh.n_elem)
      throw hoc::StatusError(HOC_EINVAL, "instance element offsets must run from 0 to n_elem");
    for (int i = 0; i < n_inst; ++i) {
      if (elem_offsets[i + 1] <= elem_offsets[i]) throw hoc::StatusError(HOC_EINVAL, "instance element offsets must increase");
      if (!(Re[i] > 1e-3) || !(dt[i] > 0.)) throw hoc::StatusError(HOC_EINVAL, "instance Re must exceed 1e-3 and dt be positive");
    }
    s.n_inst = n_inst;
    s.inst_off.assign(elem_offsets, elem_offsets + n_inst + 1);
    s.inst_re.assign(Re, Re + n_inst);
    s.inst_dt.assign(dt, dt + n_inst);
    s.ready = false;
  });
}

int hoc_setup(hoc_t h) {
  return guarded(h, [&] {
    if (!h->s.mesh_loaded) throw hoc::StatusError(HOC_ESTATE, "hoc_setup before hoc_load_mesh");
    h->s.setup();
  });
}

#define NEED_READY(h) if (!h->s.ready) throw hoc::StatusError(HOC_ESTATE, "call hoc_setup first")
#define NEED_LEN(len, want) if ((len) != (int64_t)(want)) throw hoc::StatusError(HOC_EINVAL, "length mismatch")

int hoc_set_boundary_values(hoc_t h, int32_t n_bedges, const int32_t* nval, const double* vals) {
  return guarded(h, [&] {
    Solver& s = h->s;
    if (!s.mesh_loaded) throw hoc::StatusError(HOC_ESTATE, "hoc_set_boundary_values before hoc_load_mesh");
    if (n_bedges != s.mesh.n_bedges || !nval || !vals) throw hoc::StatusError(HOC_EINVAL, "length mismatch");
    s.edge_nvel_in.assign(n_bedges, 0.);
    s.edge_has_nvel.assign(n_bedges, 0);
    for (int e = 0; e < n_bedges; ++e) {  // arrays_to_boundary (src/HO_2D.cpp:3779-3793, 3806-3814)
      if (nval[e] < 0 || nval[e] > 2) throw hoc::StatusError(HOC_EINVAL, "values per edge must be 0, 1 or 2");
      if (nval[e] == 1) s.edge_nvel_in[e] = vals[2 * (size_t)e];
      if (nval[e] == 2) s.edge_nvel_in[e] = vals[2 * (size_t)e + 1];
      s.edge_has_nvel[e] = nval[e] > 0;
    }
    s.ready = false;
  });
}

int64_t hoc_open_point_count(hoc_t h) {
  return (h && h->s.ready && h->s.cfg.problem_type == 11) ? (int64_t)h->s.open_edges.size() * h->s.K : 0;
}

int hoc_get_open_points(hoc_t h, int64_t len, double* xy) {
  return guarded(h, [&] {
    NEED_READY(h); h->s.need_hybrid(); NEED_LEN(len, 2 * hoc_open_point_count(h));
    h->s.open_points(xy);
  });
}

int hoc_set_open_velocities(hoc_t h, int64_t len, const double* uv) {
  return guarded(h, [&] {
    NEED_READY(h); h->s.need_hybrid(); NEED_LEN(len, 2 * hoc_open_point_count(h));
    h->s.set_open_velocities(uv);
  });
}

int hoc_set_open_vorticity(hoc_t h, int64_t len, const double* w) {
  return guarded(h, [&] {
    NEED_READY(h); h->s.need_hybrid(); NEED_LEN(len, hoc_open_point_count(h));
    h->s.set_open_vorticity(w);
  });
}

int hoc_set_vorticity_target(hoc_t h, int64_t len, const double* w) {
  return guarded(h, [&] {
    NEED_READY(h); h->s.need_hybrid(); NEED_LEN(len, h->s.D);
    h->s.set_vorticity_target(w);
  });
}

int hoc_set_p2g_weights(hoc_t h, int64_t len, const double* w) {
  return guarded(h, [&] {
    NEED_READY(h); h->s.need_hybrid(); NEED_LEN(len, h->s.D);
    h->s.h2d(h->s.d_hwgt.p, w, h->s.D * sizeof(double));
  });
}

int hoc_get_instance_times(hoc_t h, int32_t n_inst, double* t) {
  return guarded(h, [&] {
    NEED_READY(h);
    Solver& s = h->s;
    if (n_inst != s.n_inst || !t) throw hoc::StatusError(HOC_EINVAL, "length mismatch");
    if (s.n_inst == 1) t[0] = s.time;
    else std::copy(s.inst_time.begin(), s.inst_time.end(), t);
  });
}


int hoc_set_vorticity(hoc_t h, int64_t len, const double* w) {
  return guarded(h, [&] {
    NEED_READY(h); NEED_LEN(len, h->s.D);
    h->s.h2d(h->s.w.p, w, h->s.D * sizeof(double));
    h->s.invalidate_guess();
  });
}
int hoc_get_vorticity(hoc_t h, int64_t len, double* w) {
  return guarded(h, [&] {
    NEED_READY(h); NEED_LEN(len, h->s.D);
    h->s.d2h(w, h->s.w.p, h->s.D * sizeof(double));
  });
}
int hoc_get_streamfunction(hoc_t h, int64_t len, double* p) {
  return guarded(h, [&] {
    NEED_READY(h); NEED_LEN(len, h->s.D);
    h->s.d2h(p, h->s.psi.p, h->s.D * sizeof(double));
  });
}
int hoc_set_streamfunction(hoc_t h, int64_t len, const double* p) {
  return guarded(h, [&] {
    NEED_READY(h); NEED_LEN(len, h->s.D);
    h->s.h2d(h->s.psi.p, p, h->s.D * sizeof(double));
    h->s.invalidate_guess();
  });
}
int hoc_get_velocity(hoc_t h, int64_t len, double* uv) {
  return guarded(h, [&] {
    NEED_READY(h); NEED_LEN(len, 2 * h->s.D);
    h->s.d2h(uv, h->s.vel.p, 2 * h->s.D * sizeof(double));
  });
}
int hoc_get_solution_points(hoc_t h, int64_t len, double* xy) {
  return guarded(h, [&] {
    NEED_READY(h); NEED_LEN(len, 2 * h->s.D);
    std::vector<double> pts;
    hoc::solution_points(h->s.mesh, h->s.tab, pts);
    std::memcpy(xy, pts.data(), pts.size() * sizeof(double));
  });
}
int hoc_get_solution_areas(hoc_t h, int64_t len, double* a) {
  return guarded(h, [&] {
    NEED_READY(h); NEED_LEN(len, h->s.D);
    Solver& s = h->s;
    for (int el = 0; el < s.N; ++el)
      for (int j = 0; j < s.K; ++j)
        for (int k = 0; k < s.K; ++k) {
          size_t v = (size_t)el * s.K2 + j * s.K + k;
          a[v] = s.tab.wgt[j] * s.tab.wgt[k] * s.geo.jac[v];  // src/HO_2D.cpp:3988
        }
  });
}
int hoc_get_node_fields(hoc_t h, int32_t refresh, int32_t n_nodes, double* out) {
  return guarded(h, [&] {
    NEED_READY(h);
    Solver& s = h->s;
    if (n_nodes != s.mesh.n_nodes || !out) throw hoc::StatusError(HOC_EINVAL, "n_nodes must be the mesh's node count");
    s.node_fields(refresh != 0);
    s.d2h(out, s.d_nodeout.p, (size_t)n_nodes * 4 * sizeof(double));
  });
}

int hoc_write_smooth_vtk(hoc_t h, int32_t refresh, const char* path) {
  return guarded(h, [&] {
    NEED_READY(h);
    if (!path) throw hoc::StatusError(HOC_EINVAL, "null path");
    Solver& s = h->s;
    s.node_fields(refresh != 0);
    std::vector<double> nodes((size_t)s.mesh.n_nodes * 4);
    s.d2h(nodes.data(), s.d_nodeout.p, nodes.size() * sizeof(double));
    try {
      hoc::write_vtk_nodes(path, s.mesh, nodes.data());
    } catch (const std::runtime_error& e) {
      throw hoc::StatusError(HOC_EINVAL, e.what());
    }
  });
}

int hoc_set_sample_points(hoc_t h, int32_t n, const double* xy) {
  return guarded(h, [&] {
    NEED_READY(h);
    if (n < 0 || (n > 0 && !xy)) throw hoc::StatusError(HOC_EINVAL, "bad sample point array");
    for (int32_t i = 0; i < 2 * n; ++i)
      if (!std::isfinite(xy[i])) throw hoc::StatusError(HOC_EINVAL, "non-finite sample point coordinate");
    h->s.set_sample_points(n, xy);
  });
}

int hoc_get_sample_location(hoc_t h, int32_t n, int32_t* elem, double* local_xy) {
  return guarded(h, [&] {
    NEED_READY(h);
    Solver& s = h->s;
    if (n != s.n_sp) throw hoc::StatusError(HOC_EINVAL, "n must be the number of sample points set");
    if (elem) std::copy(s.sp_elem.begin(), s.sp_elem.end(), elem);
    if (local_xy) std::copy(s.sp_loc.begin(), s.sp_loc.end(), local_xy);
  });
}

int hoc_get_sample_values(hoc_t h, int32_t refresh, int32_t n, double* out) {
  return guarded(h, [&] {
    NEED_READY(h);
    Solver& s = h->s;
    if (n != s.n_sp || (n > 0 && !out)) throw hoc::StatusError(HOC_EINVAL, "n must be the number of sample points set");
    s.sample_values(refresh != 0, out);
  });
}

int hoc_write_samples(hoc_t h, int32_t refresh, const char* path) {
  return guarded(h, [&] {
    NEED_READY(h);
    if (!path) throw hoc::StatusError(HOC_EINVAL, "null path");
    Solver& s = h->s;
    std::vector<double> vals(3 * (size_t)s.n_sp);
    s.sample_values(refresh != 0, vals.data());
    try {
      hoc::write_samples(path, s.n_sp, s.sp_xy.data(), vals.data());
    } catch (const std::runtime_error& e) {
      throw hoc::StatusError(HOC_EINVAL, e.what());
    }
  });
}

int hoc_write_vtk(hoc_t h, const char* path) {
  return guarded(h, [&] {
    NEED_READY(h);
    if (!path) throw hoc::StatusError(HOC_EINVAL, "null path");
    Solver& s = h->s;
    std::vector<double> pts, w(s.D), p(s.D), uv(2 * s.D), u(s.D), v(s.D);
    hoc::solution_points(s.mesh, s.tab, pts);
    s.d2h(w.data(), s.w.p, s.D * sizeof(double));
    s.d2h(p.data(), s.psi.p, s.D * sizeof(double));
    s.d2h(uv.data(), s.vel.p, 2 * s.D * sizeof(double));
    for (size_t i = 0; i < s.D; ++i) { u[i] = uv[2 * i]; v[i] = uv[2 * i + 1]; }
    const char* names[4] = {"vorticity", "streamfunction", "u", "v"};
    const double* fields[4] = {w.data(), p.data(), u.data(), v.data()};
    try {
      hoc::write_vtk_sps(path, s.K, s.N, pts.data(), 4, names, fields);
    } catch (const std::runtime_error& e) {
      throw hoc::StatusError(HOC_EINVAL, e.what());
    }
  });
}

double hoc_get_time(hoc_t h) { return h ? h->s.time : 0.; }
int hoc_set_time(hoc_t h, double t) {
  return guarded(h, [&] {
    h->s.time = t;
    std::fill(h->s.inst_time.begin(), h->s.inst_time.end(), t);
  });
}

int hoc_step(hoc_t h, int32_t nsteps) {
  return guarded(h, [&] {
    NEED_READY(h);
    if (nsteps < 0) throw hoc::StatusError(HOC_EINVAL, "nsteps < 0");
    for (int i = 0; i < nsteps; ++i) h->s.rk_step();
    HIPCHK(hipStreamSynchronize(h->s.stream));
  });
}

int hoc_solveto(hoc_t h, double outer_dt, int32_t nsub, int32_t integ, double Re) {
  return guarded(h, [&] {
    NEED_READY(h);
    if (nsub <= 0 || !(outer_dt > 0.)) throw hoc::StatusError(HOC_EINVAL, "bad outer_dt/nsub");
    if (integ != 1 && integ != 2 && integ != 4) throw hoc::StatusError(HOC_EINVAL, "integ must be 1, 2 or 4");
    if (!(Re > 0.)) throw hoc::StatusError(HOC_EINVAL, "Re must be positive");
    Solver& s = h->s;
    if (s.n_inst > 1) throw hoc::StatusError(HOC_EINVAL, "hoc_solveto sets one dt and Re: not for an ensemble");
    s.cfg.dt = outer_dt / (double)nsub;   // src/HO_2D.cpp:4199
    s.cfg.time_integration_type = integ;  // :4203
    s.re_inv = 1.0 / Re;                  // :4204
    s.time_start = s.time_end;            // :4208-4211
    s.time = s.time_start;
    s.time_end += outer_dt;
    for (int i = 0; i < nsub; ++i) s.rk_step();
    HIPCHK(hipStreamSynchronize(s.stream));
  });
}

int hoc_poisson(hoc_t h, const double* omega, double t, double* psi_out) {
  return guarded(h, [&] {
    NEED_READY(h);
    Solver& s = h->s;
    s.h2d(s.kscratch.p, omega, s.D * sizeof(double));
    s.update_bcs(0., t);
    s.poisson(s.kscratch);
    s.d2h(psi_out, s.psi.p, s.D * sizeof(double));
  });
}

int hoc_stage_residual(hoc_t h, const double* omega, double t, double* k) {
  return guarded(h, [&] {
    NEED_READY(h);
    Solver& s = h->s;
    s.h2d(s.vort.p, omega, s.D * sizeof(double));
    hoc::StageOut so;
    so.k = s.kscratch;
    s.update_bcs(0., t);
    s.stage_body(s.vort, so);
    s.d2h(k, s.kscratch.p, s.D * sizeof(double));
  });
}

int64_t hoc_boundary_point_count(hoc_t h) { return h && h->s.mesh_loaded ? (int64_t)h->s.E * h->s.K : 0; }

int hoc_get_boundary_points(hoc_t h, int64_t len, double* xy) {
  return guarded(h, [&] {
    NEED_READY(h);
    NEED_LEN(len, 2 * (int64_t)h->s.E * h->s.K);
    h->s.boundary_points(xy);
  });
}

int hoc_get_wall_velocity(hoc_t h, int64_t len, double* v) {
  return guarded(h, [&] {
    NEED_READY(h);
    NEED_LEN(len, (int64_t)h->s.E * h->s.K);
    if (len) h->s.d2h(v, h->s.b_pvel0.p, len * sizeof(double));
  });
}

int hoc_set_wall_velocity(hoc_t h, int64_t len, const double* v) {
  return guarded(h, [&] {
    NEED_READY(h);
    NEED_LEN(len, (int64_t)h->s.E * h->s.K);
    for (int64_t i = 0; i < len; ++i)
      if (!std::isfinite(v[i])) throw hoc::StatusError(HOC_EINVAL, "wall velocity must be finite");
    if (len) h->s.h2d(h->s.b_pvel0.p, v, len * sizeof(double));
  });
}

int hoc_adjoint_forward(hoc_t h, int32_t nsteps, int32_t checkpoint_every) {
  return guarded(h, [&] {
    NEED_READY(h);
    h->s.adjoint_forward(nsteps, checkpoint_every);
  });
}

int hoc_adjoint_backward(hoc_t h, int64_t len, const double* wbar_final, double* wbar_initial, int64_t len_wall,
                         double* wall_velocity_bar) {
  return guarded(h, [&] {
    NEED_READY(h);
    NEED_LEN(len, h->s.D);
    if (wall_velocity_bar) NEED_LEN(len_wall, (int64_t)h->s.E * h->s.K);
    if (!wbar_final || !wbar_initial) throw hoc::StatusError(HOC_EINVAL, "null buffer");
    h->s.adjoint_backward(wbar_final, wbar_initial, wall_velocity_bar);
  });
}

int hoc_stage_residual_vjp(hoc_t h, const double* omega, double t, const double* lam, double* omega_bar,
                           double* wall_velocity_bar) {
  return guarded(h, [&] {
    NEED_READY(h);
    if (!omega || !lam || !omega_bar) throw hoc::StatusError(HOC_EINVAL, "null buffer");
    h->s.stage_vjp(omega, t, lam, omega_bar, wall_velocity_bar);
  });
}

int hoc_apply_laplacian(hoc_t h, const double* x, double* y) {
  return guarded(h, [&] {
    NEED_READY(h);
    Solver& s = h->s;
    s.h2d(s.P64.z.p, x, s.D * sizeof(double));
    s.apply_As(s.P64.z, s.P64.q);
    std::vector<double> hy(s.D);
    s.d2h(hy.data(), s.P64.q.p, s.D * sizeof(double));
    // As = -W A  =>  A x = -(As x) / W
    for (size_t v = 0; v < s.D; ++v) {
      const int m = (int)(v % s.K2);
      y[v] = -hy[v] / (s.tab.wgt[m / s.K] * s.tab.wgt[m % s.K]);
    }
  });
}

int hoc_apply_laplacian_device(hoc_t h, const double* x, double* y, void* stream) {
  return guarded(h, [&] {
    NEED_READY(h);
    Solver& s = h->s;
    if (!x || !y) throw hoc::StatusError(HOC_EINVAL, "null buffer");
    if (x == y) throw hoc::StatusError(HOC_EINVAL, "x and y must not alias");
    const size_t need = (size_t)s.N * 8 * s.K * sizeof(double);  // face traces: 4 faces x (value, derivative) x K
    if (need > hoc::kNoAllocScratchBytes)
      throw hoc::StatusError(HOC_EINVAL, "mesh too large for the no-allocation apply: its face traces need " +
                                             std::to_string(need) + " bytes, the scratch array holds " +
                                             std::to_string(hoc::kNoAllocScratchBytes));
    double* tr = hoc::noalloc_scratch(s.device);  // set at hoc_setup on this device
    if (!tr) throw hoc::StatusError(HOC_ESTATE, "scratch array not set up on this device");
    HIPCHK(hipSetDevice(s.device));
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : s.stream;
    double* xm = const_cast<double*>(x);  // read only: PRE_COPY and EPI_STORE do not write x
    HIPCHK(hoc::launch_trace(s.K, hoc::PRE_COPY, s.N, s.T, xm, nullptr, nullptr, nullptr, 0., nullptr, 0, nullptr, tr, st));
    HIPCHK(hoc::launch_lap(s.K, hoc::EPI_STORE, s.N, s.P64.T, s.P64.G, xm, tr, y, nullptr, nullptr, nullptr, 0., 0.,
                           nullptr, nullptr, st));
    hoc::unweight(st, s.D, s.K, s.tab.wgt, y);  // As = -W A  =>  A x = -(As x) / W
    HIPCHK(hipGetLastError());
  });
}

int hoc_stats_get(hoc_t h, hoc_stats* st) {
  return guarded(h, [&] {
    Solver& s = h->s;
    std::memset(st, 0, sizeof(*st));
    st->n_elem = s.N; st->n_dofs = (int64_t)s.D; st->n_bedges = s.E;
    st->n_levels = (int32_t)(s.mixed ? s.P32.lv.size() : s.P64.lv.size()) + 1;
    st->last_refinements = s.last_outer;
    st->precision = s.cfg.precision;
    st->n_nodes = s.mesh.n_nodes;
    st->launch_graphs = s.use_graphs ? 1 : 0;
    st->launch_tune_ratio = s.launch_tune_ratio;
    st->krylov = s.krylov;
    st->last_iters = s.last_iters;
    st->last_rel_residual = s.last_rel;
    st->total_iters = s.total_iters;
    st->n_solves = s.n_solves;
    st->setup_seconds = s.setup_seconds;
    st->fine_op_bytes = s.ready ? s.lap_bytes(hoc::EPI_STORE) : 0.;
  });
}

int hoc_profile_enable(hoc_t h, int32_t on) {
  return guarded(h, [&] {
    NEED_READY(h);
    Solver& s = h->s;
    s.prof_flush();
    if (on && s.prof.e0.empty()) {
      const int cap = 4096;
      s.prof.e0.resize(cap); s.prof.e1.resize(cap); s.prof.var.resize(cap); s.prof.nbytes.resize(cap);
      for (int i = 0; i < cap; ++i) { HIPCHK(hipEventCreate(&s.prof.e0[i])); HIPCHK(hipEventCreate(&s.prof.e1[i])); }
      s.prof.ran.alloc(cap);
    }
    if (on) {
      for (int v = 0; v < HOC_PROF_VARIANTS; ++v) { s.prof.ms[v] = 0.; s.prof.cnt[v] = 0; s.prof.bytes[v] = 0.; }
      s.prof.skipped = 0;
    }
    s.prof.on = on != 0;
    s.prof.every = on > 1 ? on : 1;
    s.prof.seen = 0;
  });
}

int hoc_profile_get(hoc_t h, hoc_kernel_profile* p) {
  return guarded(h, [&] {
    NEED_READY(h);
    Solver& s = h->s;
    s.prof_flush();
    static const char* names[HOC_PROF_VARIANTS] = {"lap_apply",     "lap_residual",     "lap_cheb",     "lap_cheb_first",
                                                   "lap_restrict",  "lap_apply_f32",    "lap_residual_f32", "lap_cheb_f32",
                                                   "lap_cheb_first_f32", "lap_restrict_f32"};
    std::memset(p, 0, sizeof(*p));
    for (int v = 0; v < HOC_PROF_VARIANTS; ++v) {
      p->launches[v] = s.prof.cnt[v];
      p->total_ms[v] = s.prof.ms[v];
      const int half = HOC_PROF_VARIANTS / 2;
      p->bytes_per_launch[v] = s.prof.cnt[v] > 0 ? s.prof.bytes[v] / (double)s.prof.cnt[v]
                                                 : (v < half ? s.lap_bytes(v, 8.) : s.lap_bytes(v - half, 4.));
      std::strncpy(p->names[v], names[v], sizeof(p->names[v]) - 1);
    }
    p->noop_launches = s.prof.skipped;
  });
}

const char* hoc_last_error(hoc_t h) { return h ? h->s.err.c_str() : g_err.c_str(); }

void hoc_destroy(hoc_t h) {
  if (!h) return;
  (void)hipSetDevice(h->s.device);
  if (h->s.stream) (void)hipStreamSynchronize(h->s.stream);
  delete h;
}

}  // extern "C"
