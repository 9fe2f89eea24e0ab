// hoc_fr.hpp - launchers of the FR stage kernels (hoc_fr.hip).
#pragma once

#include "hoc_kernels.hpp"

namespace hoc {

enum { PRE_COPY = 0, PRE_CHEB0 = 1, PRE_PROL = 2, PRE_PUPD = 3 };
enum { EPI_STORE = 0, EPI_RESID = 1, EPI_CHEB = 2, EPI_CHEBF = 3, EPI_RESTRICT = 4 };

// Outputs of the stage kernel.
struct StageOut {
  double* k = nullptr;       // k_s = dt (adv + diff), or nullptr
  double* next = nullptr;    // base + c k_s (next stage input), or nullptr
  double c = 0.;
  double* base = nullptr;    // omega_n (read for next; updated for final_mode)
  int final_mode = 0;        // 0 none, 1 Euler, 2 RK2 (needs k1), 4 RK4 (needs k1..k3)
  const double *k1 = nullptr, *k2 = nullptr, *k3 = nullptr;
  // linearised stage (launch_stage mode 1 records, mode 2 reads the upwind switches)
  int8_t* sig = nullptr;               // [N][4][K] upwind switch of each interior flux point
  const double* omega_adv = nullptr;   // mode 2: advected vorticity and its traces
  const double* trw_adv = nullptr;
  // mode 2 batches: nbatch independent products in one launch, probe b reading/writing its
  // arrays at b * (stride in values); stride 0 shares an array between the probes
  int nbatch = 1;
  int adv_only = 0;  // mode 2: the diffusion inputs are zero (skip the diffusion term)
  size_t bs_w = 0, bs_trw = 0, bs_U = 0, bs_tru = 0, bs_wa = 0, bs_trwa = 0, bs_k = 0, bs_diff = 0;
};

// Batch strides (in values) of the batched velocity launch.
struct VelBatch {
  size_t psi = 0, trpsi = 0, omega = 0, vel = 0, U = 0, trw = 0, tru = 0, vjump = 0, diff = 0, pvel = 0;
};

// Poisson-operator kernels in FP64 or (mixed-precision preconditioner) FP32.  einst, iact (may
// be null): element -> ensemble instance and the instances' live flags; blocks of one stopped
// instance return at once, and PRE_PUPD with scal_idx < 0 takes beta per instance, scal[einst[el]].
hipError_t launch_trace(int K, int pre, int N, const TabT<double>& T, double* x, const double* a, const double* b,
                        double* d, double coef, const double* scal, int scal_idx, const int* active, double* tr,
                        hipStream_t s, const int32_t* ev = nullptr, const int32_t* einst = nullptr,
                        const int* iact = nullptr);
hipError_t launch_trace(int K, int pre, int N, const TabT<float>& T, float* x, const float* a, const float* b,
                        float* d, double coef, const double* scal, int scal_idx, const int* active, float* tr,
                        hipStream_t s, const int32_t* ev = nullptr, const int32_t* einst = nullptr,
                        const int* iact = nullptr);
// PRE_PUPD of the FP64 Krylov method with the FP32 V-cycle: p = z + beta p with z in FP32.
hipError_t launch_trace_pupd_f32z(int K, int N, const TabT<double>& T, double* p, const float* z, const double* scal,
                                  int scal_idx, const int* active, double* tr, hipStream_t s,
                                  const int32_t* einst = nullptr, const int* iact = nullptr);
// tr: traces of x (input); tr_out (EPI_CHEB/EPI_CHEBF only, may be null): traces of the
// updated x, written for the next application.  tr and tr_out must differ.  ran (may be null):
// set to 1 when the launch worked, 0 when the active flag made it return at once.
hipError_t launch_lap(int K, int epi, int N, const TabT<double>& T, const GeoT<double>& G, double* x, const double* tr,
                      double* y, const double* r, const double* diag, double* d, double c1, double c2, double* part,
                      const int* active, hipStream_t s, double* tr_out = nullptr, int* ran = nullptr);
hipError_t launch_lap(int K, int epi, int N, const TabT<float>& T, const GeoT<float>& G, float* x, const float* tr,
                      float* y, const float* r, const float* diag, float* d, double c1, double c2, double* part,
                      const int* active, hipStream_t s, float* tr_out = nullptr, int* ran = nullptr);
hipError_t launch_rhs(int K, int N, const Tab& T, const Geo& G, const double* omega, const double* bcpsi, double* bs,
                      double* part, hipStream_t s);
hipError_t launch_velocity(int K, int N, const Tab& T, const Geo& G, const BCDev& B, const double* psi,
                           const double* trpsi, const double* omega, double* vel, double* U, double* trw,
                           double* tru, double inv_dt_re, const InstDev& I, hipStream_t s);
// nbatch velocity evaluations in one launch (probe b offset by b * the strides of vb)
hipError_t launch_velocity_batch(int K, int N, const Tab& T, const Geo& G, const BCDev& B, const double* psi,
                                 const double* trpsi, const double* omega, double* vel, double* U, double* trw,
                                 double* tru, double inv_dt_re, const InstDev& I, int nbatch, const VelBatch& vb,
                                 hipStream_t s);
hipError_t launch_stage(int K, int adv, int N, const Tab& T, const Geo& G, const BCDev& B, const double* omega,
                        const double* U, const double* trw, const double* tru, double re_inv, double dt,
                        const InstDev& I, const StageOut& so, hipStream_t s, int mode = 0);
// n = E*K boundary values; tc = time factor (per instance through I.ramp when set)
hipError_t launch_update_bcs(int n, int K, const BCDev& B, double tc, const InstDev& I, hipStream_t s);
// problem_type 11: boundary values at fraction fac1 (fac2 = 1 - fac1) of the outer step,
// and the blend w = (1 - wgt) w + wgt (fac1 w0 + fac2 w1)
hipError_t launch_update_bcs_hybrid(int n, int K, const BCDev& B, const HybridDev& H, double fac1, double fac2,
                                    hipStream_t s);
hipError_t launch_hybrid_blend(size_t n, double* w, const HybridDev& H, double fac1, double fac2, hipStream_t s);
int lap_blocks(int N);

}  // namespace hoc
