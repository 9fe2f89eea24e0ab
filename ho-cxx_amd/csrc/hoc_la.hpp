// hoc_la.hpp - vector/reduction/CSR/probing kernels of the Poisson solver (hoc_la.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace hoc {

// Device scalar slots (double) and flags (int) of the Krylov solvers.
enum { S_BB = 0, S_RR, S_RZ, S_PQ, S_ALPHA, S_BETA, S_OMEGA, S_RHO, S_TOL2, S_COUNT = 16 };
enum { F_ACTIVE = 0, F_ITER, F_ZERO, F_HALF, F_COUNT = 8 };
enum { PH_BB = 0, PH_INIT, PH_RZ, PH_PQ, PH_UPD, PH_RHO, PH_ALPHA, PH_SNORM, PH_OMEGA, PH_RNORM };
enum { VOP_PCG_UPD = 0, VOP_BICG_P, VOP_BICG_S, VOP_DOT2, VOP_COPY, VOP_BICG_X };

struct VecArgs {
  int K;
  double w[8];
};

int vec_grid(size_t n);
void vec(hipStream_t s, int op, const VecArgs& A, size_t n, double* a, double* b, const double* c, const double* d,
         double* e, const double* S, const int* F, double* part);
void bicg_half(hipStream_t s, size_t n, double* x, const double* ph, const double* S, int* F);
void zero_if_flag(hipStream_t s, size_t n, double* x, const int* F);
void finalize(hipStream_t s, int phase, const double* part, int np, double* S, int* F, int maxiter);
void csr(hipStream_t s, int mode, int n, const int32_t* ptr, const int32_t* col, const double* val, const double* x,
         const double* f, double* y, const double* M, double* xio, const int* F, int lanes_per_row);
int csr_lanes(int n, size_t nnz);  // lanes per row suited to a matrix's average row length
void scale(hipStream_t s, int n, const double* M, const double* f, double* x, const int* F);
void dense(hipStream_t s, int n, const double* Ainv, const double* f, double* x, const int* F);
void probe_set(hipStream_t s, int N, int K2, const int32_t* color, int c, int rs, double* x);
void probe_diag(hipStream_t s, int N, int K2, const int32_t* color, int c, int rs, const double* y, double* diag);
void elem_sum(hipStream_t s, int N, int K2, const double* y, double* out);
struct Tab;
void q1_gather(hipStream_t s, int nv, const int32_t* vptr, const int32_t* vinc, const double* ec, double* rc, const int* F);
void q1_probe_accum(hipStream_t s, int N, int K, const Tab& T, int rs, const double* y, double* M);

}  // namespace hoc
