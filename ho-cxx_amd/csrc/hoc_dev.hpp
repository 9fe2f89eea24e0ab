// hoc_dev.hpp - device buffers, error types and the CSR level structures shared by the
// FR time-step solver (hoc_solver.hip) and the CSR linear-solver plug-in
// (hoc_linsolver.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "hoc_host.hpp"
#include "hoc_la.hpp"

namespace hoc {

struct HipError : std::runtime_error {
  explicit HipError(const std::string& s) : std::runtime_error(s) {}
};
struct StatusError : std::runtime_error {
  int code;
  StatusError(int c, const std::string& s) : std::runtime_error(s), code(c) {}
};

#define HIPCHK(x)                                                                         \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess) throw ::hoc::HipError(std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { if (p) (void)hipFree(p); }
  void alloc(size_t cnt) {
    if (p) { (void)hipFree(p); p = nullptr; }
    n = cnt;
    if (cnt) {
      HIPCHK(hipMalloc(&p, cnt * sizeof(T)));
      HIPCHK(hipMemset(p, 0, cnt * sizeof(T)));
      HIPCHK(hipDeviceSynchronize());
    }
  }
  void upload(const std::vector<T>& v) {
    alloc(v.size());
    if (!v.empty()) HIPCHK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    HIPCHK(hipDeviceSynchronize());
  }
  // Upload with conversion (FP64 host data into an FP32 buffer).
  template <class U>
  void upload_as(const std::vector<U>& v) {
    std::vector<T> c(v.begin(), v.end());
    upload(c);
  }
  operator T*() const { return p; }
};

// A sparse level matrix on the device: ELL when its rows are many and of nearly equal
// length (coalesced thread-per-row loads), CSR otherwise.  The matrix is checked (every
// column index below its column count A.m) before upload.
template <class R>
struct DevCSR {
  static constexpr int kEllMinRows = 16384, kEllMaxWidth = 32;
  int n = 0, lanes = 1, ell_w = 0, ell_lanes = 1;
  DBuf<int32_t> ptr, col;
  DBuf<R> val;
  void upload(const CSR& A) {
    try {
      csr_check(A);
    } catch (const std::invalid_argument& e) {
      throw StatusError(HOC_EINVAL, std::string("level matrix: ") + e.what());
    }
    n = A.n;
    std::vector<int32_t> ec;
    std::vector<double> evl;
    if (ell_pack(A, kEllMinRows, kEllMaxWidth, ell_w, ell_lanes, ec, evl)) {
      ptr.alloc(0);
      col.upload(ec);
      val.upload_as(evl);
      return;
    }
    ptr.upload(A.ptr); col.upload(A.col); val.upload_as(A.val);
    lanes = csr_lanes(A.n, A.val.size());
  }
  // See k_csr for the modes; outputs must not alias x (nor f in mode 4).
  void apply(hipStream_t s, int mode, const R* x, const R* f, R* y, const R* M, R* xio, const int* F) const {
    if (ell_w) ell<R>(s, mode, n, ell_w, ell_lanes, col.p, val.p, x, f, y, M, xio, F);
    else csr<R>(s, mode, n, ptr.p, col.p, val.p, x, f, y, M, xio, F, lanes);
  }
};
template <class R>
struct DevLevel {
  DevCSR<R> A, P, Rs;
  DBuf<R> M, f, u, t;
};

// V(1,1) cycle with spai0 relaxation from a zero guess on a sparse hierarchy (level l and
// below; the last level is solved with its dense inverse): u_l = V_l(f_l).  Every step is
// out of place, so the cycle is a fixed linear operator.  All launches are predicated on
// the active flag Fp (may be null).
template <class R>
void csr_vcycle(hipStream_t s, std::vector<std::unique_ptr<DevLevel<R>>>& lv, const R* coarse_inv, int coarse_n,
                size_t l, const R* f, R* u, const int* Fp) {
  DevLevel<R>& L = *lv[l];
  if (l + 1 == lv.size()) { dense<R>(s, coarse_n, coarse_inv, f, u, Fp); return; }
  DevLevel<R>& C = *lv[l + 1];
  L.A.apply(s, 4, nullptr, f, L.t, L.M, u, Fp);           // u = M f, t = f - A u
  L.Rs.apply(s, 0, L.t, nullptr, C.f, nullptr, nullptr, Fp);
  csr_vcycle(s, lv, coarse_inv, coarse_n, l + 1, C.f.p, C.u.p, Fp);
  L.P.apply(s, 2, C.u, u, L.t, nullptr, nullptr, Fp);      // t = u + P u_c
  L.A.apply(s, 3, L.t, f, u, L.M, nullptr, Fp);            // u = t + M (f - A t)
}

}  // namespace hoc
