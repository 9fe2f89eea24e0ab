// hoc_dev.hpp - device buffers, error types and the CSR level structures shared by the
// FR time-step solver (hoc_solver.hip) and the CSR linear-solver plug-in
// (hoc_linsolver.hip).
#pragma once

#include <hip/hip_runtime.h>

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

template <class R>
struct DevCSR {
  int n = 0, lanes = 1;
  DBuf<int32_t> ptr, col;
  DBuf<R> val;
  void upload(const CSR& A) {
    n = A.n; ptr.upload(A.ptr); col.upload(A.col); val.upload_as(A.val);
    lanes = csr_lanes(A.n, A.val.size());
  }
};
template <class R>
struct DevLevel {
  DevCSR<R> A, P, Rs;
  DBuf<R> M, f, u, t;
};

// V(1,1) cycle with spai0 relaxation from a zero guess on a CSR hierarchy (level l and
// below; the last level is solved with its dense inverse): u_l = V_l(f_l).  All launches
// are predicated on the active flag Fp (may be null).
template <class R>
void csr_vcycle(hipStream_t s, std::vector<std::unique_ptr<DevLevel<R>>>& lv, const R* coarse_inv, int coarse_n,
                size_t l, const R* f, R* u, const int* Fp) {
  DevLevel<R>& L = *lv[l];
  if (l + 1 == lv.size()) { dense<R>(s, coarse_n, coarse_inv, f, u, Fp); return; }
  DevLevel<R>& C = *lv[l + 1];
  // pre-relaxation from zero (u = M f) fused with its residual t = f - A u
  csr<R>(s, 4, L.A.n, L.A.ptr, L.A.col, L.A.val, nullptr, f, L.t, L.M, u, Fp, L.A.lanes);
  csr<R>(s, 0, C.A.n, L.Rs.ptr, L.Rs.col, L.Rs.val, L.t, nullptr, C.f, nullptr, nullptr, Fp, L.Rs.lanes);
  csr_vcycle(s, lv, coarse_inv, coarse_n, l + 1, C.f.p, C.u.p, Fp);
  csr<R>(s, 2, L.A.n, L.P.ptr, L.P.col, L.P.val, C.u, nullptr, u, nullptr, nullptr, Fp, L.P.lanes);
  csr<R>(s, 3, L.A.n, L.A.ptr, L.A.col, L.A.val, u, f, nullptr, L.M, u, Fp, L.A.lanes);
}

}  // namespace hoc
