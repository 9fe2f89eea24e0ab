/*
 * ho_oracle.cpp - FP64 CPU oracle for the HO-CXX time-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  Plain,
 * single-threaded C++ that restates the reference solver's arithmetic loop by
 * loop and in the reference's operation order, so the GPU path can be compared
 * against it field by field.  Every function cites the reference lines it
 * follows (paths relative to the reference repository root).
 *
 * The reference delegates the Poisson solve to AMGCL 1.4.0 (not vendored, not
 * installed): make_solver<amg<builtin<double>, smoothed_aggregation, spai0>,
 * bicgstab> with tol/maxiter set at src/HO_2D.cpp:3066-3072 and all other
 * parameters at their defaults.  Section "AMGCL restatement" below restates that
 * published algorithm (plain aggregation with eps_strong=0.08 halved per level,
 * prolongator smoothing with omega=2/3 on the filtered matrix, Galerkin coarse
 * operators, V(1,1), skyline-LU direct solve once n <= 3000, right-preconditioned
 * BiCGSTAB).  With the default spai0 relaxation that solver stalls for K >= 5
 * (the FR Laplacian has lambda_max(D^-1 A) ~ 3.5 there), so by default the oracle
 * relaxes with AMGCL's relaxation::chebyshev (degree 3, interval [lmax/30, lmax])
 * and applies the hierarchy to the row-scaled system (W A) psi = W b, W = w_i w_j,
 * which is symmetric on affine meshes; the stopping test stays on the unscaled
 * residual ||b - A psi|| / ||b||.  Environment knobs reproduce the default
 * configuration: HO_ORACLE_RELAX=0 (spai0), HO_ORACLE_UNSCALED=1,
 * HO_ORACLE_SA_RHO=1 (omega = 4/3 / rho), HO_ORACLE_CHEB_DEG.  A converged solve of
 * A psi = b does not depend on the preconditioner beyond the tolerance.
 *
 * Pins: every function is checked against what the mathematics fixes (tests/test_oracle_pins.py:
 * tables, metrics, Laplacian, manufactured Poisson, RK4; tests/test_oracle_pins_flow.py: advective
 * conservation, upwinding and polynomial exactness, the wall-circulation identity, Neumann data,
 * inflow and hybrid boundary programmes) or against the reference's own src/Mesh.cpp (connectivity).
 * Parity unpinned: only AMGCL's exact hierarchy, which a converged solve does not depend on.
 */
#include "ho_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <map>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include <omp.h>

namespace {

enum { SOUTH = 0, EAST = 1, NORTH = 2, WEST = 3 };
enum { DIRICHLET = 0, NEUMANN = 1 };

// src/MathHelpers.hpp:23-27
const int i2f[4] = {WEST, EAST, SOUTH, NORTH};
const int f2i[4] = {NORTH, EAST, WEST, SOUTH};
const double sgn_tab[4] = {-1., 1., -1., 1.};
const double RK4_coeff[4] = {0., 0.5, 0.5, 1.};
const double RK2_coeff[2] = {0., 1.};

std::string g_last_error;

inline double minus_one_to_power(int n) { return (n % 2 == 0 ? 1. : -1.); }

// src/MathHelpers.hpp:33-51
double Legendre(int n, double x) {
  double LegendreP;
  if (!n) LegendreP = 1.;
  else if (n == 1) LegendreP = x;
  else {
    double xn = -1., xnm2 = 1., xnm1 = x;
    for (int i = 2; i <= n; ++i) {
      xn = ((2. * i - 1.) * x * xnm1 - (i - 1.) * xnm2) / i;
      xnm2 = xnm1;
      xnm1 = xn;
    }
    LegendreP = xn;
  }
  return LegendreP;
}

// src/MathHelpers.hpp:53-74: Gauss elimination without pivoting (modifies LHS/RHS).
void Gauss_solver(int n, std::vector<double>& LHS, std::vector<double>& RHS, double* x) {
  for (int k = 0; k < n - 1; ++k) {
    for (int i = k + 1; i < n; ++i) {
      double c = LHS[i * n + k] / LHS[k * n + k];
      LHS[i * n + k] = 0.;
      RHS[i] = RHS[i] - c * RHS[k];
      for (int j = k + 1; j < n; ++j) LHS[i * n + j] -= c * LHS[k * n + j];
    }
  }
  x[n - 1] = RHS[n - 1] / LHS[(n - 1) * n + n - 1];
  for (int i = n - 2; i >= 0; i--) {
    double c = 0.;
    for (int j = i + 1; j < n; ++j) c += LHS[i * n + j] * x[j];
    x[i] = (RHS[i] - c) / LHS[i * n + i];
  }
}

// Dense inverse by Gauss-Jordan with partial pivoting (stands where the reference
// calls Eigen::MatrixXd::inverse(), src/HO_2D.cpp:1669).
std::vector<double> dense_inverse(int n, std::vector<double> a) {
  std::vector<double> inv(n * n, 0.);
  for (int i = 0; i < n; ++i) inv[i * n + i] = 1.;
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(a[r * n + c]) > std::fabs(a[p * n + c])) p = r;
    if (a[p * n + c] == 0.) throw std::runtime_error("singular Neumann matrix");
    if (p != c)
      for (int j = 0; j < n; ++j) { std::swap(a[p * n + j], a[c * n + j]); std::swap(inv[p * n + j], inv[c * n + j]); }
    double d = 1. / a[c * n + c];
    for (int j = 0; j < n; ++j) { a[c * n + j] *= d; inv[c * n + j] *= d; }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      double f = a[r * n + c];
      if (f == 0.) continue;
      for (int j = 0; j < n; ++j) { a[r * n + j] -= f * a[c * n + j]; inv[r * n + j] -= f * inv[c * n + j]; }
    }
  }
  return inv;
}

// Gauss-Legendre nodes/weights: closed forms of src/HO_2D.cpp:488-521 for K<=5;
// Newton iteration on P_K for K>=6 (beyond the reference's limit at :522-525).
void gauss_legendre(int K, std::vector<double>& x, std::vector<double>& w) {
  x.assign(K, 0.);
  w.assign(K, 0.);
  if (K == 1) { x[0] = 0.; w[0] = 2.; }
  else if (K == 2) { x[1] = 1. / std::sqrt(3.); x[0] = -x[1]; w[0] = w[1] = 1.0; }
  else if (K == 3) { x[2] = std::sqrt(0.6); x[1] = 0.; x[0] = -x[2]; w[2] = w[0] = 5.0 / 9.0; w[1] = 8.0 / 9.0; }
  else if (K == 4) {
    x[3] = std::sqrt((15. + 2. * std::sqrt(30.)) / 35.);
    x[2] = std::sqrt((15. - 2. * std::sqrt(30.)) / 35.);
    x[1] = -x[2]; x[0] = -x[3];
    w[3] = w[0] = (18. - std::sqrt(30.)) / 36.;
    w[2] = w[1] = (18. + std::sqrt(30.)) / 36.;
  } else if (K == 5) {
    x[4] = std::sqrt(5. + 2. * std::sqrt(70.) / 7.) / 3.;
    x[3] = std::sqrt(5. - 2. * std::sqrt(70.) / 7.) / 3.;
    x[2] = 0.; x[1] = -x[3]; x[0] = -x[4];
    w[0] = w[4] = (322. - 13. * std::sqrt(70.)) / 900.;
    w[1] = w[3] = (322. + 13. * std::sqrt(70.)) / 900.;
    w[2] = 128. / 225.;
  } else {
    for (int i = 0; i < K; ++i) {
      double z = std::cos(M_PI * (i + 0.75) / (K + 0.5));
      for (int it = 0; it < 100; ++it) {
        double p0 = 1., p1 = z;
        for (int n = 2; n <= K; ++n) { double p2 = ((2. * n - 1.) * z * p1 - (n - 1.) * p0) / n; p0 = p1; p1 = p2; }
        double dp = K * (z * p1 - p0) / (z * z - 1.);
        double dz = p1 / dp;
        z -= dz;
        if (std::fabs(dz) < 1e-17) break;
      }
      double p0 = 1., p1 = z;
      for (int n = 2; n <= K; ++n) { double p2 = ((2. * n - 1.) * z * p1 - (n - 1.) * p0) / n; p0 = p1; p1 = p2; }
      double dp = K * (z * p1 - p0) / (z * z - 1.);
      x[K - 1 - i] = z;
      w[K - 1 - i] = 2. / ((1. - z * z) * dp * dp);
    }
    for (int i = 0; i < K / 2; ++i) {  // enforce exact symmetry
      double a = 0.5 * (x[K - 1 - i] - x[i]); x[K - 1 - i] = a; x[i] = -a;
      double b = 0.5 * (w[K - 1 - i] + w[i]); w[K - 1 - i] = b; w[i] = b;
    }
    if (K % 2) x[K / 2] = 0.;
  }
}

// --------------------------------------------------------------------------------------------
// CSR matrix + AMGCL restatement
// --------------------------------------------------------------------------------------------
struct CSR {
  int n = 0, m = 0;
  std::vector<int> ptr, col;
  std::vector<double> val;
};

void spmv(const CSR& A, const double* x, double* y) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < A.n; ++i) {
    double s = 0.;
    for (int j = A.ptr[i]; j < A.ptr[i + 1]; ++j) s += A.val[j] * x[A.col[j]];
    y[i] = s;
  }
}

CSR transpose(const CSR& A) {
  CSR T;
  T.n = A.m; T.m = A.n;
  T.ptr.assign(T.n + 1, 0);
  for (int c : A.col) T.ptr[c + 1]++;
  for (int i = 0; i < T.n; ++i) T.ptr[i + 1] += T.ptr[i];
  T.col.resize(A.col.size()); T.val.resize(A.val.size());
  std::vector<int> pos(T.ptr.begin(), T.ptr.end() - 1);
  for (int i = 0; i < A.n; ++i)
    for (int j = A.ptr[i]; j < A.ptr[i + 1]; ++j) {
      int p = pos[A.col[j]]++;
      T.col[p] = i; T.val[p] = A.val[j];
    }
  return T;
}

// C = A*B (row-wise accumulation with a dense marker; columns sorted).
CSR multiply(const CSR& A, const CSR& B) {
  CSR C;
  C.n = A.n; C.m = B.m;
  C.ptr.assign(C.n + 1, 0);
  std::vector<int> marker(B.m, -1);
  std::vector<double> acc(B.m, 0.);
  std::vector<int> cols;
  for (int i = 0; i < A.n; ++i) {
    cols.clear();
    for (int ja = A.ptr[i]; ja < A.ptr[i + 1]; ++ja) {
      int k = A.col[ja]; double va = A.val[ja];
      for (int jb = B.ptr[k]; jb < B.ptr[k + 1]; ++jb) {
        int c = B.col[jb];
        if (marker[c] != i) { marker[c] = i; acc[c] = 0.; cols.push_back(c); }
        acc[c] += va * B.val[jb];
      }
    }
    std::sort(cols.begin(), cols.end());
    for (int c : cols) { C.col.push_back(c); C.val.push_back(acc[c]); }
    C.ptr[i + 1] = (int)C.col.size();
  }
  return C;
}

struct AMGLevel {
  CSR A, P, R;
  std::vector<double> M;            // spai0 weights
  std::vector<double> dinv;         // 1/a_ii (Chebyshev relaxation)
  double lmax = 0.;                 // lambda_max(D^-1 A) by power iteration
  std::vector<double> f, u, t;      // work vectors
};

struct AMG {
  int omega_mode = 0;
  int relax_mode = 1;   // 0 spai0 (AMGCL default; stalls for K >= 5), 1 Chebyshev
  int cheb_degree = 3;
  std::vector<AMGLevel> levels;
  // Coarse direct solve (AMGCL's builtin skyline_lu): reverse Cuthill-McKee ordering, then LU
  // without pivoting on the symmetric envelope (skyline) of the reordered matrix.
  int nc = 0;
  std::vector<int> perm, iperm;            // perm[new] = old
  std::vector<int> first;                  // first column of row i in L == first row of column i in U
  std::vector<size_t> off;
  std::vector<double> Lv, Uv, Dv;          // strictly lower rows, strictly upper columns, diagonal

  void rcm(const CSR& A) {
    const int n = A.n;
    std::vector<std::vector<int>> adj(n);
    for (int i = 0; i < n; ++i)
      for (int j = A.ptr[i]; j < A.ptr[i + 1]; ++j)
        if (A.col[j] != i) { adj[i].push_back(A.col[j]); adj[A.col[j]].push_back(i); }
    for (auto& a : adj) { std::sort(a.begin(), a.end()); a.erase(std::unique(a.begin(), a.end()), a.end()); }
    std::vector<char> seen(n, 0);
    perm.clear();
    while ((int)perm.size() < n) {
      int start = -1;
      for (int i = 0; i < n; ++i)
        if (!seen[i] && (start < 0 || adj[i].size() < adj[start].size())) start = i;
      std::vector<int> q{start};
      seen[start] = 1;
      for (size_t h = 0; h < q.size(); ++h) {
        std::vector<int> nb;
        for (int u : adj[q[h]]) if (!seen[u]) { seen[u] = 1; nb.push_back(u); }
        std::sort(nb.begin(), nb.end(), [&](int a, int b) { return adj[a].size() < adj[b].size(); });
        q.insert(q.end(), nb.begin(), nb.end());
      }
      perm.insert(perm.end(), q.begin(), q.end());
    }
    std::reverse(perm.begin(), perm.end());
    iperm.assign(n, 0);
    for (int i = 0; i < n; ++i) iperm[perm[i]] = i;
  }

  void factor_coarse(const CSR& A) {
    nc = A.n;
    rcm(A);
    first.assign(nc, 0);
    for (int i = 0; i < nc; ++i) first[i] = i;
    for (int io = 0; io < nc; ++io)
      for (int j = A.ptr[io]; j < A.ptr[io + 1]; ++j) {
        int r = iperm[io], c = iperm[A.col[j]];
        if (c < r) first[r] = std::min(first[r], c);
        else if (r < c) first[c] = std::min(first[c], r);
      }
    off.assign(nc + 1, 0);
    for (int i = 0; i < nc; ++i) off[i + 1] = off[i] + (i - first[i]);
    Lv.assign(off[nc], 0.); Uv.assign(off[nc], 0.); Dv.assign(nc, 0.);
    for (int io = 0; io < nc; ++io)
      for (int j = A.ptr[io]; j < A.ptr[io + 1]; ++j) {
        int r = iperm[io], c = iperm[A.col[j]];
        if (c < r) Lv[off[r] + (c - first[r])] += A.val[j];
        else if (r < c) Uv[off[c] + (r - first[c])] += A.val[j];
        else Dv[r] += A.val[j];
      }
    // Doolittle LU on the envelope: step i finishes row i of L, column i of U and U(i,i)
    for (int i = 0; i < nc; ++i) {
      const int fi = first[i];
      double* Li = &Lv[off[i]];
      double* Ui = &Uv[off[i]];
      for (int c = fi; c < i; ++c) {
        const int fc = first[c];
        const int k0 = std::max(fi, fc);
        const double* Uc = &Uv[off[c]];
        const double* Lc = &Lv[off[c]];
        double sl = Li[c - fi], su = Ui[c - fi];
        for (int k = k0; k < c; ++k) { sl -= Li[k - fi] * Uc[k - fc]; su -= Lc[k - fc] * Ui[k - fi]; }
        Li[c - fi] = Dv[c] != 0. ? sl / Dv[c] : 0.;
        Ui[c - fi] = su;
      }
      double d = Dv[i];
      for (int k = fi; k < i; ++k) d -= Li[k - fi] * Ui[k - fi];
      Dv[i] = d;
    }
  }

  void coarse_solve(const std::vector<double>& b, std::vector<double>& x) {
    std::vector<double> y(nc);
    for (int i = 0; i < nc; ++i) {
      double s = b[perm[i]];
      for (int k = first[i]; k < i; ++k) s -= Lv[off[i] + (k - first[i])] * y[k];
      y[i] = s;
    }
    for (int i = nc - 1; i >= 0; --i) {
      double yi = Dv[i] != 0. ? y[i] / Dv[i] : 0.;
      y[i] = yi;
      for (int k = first[i]; k < i; ++k) y[k] -= Uv[off[i] + (k - first[i])] * yi;
    }
    for (int i = 0; i < nc; ++i) x[perm[i]] = y[i];
  }


  void build(const CSR& A0) {
    const int coarse_enough = 3000;         // skyline_lu::coarse_enough()
    double eps_strong = 0.08;               // plain_aggregates::params default
    CSR A = A0;
    levels.clear();
    while (A.n > coarse_enough) {
      AMGLevel L;
      L.A = A;
      // spai0: M_i = a_ii / sum_j a_ij^2
      L.M.assign(A.n, 0.);
      for (int i = 0; i < A.n; ++i) {
        double num = 0., den = 0.;
        for (int j = A.ptr[i]; j < A.ptr[i + 1]; ++j) {
          if (A.col[j] == i) num += A.val[j];
          den += A.val[j] * A.val[j];
        }
        L.M[i] = num / den;
      }
      if (relax_mode == 1) {  // AMGCL relaxation::chebyshev: interval [lmax/30, lmax] of D^-1 A
        L.dinv.assign(A.n, 0.);
        for (int i = 0; i < A.n; ++i)
          for (int j = A.ptr[i]; j < A.ptr[i + 1]; ++j) if (A.col[j] == i) L.dinv[i] = 1. / A.val[j];
        std::vector<double> x(A.n), y(A.n);
        for (int i = 0; i < A.n; ++i) x[i] = 1. + 0.37 * std::sin(1.7 * i + 0.3);
        double lam = 1.;
        for (int it = 0; it < 25; ++it) {
          spmv(A, x.data(), y.data());
          double ny = 0., xy = 0., xx = 0.;
          for (int i = 0; i < A.n; ++i) { y[i] *= L.dinv[i]; ny += y[i] * y[i]; xy += x[i] * y[i]; xx += x[i] * x[i]; }
          lam = std::sqrt(ny / xx);
          for (int i = 0; i < A.n; ++i) x[i] = y[i] / std::sqrt(ny);
        }
        L.lmax = 1.1 * lam;
      }
      // plain aggregation
      std::vector<double> dia(A.n, 0.);
      for (int i = 0; i < A.n; ++i)
        for (int j = A.ptr[i]; j < A.ptr[i + 1]; ++j) if (A.col[j] == i) dia[i] += A.val[j];
      std::vector<char> strong(A.col.size(), 0);
      const double eps2 = eps_strong * eps_strong;
      for (int i = 0; i < A.n; ++i)
        for (int j = A.ptr[i]; j < A.ptr[i + 1]; ++j) {
          int c = A.col[j]; double v = A.val[j];
          strong[j] = (c != i) && (eps2 * dia[i] * dia[c] < v * v);
        }
      const int undefined = -1, removed = -2;
      std::vector<int> id(A.n, undefined);
      for (int i = 0; i < A.n; ++i) {
        bool any = false;
        for (int j = A.ptr[i]; j < A.ptr[i + 1]; ++j) if (strong[j]) { any = true; break; }
        if (!any) id[i] = removed;
      }
      int count = 0;
      std::vector<int> neib;
      for (int i = 0; i < A.n; ++i) {
        if (id[i] != undefined) continue;
        int cur = count++;
        id[i] = cur;
        neib.clear();
        for (int j = A.ptr[i]; j < A.ptr[i + 1]; ++j) {
          int c = A.col[j];
          if (strong[j] && id[c] != removed) { id[c] = cur; neib.push_back(c); }
        }
        for (int c : neib)
          for (int j = A.ptr[c]; j < A.ptr[c + 1]; ++j) {
            int cc = A.col[j];
            if (strong[j] && id[cc] == undefined) id[cc] = cur;
          }
      }
      if (!count) break;
      // renumber aggregates that vanished during expansion
      std::vector<int> cnt(count, 0);
      for (int i = 0; i < A.n; ++i) if (id[i] >= 0) cnt[id[i]]++;
      std::vector<int> newid(count, -1);
      int nc2 = 0;
      for (int a = 0; a < count; ++a) if (cnt[a]) newid[a] = nc2++;
      for (int i = 0; i < A.n; ++i) if (id[i] >= 0) id[i] = newid[id[i]];
      count = nc2;
      eps_strong *= 0.5;
      // smoothed prolongation P = (I - omega Df^-1 Af) Ptent.  AMGCL's default is omega = 2/3;
      // with estimate_spectral_radius it is 4/3 / rho(Df^-1 Af), rho from power iteration.
      double omega = 2. / 3.;
      if (omega_mode == 1) {
        std::vector<double> df(A.n, 0.);
        for (int i = 0; i < A.n; ++i)
          for (int j = A.ptr[i]; j < A.ptr[i + 1]; ++j) if (A.col[j] == i || !strong[j]) df[i] += A.val[j];
        std::vector<double> x(A.n), y(A.n);
        for (int i = 0; i < A.n; ++i) x[i] = 1. + 0.37 * std::sin(1.7 * i);
        double rho = 1.;
        for (int it = 0; it < 20; ++it) {
          for (int i = 0; i < A.n; ++i) {
            double sacc = 0.;
            for (int j = A.ptr[i]; j < A.ptr[i + 1]; ++j)
              if (A.col[j] == i) sacc += df[i] * x[i];
              else if (strong[j]) sacc += A.val[j] * x[A.col[j]];
            y[i] = df[i] != 0. ? sacc / df[i] : 0.;
          }
          double nx = 0., ny = 0.;
          for (int i = 0; i < A.n; ++i) { nx += x[i] * x[i]; ny += y[i] * y[i]; }
          rho = std::sqrt(ny / nx);
          for (int i = 0; i < A.n; ++i) x[i] = y[i] / std::sqrt(ny);
        }
        omega = 4. / 3. / rho;
      }
      CSR P;
      P.n = A.n; P.m = count;
      P.ptr.assign(A.n + 1, 0);
      std::vector<int> marker(count, -1);
      std::vector<double> acc(count, 0.);
      std::vector<int> cols;
      for (int i = 0; i < A.n; ++i) {
        double d = 0.;
        for (int j = A.ptr[i]; j < A.ptr[i + 1]; ++j)
          if (A.col[j] == i || !strong[j]) d += A.val[j];
        double s = (d != 0.) ? -omega / d : 0.;
        cols.clear();
        for (int j = A.ptr[i]; j < A.ptr[i + 1]; ++j) {
          int c = A.col[j];
          if (c != i && !strong[j]) continue;
          int g = id[c];
          if (g < 0) continue;
          double v = (c == i) ? 1. - omega : s * A.val[j];
          if (marker[g] != i) { marker[g] = i; acc[g] = 0.; cols.push_back(g); }
          acc[g] += v;
        }
        std::sort(cols.begin(), cols.end());
        for (int g : cols) { P.col.push_back(g); P.val.push_back(acc[g]); }
        P.ptr[i + 1] = (int)P.col.size();
      }
      L.P = P;
      L.R = transpose(P);
      CSR Ac = multiply(L.R, multiply(A, P));
      L.f.assign(A.n, 0.); L.u.assign(A.n, 0.); L.t.assign(A.n, 0.);
      levels.push_back(std::move(L));
      A = std::move(Ac);
    }
    // coarsest level: direct solve
    AMGLevel L;
    L.A = A;
    L.f.assign(A.n, 0.); L.u.assign(A.n, 0.); L.t.assign(A.n, 0.);
    levels.push_back(std::move(L));
    factor_coarse(levels.back().A);
  }

  // V(1,1) cycle started from x = 0 (amg::apply with pre_cycles = 1).
  void cycle(size_t l, const std::vector<double>& rhs, std::vector<double>& x) {
    AMGLevel& L = levels[l];
    if (l + 1 == levels.size()) { coarse_solve(rhs, x); return; }
    const int n = L.A.n;
    // pre-relaxation from x = 0
    std::fill(x.begin(), x.end(), 0.);
    relax(L, rhs, x);
    spmv(L.A, x.data(), L.t.data());
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) L.t[i] = rhs[i] - L.t[i];
    AMGLevel& C = levels[l + 1];
    spmv(L.R, L.t.data(), C.f.data());
    std::fill(C.u.begin(), C.u.end(), 0.);
    cycle(l + 1, C.f, C.u);
    spmv(L.P, C.u.data(), L.t.data());
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) x[i] += L.t[i];
    relax(L, rhs, x);  // post-relaxation
  }

  // One relaxation sweep x <- x + S (rhs - A x): spai0, or Chebyshev of degree cheb_degree.
  void relax(AMGLevel& L, const std::vector<double>& rhs, std::vector<double>& x) {
    const int n = L.A.n;
    std::vector<double> r(n), d(n);
    spmv(L.A, x.data(), r.data());
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) r[i] = rhs[i] - r[i];
    if (relax_mode == 0) {
#pragma omp parallel for schedule(static)
      for (int i = 0; i < n; ++i) x[i] += L.M[i] * r[i];
      return;
    }
    const double lmax = L.lmax, lmin = lmax / 30.;
    const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin), sigma = theta / delta;
    double rho = 1. / sigma;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) { d[i] = L.dinv[i] * r[i] / theta; x[i] += d[i]; }
    for (int k = 1; k < cheb_degree; ++k) {
      spmv(L.A, x.data(), r.data());
      const double rho_new = 1. / (2. * sigma - rho);
#pragma omp parallel for schedule(static)
      for (int i = 0; i < n; ++i) {
        d[i] = rho_new * rho * d[i] + 2. * rho_new / delta * L.dinv[i] * (rhs[i] - r[i]);
        x[i] += d[i];
      }
      rho = rho_new;
    }
  }
  void apply(const std::vector<double>& rhs, std::vector<double>& x) { cycle(0, rhs, x); }
};

double dot(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0.;
  const long n = (long)a.size();
#pragma omp parallel for schedule(static) reduction(+ : s)
  for (long i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

// Right-preconditioned BiCGSTAB (AMGCL solver::bicgstab, pside = right).
// Returns iterations; rel = ||r||/||b||.  Zero rhs clears x (as AMGCL does).
// With wts != nullptr the system is the row-scaled one (W A) x = W b and the stopping test uses
// the residual of the unscaled system, ||W^-1 r|| / ||W^-1 b||.
double wnorm(const std::vector<double>& r, const std::vector<double>* wts) {
  if (!wts) return std::sqrt(dot(r, r));
  double s = 0.;
  const long n = (long)r.size();
#pragma omp parallel for schedule(static) reduction(+ : s)
  for (long i = 0; i < n; ++i) { double u = r[i] / (*wts)[i]; s += u * u; }
  return std::sqrt(s);
}

int bicgstab(const CSR& A, AMG& M, const std::vector<double>& b, std::vector<double>& x,
             double tol, int maxiter, double& rel, const std::vector<double>* wts = nullptr) {
  const int n = A.n;
  double norm_b = wnorm(b, wts);
  if (norm_b < 2. * 2.220446049250313e-16) { std::fill(x.begin(), x.end(), 0.); rel = norm_b; return 0; }
  double eps = tol * norm_b;
  std::vector<double> r(n), rh(n), p(n), v(n), ph(n), sh(n), t(n);
  spmv(A, x.data(), r.data());
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) r[i] = b[i] - r[i];
  rh = r;
  double rho1 = 0., rho2 = 0., alpha = 0., omega = 0.;
  double res = wnorm(r, wts);
  int iter = 0;
  for (; iter < maxiter && res > eps; ++iter) {
    rho1 = dot(r, rh);
    if (rho1 == 0.) break;
    if (iter) {
      double beta = (rho1 * alpha) / (rho2 * omega);
#pragma omp parallel for schedule(static)
      for (int i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - omega * v[i]);
    } else {
      p = r;
    }
    M.apply(p, ph);
    spmv(A, ph.data(), v.data());
    double rv = dot(rh, v);
    if (rv == 0.) break;
    alpha = rho1 / rv;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) r[i] -= alpha * v[i];
    res = wnorm(r, wts);
    if (res > eps) {
      M.apply(r, sh);
      spmv(A, sh.data(), t.data());
      double tt = dot(t, t);
      omega = tt != 0. ? dot(t, r) / tt : 0.;
#pragma omp parallel for schedule(static)
      for (int i = 0; i < n; ++i) {
        x[i] += alpha * ph[i] + omega * sh[i];
        r[i] -= omega * t[i];
      }
      res = wnorm(r, wts);
      if (omega == 0.) { ++iter; break; }
    } else {
#pragma omp parallel for schedule(static)
      for (int i = 0; i < n; ++i) x[i] += alpha * ph[i];
    }
    rho2 = rho1;
  }
  rel = res / norm_b;
  return iter;
}

}  // namespace

// --------------------------------------------------------------------------------------------
// The solver state (mirrors the HO_2D members of src/HO_2D.hpp:72-181)
// --------------------------------------------------------------------------------------------
struct ho_oracle {
  // mesh (src/Mesh.hpp:155-215)
  int Lnod = 2, Lnod_in = 1, N_el = 0, N_nodes = 0, N_eb = 0, N_Gb = 0;
  std::vector<double> nx, ny;
  std::vector<int> elem_nodes;                      // N_el x Lnod^2
  std::vector<std::vector<int>> edges;              // node lists
  std::vector<int> elem_edges;                      // N_el x 4
  std::vector<std::string> bnames;
  std::vector<std::vector<int>> bedges;             // per boundary, boundary edge indices
  std::vector<int> nbr, nbr_side, is_bnd, bnd_index;  // N_el x 4 (by cell side)
  std::vector<int> be_elem, be_side;                // boundary_edges

  // parameters
  int K = 2, problem_type = 1, time_integration_type = 4, advection_type = 1;
  double Reyn_inv = 0.01, dt = 1e-3, tol = 1e-15, current_time = 0.;
  int maxiter = 150;
  double g_prime[2];

  // bases (src/HO_2D.hpp:103-113), [k][LR] etc. stored flat
  std::vector<double> sps, wgt, gps;
  std::vector<double> sbb, sbgb, ssgb;   // sps_boundary_basis[K][2], sps_boundary_grad_basis[K][2], sps_sps_grad_basis[K][K]
  std::vector<double> gbb, gbgb, gsb, sgb, gsgb;  // gps_boundary_basis[L][2], gps_boundary_grad_basis[L][2], gps_sps_basis[L][K], sps_gps_basis[K][L], gps_sps_grad_basis[L][K]
  std::vector<double> srad, sgrad;       // sps_radau[K][2], sps_grad_radau[K][2]

  // metrics
  std::vector<double> dxdxi, dxdeta, dydxi, dydeta, vol_jac, dxsidx, G;  // dxsidx,G: [el][j][i][2][2]
  std::vector<double> f_dxdxi, f_dxdeta, f_dydxi, f_dydeta, face_jac, face_A, face_B, face_An;  // [el][4][K]
  std::vector<double> GB;  // [el][2][2][K][2]

  // state
  std::vector<double> vorticity, stream_function, vel;  // vel: [el][j][i][2]
  std::vector<double> RHS_adv, RHS_diff, k1, k2, k3, k4, vort;
  std::vector<unsigned char> BC_no_slip, BC_switch_Poisson, BC_switch_diffusion;
  std::vector<double> BC_parl_vel, BC_normal_vel, BC_Poisson, BC_advection, BC_diffusion, BC_vorticity, velocity_jump;
  std::vector<double> boundary_source;  // [eb][K2][K]

  // hybrid operation (problem_type 11; src/HO_2D.cpp:893-1134, 2104-2191, 3694-4278)
  std::vector<double> edge_normvel, edge_parlvel;  // [N_eb] values passed with the boundary
  std::vector<int> edge_nval;                      // 0, 1 (normal) or 2 (tangential, normal)
  std::vector<double> VelNorm_start, VelNorm_end, VelParl_start, VelParl_end, BVort_start, BVort_end;  // [N_eb][K]
  std::vector<double> Vort_start, Vort_end, Vort_wgt;  // [D]
  double time_start = 0., time_end = 0.;
  int ti = 0;

  // Poisson
  CSR A;                  // the reference's matrix (src/HO_2D.cpp:3011-3063)
  CSR As;                 // W A, rows scaled by the Gauss weights w_i w_j (solve_scaled)
  std::vector<double> Wd; // w_i w_j per dof
  bool solve_scaled = true;
  AMG amg;
  int last_iters = 0;
  double last_rel = 0.;
  int64_t n_solves = 0, total_iters = 0;
  std::string err;

  int K2() const { return K * K; }
  int v3(int el, int j, int i) const { return (el * K + j) * K + i; }
  int f3(int el, int ijp, int k) const { return (el * 4 + ijp) * K + k; }
  int b2(int eb, int k) const { return eb * K + k; }

  int tensor2FEM(int i) const {  // src/Mesh.cpp:27-32
    if (!i) return 0;
    else if (i == Lnod_in) return 1;
    else return i + 1;
  }
  int tensor2FEM(int i, int j) const {  // src/Mesh.cpp:34-91
    int t2f = -1;
    if (Lnod_in == 1) t2f = i + 3 * j - 2 * i * j;
    else if (Lnod_in == 2) {
      static const int m[3][3] = {{0, 4, 1}, {7, 8, 5}, {3, 6, 2}};
      t2f = m[j][i];
    } else if (Lnod_in == 3) {
      static const int m[4][4] = {{0, 4, 5, 1}, {11, 12, 13, 6}, {10, 15, 14, 7}, {3, 9, 8, 2}};
      t2f = m[j][i];
    } else throw std::runtime_error("only up to cubic geometry (Lnod<=4) is supported");
    return t2f;
  }

  // --- src/Mesh.cpp:342-439 (process_mesh), with hashed edge/neighbour search -------------------
  void process_mesh() {
    N_nodes = (int)nx.size();
    N_eb = (int)edges.size();
    elem_edges.assign(N_el * 4, -1);
    // edges: search existing edge with identical node list or reversed end nodes; the first
    // match in index order wins (src/Mesh.cpp:380-389); the found edge takes this element's
    // orientation (src/Mesh.cpp:385).
    std::unordered_map<uint64_t, std::vector<int>> by_pair;
    auto key = [](int a, int b) { if (a > b) std::swap(a, b); return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b; };
    for (int e = 0; e < (int)edges.size(); ++e) by_pair[key(edges[e][0], edges[e][1])].push_back(e);
    const int ne = Lnod;  // nodes per edge
    for (int el = 0; el < N_el; ++el) {
      for (int s = SOUTH; s <= WEST; ++s) {
        std::vector<int> ed(ne);
        const int* en = &elem_nodes[(size_t)el * Lnod * Lnod];
        ed[0] = en[s];
        ed[1] = en[(s + 1) % 4];
        for (int q = 0; q < ne - 2; ++q) ed[2 + q] = en[s * (ne - 2) + 4 + q];
        int found = -1;
        auto it = by_pair.find(key(ed[0], ed[1]));
        if (it != by_pair.end()) {
          for (int cand : it->second) {  // ascending index order
            const std::vector<int>& en2 = edges[cand];
            if (en2 == ed || (en2[1] == ed[0] && en2[0] == ed[1])) { found = cand; break; }
          }
        }
        if (found >= 0) {
          elem_edges[el * 4 + s] = found;
          edges[found] = ed;
        } else {
          edges.push_back(ed);
          int idx = (int)edges.size() - 1;
          by_pair[key(ed[0], ed[1])].push_back(idx);
          elem_edges[el * 4 + s] = idx;
        }
      }
    }
    // neighbours: lowest-index other element sharing the edge, its first matching side
    // (src/Mesh.cpp:402-424)
    std::vector<std::vector<std::pair<int, int>>> inc(edges.size());
    for (int el = 0; el < N_el; ++el)
      for (int s = 0; s < 4; ++s) inc[elem_edges[el * 4 + s]].push_back({el, s});
    nbr.assign(N_el * 4, -1); nbr_side.assign(N_el * 4, -1); is_bnd.assign(N_el * 4, 0); bnd_index.assign(N_el * 4, -1);
    for (int el = 0; el < N_el; ++el)
      for (int s = 0; s < 4; ++s) {
        int e = elem_edges[el * 4 + s];
        int best = -1, bside = -1;
        for (auto& pr : inc[e]) {
          if (pr.first == el) continue;
          if (best < 0 || pr.first < best || (pr.first == best && pr.second < bside)) { best = pr.first; bside = pr.second; }
        }
        if (best >= 0) { nbr[el * 4 + s] = best; nbr_side[el * 4 + s] = bside; }
        else {
          is_bnd[el * 4 + s] = 1;
          bnd_index[el * 4 + s] = e;
          if (e >= N_eb) throw std::runtime_error("process_mesh: an element side on the boundary is not a listed boundary edge");
        }
      }
    be_elem.assign(N_eb, -1); be_side.assign(N_eb, -1);
    for (int el = 0; el < N_el; ++el)
      for (int s = 0; s < 4; ++s)
        if (is_bnd[el * 4 + s]) { int e = bnd_index[el * 4 + s]; be_elem[e] = el; be_side[e] = s; }
    for (int e = 0; e < N_eb; ++e)
      if (be_elem[e] < 0) throw std::runtime_error("process_mesh: boundary edge not attached to any element");
  }

  // --- src/HO_2D.cpp:486-552 ------------------------------------------------------------------
  void setup_sps_gps() {
    gauss_legendre(K, sps, wgt);
    gps.assign(Lnod, 0.);
    if (Lnod == 2) { gps[1] = 1.; gps[0] = -1.; }
    else if (Lnod == 3) { gps[2] = 1.0; gps[1] = 0.0; gps[0] = -1.; }
    else if (Lnod == 4) { gps[3] = 1.0; gps[2] = 1. / 3.; gps[1] = -1. / 3.; gps[0] = -1.; }
    else throw std::runtime_error("Only up to cubic elements (4 nodes on each edge) is allowed!");
    g_prime[0] = -0.5 * K * (K + 1.);
    g_prime[1] = 0.5 * K * (K + 1.);
  }

  // --- src/HO_2D.cpp:554-673 ------------------------------------------------------------------
  void form_bases() {
    const int L = Lnod;
    sbb.assign(K * 2, 0.); sbgb.assign(K * 2, 0.); ssgb.assign(K * K, 0.); sgb.assign(K * L, 0.);
    gbb.assign(L * 2, 0.); gbgb.assign(L * 2, 0.); gsb.assign(L * K, 0.); gsgb.assign(L * K, 0.);
    srad.assign(K * 2, 0.); sgrad.assign(K * 2, 0.);
    std::vector<double> sgn_(K);
    double gn[2];
    for (int k = 0; k < K; ++k) {
      sbb[k * 2 + 0] = sbb[k * 2 + 1] = 1.;
      sbgb[k * 2 + 0] = sbgb[k * 2 + 1] = 0.;
      for (int m = 0; m < K; ++m) ssgb[k * K + m] = 0.;
      for (int m = 0; m < L; ++m) sgb[k * L + m] = 1.0;
      double den = 1.;
      for (int j = 0; j < K; ++j) {
        if (j == k) continue;
        den *= sps[k] - sps[j];
        sbb[k * 2 + 0] *= -1.0 - sps[j];
        sbb[k * 2 + 1] *= 1.0 - sps[j];
        gn[0] = gn[1] = 1.;
        for (int m = 0; m < K; ++m) sgn_[m] = 1.;
        for (int m = 0; m < L; ++m) sgb[k * L + m] *= gps[m] - sps[j];
        for (int i = 0; i < K; ++i) {
          if (i == k || i == j) continue;
          gn[0] *= -1.0 - sps[i];
          gn[1] *= 1.0 - sps[i];
          for (int m = 0; m < K; ++m) sgn_[m] *= sps[m] - sps[i];
        }
        for (int LR = 0; LR < 2; ++LR) sbgb[k * 2 + LR] += gn[LR];
        for (int m = 0; m < K; ++m) ssgb[k * K + m] += sgn_[m];
      }
      for (int LR = 0; LR < 2; ++LR) { sbb[k * 2 + LR] /= den; sbgb[k * 2 + LR] /= den; }
      for (int m = 0; m < K; ++m) ssgb[k * K + m] /= den;
      for (int m = 0; m < L; ++m) sgb[k * L + m] /= den;
    }
    for (int k = 0; k < L; ++k) {
      gbb[k * 2 + 0] = gbb[k * 2 + 1] = 1.;
      gbgb[k * 2 + 0] = gbgb[k * 2 + 1] = 0.;
      for (int m = 0; m < K; ++m) gsgb[k * K + m] = 0.0;
      for (int m = 0; m < K; ++m) gsb[k * K + m] = 1;
      double den = 1.;
      for (int j = 0; j < L; ++j) {
        if (j == k) continue;
        den *= gps[k] - gps[j];
        gbb[k * 2 + 1] *= 1. - gps[j];
        gbb[k * 2 + 0] *= -1. - gps[j];
        for (int m = 0; m < K; ++m) gsb[k * K + m] *= sps[m] - gps[j];
        gn[0] = gn[1] = 1.;
        for (int m = 0; m < K; ++m) sgn_[m] = 1.;
        for (int i = 0; i < L; ++i) {
          if (i == k || i == j) continue;
          gn[0] *= -1.0 - gps[i];
          gn[1] *= 1.0 - gps[i];
          for (int m = 0; m < K; ++m) sgn_[m] *= sps[m] - gps[i];
        }
        for (int LR = 0; LR < 2; ++LR) gbgb[k * 2 + LR] += gn[LR];
        for (int m = 0; m < K; ++m) gsgb[k * K + m] += sgn_[m];
      }
      for (int LR = 0; LR < 2; ++LR) gbb[k * 2 + LR] /= den;
      for (int LR = 0; LR < 2; ++LR) gbgb[k * 2 + LR] /= den;
      for (int m = 0; m < K; ++m) gsgb[k * K + m] /= den;
      for (int m = 0; m < K; ++m) gsb[k * K + m] /= den;
    }
    if (K == 1) {
      srad[0] = srad[1] = 0.5;
      sgrad[1] = 0.5;
      sgrad[0] = -0.5;
    } else {
      double coef = 0.5 * minus_one_to_power(K);
      double coefd = 0.5 * K * minus_one_to_power(K);
      for (int i = 0; i < K; ++i) {
        srad[i * 2 + 0] = coef * (Legendre(K, sps[i]) - Legendre(K - 1, sps[i]));
        srad[i * 2 + 1] = coef * (Legendre(K, -sps[i]) - Legendre(K - 1, -sps[i]));
        sgrad[i * 2 + 0] = coefd * (Legendre(K, sps[i]) + Legendre(K - 1, sps[i])) / (1. + sps[i]);
        sgrad[i * 2 + 1] = -coefd * (Legendre(K, -sps[i]) + Legendre(K - 1, -sps[i])) / (1. - sps[i]);
      }
    }
  }

  // --- src/HO_2D.cpp:1240-1388 ----------------------------------------------------------------
  void form_metrics() {
    const int L = Lnod, K2_ = K2();
    dxdxi.assign(N_el * K2_, 0.); dxdeta.assign(N_el * K2_, 0.); dydxi.assign(N_el * K2_, 0.); dydeta.assign(N_el * K2_, 0.);
    vol_jac.assign(N_el * K2_, 0.); dxsidx.assign(N_el * K2_ * 4, 0.); G.assign(N_el * K2_ * 4, 0.);
    f_dxdxi.assign(N_el * 4 * K, 0.); f_dxdeta.assign(N_el * 4 * K, 0.); f_dydxi.assign(N_el * 4 * K, 0.); f_dydeta.assign(N_el * 4 * K, 0.);
    face_jac.assign(N_el * 4 * K, 0.); face_A.assign(N_el * 4 * K, 0.); face_B.assign(N_el * 4 * K, 0.); face_An.assign(N_el * 4 * K, 0.);
    GB.assign(N_el * 2 * 2 * K * 2, 0.);
    std::vector<double> lx(L * L), ly(L * L);
    auto GBi = [&](int el, int d, int t, int k, int r) { return (((el * 2 + d) * 2 + t) * K + k) * 2 + r; };
    for (int el = 0; el < N_el; ++el) {
      for (int j = 0; j < L; ++j)
        for (int i = 0; i < L; ++i) {
          int nd = elem_nodes[(size_t)el * L * L + tensor2FEM(i, j)];
          lx[j * L + i] = nx[nd]; ly[j * L + i] = ny[nd];
        }
      for (int j = 0; j < K; ++j)
        for (int i = 0; i < K; ++i) {
          int v = v3(el, j, i);
          for (int n = 0; n < L; ++n)
            for (int m = 0; m < L; ++m) {
              dxdxi[v] += gsgb[m * K + i] * gsb[n * K + j] * lx[n * L + m];
              dydxi[v] += gsgb[m * K + i] * gsb[n * K + j] * ly[n * L + m];
              dxdeta[v] += gsgb[n * K + j] * gsb[m * K + i] * lx[n * L + m];
              dydeta[v] += gsgb[n * K + j] * gsb[m * K + i] * ly[n * L + m];
            }
          vol_jac[v] = dxdxi[v] * dydeta[v] - dxdeta[v] * dydxi[v];
          const double oovj = 1.0 / vol_jac[v];
          dxsidx[v * 4 + 0] = oovj * dydeta[v];
          dxsidx[v * 4 + 1] = -oovj * dxdeta[v];
          dxsidx[v * 4 + 2] = -oovj * dydxi[v];
          dxsidx[v * 4 + 3] = oovj * dxdxi[v];
          double g[2][2] = {{-dxdeta[v], dydeta[v]}, {dxdxi[v], -dydxi[v]}};
          for (int r = 0; r < 2; ++r)
            for (int s = 0; s < 2; ++s) G[v * 4 + r * 2 + s] = (g[r][0] * g[s][0] + g[r][1] * g[s][1]) / vol_jac[v];
        }
      for (int k = 0; k < K; ++k) {
        for (int side = 0; side < 4; ++side) {
          double dx_dxsi = 0., dx_deta = 0., dy_dxsi = 0., dy_deta = 0.;
          for (int j = 0; j < L; ++j)
            for (int i = 0; i < L; ++i) {
              double X = lx[j * L + i], Y = ly[j * L + i];
              if (side < 2) {  // west (xsi=-1) / east (xsi=+1)
                int LR = side;
                dx_dxsi += gbgb[i * 2 + LR] * gsb[j * K + k] * X;
                dy_dxsi += gbgb[i * 2 + LR] * gsb[j * K + k] * Y;
                dx_deta += gbb[i * 2 + LR] * gsgb[j * K + k] * X;
                dy_deta += gbb[i * 2 + LR] * gsgb[j * K + k] * Y;
              } else {  // south (eta=-1) / north (eta=+1)
                int LR = side - 2;
                dx_dxsi += gbb[j * 2 + LR] * gsgb[i * K + k] * X;
                dy_dxsi += gbb[j * 2 + LR] * gsgb[i * K + k] * Y;
                dx_deta += gbgb[j * 2 + LR] * gsb[i * K + k] * X;
                dy_deta += gbgb[j * 2 + LR] * gsb[i * K + k] * Y;
              }
            }
          int f = f3(el, side, k);
          f_dxdxi[f] = dx_dxsi; f_dxdeta[f] = dx_deta; f_dydxi[f] = dy_dxsi; f_dydeta[f] = dy_deta;
          face_jac[f] = dx_dxsi * dy_deta - dx_deta * dy_dxsi;
          if (side < 2) face_A[f] = dx_deta * dx_deta + dy_deta * dy_deta;
          else face_A[f] = dx_dxsi * dx_dxsi + dy_dxsi * dy_dxsi;
          face_B[f] = dx_dxsi * dx_deta + dy_dxsi * dy_deta;
          face_An[f] = std::sqrt(face_A[f]);
          double g[2][2] = {{-dx_deta, dy_deta}, {dx_dxsi, -dy_dxsi}};
          int d = side / 2, t = side % 2;
          for (int r = 0; r < 2; ++r) GB[GBi(el, d, t, k, r)] = (g[d][0] * g[r][0] + g[d][1] * g[r][1]) / face_jac[f];
        }
      }
    }
  }

  int find_boundary(const char* name) const {
    for (int b = 0; b < N_Gb; ++b) if (bnames[b] == name) return b;
    return -1;
  }

  // y at flux point k of boundary edge e (src/HO_2D.cpp:752)
  double edge_y(int e, int k) const {
    double y = 0.;
    for (int j = 0; j < Lnod; ++j) y += gsb[j * K + k] * ny[edges[e][tensor2FEM(j)]];
    return y;
  }

  void inlet_yrange(double& y_min, double& y_max) const {  // src/HO_2D.cpp:730-741
    y_min = 0; y_max = 0;
    for (int b = 0; b < N_Gb; ++b) {
      if (bnames[b] != "inlet") continue;
      y_min = ny[edges[bedges[b][0]][0]];
      y_max = ny[edges[bedges[b][0]][0]];
      for (int e : bedges[b]) {
        y_min = std::min(ny[edges[e][0]], y_min);
        y_min = std::min(ny[edges[e][1]], y_min);
        y_max = std::max(ny[edges[e][0]], y_max);
        y_max = std::max(ny[edges[e][1]], y_max);
      }
    }
  }

  // --- src/HO_2D.cpp:675-1238 (problem types 1, 2, 3) ---------------------------------------
  void setup_IC_BC_SRC() {
    const int D = N_el * K2();
    vorticity.assign(D, 0.); stream_function.assign(D, 0.);
    BC_no_slip.assign(N_Gb, 0); BC_switch_Poisson.assign(N_eb, 0); BC_switch_diffusion.assign(N_eb, 0);
    BC_parl_vel.assign(N_eb * K, 0.); BC_normal_vel.assign(N_eb * K, 0.); BC_Poisson.assign(N_eb * K, 0.);
    BC_advection.assign(N_eb * K, 0.); BC_diffusion.assign(N_eb * K, 0.); BC_vorticity.assign(N_eb * K, 0.);
    velocity_jump.assign(N_eb * K, 0.);
    if (problem_type == 1) {
      for (int b = 0; b < N_Gb; ++b) BC_no_slip[b] = 1;
      std::fill(BC_switch_Poisson.begin(), BC_switch_Poisson.end(), DIRICHLET);
      std::fill(BC_switch_diffusion.begin(), BC_switch_diffusion.end(), NEUMANN);
      for (int b = 0; b < N_Gb; ++b)
        if (bnames[b] == "top")
          for (int e : bedges[b]) for (int i = 0; i < K; ++i) BC_parl_vel[b2(e, i)] = -1.;
    } else if (problem_type == 2) {
      double y_min, y_max;
      inlet_yrange(y_min, y_max);
      for (int b = 0; b < N_Gb; ++b) {
        const std::string& nm = bnames[b];
        if (nm == "inlet") {
          BC_no_slip[b] = 0;
          for (int e : bedges[b]) {
            BC_switch_diffusion[e] = DIRICHLET; BC_switch_Poisson[e] = DIRICHLET;
            for (int k = 0; k < K; ++k) {
              double y = edge_y(e, k);
              BC_Poisson[b2(e, k)] = std::pow(y - y_min, 2.) * (3. * y_max - y_min - 2. * y) / std::pow(y_max - y_min, 2.);
              BC_diffusion[b2(e, k)] = 6. * (2. * y - y_min - y_max) / std::pow(y_max - y_min, 2.);
              BC_normal_vel[b2(e, k)] = 6. * (y - y_min) * (y - y_max) / std::pow(y_max - y_min, 2.);
            }
          }
        } else if (nm == "outlet") {
          BC_no_slip[b] = 0;
          for (int e : bedges[b]) {
            BC_switch_diffusion[e] = NEUMANN; BC_switch_Poisson[e] = NEUMANN;
            for (int k = 0; k < K; ++k) { BC_diffusion[b2(e, k)] = 0.; BC_Poisson[b2(e, k)] = 0.; }
          }
        } else if (nm == "bottom") {
          BC_no_slip[b] = 1;
          for (int e : bedges[b]) {
            BC_switch_Poisson[e] = DIRICHLET; BC_switch_diffusion[e] = NEUMANN;
            for (int k = 0; k < K; ++k) { BC_Poisson[b2(e, k)] = 0.; BC_normal_vel[b2(e, k)] = 0.; BC_parl_vel[b2(e, k)] = 0.; }
          }
        } else if (nm == "top") {
          BC_no_slip[b] = 1;
          for (int e : bedges[b]) {
            BC_switch_Poisson[e] = DIRICHLET; BC_switch_diffusion[e] = NEUMANN;
            for (int k = 0; k < K; ++k) { BC_Poisson[b2(e, k)] = y_max - y_min; BC_normal_vel[b2(e, k)] = 0.; BC_parl_vel[b2(e, k)] = 0.; }
          }
        }
      }
    } else if (problem_type == 3) {
      double y_min, y_max;
      inlet_yrange(y_min, y_max);
      for (int b = 0; b < N_Gb; ++b) {
        const std::string& nm = bnames[b];
        if (nm == "inlet") {
          BC_no_slip[b] = 0;
          for (int e : bedges[b]) {
            BC_switch_diffusion[e] = DIRICHLET; BC_switch_Poisson[e] = DIRICHLET;
            for (int k = 0; k < K; ++k) {
              double y = edge_y(e, k);
              BC_Poisson[b2(e, k)] = y - y_min; BC_diffusion[b2(e, k)] = 0.; BC_normal_vel[b2(e, k)] = -1.;
            }
          }
        } else if (nm == "outlet") {
          BC_no_slip[b] = 0;
          for (int e : bedges[b]) {
            BC_switch_diffusion[e] = NEUMANN; BC_switch_Poisson[e] = NEUMANN;
            for (int k = 0; k < K; ++k) { BC_diffusion[b2(e, k)] = 0.; BC_Poisson[b2(e, k)] = 0.; }
          }
        } else if (nm == "top") {
          BC_no_slip[b] = 0;
          for (int e : bedges[b]) {
            BC_switch_Poisson[e] = DIRICHLET; BC_switch_diffusion[e] = NEUMANN;
            for (int k = 0; k < K; ++k) { BC_Poisson[b2(e, k)] = y_max - y_min; BC_normal_vel[b2(e, k)] = 0.; BC_parl_vel[b2(e, k)] = -1.; }
          }
        } else if (nm == "bottom") {
          BC_no_slip[b] = 0;
          for (int e : bedges[b]) {
            BC_switch_Poisson[e] = DIRICHLET; BC_switch_diffusion[e] = NEUMANN;
            for (int k = 0; k < K; ++k) { BC_Poisson[b2(e, k)] = 0.; BC_normal_vel[b2(e, k)] = 0.; BC_parl_vel[b2(e, k)] = 1.; }
          }
        } else if (nm == "object") {
          BC_no_slip[b] = 1;
          for (int e : bedges[b]) {
            BC_switch_Poisson[e] = DIRICHLET; BC_switch_diffusion[e] = NEUMANN;
            for (int k = 0; k < K; ++k) { BC_Poisson[b2(e, k)] = 0.5 * (y_max - y_min); BC_normal_vel[b2(e, k)] = 0.; BC_parl_vel[b2(e, k)] = 0.; }
          }
        }
      }
    } else if (problem_type == 11) {
      setup_hybrid_BCs();
    } else {
      throw std::runtime_error("problem_type unsupported (1 cavity, 2 BFS, 3 cylinder, 11 hybrid)");
    }
    // src/HO_2D.cpp:1227-1237
    for (int eb = 0; eb < N_eb; ++eb) {
      const int ijp = f2i[be_side[eb]];
      const double sign = 2. * (ijp % 2) - 1.;
      for (int k = 0; k < K; ++k) {
        BC_normal_vel[b2(eb, k)] *= sign;
        if (BC_switch_diffusion[eb] == NEUMANN) BC_diffusion[b2(eb, k)] *= sign;
      }
    }
  }

  // --- src/HO_2D.cpp:1533-1804 (assembly) + 3011-3063 (CSR with |v|>1e-14 filter) -------------
  void form_Laplace_operator_matrix() {
    const int Ksq = K * K, K4 = Ksq * Ksq;
    auto Gv = [&](int el, int j, int i, int r, int s) { return G[(v3(el, j, i)) * 4 + r * 2 + s]; };
    auto GBv = [&](int el, int d, int t, int k, int r) { return GB[((((el * 2 + d) * 2 + t) * K + k) * 2) + r]; };
    auto SBB = [&](int k, int lr) { return sbb[k * 2 + lr]; };
    auto SBGB = [&](int k, int lr) { return sbgb[k * 2 + lr]; };
    auto SSGB = [&](int k, int m) { return ssgb[k * K + m]; };
    auto SGR = [&](int k, int lr) { return sgrad[k * 2 + lr]; };
    boundary_source.assign((size_t)N_eb * Ksq * K, 0.);
    auto BS = [&](int eb, int ij, int a) -> double& { return boundary_source[((size_t)eb * Ksq + ij) * K + a]; };
    std::vector<double> lc(K4), ln(4 * K4);
    std::vector<std::vector<int>> columns((size_t)N_el * Ksq);
    std::vector<std::vector<double>> values((size_t)N_el * Ksq);
    std::vector<std::vector<double>> nbr_blocks((size_t)N_el);  // 4*K4 per element
    std::vector<double> NeuOrig(K * K), Neu;
    double tmp2, tmp3, coeff;
    for (int el = 0; el < N_el; ++el) {
      std::fill(lc.begin(), lc.end(), 0.);
      std::fill(ln.begin(), ln.end(), 0.);
      auto LC = [&](int ij, int rs) -> double& { return lc[ij * Ksq + rs]; };
      auto LN = [&](int side, int ij, int rs) -> double& { return ln[(side * Ksq + ij) * Ksq + rs]; };
      for (int j = 0; j < K; ++j)
        for (int i = 0; i < K; ++i) {
          const int ij = j * K + i;
          for (int m = 0; m < K; ++m)
            for (int r = 0; r < 2; r++) {
              const int a = m * (1 - r) + i * r;
              const int b = m * r + j * (1 - r);
              tmp3 = 0.;
              for (int s = 0; s < 2; s++) tmp3 += SBB(m, s) * SGR(j * r + i * (1 - r), s);
              double SNGLB1 = SSGB(m, i * (1 - r) + j * r);
              for (int d = 0; d < 2; ++d) {
                const double A_ = Gv(el, b, a, r, d);
                const int tmp1 = a * (1 - d) + b * d;
                for (int p = 0; p < K; ++p) {
                  const int j1 = p * d + b * (1 - d);
                  const int i1 = p * (1 - d) + a * d;
                  const int rs = j1 * K + i1;
                  tmp2 = 0.;
                  for (int t = 0; t < 2; ++t) tmp2 += 0.5 * SBB(p, t) * SGR(tmp1, t);
                  LC(ij, rs) += A_ * SNGLB1 * (SSGB(p, tmp1) - tmp2);
                  LC(ij, rs) += A_ * tmp3 * (-SSGB(p, tmp1) + tmp2);
                }
              }
            }
        }
      for (int d = 0; d < 2; ++d)
        for (int t = 0; t < 2; ++t) {
          const int ijp = 2 * d + t;
          const int elem_side = i2f[ijp];
          if (is_bnd[el * 4 + elem_side]) {
            const int edge_index = bnd_index[el * 4 + elem_side];
            if (BC_switch_Poisson[edge_index] == DIRICHLET) {
              for (int j = 0; j < K; ++j)
                for (int i = 0; i < K; ++i) {
                  const int ij = j * K + i;
                  for (int m = 0; m < K; ++m)
                    for (int r = 0; r < 2; ++r) {
                      const int a = m * (1 - r) + i * r;
                      const int b = m * r + j * (1 - r);
                      const int q = a * d + b * (1 - d);
                      double SNGLB1 = SSGB(m, i * (1 - r) + j * r);
                      double NGR1 = SGR(a * (1 - d) + b * d, t);
                      double A_ = Gv(el, b, a, r, d);
                      tmp2 = 0.;
                      for (int s = 0; s < 2; s++) tmp2 += SBB(m, s) * SGR(j * r + i * (1 - r), s);
                      for (int p = 0; p < K; ++p) {
                        const int j1 = p * d + b * (1 - d);
                        const int i1 = p * (1 - d) + a * d;
                        const int rs = j1 * K + i1;
                        LC(ij, rs) -= 0.5 * SNGLB1 * A_ * SBB(p, t) * NGR1;
                        LC(ij, rs) += 0.5 * tmp2 * A_ * SBB(p, t) * NGR1;
                      }
                      BS(edge_index, ij, (elem_side / 2) * (K - 2 * q - 1) + q) += SNGLB1 * A_ * NGR1;
                      BS(edge_index, ij, (elem_side / 2) * (K - 2 * q - 1) + q) -= tmp2 * A_ * NGR1;
                    }
                }
              for (int j = 0; j < K; ++j)
                for (int i = 0; i < K; ++i) {
                  const int ij = j * K + i;
                  const int alpha = i * d + j * (1 - d);
                  const double A_ = GBv(el, d, t, alpha, d);
                  const double NGR1 = SGR(j * d + i * (1 - d), t);
                  for (int m = 0; m < K; ++m) {
                    const int j1 = m * d + j * (1 - d);
                    const int i1 = m * (1 - d) + i * d;
                    const int rs = j1 * K + i1;
                    LC(ij, rs) += A_ * (SBGB(m, t) - g_prime[t] * SBB(m, t)) * NGR1;
                  }
                  BS(edge_index, ij, (elem_side / 2) * (K - 2 * alpha - 1) + alpha) += A_ * g_prime[t] * NGR1;
                  const double B_ = GBv(el, d, t, alpha, 1 - d);
                  for (int m = 0; m < K; ++m)
                    BS(edge_index, ij, (elem_side / 2) * (K - 2 * m - 1) + m) += B_ * SSGB(m, alpha) * NGR1;
                }
            } else if (BC_switch_Poisson[edge_index] == NEUMANN) {
              for (int i = 0; i < K; ++i) {
                for (int j = 0; j < K; ++j) NeuOrig[i * K + j] = -GBv(el, d, t, i, 1 - d) * SSGB(j, i);
                NeuOrig[i * K + i] += -GBv(el, d, t, i, d) * g_prime[t];
              }
              Neu = dense_inverse(K, NeuOrig);
              const double sign = t ? 1. : -1.;
              for (int j = 0; j < K; ++j)
                for (int i = 0; i < K; ++i) {
                  const int ij = j * K + i;
                  const int alpha = i * d + j * (1 - d);
                  BS(edge_index, ij, (elem_side / 2) * (K - 2 * alpha - 1) + alpha) +=
                      sign * SGR(j * d + i * (1 - d), t) * std::sqrt(GBv(el, d, t, alpha, d) * face_jac[f3(el, ijp, alpha)]);
                  for (int m = 0; m < K; ++m)
                    for (int r = 0; r < 2; r++) {
                      const int a = m * (1 - r) + i * r;
                      const int b = m * r + j * (1 - r);
                      const int q = a * d + b * (1 - d);
                      double SNGLB1 = SSGB(m, i * (1 - r) + j * r);
                      double NGR1 = SGR(a * (1 - d) + b * d, t);
                      double A_ = Gv(el, b, a, r, d);
                      tmp2 = 0.;
                      for (int s = 0; s < 2; s++) tmp2 += SBB(m, s) * SGR(j * r + i * (1 - r), s);
                      for (int p = 0; p < K; ++p) {
                        const int j1 = p * d + b * (1 - d);
                        const int i1 = p * (1 - d) + a * d;
                        const int rs = j1 * K + i1;
                        LC(ij, rs) -= 0.5 * SNGLB1 * A_ * SBB(p, t) * NGR1;
                        LC(ij, rs) += 0.5 * tmp2 * A_ * SBB(p, t) * NGR1;
                        BS(edge_index, ij, (elem_side / 2) * (K - 2 * p - 1) + p) += A_ * NGR1 * sign * Neu[q * K + p] *
                            std::sqrt(GBv(el, d, t, p, d) * face_jac[f3(el, ijp, p)]) * (tmp2 - SNGLB1);
                        double tmp4 = SBGB(p, t) - g_prime[t] * SBB(p, t);
                        for (int c = 0; c < K; ++c) {
                          const int j2 = p * d + c * (1 - d);
                          const int i2 = p * (1 - d) + c * d;
                          const int rs2 = j2 * K + i2;
                          LC(ij, rs2) += SNGLB1 * A_ * NGR1 * tmp4 * Neu[q * K + c] * GBv(el, d, t, c, d);
                          LC(ij, rs2) -= tmp2 * A_ * NGR1 * tmp4 * Neu[q * K + c] * GBv(el, d, t, c, d);
                        }
                      }
                    }
                }
            }
          } else {
            const int eln = nbr[el * 4 + elem_side];
            const int ijpm = f2i[nbr_side[el * 4 + elem_side]];
            const int dn = ijpm / 2, tn = ijpm % 2;
            for (int j = 0; j < K; ++j)
              for (int i = 0; i < K; ++i) {
                int ij = j * K + i;
                for (int m = 0; m < K; ++m)
                  for (int r = 0; r < 2; r++) {
                    const int a = m * (1 - r) + i * r;
                    const int b = m * r + j * (1 - r);
                    const int q = a * d + b * (1 - d);
                    const int Mq = ((d + t + dn + tn + 1) % 2) * (K - 2 * q - 1) + q;
                    double SNGLB1 = SSGB(m, i * (1 - r) + j * r);
                    double NGR1 = SGR(a * (1 - d) + b * d, t);
                    double A_ = Gv(el, b, a, r, d);
                    tmp2 = 0.;
                    for (int s = 0; s < 2; s++) tmp2 += SBB(m, s) * SGR(j * r + i * (1 - r), s);
                    for (int p = 0; p < K; ++p) {
                      const int j1 = p * dn + Mq * (1 - dn);
                      const int i1 = p * (1 - dn) + Mq * dn;
                      const int rs = j1 * K + i1;
                      LN(elem_side, ij, rs) += 0.5 * SNGLB1 * A_ * SBB(p, tn) * NGR1;
                      LN(elem_side, ij, rs) -= 0.5 * tmp2 * A_ * SBB(p, tn) * NGR1;
                    }
                  }
              }
            for (int j = 0; j < K; ++j)
              for (int i = 0; i < K; ++i) {
                const int ij = j * K + i;
                int alpha = i * d + j * (1 - d);
                int Malpha = ((d + t + dn + tn + 1) % 2) * (K - 2 * alpha - 1) + alpha;
                double A_ = GBv(el, d, t, alpha, d);
                double B_ = GBv(el, d, t, alpha, 1 - d);
                double An = GBv(eln, dn, tn, Malpha, dn);
                double Bn = GBv(eln, dn, tn, Malpha, 1 - dn);
                double NGR1 = SGR(j * d + i * (1 - d), t);
                double sign = (t == tn) ? -1. : 1.;
                for (int m = 0; m < K; ++m) {
                  int Mm = ((d + t + dn + tn + 1) % 2) * (K - 2 * m - 1) + m;
                  int j1 = m * d + j * (1 - d);
                  int i1 = m * (1 - d) + i * d;
                  int rs = j1 * K + i1;
                  LC(ij, rs) += (0.5 * A_ * SBGB(m, t) - 0.25 * g_prime[t] * A_ * SBB(m, t) +
                                 0.25 * sign * g_prime[tn] * An * SBB(m, t)) * NGR1;
                  j1 = m * dn + Malpha * (1 - dn);
                  i1 = m * (1 - dn) + Malpha * dn;
                  rs = j1 * K + i1;
                  LN(elem_side, ij, rs) += (0.5 * sign * An * SBGB(m, tn) - 0.25 * sign * g_prime[tn] * An * SBB(m, tn) +
                                            0.25 * g_prime[t] * A_ * SBB(m, tn)) * NGR1;
                  for (int p = 0; p < K; ++p) {
                    j1 = p * d + m * (1 - d);
                    i1 = p * (1 - d) + m * d;
                    rs = j1 * K + i1;
                    coeff = 0.25 * (B_ * SSGB(m, alpha) + sign * Bn * SSGB(Mm, Malpha)) * NGR1;
                    LC(ij, rs) += coeff * SBB(p, t);
                    j1 = p * dn + Mm * (1 - dn);
                    i1 = p * (1 - dn) + Mm * dn;
                    rs = j1 * K + i1;
                    LN(elem_side, ij, rs) += coeff * SBB(p, tn);
                  }
                }
              }
          }
        }
      // CSR rows (src/HO_2D.cpp:3023-3049): centre block first, then neighbours south..west
      for (int ij = 0; ij < Ksq; ++ij) {
        size_t row = (size_t)el * Ksq + ij;
        for (int rs = 0; rs < Ksq; ++rs)
          if (std::fabs(lc[ij * Ksq + rs]) > 1.e-14) { columns[row].push_back(el * Ksq + rs); values[row].push_back(lc[ij * Ksq + rs]); }
      }
      nbr_blocks[el] = ln;
    }
    for (int el = 0; el < N_el; ++el)
      for (int side = SOUTH; side <= WEST; ++side)
        if (!is_bnd[el * 4 + side]) {
          const int eln = nbr[el * 4 + side];
          for (int ij = 0; ij < Ksq; ++ij) {
            size_t row = (size_t)el * Ksq + ij;
            for (int rs = 0; rs < Ksq; ++rs) {
              double v = nbr_blocks[el][(side * Ksq + ij) * Ksq + rs];
              if (std::fabs(v) > 1.e-14) { columns[row].push_back(eln * Ksq + rs); values[row].push_back(v); }
            }
          }
        }
    A.n = A.m = N_el * Ksq;
    A.ptr.assign(A.n + 1, 0);
    A.col.clear(); A.val.clear();
    for (int r = 0; r < A.n; ++r) {
      A.col.insert(A.col.end(), columns[r].begin(), columns[r].end());
      A.val.insert(A.val.end(), values[r].begin(), values[r].end());
      A.ptr[r + 1] = (int)A.col.size();
    }
    amg.omega_mode = getenv("HO_ORACLE_SA_RHO") ? atoi(getenv("HO_ORACLE_SA_RHO")) : 0;
    amg.relax_mode = getenv("HO_ORACLE_RELAX") ? atoi(getenv("HO_ORACLE_RELAX")) : 1;
    if (getenv("HO_ORACLE_CHEB_DEG")) amg.cheb_degree = atoi(getenv("HO_ORACLE_CHEB_DEG"));
    if (getenv("HO_ORACLE_UNSCALED")) solve_scaled = false;
    Wd.assign(A.n, 0.);
    for (int el = 0; el < N_el; ++el)
      for (int j = 0; j < K; ++j)
        for (int i = 0; i < K; ++i) Wd[v3(el, j, i)] = wgt[j] * wgt[i];
    if (solve_scaled) {
      As = A;
      for (int r = 0; r < As.n; ++r)
        for (int j = As.ptr[r]; j < As.ptr[r + 1]; ++j) As.val[j] *= Wd[r];
      amg.build(As);
    } else {
      amg.build(A);
    }
  }

  // --- src/HO_2D.cpp:1979-2217 (problem types 1, 2, 3) ---------------------------------------
  // --- src/HO_2D.cpp:893-1134: boundary switches by name, then psi on inlet and wall edges
  // marched along the closed boundary loop (requires Lnod = 2, as the reference asserts) ---
  void setup_hybrid_BCs() {
    const int D = N_el * K2();
    VelNorm_start.assign(N_eb * K, 0.); VelNorm_end.assign(N_eb * K, 0.);
    VelParl_start.assign(N_eb * K, 0.); VelParl_end.assign(N_eb * K, 0.);
    BVort_start.assign(N_eb * K, 0.); BVort_end.assign(N_eb * K, 0.);
    Vort_start.assign(D, 0.); Vort_end.assign(D, 0.); Vort_wgt.assign(D, 0.);
    if (edge_nval.empty()) { edge_nval.assign(N_eb, 0); edge_normvel.assign(N_eb, 0.); edge_parlvel.assign(N_eb, 0.); }
    bool have_inlets = false;
    for (int b = 0; b < N_Gb; ++b) {
      const std::string& nm = bnames[b];
      if (nm == "open") {
        BC_no_slip[b] = 0;
        for (int e : bedges[b]) {
          BC_switch_Poisson[e] = NEUMANN; BC_switch_diffusion[e] = DIRICHLET;
          for (int k = 0; k < K; ++k) BC_Poisson[b2(e, k)] = 0.;
        }
      } else if (nm == "wall") {
        BC_no_slip[b] = 1;
        for (int e : bedges[b]) {
          BC_switch_Poisson[e] = DIRICHLET; BC_switch_diffusion[e] = NEUMANN;
          for (int k = 0; k < K; ++k) { BC_normal_vel[b2(e, k)] = 0.; BC_parl_vel[b2(e, k)] = 0.; }
        }
      } else if (nm == "inlet") {
        have_inlets = true;
        BC_no_slip[b] = 0;
        for (int e : bedges[b]) { BC_switch_Poisson[e] = DIRICHLET; BC_switch_diffusion[e] = DIRICHLET; }
      } else if (nm == "outlet") {
        BC_no_slip[b] = 0;
        for (int e : bedges[b]) {
          BC_switch_Poisson[e] = NEUMANN; BC_switch_diffusion[e] = NEUMANN;
          for (int k = 0; k < K; ++k) { BC_Poisson[b2(e, k)] = 0.; BC_diffusion[b2(e, k)] = 0.; }
        }
      } else {
        throw std::runtime_error("hybrid boundaries must be named wall, open, inlet or outlet");
      }
    }
    if (!have_inlets) return;
    if (Lnod != 2) throw std::runtime_error("hybrid inlets need Lnod = 2 (src/HO_2D.cpp:1049)");
    std::vector<int> bdry_of(N_eb, -1);
    for (int b = 0; b < N_Gb; ++b) for (int e : bedges[b]) bdry_of[e] = b;
    auto is = [&](int e, const char* nm) { return bnames[bdry_of[e]] == nm; };
    std::vector<int> firstnode(N_nodes, 0);
    for (int b = 0; b < N_Gb; ++b)
      if (bnames[b] != "open")
        for (int e : bedges[b]) firstnode[edges[e][0]] = e;
    size_t istart = 0;
    while (is((int)istart, "inlet") || is((int)istart, "outlet")) ++istart;
    std::vector<int> ordered{(int)istart};
    for (;;) {
      const int next = firstnode[edges[ordered.back()][1]];
      if (next == ordered[0]) break;
      ordered.push_back(next);
      if (ordered.size() > (size_t)N_eb) throw std::runtime_error("hybrid boundary edges do not form a loop");
    }
    std::vector<double> len(ordered.size());
    for (size_t q = 0; q < ordered.size(); ++q) {
      const auto& nd = edges[ordered[q]];
      len[q] = std::hypot(nx[nd[tensor2FEM(0)]] - nx[nd[tensor2FEM(1)]], ny[nd[tensor2FEM(0)]] - ny[nd[tensor2FEM(1)]]);
    }
    const size_t M = ordered.size();
    double psi_end = 0.;
    for (size_t q = 0; q < M; ++q) {
      const int e = ordered[q];
      const double psi_start = psi_end;
      if (is(e, "inlet")) {
        if (edge_nval[e] < 1) throw std::runtime_error("inlet edge without a normal velocity");
        const double vn = edge_normvel[e];
        psi_end -= len[q] * vn;
        const int qp = q == 0 ? (int)M - 1 : (int)q - 1, qn = q == M - 1 ? 0 : (int)q + 1;
        const double vp = is(ordered[qp], "inlet") ? edge_normvel[ordered[qp]] : 0.;
        const double vnx = is(ordered[qn], "inlet") ? edge_normvel[ordered[qn]] : 0.;
        const double w = (vnx - vn) / (len[qn] + len[q]) + (vn - vp) / (len[q] + len[qp]);
        for (int k = 0; k < K; ++k) {
          BC_Poisson[b2(e, k)] = gsb[0 * K + k] * psi_start + gsb[1 * K + k] * psi_end;
          BC_diffusion[b2(e, k)] = w;
          BC_normal_vel[b2(e, k)] = -vn;
        }
      } else if (is(e, "outlet")) {
        if (edge_nval[e] < 1) throw std::runtime_error("outlet edge without a normal velocity");
        psi_end -= len[q] * edge_normvel[e];
      } else if (is(e, "wall")) {
        for (int k = 0; k < K; ++k) BC_Poisson[b2(e, k)] = psi_end;
      }
    }
  }

  // --- src/HO_2D.cpp:2104-2191: blend the hybrid values across the outer step -------------
  void update_hybrid_BCs(double this_time) {
    const double fac1 = (this_time - time_start) / (time_end - time_start);
    const double fac2 = 1.0 - fac1;
    for (int b = 0; b < N_Gb; ++b) {
      const std::string& nm = bnames[b];
      const auto& eb = bedges[b];
      for (size_t i = 0; i < eb.size(); ++i) {
        const int e = eb[i];
        for (int j = 0; j < K; ++j) {
          if (nm == "open") {
            BC_normal_vel[b2(e, j)] = -(fac1 * VelNorm_start[i * K + j] + fac2 * VelNorm_end[i * K + j]);
            BC_parl_vel[b2(e, j)] = -(fac1 * VelParl_start[i * K + j] + fac2 * VelParl_end[i * K + j]);
            BC_Poisson[b2(e, j)] = BC_parl_vel[b2(e, j)];
            BC_diffusion[b2(e, j)] = fac1 * BVort_start[i * K + j] + fac2 * BVort_end[i * K + j];
          } else if (nm == "wall") {
            BC_normal_vel[b2(e, j)] = 0.; BC_parl_vel[b2(e, j)] = 0.;
          } else if (nm == "inlet") {
            BC_normal_vel[b2(e, j)] = -edge_normvel[e];
          }
        }
      }
    }
    const int D = N_el * K2();
    for (int v = 0; v < D; ++v)
      vorticity[v] = (1.0 - Vort_wgt[v]) * vorticity[v] + Vort_wgt[v] * (fac1 * Vort_start[v] + fac2 * Vort_end[v]);
  }

  // --- src/HO_2D.cpp:3994-4151: the hybrid driver's setters -------------------------------
  int open_boundary() const {
    for (int b = 0; b < N_Gb; ++b) if (bnames[b] == "open") return b;
    return -1;
  }
  void open_points(double* xy) const {  // getopenpts_d (:4007-4046)
    const int b = open_boundary();
    if (b < 0) return;
    for (size_t i = 0; i < bedges[b].size(); ++i) {
      const auto& nd = edges[bedges[b][i]];
      for (int j = 0; j < K; ++j) {
        double x = 0., y = 0.;
        for (int l = 0; l < Lnod; ++l) {
          x += gsb[l * K + j] * nx[nd[tensor2FEM(l)]];
          y += gsb[l * K + j] * ny[nd[tensor2FEM(l)]];
        }
        xy[2 * (i * K + j)] = x; xy[2 * (i * K + j) + 1] = y;
      }
    }
  }
  void set_open_vels(const double* xyvel) {  // setopenvels_d (:4050-4117)
    VelNorm_start = VelNorm_end; VelParl_start = VelParl_end;
    const int b = open_boundary();
    if (b < 0) return;
    for (size_t i = 0; i < bedges[b].size(); ++i) {
      const int e = bedges[b][i];
      const int el = be_elem[e], ijp = f2i[be_side[e]];
      const int t = ijp % 2, d = ijp / 2;
      const double coeff = 2. * t - 1.;
      for (int j = 0; j < K; ++j) {
        const double u = xyvel[2 * (i * K + j)], v = xyvel[2 * (i * K + j) + 1];
        double nxv, nyv;
        if (d == 1) { nxv = -f_dydxi[f3(el, ijp, j)]; nyv = f_dxdxi[f3(el, ijp, j)]; }
        else { nxv = f_dydeta[f3(el, ijp, j)]; nyv = -f_dxdeta[f3(el, ijp, j)]; }
        nxv *= coeff / face_An[f3(el, ijp, j)];
        nyv *= coeff / face_An[f3(el, ijp, j)];
        VelNorm_end[i * K + j] = nxv * u + nyv * v;
        VelParl_end[i * K + j] = nxv * v - nyv * u;
      }
    }
  }
  void set_open_vort(const double* w) {  // setopenvort_d (:4120-4141)
    const int b = open_boundary();
    if (b < 0) return;
    for (size_t i = 0; i < bedges[b].size(); ++i)
      for (int j = 0; j < K; ++j) { BVort_start[i * K + j] = BVort_end[i * K + j]; BVort_end[i * K + j] = w[K * i + j]; }
  }
  void set_soln_vort(const double* w) {  // setsolnvort_d (:4144-4166)
    Vort_start = Vort_end;
    Vort_end.assign(w, w + N_el * K2());
  }
  void set_p2g_weights(const double* w) { Vort_wgt.assign(w, w + N_el * K2()); }  // setptogweights_d (:4170-4184)
  void solveto(double outer_dt, int nstep, int integ, double Re) {  // solveto_d (:4191-4233)
    dt = outer_dt / (double)nstep;
    time_integration_type = integ;
    Reyn_inv = 1.0 / Re;
    time_start = time_end;
    current_time = time_start;
    time_end += outer_dt;
    for (int it = 0; it < nstep; ++it) {
      solve_advection_diffusion();
      ++ti;
      current_time += dt;
    }
  }

  void update_BCs(double this_time) {
    if (problem_type == 1) {
      // no time dependence (is_oscillating = false at :1985)
    } else if (problem_type == 2) {
      double y_min, y_max;
      inlet_yrange(y_min, y_max);
      const double time_coeff = 2. * std::fabs(std::sin(M_PI / 40. * this_time));
      for (int b = 0; b < N_Gb; ++b) {
        if (bnames[b] == "inlet") {
          for (int e : bedges[b])
            for (int k = 0; k < K; ++k) {
              double y = edge_y(e, k);
              BC_Poisson[b2(e, k)] = time_coeff * std::pow(y - y_min, 2.) * (3. * y_max - y_min - 2. * y) / std::pow(y_max - y_min, 2.);
              BC_diffusion[b2(e, k)] = time_coeff * 6. * (2. * y - y_min - y_max) / std::pow(y_max - y_min, 2.);
              BC_normal_vel[b2(e, k)] = time_coeff * 6. * (y - y_min) * (y - y_max) / std::pow(y_max - y_min, 2.);
            }
        } else if (bnames[b] == "top") {
          for (int e : bedges[b]) for (int k = 0; k < K; ++k) BC_Poisson[b2(e, k)] = time_coeff * (y_max - y_min);
        }
      }
    } else if (problem_type == 3) {
      double y_min, y_max;
      inlet_yrange(y_min, y_max);
      const double time_coeff = 2. * std::fabs(std::sin(M_PI / 40. * this_time));
      for (int b = 0; b < N_Gb; ++b) {
        const std::string& nm = bnames[b];
        if (nm == "inlet") {
          for (int e : bedges[b])
            for (int k = 0; k < K; ++k) {
              double y = edge_y(e, k);
              BC_Poisson[b2(e, k)] = time_coeff * (y - y_min);
              BC_normal_vel[b2(e, k)] = -time_coeff;
            }
        } else if (nm == "top") {
          for (int e : bedges[b])
            for (int k = 0; k < K; ++k) { BC_Poisson[b2(e, k)] = time_coeff * (y_max - y_min); BC_diffusion[b2(e, k)] = 0.; BC_parl_vel[b2(e, k)] = -time_coeff; }
        } else if (nm == "bottom") {
          for (int e : bedges[b]) for (int k = 0; k < K; ++k) BC_parl_vel[b2(e, k)] = time_coeff;
        } else if (nm == "object") {
          for (int e : bedges[b]) for (int k = 0; k < K; ++k) BC_Poisson[b2(e, k)] = time_coeff * 0.5 * (y_max - y_min);
        }
      }
    } else if (problem_type == 11) {
      update_hybrid_BCs(this_time);
    }
    for (int eb = 0; eb < N_eb; ++eb) {
      const int ijp = f2i[be_side[eb]];
      const double sign = 2. * (ijp % 2) - 1.;
      for (int k = 0; k < K; ++k) {
        BC_normal_vel[b2(eb, k)] *= sign;
        if (BC_switch_diffusion[eb] == NEUMANN) BC_diffusion[b2(eb, k)] *= sign;
      }
    }
  }

  // --- src/HO_2D.cpp:1841-1977 ----------------------------------------------------------------
  void solve_Poisson(const double* w) {
    const int Ksq = K * K, D = N_el * Ksq;
    std::vector<double> RHS(D);
    for (int el = 0; el < N_el; ++el)
      for (int j = 0; j < K; ++j)
        for (int i = 0; i < K; ++i) RHS[el * Ksq + j * K + i] = -w[v3(el, j, i)] * vol_jac[v3(el, j, i)];
    for (int eb = 0; eb < N_eb; ++eb)
      for (int j = 0; j < K; ++j)
        for (int i = 0; i < K; ++i) {
          const int ij = j * K + i;
          const int el = be_elem[eb];
          for (int alpha = 0; alpha < K; ++alpha)
            RHS[el * Ksq + ij] -= boundary_source[((size_t)eb * Ksq + ij) * K + alpha] * BC_Poisson[b2(eb, alpha)];
        }
    std::vector<double> x(stream_function);
    if (solve_scaled) {
      for (int i = 0; i < D; ++i) RHS[i] *= Wd[i];
      last_iters = bicgstab(As, amg, RHS, x, tol, maxiter, last_rel, &Wd);
    } else {
      last_iters = bicgstab(A, amg, RHS, x, tol, maxiter, last_rel);
    }
    stream_function = x;
    ++n_solves;
    total_iters += last_iters;
  }

  // --- src/HO_2D.cpp:2833-2870 ----------------------------------------------------------------
  void internal_comm(int el, int ijp, int ijpm, int ibnd, const double* v, const double* vn, double* Dv, double* com) {
    std::vector<double> cross(K);
    int revert_coeff = (ijp + ijpm - ijp / 2 - ijpm / 2 + 1) % 2;
    for (int k = 0; k < K; ++k) com[k] = 0.50 * (v[k] + vn[revert_coeff * (K - 2 * k - 1) + k]);
    for (int k = 0; k < K; ++k) Dv[k] += (com[k] - v[k]) * g_prime[ibnd];
    for (int k = 0; k < K; ++k) {
      cross[k] = 0.;
      for (int d = 0; d < K; ++d) cross[k] += com[d] * ssgb[d * K + k];
    }
    for (int k = 0; k < K; ++k) {
      int f = f3(el, ijp, k);
      Dv[k] = (Dv[k] * face_A[f] - cross[k] * face_B[f]) / face_jac[f];
    }
  }

  // --- src/HO_2D.cpp:2872-2927 ----------------------------------------------------------------
  void boundary_comm(int el, int ijp, int ibnd, const double* v, double* Dv, double* com, double* com_grad,
                     int BC_type, const double* BC_vals) {
    const int elem_side = i2f[ijp];
    std::vector<double> cross(K), NeuRHS(K), NeuM(K * K);
    if (BC_type == DIRICHLET) {
      for (int m = 0; m < K; ++m) com[m] = BC_vals[(elem_side / 2) * (K - 2 * m - 1) + m];
      for (int k = 0; k < K; ++k) Dv[k] += (com[k] - v[k]) * g_prime[ibnd];
      for (int k = 0; k < K; ++k) {
        cross[k] = 0.;
        for (int m = 0; m < K; ++m) cross[k] += com[m] * ssgb[m * K + k];
      }
      for (int k = 0; k < K; ++k) {
        int f = f3(el, ijp, k);
        com_grad[k] = (Dv[k] * face_A[f] - cross[k] * face_B[f]) / face_jac[f];
      }
    } else if (BC_type == NEUMANN) {
      for (int k = 0; k < K; ++k) com_grad[k] = BC_vals[(elem_side / 2) * (K - 2 * k - 1) + k] * face_An[f3(el, ijp, k)];
      for (int k = 0; k < K; ++k) {
        int f = f3(el, ijp, k);
        NeuRHS[k] = face_jac[f] * com_grad[k] - face_A[f] * (Dv[k] - v[k] * g_prime[ibnd]);
      }
      for (int k = 0; k < K; ++k) {
        int f = f3(el, ijp, k);
        for (int i = 0; i < K; ++i) NeuM[k * K + i] = -face_B[f] * ssgb[i * K + k];
        NeuM[k * K + k] += face_A[f] * g_prime[ibnd];
      }
      Gauss_solver(K, NeuM, NeuRHS, com);
    }
  }

  // Face traces of u: value and ξ/η derivative on the four faces (src/HO_2D.cpp:2695-2713).
  void traces(const double* u, std::vector<double>& bv, std::vector<double>& bg) {
    bv.assign(N_el * 4 * K, 0.); bg.assign(N_el * 4 * K, 0.);
    for (int el = 0; el < N_el; el++)
      for (int rc = 0; rc < K; ++rc) {
        int ijp = 0;
        for (int idir = 0; idir < 2; ++idir)
          for (int ibnd = 0; ibnd < 2; ++ibnd) {
            double s = 0., g = 0.;
            for (int m = 0; m < K; m++) {
              double lv = idir == 0 ? u[v3(el, rc, m)] : u[v3(el, m, rc)];
              s += lv * sbb[m * 2 + ibnd];
            }
            for (int m = 0; m < K; ++m) {
              double lv = idir == 0 ? u[v3(el, rc, m)] : u[v3(el, m, rc)];
              g += lv * sbgb[m * 2 + ibnd];
            }
            bv[f3(el, ijp, rc)] = s;
            bg[f3(el, ijp, rc)] = g;
            ijp++;
          }
      }
  }

  // Second-order FR operator shared by calc_RHS_diffusion (src/HO_2D.cpp:2666-2831) and the
  // matrix-free Laplacian: out = div(common-flux corrected gradient), before Re^-1/J scaling.
  // bc_switch/bc_vals per boundary edge; when bc_vort_out != nullptr, the boundary common
  // values are stored into it (BC_vorticity, :2726-2727).
  void fr_second_order(const double* u, const unsigned char* bc_switch, const double* bc_vals,
                       double* bc_vort_out, double* out) {
    std::vector<double> bv, bg;
    traces(u, bv, bg);
    std::vector<double> cv(N_el * 4 * K, 0.), cg(N_el * 4 * K, 0.);
    for (int el = 0; el < N_el; el++) {
      int ijp = 0;
      for (int idir = 0; idir < 2; ++idir)
        for (int ibnd = 0; ibnd < 2; ++ibnd) {
          const int elem_side = i2f[ijp];
          if (is_bnd[el * 4 + elem_side]) {
            const int el_b = bnd_index[el * 4 + elem_side];
            boundary_comm(el, ijp, ibnd, &bv[f3(el, ijp, 0)], &bg[f3(el, ijp, 0)], &cv[f3(el, ijp, 0)], &cg[f3(el, ijp, 0)],
                          bc_switch[el_b], bc_vals + (size_t)el_b * K);
            if (bc_vort_out)
              for (int m = 0; m < K; ++m) bc_vort_out[b2(el_b, (elem_side / 2) * (K - 2 * m - 1) + m)] = cv[f3(el, ijp, m)];
          } else {
            const int eln = nbr[el * 4 + elem_side];
            const int ijpm = f2i[nbr_side[el * 4 + elem_side]];
            internal_comm(el, ijp, ijpm, ibnd, &bv[f3(el, ijp, 0)], &bv[f3(eln, ijpm, 0)], &bg[f3(el, ijp, 0)], &cv[f3(el, ijp, 0)]);
          }
          ijp++;
        }
    }
    for (int el = 0; el < N_el; el++)
      for (int ijp = 0; ijp < 4; ++ijp) {
        int elem_side = i2f[ijp];
        if (!is_bnd[el * 4 + elem_side]) {
          const int eln = nbr[el * 4 + elem_side];
          int ijpm = f2i[nbr_side[el * 4 + elem_side]];
          double sign = 2. * ((ijp + ijpm) % 2) - 1.;
          int revert_coeff = (ijp + ijpm - ijp / 2 - ijpm / 2 + 1) % 2;
          for (int k = 0; k < K; ++k)
            cg[f3(el, ijp, k)] = 0.5 * (bg[f3(el, ijp, k)] + sign * bg[f3(eln, ijpm, revert_coeff * (K - 2 * k - 1) + k)]);
        }
      }
    std::vector<double> ft(K * K), gt(K * K), ftB(2 * K), gtB(2 * K);
    for (int el = 0; el < N_el; ++el) {
      for (int ky = 0; ky < K; ++ky)
        for (int kx = 0; kx < K; ++kx) {
          double dw_dxsi = 0., dw_deta = 0.;
          for (int d = 0; d < K; ++d) dw_dxsi += u[v3(el, ky, d)] * ssgb[d * K + kx];
          dw_dxsi += (cv[f3(el, 0, ky)] - bv[f3(el, 0, ky)]) * sgrad[kx * 2 + 0];
          dw_dxsi += (cv[f3(el, 1, ky)] - bv[f3(el, 1, ky)]) * sgrad[kx * 2 + 1];
          for (int d = 0; d < K; ++d) dw_deta += u[v3(el, d, kx)] * ssgb[d * K + ky];
          dw_deta += (cv[f3(el, 2, kx)] - bv[f3(el, 2, kx)]) * sgrad[ky * 2 + 0];
          dw_deta += (cv[f3(el, 3, kx)] - bv[f3(el, 3, kx)]) * sgrad[ky * 2 + 1];
          int v = v3(el, ky, kx);
          ft[ky * K + kx] = (dw_dxsi * (dxdeta[v] * dxdeta[v] + dydeta[v] * dydeta[v]) -
                             dw_deta * (dxdxi[v] * dxdeta[v] + dydxi[v] * dydeta[v])) / vol_jac[v];
          gt[ky * K + kx] = (dw_deta * (dxdxi[v] * dxdxi[v] + dydxi[v] * dydxi[v]) -
                             dw_dxsi * (dxdxi[v] * dxdeta[v] + dydxi[v] * dydeta[v])) / vol_jac[v];
        }
      for (int ibnd = 0; ibnd < 2; ++ibnd)
        for (int k = 0; k < K; k++) {
          ftB[ibnd * K + k] = cg[f3(el, ibnd, k)];
          for (int d = 0; d < K; ++d) ftB[ibnd * K + k] -= ft[k * K + d] * sbb[d * 2 + ibnd];
          gtB[ibnd * K + k] = cg[f3(el, ibnd + 2, k)];
          for (int d = 0; d < K; ++d) gtB[ibnd * K + k] -= gt[d * K + k] * sbb[d * 2 + ibnd];
        }
      for (int ky = 0; ky < K; ++ky)
        for (int kx = 0; kx < K; ++kx) {
          double r = 0.;
          for (int d = 0; d < K; ++d) r += ft[ky * K + d] * ssgb[d * K + kx];
          r += ftB[0 * K + ky] * sgrad[kx * 2 + 0] + ftB[1 * K + ky] * sgrad[kx * 2 + 1];
          for (int d = 0; d < K; ++d) r += gt[d * K + kx] * ssgb[d * K + ky];
          r += gtB[0 * K + kx] * sgrad[ky * 2 + 0] + gtB[1 * K + kx] * sgrad[ky * 2 + 1];
          out[v3(el, ky, kx)] = r;
        }
    }
  }

  // --- src/HO_2D.cpp:2666-2831 ----------------------------------------------------------------
  void calc_RHS_diffusion(const double* w) {
    fr_second_order(w, BC_switch_diffusion.data(), BC_diffusion.data(), BC_vorticity.data(), RHS_diff.data());
    for (int v = 0; v < N_el * K2(); ++v) RHS_diff[v] *= Reyn_inv / vol_jac[v];
  }

  // --- src/HO_2D.cpp:3174-3366 ----------------------------------------------------------------
  void calc_velocity() {
    const double* psi = stream_function.data();
    std::vector<double> bv, bg;
    traces(psi, bv, bg);
    std::vector<double> cv(N_el * 4 * K, 0.), cg(N_el * 4 * K, 0.), tmp(K);
    for (int el = 0; el < N_el; el++) {
      int ijp = 0;
      for (int idir = 0; idir < 2; ++idir)
        for (int ibnd = 0; ibnd < 2; ++ibnd) {
          const int elem_side = i2f[ijp];
          if (is_bnd[el * 4 + elem_side]) {
            const int el_b = bnd_index[el * 4 + elem_side];
            if (BC_switch_Poisson[el_b] == DIRICHLET) {
              boundary_comm(el, ijp, ibnd, &bv[f3(el, ijp, 0)], &bg[f3(el, ijp, 0)], &cv[f3(el, ijp, 0)], &cg[f3(el, ijp, 0)],
                            DIRICHLET, &BC_Poisson[b2(el_b, 0)]);
              for (int k = 0; k < K; ++k) {
                const int bfpi = (elem_side / 2) * (K - 2 * k - 1) + k;
                velocity_jump[b2(el_b, bfpi)] = cg[f3(el, ijp, k)] / face_An[f3(el, ijp, k)];
              }
            } else if (BC_switch_Poisson[el_b] == NEUMANN) {
              const double psign = 2. * (ijp % 2) - 1.;
              for (int k = 0; k < K; ++k) tmp[k] = psign * BC_Poisson[b2(el_b, k)];
              boundary_comm(el, ijp, ibnd, &bv[f3(el, ijp, 0)], &bg[f3(el, ijp, 0)], &cv[f3(el, ijp, 0)], &cg[f3(el, ijp, 0)],
                            NEUMANN, tmp.data());
              const double vsign = 1. - 2. * (ijp / 2);
              for (int k = 0; k < K; ++k) {
                const int bfpi = (elem_side / 2) * (K - 2 * k - 1) + k;
                double s = 0.;
                for (int d = 0; d < K; ++d) s += vsign * cv[f3(el, ijp, d)] * ssgb[d * K + k];
                BC_normal_vel[b2(el_b, bfpi)] = s / face_An[f3(el, ijp, k)];
              }
            }
          } else {
            const int eln = nbr[el * 4 + elem_side];
            const int ijpm = f2i[nbr_side[el * 4 + elem_side]];
            internal_comm(el, ijp, ijpm, ibnd, &bv[f3(el, ijp, 0)], &bv[f3(eln, ijpm, 0)], &bg[f3(el, ijp, 0)], &cv[f3(el, ijp, 0)]);
          }
          ijp++;
        }
    }
    for (int el = 0; el < N_el; el++)
      for (int ijp = 0; ijp < 4; ++ijp) {
        int elem_side = i2f[ijp];
        if (!is_bnd[el * 4 + elem_side]) {
          const int eln = nbr[el * 4 + elem_side];
          int ijpm = f2i[nbr_side[el * 4 + elem_side]];
          double sign = 2. * ((ijp + ijpm) % 2) - 1.;
          int revert_coeff = (ijp + ijpm - ijp / 2 - ijpm / 2 + 1) % 2;
          for (int k = 0; k < K; ++k)
            cg[f3(el, ijp, k)] = 0.5 * (bg[f3(el, ijp, k)] + sign * bg[f3(eln, ijpm, revert_coeff * (K - 2 * k - 1) + k)]);
        }
      }
    std::vector<double> ft(K * K), gt(K * K), ftB(2 * K), gtB(2 * K);
    for (int el = 0; el < N_el; ++el) {
      for (int ky = 0; ky < K; ++ky)
        for (int kx = 0; kx < K; ++kx) {
          double dpx = 0., dpe = 0.;
          for (int d = 0; d < K; ++d) dpx += psi[v3(el, ky, d)] * ssgb[d * K + kx];
          dpx += (cv[f3(el, 0, ky)] - bv[f3(el, 0, ky)]) * sgrad[kx * 2 + 0];
          dpx += (cv[f3(el, 1, ky)] - bv[f3(el, 1, ky)]) * sgrad[kx * 2 + 1];
          for (int d = 0; d < K; ++d) dpe += psi[v3(el, d, kx)] * ssgb[d * K + ky];
          dpe += (cv[f3(el, 2, kx)] - bv[f3(el, 2, kx)]) * sgrad[ky * 2 + 0];
          dpe += (cv[f3(el, 3, kx)] - bv[f3(el, 3, kx)]) * sgrad[ky * 2 + 1];
          int v = v3(el, ky, kx);
          ft[ky * K + kx] = (dpx * (dxdeta[v] * dxdeta[v] + dydeta[v] * dydeta[v]) - dpe * (dxdxi[v] * dxdeta[v] + dydxi[v] * dydeta[v])) / vol_jac[v];
          gt[ky * K + kx] = (dpe * (dxdxi[v] * dxdxi[v] + dydxi[v] * dydxi[v]) - dpx * (dxdxi[v] * dxdeta[v] + dydxi[v] * dydeta[v])) / vol_jac[v];
        }
      for (int ibnd = 0; ibnd < 2; ++ibnd)
        for (int k = 0; k < K; k++) {
          ftB[ibnd * K + k] = cg[f3(el, ibnd, k)];
          for (int d = 0; d < K; ++d) ftB[ibnd * K + k] -= ft[k * K + d] * sbb[d * 2 + ibnd];
          gtB[ibnd * K + k] = cg[f3(el, ibnd + 2, k)];
          for (int d = 0; d < K; ++d) gtB[ibnd * K + k] -= gt[d * K + k] * sbb[d * 2 + ibnd];
        }
      for (int ky = 0; ky < K; ++ky)
        for (int kx = 0; kx < K; ++kx) {
          int v = v3(el, ky, kx);
          double tempx = ft[ky * K + kx] + ftB[0 * K + ky] * srad[kx * 2 + 0] + ftB[1 * K + ky] * srad[kx * 2 + 1];
          double tempy = gt[ky * K + kx] + gtB[0 * K + kx] * srad[ky * 2 + 0] + gtB[1 * K + kx] * srad[ky * 2 + 1];
          double dpsi_dx = (tempx * dxdxi[v] + tempy * dxdeta[v]) / vol_jac[v];
          double dpsi_dy = (tempx * dydxi[v] + tempy * dydeta[v]) / vol_jac[v];
          // get_curl = true: u = dpsi/dy, v = -dpsi/dx (src/HO_2D.cpp:3186, 3348-3349)
          vel[v * 2 + 0] = 0. * dpsi_dx + 1. * dpsi_dy;
          vel[v * 2 + 1] = -1. * dpsi_dx + 0. * dpsi_dy;
        }
    }
  }

  // --- src/HO_2D.cpp:2647-2664 ----------------------------------------------------------------
  void update_diffusion_BC() {
    for (int gb = 0; gb < N_Gb; ++gb) {
      if (!BC_no_slip[gb]) continue;
      for (int e : bedges[gb]) {
        double sign = (f2i[be_side[e]] % 2) * 2. - 1.;
        for (int m = 0; m < K; ++m) velocity_jump[b2(e, m)] += sign * BC_parl_vel[b2(e, m)];
        for (int m = 0; m < K; ++m) BC_diffusion[b2(e, m)] = velocity_jump[b2(e, m)] / (dt * Reyn_inv);
      }
    }
  }

  // --- src/HO_2D.cpp:2219-2243 ----------------------------------------------------------------
  void update_advection_BC() {
    for (int e = 0; e < N_eb; ++e) {
      const int el = be_elem[e], side = be_side[e], ijp = f2i[side];
      for (int i = 0; i < K; ++i) {
        const int fp = (side / 2) * (K - 2 * i - 1) + i;
        BC_advection[b2(e, fp)] = BC_vorticity[b2(e, fp)] * BC_normal_vel[b2(e, fp)] * face_An[f3(el, ijp, i)];
      }
    }
  }

  // --- src/HO_2D.cpp:2245-2437 (advection_type 1) -------------------------------------------
  void calc_RHS_advection(const double* w) {
    std::vector<double> bvort(N_el * 4 * K, 0.), bflx(N_el * 4 * K, 0.), upw(N_el * 4 * K, 0.);
    std::vector<double> disc((size_t)N_el * 2 * K * K, 0.);
    std::vector<double> lv(K * 2), lu(K * 2);
    auto DF = [&](int el, int idir, int rc, int i) -> double& { return disc[(((size_t)el * 2 + idir) * K + rc) * K + i]; };
    for (int el = 0; el < N_el; el++) {
      for (int rc = 0; rc < K; ++rc) {
        for (int ijp = 0; ijp < 4; ++ijp) bvort[f3(el, ijp, rc)] = bflx[f3(el, ijp, rc)] = upw[f3(el, ijp, rc)] = 0.;
        for (int i = 0; i < K; i++) {
          lv[i * 2 + 0] = w[v3(el, rc, i)];
          lv[i * 2 + 1] = w[v3(el, i, rc)];
          int v1 = v3(el, rc, i), v2 = v3(el, i, rc);
          const double vj1 = vol_jac[v1];
          lu[i * 2 + 0] = vel[v1 * 2 + 0] * dxsidx[v1 * 4 + 0] * vj1;
          lu[i * 2 + 0] += vel[v1 * 2 + 1] * dxsidx[v1 * 4 + 1] * vj1;
          const double vj2 = vol_jac[v2];
          lu[i * 2 + 1] = vel[v2 * 2 + 0] * dxsidx[v2 * 4 + 2] * vj2;
          lu[i * 2 + 1] += vel[v2 * 2 + 1] * dxsidx[v2 * 4 + 3] * vj2;
        }
        int ijp = 0;
        for (int idir = 0; idir < 2; ++idir) {
          for (int i = 0; i < K; ++i) DF(el, idir, rc, i) = lu[i * 2 + idir] * lv[i * 2 + idir];
          for (int ibnd = 0; ibnd < 2; ++ibnd) {
            double& bvr = bvort[f3(el, ijp, rc)];
            for (int m = 0; m < K; m++) bvr += lv[m * 2 + idir] * sbb[m * 2 + ibnd];
            const int elem_side = i2f[ijp];
            if (is_bnd[el * 4 + elem_side]) {
              int e = bnd_index[el * 4 + elem_side];
              int fp = (elem_side / 2) * (K - 2 * rc - 1) + rc;
              bflx[f3(el, ijp, rc)] = BC_advection[b2(e, fp)];
            } else {
              double& bf = bflx[f3(el, ijp, rc)];
              for (int m = 0; m < K; ++m) bf += lu[m * 2 + idir] * sbb[m * 2 + ibnd];
              bf *= bvr;
            }
            ijp++;
          }
        }
      }
    }
    for (int el = 0; el < N_el; el++)
      for (int ijp = 0; ijp < 4; ++ijp) {
        const int elem_side = i2f[ijp];
        if (is_bnd[el * 4 + elem_side])
          for (int j = 0; j < K; ++j) upw[f3(el, ijp, j)] = bflx[f3(el, ijp, j)];
        else {
          const int eln = nbr[el * 4 + elem_side];
          const int ijpm = f2i[nbr_side[el * 4 + elem_side]];
          const double sign = 2. * ((ijp + ijpm) % 2) - 1.;
          const int revert_coeff = (ijp + ijpm - ijp / 2 - ijpm / 2 + 1) % 2;
          for (int j = 0; j < K; ++j) {
            const int jn = revert_coeff * (K - 2 * j - 1) + j;
            upw[f3(el, ijp, j)] = 0.5 * (bflx[f3(el, ijp, j)] + sign * bflx[f3(eln, ijpm, jn)]);
            double vort_diff = bvort[f3(el, ijp, j)] - bvort[f3(eln, ijpm, jn)];
            if (std::fabs(vort_diff) > 1.e-4) {
              const double bndr_vel = (bflx[f3(el, ijp, j)] - sign * bflx[f3(eln, ijpm, jn)]) / vort_diff;
              upw[f3(el, ijp, j)] += 0.5 * sgn_tab[ijp] * std::fabs(bndr_vel) * vort_diff;
            }
          }
        }
      }
    std::vector<double> bdf(4 * K);
    for (int el = 0; el < N_el; el++) {
      int ijp = 0;
      for (int idir = 0; idir < 2; ++idir)
        for (int ibnd = 0; ibnd < 2; ++ibnd) {
          for (int j = 0; j < K; ++j) {
            bdf[ijp * K + j] = 0.;
            for (int m = 0; m < K; ++m) bdf[ijp * K + j] -= DF(el, idir, j, m) * sbb[m * 2 + ibnd];
            bdf[ijp * K + j] += upw[f3(el, ijp, j)];
          }
          ijp++;
        }
      for (int j = 0; j < K; ++j)
        for (int i = 0; i < K; ++i) {
          double r = 0.;
          for (int d = 0; d < K; ++d) r -= DF(el, 0, j, d) * ssgb[d * K + i];
          for (int d = 0; d < K; ++d) r -= DF(el, 1, i, d) * ssgb[d * K + j];
          r -= bdf[0 * K + j] * sgrad[i * 2 + 0] + bdf[1 * K + j] * sgrad[i * 2 + 1];
          r -= bdf[2 * K + i] * sgrad[j * 2 + 0] + bdf[3 * K + i] * sgrad[j * 2 + 1];
          r /= vol_jac[v3(el, j, i)];
          RHS_adv[v3(el, j, i)] = r;
        }
    }
  }

  // --- src/HO_2D.cpp:2439-2645 (advection_type 2) -------------------------------------------
  void calc_RHS_advection_continuous(const double* w) {
    const int NF = N_el * 4 * K;
    std::vector<double> bvort(NF, 0.), bflx(NF, 0.), upw(NF, 0.), bvf(NF, 0.), cvf(NF, 0.);
    std::vector<double> dvf((size_t)N_el * K * K * 2), cont((size_t)N_el * K * K * 2);
    std::vector<double> disc((size_t)N_el * 2 * K * K, 0.);
    auto DF = [&](int el, int idir, int rc, int i) -> double& { return disc[(((size_t)el * 2 + idir) * K + rc) * K + i]; };
    for (int el = 0; el < N_el; el++)
      for (int j = 0; j < K; ++j)
        for (int i = 0; i < K; ++i) {
          int v = v3(el, j, i);
          dvf[v * 2 + 0] = vel[v * 2 + 0] * dydeta[v] - vel[v * 2 + 1] * dxdeta[v];
          dvf[v * 2 + 1] = -vel[v * 2 + 0] * dydxi[v] + vel[v * 2 + 1] * dxdxi[v];
        }
    for (int el = 0; el < N_el; el++) {
      int ijp = 0;
      for (int idir = 0; idir < 2; ++idir)
        for (int ibnd = 0; ibnd < 2; ++ibnd) {
          for (int a = 0; a < K; ++a) {
            double s = 0.;
            for (int m = 0; m < K; ++m) s += dvf[v3(el, a * (1 - idir) + m * idir, a * idir + m * (1 - idir)) * 2 + idir] * sbb[m * 2 + ibnd];
            bvf[f3(el, ijp, a)] = s;
          }
          ijp++;
        }
    }
    for (int el = 0; el < N_el; ++el) {
      int ijp = 0;
      for (int idir = 0; idir < 2; idir++)
        for (int ibnd = 0; ibnd < 2; ibnd++) {
          int elem_side = i2f[ijp];
          if (is_bnd[el * 4 + elem_side]) {
            int e = bnd_index[el * 4 + elem_side];
            for (int m = 0; m < K; ++m) {
              int fp = (elem_side / 2) * (K - 2 * m - 1) + m;
              cvf[f3(el, ijp, m)] = BC_normal_vel[b2(e, fp)] * face_An[f3(el, ijp, m)];
            }
          } else {
            const int eln = nbr[el * 4 + elem_side];
            const int ijpm = f2i[nbr_side[el * 4 + elem_side]];
            double sign = 2. * ((ijp + ijpm) % 2) - 1.;
            int revert_coeff = (ijp + ijpm - ijp / 2 - ijpm / 2 + 1) % 2;
            for (int m = 0; m < K; ++m) {
              int mn = revert_coeff * (K - 2 * m - 1) + m;
              cvf[f3(el, ijp, m)] = 0.5 * (bvf[f3(el, ijp, m)] + sign * bvf[f3(eln, ijpm, mn)]);
            }
          }
          ijp++;
        }
    }
    for (int el = 0; el < N_el; ++el)
      for (int j = 0; j < K; ++j)
        for (int i = 0; i < K; ++i) {
          int v = v3(el, j, i);
          cont[v * 2 + 0] = dvf[v * 2 + 0] + (cvf[f3(el, 0, j)] - bvf[f3(el, 0, j)]) * srad[i * 2 + 0] +
                            (cvf[f3(el, 1, j)] - bvf[f3(el, 1, j)]) * srad[i * 2 + 1];
          cont[v * 2 + 1] = dvf[v * 2 + 1] + (cvf[f3(el, 2, i)] - bvf[f3(el, 2, i)]) * srad[j * 2 + 0] +
                            (cvf[f3(el, 3, i)] - bvf[f3(el, 3, i)]) * srad[j * 2 + 1];
        }
    std::vector<double> lv(K * 2), lu(K * 2);
    for (int el = 0; el < N_el; el++)
      for (int rc = 0; rc < K; ++rc) {
        for (int ijp = 0; ijp < 4; ++ijp) bvort[f3(el, ijp, rc)] = upw[f3(el, ijp, rc)] = 0.;
        for (int i = 0; i < K; i++) {
          lv[i * 2 + 0] = w[v3(el, rc, i)];
          lv[i * 2 + 1] = w[v3(el, i, rc)];
          lu[i * 2 + 0] = cont[v3(el, rc, i) * 2 + 0];
          lu[i * 2 + 1] = cont[v3(el, i, rc) * 2 + 1];
        }
        int ijp = 0;
        for (int idir = 0; idir < 2; ++idir) {
          for (int i = 0; i < K; ++i) DF(el, idir, rc, i) = lu[i * 2 + idir] * lv[i * 2 + idir];
          for (int ibnd = 0; ibnd < 2; ++ibnd) {
            double& bvr = bvort[f3(el, ijp, rc)];
            for (int m = 0; m < K; m++) bvr += lv[m * 2 + idir] * sbb[m * 2 + ibnd];
            int elem_side = i2f[ijp];
            if (is_bnd[el * 4 + elem_side]) {
              int e = bnd_index[el * 4 + elem_side];
              int fp = (elem_side / 2) * (K - 2 * rc - 1) + rc;
              bflx[f3(el, ijp, rc)] = BC_advection[b2(e, fp)];
            } else {
              bflx[f3(el, ijp, rc)] = cvf[f3(el, ijp, rc)] * bvr;
            }
            ijp++;
          }
        }
      }
    for (int el = 0; el < N_el; el++)
      for (int ijp = 0; ijp < 4; ++ijp) {
        const int elem_side = i2f[ijp];
        if (is_bnd[el * 4 + elem_side])
          for (int j = 0; j < K; ++j) upw[f3(el, ijp, j)] = bflx[f3(el, ijp, j)];
        else {
          const double t = (double)(ijp % 2);
          const int eln = nbr[el * 4 + elem_side];
          const int ijpm = f2i[nbr_side[el * 4 + elem_side]];
          const int revert_coeff = (ijp + ijpm - ijp / 2 - ijpm / 2 + 1) % 2;
          for (int j = 0; j < K; ++j) {
            const int jn = revert_coeff * (K - 2 * j - 1) + j;
            if (cvf[f3(el, ijp, j)] > 0.)
              upw[f3(el, ijp, j)] = cvf[f3(el, ijp, j)] * (t * bvort[f3(el, ijp, j)] + (1. - t) * bvort[f3(eln, ijpm, jn)]);
            else
              upw[f3(el, ijp, j)] = cvf[f3(el, ijp, j)] * ((1. - t) * bvort[f3(el, ijp, j)] + t * bvort[f3(eln, ijpm, jn)]);
          }
        }
      }
    std::vector<double> bdf(4 * K);
    for (int el = 0; el < N_el; el++) {
      int ijp = 0;
      for (int idir = 0; idir < 2; ++idir)
        for (int ibnd = 0; ibnd < 2; ++ibnd) {
          for (int j = 0; j < K; ++j) {
            bdf[ijp * K + j] = 0.;
            for (int m = 0; m < K; ++m) {
              int jj = j * (1 - idir) + m * idir, ii = j * idir + m * (1 - idir);
              bdf[ijp * K + j] -= cont[v3(el, jj, ii) * 2 + idir] * w[v3(el, jj, ii)] * sbb[m * 2 + ibnd];
            }
            bdf[ijp * K + j] += upw[f3(el, ijp, j)];
          }
          ijp++;
        }
      for (int j = 0; j < K; ++j)
        for (int i = 0; i < K; ++i) {
          double r = 0.;
          for (int d = 0; d < K; ++d) {
            r -= DF(el, 0, j, d) * ssgb[d * K + i];
            r -= DF(el, 1, i, d) * ssgb[d * K + j];
          }
          r -= bdf[0 * K + j] * sgrad[i * 2 + 0] + bdf[1 * K + j] * sgrad[i * 2 + 1];
          r -= bdf[2 * K + i] * sgrad[j * 2 + 0] + bdf[3 * K + i] * sgrad[j * 2 + 1];
          r /= vol_jac[v3(el, j, i)];
          RHS_adv[v3(el, j, i)] = r;
        }
    }
  }

  // --- src/HO_2D.cpp:1806-1839 ----------------------------------------------------------------
  void Euler_time_integrate(const double* vin, double* vout, double coeff) {
    update_BCs(current_time + coeff * dt);
    solve_Poisson(vin);
    calc_velocity();
    update_diffusion_BC();
    calc_RHS_diffusion(vin);
    update_advection_BC();
    if (advection_type == 1) calc_RHS_advection(vin);
    else calc_RHS_advection_continuous(vin);
    const int D = N_el * K2();
    for (int i = 0; i < D; ++i) vout[i] = dt * (RHS_adv[i] + RHS_diff[i]);
  }

  // --- src/HO_2D.cpp:1465-1531 ----------------------------------------------------------------
  void solve_advection_diffusion() {
    const int D = N_el * K2();
    if (time_integration_type == 1) {
      Euler_time_integrate(vorticity.data(), k1.data(), 0.);
      for (int i = 0; i < D; ++i) vorticity[i] += k1[i];
    } else if (time_integration_type == 2) {
      Euler_time_integrate(vorticity.data(), k1.data(), RK2_coeff[0]);
      for (int i = 0; i < D; ++i) vort[i] = vorticity[i] + RK2_coeff[1] * k1[i];
      Euler_time_integrate(vort.data(), k2.data(), RK2_coeff[1]);
      for (int i = 0; i < D; ++i) vorticity[i] += 0.5 * (k1[i] + k2[i]);
    } else if (time_integration_type == 4) {
      Euler_time_integrate(vorticity.data(), k1.data(), RK4_coeff[0]);
      for (int i = 0; i < D; ++i) vort[i] = vorticity[i] + RK4_coeff[1] * k1[i];
      Euler_time_integrate(vort.data(), k2.data(), RK4_coeff[1]);
      for (int i = 0; i < D; ++i) vort[i] = vorticity[i] + RK4_coeff[2] * k2[i];
      Euler_time_integrate(vort.data(), k3.data(), RK4_coeff[2]);
      for (int i = 0; i < D; ++i) vort[i] = vorticity[i] + RK4_coeff[3] * k3[i];
      Euler_time_integrate(vort.data(), k4.data(), RK4_coeff[3]);
      for (int i = 0; i < D; ++i) vorticity[i] += (k1[i] + 2. * k2[i] + 2. * k3[i] + k4[i]) / 6.;
    } else throw std::runtime_error("time_integration_type must be 1, 2 or 4");
  }

  // --- src/HO_2D.cpp:3369-3531 (save_smooth_vtk without the file) --------------------------
  // Node-averaged fields: out[node][4] = vorticity, stream function, u, v.
  void smooth_fields(double* out) const {
    const int L = Lnod, L2 = L * L;
    const size_t NL = (size_t)N_el * L2;
    std::vector<double> om(NL, 0.), ps(NL, 0.), ux(NL, 0.), uy(NL, 0.);
    for (int el = 0; el < N_el; ++el)  // :3413-3428, sps extrapolated to the geometry nodes
      for (int Ly = 0; Ly < L; ++Ly)
        for (int Lx = 0; Lx < L; ++Lx) {
          const size_t g = (size_t)el * L2 + tensor2FEM(Lx, Ly);
          for (int ky = 0; ky < K; ++ky)
            for (int kx = 0; kx < K; ++kx) {
              const double c = sgb[ky * L + Ly] * sgb[kx * L + Lx];
              const int v = v3(el, ky, kx);
              om[g] += c * vorticity[v];
              ps[g] += c * stream_function[v];
              ux[g] += c * vel[2 * v];
              uy[g] += c * vel[2 * v + 1];
            }
        }
    // :3431-3478, Cartesian boundary velocity from the normal and tangential values
    std::vector<double> cx((size_t)N_eb * K, 0.), cy((size_t)N_eb * K, 0.);
    for (int eb = 0; eb < N_eb; ++eb) {
      const int el = be_elem[eb], side = be_side[eb], ijp = f2i[side], dir = ijp / 2;
      const double parl_sign = -1. + 2. * (side / 2);
      for (int k = 0; k < K; ++k) {
        const int f = f3(el, ijp, k);
        const double dxx = f_dxdxi[f], dxe = f_dxdeta[f], dyx = f_dydxi[f], dye = f_dydeta[f];
        const double g1 = std::sqrt(dxx * dxx + dyx * dyx), g2 = std::sqrt(dxe * dxe + dye * dye);
        const int idx = b2(eb, (side / 2) * (K - 2 * k - 1) + k);
        if (dir == 0) {
          const double un = BC_normal_vel[idx], ut = parl_sign * BC_parl_vel[idx];
          cx[idx] = (1. / g2) * (un * dye + ut * dxe);
          cy[idx] = (1. / g2) * (-un * dxe + ut * dye);
        } else {
          const double un = BC_normal_vel[idx], ut = parl_sign * BC_parl_vel[idx];
          cx[idx] = (1. / g1) * (-un * dyx + ut * dxx);
          cy[idx] = (1. / g1) * (un * dxx + ut * dyx);
        }
      }
    }
    for (int eb = 0; eb < N_eb; ++eb) {  // :3485-3506, boundary faces take the boundary values
      const int el = be_elem[eb], side = be_side[eb], ijp = f2i[side], d = ijp / 2, t = ijp % 2;
      for (int l = 0; l < L; ++l) {
        const int su2pv = l * ((1 - d) * L + d) + t * (L - 1) * (d * (L - 1) + 1);
        const size_t g = (size_t)el * L2 + tensor2FEM(su2pv % L, su2pv / L);
        om[g] = 0.; ux[g] = 0.; uy[g] = 0.;
        for (int m = 0; m < K; ++m) {
          const double c = sgb[m * L + l];
          const int nidx = b2(eb, (side / 2) * (K - 2 * m - 1) + m);
          om[g] += c * BC_vorticity[nidx];
          ux[g] += c * cx[nidx];
          uy[g] += c * cy[nidx];
        }
      }
    }
    std::vector<double> sw(N_nodes, 0.), sp(N_nodes, 0.), sx(N_nodes, 0.), sy(N_nodes, 0.);
    std::vector<int> rep(N_nodes, 0);
    for (int el = 0; el < N_el; ++el)  // :3511-3528, averages over the elements sharing a node
      for (int l = 0; l < L2; ++l) {
        const int nd = elem_nodes[(size_t)el * L2 + l];
        const size_t g = (size_t)el * L2 + l;
        sx[nd] += 1. * ux[g]; sy[nd] += 1. * uy[g];
        sw[nd] += om[g]; sp[nd] += ps[g];
        rep[nd]++;
      }
    for (int nd = 0; nd < N_nodes; ++nd) {
      out[4 * nd] = sw[nd] / rep[nd];
      out[4 * nd + 1] = sp[nd] / rep[nd];
      const double r = 1. / rep[nd];
      out[4 * nd + 2] = sx[nd] * r;
      out[4 * nd + 3] = sy[nd] * r;
    }
  }

  // --- src/HO_2D.cpp:319-474 (read_process_sample_points on given points) ------------------
  // The reference casts its point-in-element ray with a slope drawn by std::rand() (seeded by
  // the clock); this restatement uses the fixed slope below.  The Newton iteration keeps the
  // reference's stopping test and its use of the shape functions of the last iterate before
  // the final update (:405-472); it is capped at 100 iterations (the reference has no cap).
  static constexpr double kSampleSlope = 0.3819660112501051;
  std::vector<int> sp_elem;
  std::vector<double> sp_xy, sp_loc, sp_shape;  // sp_shape: [sp][L2] in Gmsh node order
  static int ray_hits(double px, double py, double slope, double x1, double y1, double x2, double y2) {
    // are_intersecting (src/MathHelpers.hpp:99-114)
    const double dx = x2 - x1, dy = y2 - y1;
    const double s = (slope * (x1 - px) + py - y1) / (dy - slope * dx);
    const double t = x1 - px + s * dx;
    return (t > -1.e-8 && s > -1.e-8 && s < 1. + 1.e-8) ? 1 : 0;
  }
  void set_sample_points(int n, const double* xy) {
    const int L = Lnod, L2 = L * L;
    sp_xy.assign(xy, xy + 2 * n);
    sp_elem.assign(n, -1);
    sp_loc.assign(2 * (size_t)n, 0.);
    sp_shape.assign((size_t)n * L2, 0.);
    std::vector<double> mnx(N_el), mxx(N_el), mny(N_el), mxy(N_el);
    for (int el = 0; el < N_el; ++el) {  // :342-354, bounding boxes of the element edges' nodes
      const int n0 = elem_nodes[(size_t)el * L2];
      mnx[el] = mxx[el] = nx[n0];
      mny[el] = mxy[el] = ny[n0];
      for (int s = 0; s < 4; ++s)
        for (int nd : edges[elem_edges[el * 4 + s]]) {
          mnx[el] = std::min(mnx[el], nx[nd]); mxx[el] = std::max(mxx[el], nx[nd]);
          mny[el] = std::min(mny[el], ny[nd]); mxy[el] = std::max(mxy[el], ny[nd]);
        }
    }
    for (int sp = 0; sp < n; ++sp) {  // :357-395
      const double px = xy[2 * sp], py = xy[2 * sp + 1];
      bool found = false;
      for (int el = 0; el < N_el && !found; ++el) {
        if (px < mnx[el] || px > mxx[el] || py < mny[el] || py > mxy[el]) continue;
        int hits = 0;
        for (int s = 0; s < 4; ++s) {
          const std::vector<int>& en = edges[elem_edges[el * 4 + s]];
          for (int seg = 0; seg < Lnod_in; ++seg) {
            const int a = en[tensor2FEM(seg)], b = en[tensor2FEM(seg + 1)];
            hits += ray_hits(px, py, kSampleSlope, nx[a], ny[a], nx[b], ny[b]);
          }
        }
        if (hits & 1) { sp_elem[sp] = el; found = true; }
      }
      if (!found) {
        const int n0 = elem_nodes[0];
        double dist = std::sqrt((px - nx[n0]) * (px - nx[n0]) + (py - ny[n0]) * (py - ny[n0])) + 1.;
        for (int el = 0; el < N_el; ++el)
          for (int i = 0; i < L2; ++i) {
            const int nd = elem_nodes[(size_t)el * L2 + i];
            const double d = std::sqrt((px - nx[nd]) * (px - nx[nd]) + (py - ny[nd]) * (py - ny[nd]));
            if (dist > d) { dist = d; sp_elem[sp] = el; }
          }
      }
      // :405-472, local coordinates by Newton's method
      const int el = sp_elem[sp];
      std::vector<double> bx(L), by(L), gx(L), gy(L);
      double xsi = 0.08, eta = 0.08, dxi = 0., det_ = 0.;
      int it = 0;
      do {
        for (int k = 0; k < L; ++k) {
          gx[k] = gy[k] = 0.;
          bx[k] = by[k] = 1.;
          double den = 1.;
          for (int j = 0; j < L; ++j) {
            if (j == k) continue;
            den *= gps[k] - gps[j];
            bx[k] *= xsi - gps[j];
            by[k] *= eta - gps[j];
            double nxi = 1., net = 1.;
            for (int i = 0; i < L; ++i) {
              if (i == k || i == j) continue;
              nxi *= xsi - gps[i];
              net *= eta - gps[i];
            }
            gx[k] += nxi;
            gy[k] += net;
          }
          gx[k] /= den; bx[k] /= den; gy[k] /= den; by[k] /= den;
        }
        double f = px, g = py, fx = 0., fe = 0., gxx = 0., ge = 0.;
        for (int j = 0; j < L; ++j)
          for (int i = 0; i < L; ++i) {
            const int nd = elem_nodes[(size_t)el * L2 + tensor2FEM(i, j)];
            f -= bx[i] * by[j] * nx[nd];
            g -= bx[i] * by[j] * ny[nd];
            fx -= by[j] * gx[i] * nx[nd];
            gxx -= by[j] * gx[i] * ny[nd];
            fe -= gy[j] * bx[i] * nx[nd];
            ge -= gy[j] * bx[i] * ny[nd];
          }
        dxi = -1. / (fx * ge - fe * gxx) * (f * ge - g * fe);
        det_ = -1. / (fx * ge - fe * gxx) * (g * fx - f * gxx);
        xsi += dxi;
        eta += det_;
      } while ((std::fabs(dxi) > 0.001 || std::fabs(det_) > 0.001) && ++it < 100);
      sp_loc[2 * sp] = xsi;
      sp_loc[2 * sp + 1] = eta;
      for (int j = 0; j < L; ++j)
        for (int i = 0; i < L; ++i) sp_shape[(size_t)sp * L2 + tensor2FEM(i, j)] = bx[i] * by[j];
    }
  }
  // src/HO_2D.cpp:3593-3612: out[sp][3] = vorticity, u, v from the node-averaged fields.
  void sample_values(double* out) const {
    const int L2 = Lnod * Lnod, n = (int)sp_elem.size();
    std::vector<double> nodes((size_t)N_nodes * 4);
    smooth_fields(nodes.data());
    for (int sp = 0; sp < n; ++sp) {
      double w = 0., u = 0., v = 0.;
      const int el = sp_elem[sp];
      for (int l = 0; l < L2; ++l) {
        const int nd = elem_nodes[(size_t)el * L2 + l];
        const double c = sp_shape[(size_t)sp * L2 + l];
        w += c * nodes[4 * nd];
        u += c * nodes[4 * nd + 2];
        v += c * nodes[4 * nd + 3];
      }
      out[3 * sp] = w; out[3 * sp + 1] = u; out[3 * sp + 2] = v;
    }
  }
};

static int fail(ho_oracle* h, const std::exception& e) {
  if (h) h->err = e.what();
  g_last_error = e.what();
  return -1;
}

extern "C" {

ho_oracle* ho_oracle_create(int K, int Lnod, int n_nodes, const double* xy, int n_elem, const int32_t* elem_nodes,
                            int n_bedges, const int32_t* bedge_nodes, const int32_t* bedge_bdry, int n_bdry,
                            const char* bdry_names, int problem_type, double Re, double dt,
                            int time_integration_type, int advection_type, double tol, int maxiter) {
  return ho_oracle_create_ex(K, Lnod, n_nodes, xy, n_elem, elem_nodes, n_bedges, bedge_nodes, bedge_bdry, n_bdry,
                             bdry_names, problem_type, Re, dt, time_integration_type, advection_type, tol, maxiter,
                             nullptr, nullptr);
}

ho_oracle* ho_oracle_create_ex(int K, int Lnod, int n_nodes, const double* xy, int n_elem, const int32_t* elem_nodes,
                               int n_bedges, const int32_t* bedge_nodes, const int32_t* bedge_bdry, int n_bdry,
                               const char* bdry_names, int problem_type, double Re, double dt,
                               int time_integration_type, int advection_type, double tol, int maxiter,
                               const int32_t* edge_nval, const double* edge_vals) {
  ho_oracle* h = new ho_oracle();
  try {
    h->edge_nval.assign(n_bedges, 0); h->edge_normvel.assign(n_bedges, 0.); h->edge_parlvel.assign(n_bedges, 0.);
    if (edge_nval && edge_vals)
      for (int e = 0; e < n_bedges; ++e) {  // arrays_to_boundary (src/HO_2D.cpp:3779-3793)
        h->edge_nval[e] = edge_nval[e];
        if (edge_nval[e] == 1) h->edge_normvel[e] = edge_vals[2 * e];
        else if (edge_nval[e] == 2) { h->edge_parlvel[e] = edge_vals[2 * e]; h->edge_normvel[e] = edge_vals[2 * e + 1]; }
      }
    if (K < 1 || K > 8) throw std::runtime_error("K must be in 1..8");
    if (Lnod < 2 || Lnod > 4) throw std::runtime_error("Lnod must be in 2..4");
    h->K = K; h->Lnod = Lnod; h->Lnod_in = Lnod - 1;
    h->problem_type = problem_type;
    h->Reyn_inv = Re > 1.e-3 ? 1. / Re : -1.;  // src/HO_2D.cpp:255
    h->dt = dt; h->time_integration_type = time_integration_type; h->advection_type = advection_type;
    h->tol = tol; h->maxiter = maxiter;
    h->nx.resize(n_nodes); h->ny.resize(n_nodes);
    for (int i = 0; i < n_nodes; ++i) { h->nx[i] = xy[2 * i]; h->ny[i] = xy[2 * i + 1]; }
    h->N_el = n_elem;
    h->elem_nodes.assign(elem_nodes, elem_nodes + (size_t)n_elem * Lnod * Lnod);
    h->edges.resize(n_bedges);
    for (int e = 0; e < n_bedges; ++e) h->edges[e].assign(bedge_nodes + (size_t)e * Lnod, bedge_nodes + (size_t)(e + 1) * Lnod);
    h->N_Gb = n_bdry;
    h->bnames.clear();
    std::string s(bdry_names ? bdry_names : "");
    size_t p0 = 0;
    for (int b = 0; b < n_bdry; ++b) {
      size_t p1 = s.find('\n', p0);
      h->bnames.push_back(s.substr(p0, p1 == std::string::npos ? std::string::npos : p1 - p0));
      p0 = p1 == std::string::npos ? s.size() : p1 + 1;
    }
    h->bedges.assign(n_bdry, {});
    for (int e = 0; e < n_bedges; ++e) {
      if (bedge_bdry[e] < 0 || bedge_bdry[e] >= n_bdry) throw std::runtime_error("bad boundary id");
      h->bedges[bedge_bdry[e]].push_back(e);
    }
    h->process_mesh();
    h->setup_sps_gps();
    h->form_bases();
    h->form_metrics();
    h->setup_IC_BC_SRC();
    h->form_Laplace_operator_matrix();
    const int D = h->N_el * h->K2();
    h->vel.assign(2 * D, 0.);
    h->RHS_adv.assign(D, 0.); h->RHS_diff.assign(D, 0.);
    h->k1.assign(D, 0.); h->k2.assign(D, 0.); h->k3.assign(D, 0.); h->k4.assign(D, 0.); h->vort.assign(D, 0.);
  } catch (const std::exception& e) {
    g_last_error = e.what();
    delete h;
    return nullptr;
  }
  return h;
}

void ho_oracle_destroy(ho_oracle* h) { delete h; }
const char* ho_oracle_last_error(ho_oracle* h) { return h ? h->err.c_str() : g_last_error.c_str(); }

int ho_oracle_open_points(ho_oracle* h, double* xy) {
  const int b = h->open_boundary();
  const int n = b < 0 ? 0 : (int)h->bedges[b].size() * h->K;
  if (xy) h->open_points(xy);
  return n;
}
int ho_oracle_set_open_vels(ho_oracle* h, const double* xyvel) {
  try { h->set_open_vels(xyvel); } catch (const std::exception& e) { return fail(h, e); }
  return 0;
}
int ho_oracle_set_open_vort(ho_oracle* h, const double* w) {
  try { h->set_open_vort(w); } catch (const std::exception& e) { return fail(h, e); }
  return 0;
}
int ho_oracle_set_soln_vort(ho_oracle* h, const double* w) {
  try { h->set_soln_vort(w); } catch (const std::exception& e) { return fail(h, e); }
  return 0;
}
int ho_oracle_set_p2g_weights(ho_oracle* h, const double* w) {
  try { h->set_p2g_weights(w); } catch (const std::exception& e) { return fail(h, e); }
  return 0;
}
int ho_oracle_solveto(ho_oracle* h, double outer_dt, int nstep, int integ, double Re) {
  try { h->solveto(outer_dt, nstep, integ, Re); } catch (const std::exception& e) { return fail(h, e); }
  return 0;
}

int ho_oracle_set_threads(int n) {
  const int prev = omp_get_max_threads();
  if (n > 0) omp_set_num_threads(n);
  return prev;
}
int ho_oracle_dofs(ho_oracle* h) { return h->N_el * h->K2(); }
int ho_oracle_n_bedges(ho_oracle* h) { return h->N_eb; }

int ho_oracle_set_vorticity(ho_oracle* h, const double* w) { std::copy(w, w + h->N_el * h->K2(), h->vorticity.begin()); return 0; }
int ho_oracle_get_vorticity(ho_oracle* h, double* w) { std::copy(h->vorticity.begin(), h->vorticity.end(), w); return 0; }
int ho_oracle_get_streamfunction(ho_oracle* h, double* p) { std::copy(h->stream_function.begin(), h->stream_function.end(), p); return 0; }
int ho_oracle_set_streamfunction(ho_oracle* h, const double* p) { std::copy(p, p + h->N_el * h->K2(), h->stream_function.begin()); return 0; }
int ho_oracle_get_velocity(ho_oracle* h, double* uv) { std::copy(h->vel.begin(), h->vel.end(), uv); return 0; }
double ho_oracle_time(ho_oracle* h) { return h->current_time; }

int ho_oracle_step(ho_oracle* h, int nsteps) {
  try {
    for (int s = 0; s < nsteps; ++s) { h->solve_advection_diffusion(); h->current_time += h->dt; }
  } catch (const std::exception& e) { return fail(h, e); }
  return 0;
}

int ho_oracle_stage(ho_oracle* h, const double* w, double t, double* k_out) {
  try {
    double save = h->current_time;
    h->current_time = t;
    h->Euler_time_integrate(w, k_out, 0.);
    h->current_time = save;
  } catch (const std::exception& e) { return fail(h, e); }
  return 0;
}

int ho_oracle_poisson(ho_oracle* h, const double* w, double t, double* psi) {
  try {
    h->update_BCs(t);
    h->solve_Poisson(w);
    std::copy(h->stream_function.begin(), h->stream_function.end(), psi);
  } catch (const std::exception& e) { return fail(h, e); }
  return 0;
}

int ho_oracle_last_solve(ho_oracle* h, int* iters, double* rel) { *iters = h->last_iters; *rel = h->last_rel; return 0; }
int ho_oracle_solve_counts(ho_oracle* h, int64_t* n_solves, int64_t* total_iters) {
  *n_solves = h->n_solves;
  *total_iters = h->total_iters;
  return 0;
}

int64_t ho_oracle_nnz(ho_oracle* h) { return (int64_t)h->A.val.size(); }
int ho_oracle_get_csr(ho_oracle* h, int32_t* rowptr, int32_t* col, double* val) {
  std::copy(h->A.ptr.begin(), h->A.ptr.end(), rowptr);
  std::copy(h->A.col.begin(), h->A.col.end(), col);
  std::copy(h->A.val.begin(), h->A.val.end(), val);
  return 0;
}
int ho_oracle_get_boundary_source(ho_oracle* h, double* out) {
  std::copy(h->boundary_source.begin(), h->boundary_source.end(), out);
  return 0;
}

int64_t ho_oracle_table(ho_oracle* h, const char* name, double* out, int64_t cap) {
  std::string n(name);
  std::vector<double> tmp;
  auto from_int = [&](const std::vector<int>& v) { tmp.assign(v.begin(), v.end()); };
  auto from_uc = [&](const std::vector<unsigned char>& v) { tmp.assign(v.begin(), v.end()); };
  if (n == "sps") tmp = h->sps;
  else if (n == "wgt") tmp = h->wgt;
  else if (n == "sps_boundary_basis") tmp = h->sbb;
  else if (n == "sps_boundary_grad_basis") tmp = h->sbgb;
  else if (n == "sps_sps_grad_basis") tmp = h->ssgb;
  else if (n == "sps_radau") tmp = h->srad;
  else if (n == "sps_grad_radau") tmp = h->sgrad;
  else if (n == "gps_sps_basis") tmp = h->gsb;
  else if (n == "gps_sps_grad_basis") tmp = h->gsgb;
  else if (n == "vol_jac") tmp = h->vol_jac;
  else if (n == "G") tmp = h->G;
  else if (n == "GB") tmp = h->GB;
  else if (n == "face_jac") tmp = h->face_jac;
  else if (n == "face_Acoef") tmp = h->face_A;
  else if (n == "face_Bcoef") tmp = h->face_B;
  else if (n == "face_Anorm") tmp = h->face_An;
  else if (n == "dxdxi") tmp = h->dxdxi;
  else if (n == "dxdeta") tmp = h->dxdeta;
  else if (n == "dydxi") tmp = h->dydxi;
  else if (n == "dydeta") tmp = h->dydeta;
  else if (n == "BC_Poisson") tmp = h->BC_Poisson;
  else if (n == "BC_normal_vel") tmp = h->BC_normal_vel;
  else if (n == "BC_parl_vel") tmp = h->BC_parl_vel;
  else if (n == "BC_diffusion") tmp = h->BC_diffusion;
  else if (n == "BC_vorticity") tmp = h->BC_vorticity;
  else if (n == "velocity_jump") tmp = h->velocity_jump;
  else if (n == "BC_switch_Poisson") from_uc(h->BC_switch_Poisson);
  else if (n == "BC_switch_diffusion") from_uc(h->BC_switch_diffusion);
  else if (n == "nbr") from_int(h->nbr);
  else if (n == "nbr_side") from_int(h->nbr_side);
  else if (n == "bnd_index") from_int(h->bnd_index);
  else if (n == "be_elem") from_int(h->be_elem);
  else if (n == "be_side") from_int(h->be_side);
  else if (n == "RHS_advective") tmp = h->RHS_adv;
  else if (n == "RHS_diffusive") tmp = h->RHS_diff;
  else if (n == "edge_nodes") { for (int e = 0; e < h->N_eb; ++e) for (int v : h->edges[e]) tmp.push_back(v); }
  else if (n == "sps_xy") {  // src/HO_2D.cpp:3946-3980
    const int L = h->Lnod, K = h->K;
    for (int el = 0; el < h->N_el; ++el)
      for (int j = 0; j < K; ++j)
        for (int k = 0; k < K; ++k) {
          double x = 0., y = 0.;
          for (int l = 0; l < L; ++l)
            for (int m = 0; m < L; ++m) {
              int nid = h->elem_nodes[(size_t)el * L * L + h->tensor2FEM(m, l)];
              x += h->gsb[l * K + j] * h->gsb[m * K + k] * h->nx[nid];
              y += h->gsb[l * K + j] * h->gsb[m * K + k] * h->ny[nid];
            }
          tmp.push_back(x); tmp.push_back(y);
        }
  } else if (n == "edge_xy") {  // flux points of each boundary edge, edge order (src/HO_2D.cpp:752)
    const int K = h->K;
    for (int e = 0; e < h->N_eb; ++e)
      for (int k = 0; k < K; ++k) {
        double x = 0., y = 0.;
        for (int j = 0; j < h->Lnod; ++j) {
          int nd = h->edges[e][h->tensor2FEM(j)];
          x += h->gsb[j * K + k] * h->nx[nd];
          y += h->gsb[j * K + k] * h->ny[nd];
        }
        tmp.push_back(x); tmp.push_back(y);
      }
  } else if (n == "areas") {  // src/HO_2D.cpp:3982-3992
    const int K = h->K;
    for (int el = 0; el < h->N_el; ++el)
      for (int j = 0; j < K; ++j)
        for (int k = 0; k < K; ++k) tmp.push_back(h->wgt[j] * h->wgt[k] * h->vol_jac[h->v3(el, j, k)]);
  } else return -1;
  int64_t c = std::min<int64_t>(cap, (int64_t)tmp.size());
  if (out) std::copy(tmp.begin(), tmp.begin() + c, out);
  return (int64_t)tmp.size();
}

int ho_oracle_set_bc(ho_oracle* h, const char* name, const double* vals) {
  std::string n(name);
  std::vector<double>* v = nullptr;
  if (n == "BC_Poisson") v = &h->BC_Poisson;
  else if (n == "BC_diffusion") v = &h->BC_diffusion;
  else if (n == "BC_normal_vel") v = &h->BC_normal_vel;
  else if (n == "BC_parl_vel") v = &h->BC_parl_vel;
  else return -1;
  std::copy(vals, vals + v->size(), v->begin());
  return 0;
}

int ho_oracle_n_nodes(ho_oracle* h) { return h->N_nodes; }
int ho_oracle_smooth_fields(ho_oracle* h, double* out) {
  try { h->smooth_fields(out); } catch (const std::exception& e) { return fail(h, e); }
  return 0;
}
int ho_oracle_set_sample_points(ho_oracle* h, int n, const double* xy) {
  try { h->set_sample_points(n, xy); } catch (const std::exception& e) { return fail(h, e); }
  return 0;
}
int ho_oracle_sample_location(ho_oracle* h, int32_t* elem, double* loc) {
  std::copy(h->sp_elem.begin(), h->sp_elem.end(), elem);
  std::copy(h->sp_loc.begin(), h->sp_loc.end(), loc);
  return 0;
}
int ho_oracle_sample_values(ho_oracle* h, double* out) {
  try { h->sample_values(out); } catch (const std::exception& e) { return fail(h, e); }
  return 0;
}
int ho_oracle_refresh(ho_oracle* h) {  // src/HO_2D.cpp:1411-1414, before each dump
  try {
    h->update_BCs(h->current_time);
    h->solve_Poisson(h->vorticity.data());
    h->calc_velocity();
  } catch (const std::exception& e) { return fail(h, e); }
  return 0;
}

int ho_oracle_apply_laplacian_bc(ho_oracle* h, const double* x, const double* bcvals, double* y) {
  try {
    h->fr_second_order(x, h->BC_switch_Poisson.data(), bcvals, nullptr, y);
  } catch (const std::exception& e) { return fail(h, e); }
  return 0;
}

int ho_oracle_solve_poisson_bc(ho_oracle* h, const double* w, double* psi) {
  try {
    h->solve_Poisson(w);
    std::copy(h->stream_function.begin(), h->stream_function.end(), psi);
  } catch (const std::exception& e) { return fail(h, e); }
  return 0;
}

int ho_oracle_velocity_from_psi(ho_oracle* h, double* uv) {
  try {
    h->calc_velocity();
    std::copy(h->vel.begin(), h->vel.end(), uv);
  } catch (const std::exception& e) { return fail(h, e); }
  return 0;
}

int ho_oracle_set_time(ho_oracle* h, double t) { h->current_time = t; return 0; }

int ho_oracle_update_bcs(ho_oracle* h, double t) {
  try {
    h->update_BCs(t);
  } catch (const std::exception& e) { return fail(h, e); }
  return 0;
}

int ho_oracle_advection(ho_oracle* h, const double* w, const double* uv, const double* bc_vort,
                        const double* bc_nvel, double* out) {
  try {
    std::copy(uv, uv + h->vel.size(), h->vel.begin());
    std::copy(bc_vort, bc_vort + h->BC_vorticity.size(), h->BC_vorticity.begin());
    std::copy(bc_nvel, bc_nvel + h->BC_normal_vel.size(), h->BC_normal_vel.begin());
    h->update_advection_BC();
    if (h->advection_type == 1) h->calc_RHS_advection(w);
    else h->calc_RHS_advection_continuous(w);
    std::copy(h->RHS_adv.begin(), h->RHS_adv.end(), out);
  } catch (const std::exception& e) { return fail(h, e); }
  return 0;
}

int ho_oracle_apply_laplacian_mf(ho_oracle* h, const double* x, double* y) {
  try {
    std::vector<double> zeros((size_t)h->N_eb * h->K, 0.);
    h->fr_second_order(x, h->BC_switch_Poisson.data(), zeros.data(), nullptr, y);
  } catch (const std::exception& e) { return fail(h, e); }
  return 0;
}

}  // extern "C"
