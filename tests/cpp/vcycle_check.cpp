// vcycle_check.cpp - CPU check of the sparse-level data the device V-cycle consumes.
//
// Builds the smoothed-aggregation hierarchy (hoc_host.cpp) of the square of the Q1 (bilinear
// FEM) stiffness matrix on a g x g vertex grid (a 25-point stencil, the width of the FR
// solver's Q1 vertex operator), then for every level matrix A, P, R:
//   * csr_check passes and ell_pack stores only column indices below the column count;
//   * the ELL product equals the CSR product;
// and replays the device V-cycle's five sparse steps per level (modes 4, 0, 2, 3 and the
// dense coarsest solve, in the order of csr_vcycle in hoc_dev.hpp) with bounds-checked
// vectors sized like the device buffers, checking that no step writes a vector it reads at
// other rows, and that the resulting preconditioner is linear and symmetric.
// Prints "OK" and exits 0 on success.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>
#include <stdexcept>
#include <vector>

#include "hoc_host.hpp"

using namespace hoc;
using Vec = std::vector<double>;

static CSR q1_stiffness(int g) {
  // K = kron(T, M) + kron(M, T): T = tridiag(-1, 2, -1), M = tridiag(1, 4, 1) / 6 (Dirichlet)
  const double T[3] = {-1., 2., -1.}, M[3] = {1. / 6., 4. / 6., 1. / 6.};
  CSR A;
  A.n = A.m = g * g;
  A.ptr.push_back(0);
  for (int y = 0; y < g; ++y)
    for (int x = 0; x < g; ++x) {
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int yy = y + dy, xx = x + dx;
          if (yy < 0 || yy >= g || xx < 0 || xx >= g) continue;
          A.col.push_back(yy * g + xx);
          A.val.push_back(T[dy + 1] * M[dx + 1] + M[dy + 1] * T[dx + 1]);
        }
      A.ptr.push_back((int)A.col.size());
    }
  return A;
}

static CSR square(const CSR& A) {  // A A for a symmetric A with sorted rows
  CSR C;
  C.n = C.m = A.n;
  C.ptr.push_back(0);
  std::vector<double> acc(A.n, 0.);
  std::vector<int> mark(A.n, -1), cols;
  for (int i = 0; i < A.n; ++i) {
    cols.clear();
    for (int j = A.ptr[i]; j < A.ptr[i + 1]; ++j)
      for (int l = A.ptr[A.col[j]]; l < A.ptr[A.col[j] + 1]; ++l) {
        const int c = A.col[l];
        if (mark[c] != i) { mark[c] = i; acc[c] = 0.; cols.push_back(c); }
        acc[c] += A.val[j] * A.val[l];
      }
    std::sort(cols.begin(), cols.end());
    for (int c : cols) { C.col.push_back(c); C.val.push_back(acc[c]); }
    C.ptr.push_back((int)C.col.size());
  }
  return C;
}

struct Dev {  // a level matrix as the device holds it
  CSR A;
  int w = 0, lanes = 1;
  std::vector<int32_t> ec;
  Vec ev;
};

static Dev make_dev(const CSR& A) {
  csr_check(A);
  Dev d;
  d.A = A;
  if (ell_pack(A, 16384, 32, d.w, d.lanes, d.ec, d.ev)) {
    for (int32_t c : d.ec)
      if (c < 0 || c >= A.m) throw std::runtime_error("ELL column index out of range");
    if (d.w % d.lanes || d.ec.size() != (size_t)d.w * A.n) throw std::runtime_error("ELL shape");
    std::vector<char> seen(d.ec.size(), 0);  // the index map is a bijection onto the storage
    for (int i = 0; i < A.n; ++i)
      for (int k = 0; k < d.w; ++k) {
        const size_t q = ell_index(A.n, d.lanes, i, k);
        if (q >= seen.size() || seen[q]) throw std::runtime_error("ELL index map");
        seen[q] = 1;
      }
  }
  return d;
}

// The sparse modes of k_csr / k_ell: 0 y = A x; 1 y = f - A x; 2 y = f + A x;
// 3 y = x + M (f - A x); 4 xio = M f, y = f - A (M f).  Vectors are accessed with at().
static void apply(const Dev& d, int mode, const Vec* x, const Vec* f, Vec* y, const Vec* M, Vec* xio) {
  const int n = d.A.n;
  if (y == x || (mode == 4 && (y == f || xio == f))) throw std::runtime_error("output aliases an input");
  if (y->size() != (size_t)n) throw std::runtime_error("output size");
  for (int i = 0; i < n; ++i) {
    double s = 0.;
    const int len = d.w ? d.w : d.A.ptr[i + 1] - d.A.ptr[i];
    for (int k = 0; k < len; ++k) {
      const size_t q = d.w ? ell_index(n, d.lanes, i, k) : (size_t)d.A.ptr[i] + k;
      const int c = d.w ? d.ec.at(q) : d.A.col.at(q);
      const double v = d.w ? d.ev.at(q) : d.A.val.at(q);
      s += v * (mode == 4 ? M->at(c) * f->at(c) : x->at(c));
    }
    if (mode == 0) y->at(i) = s;
    else if (mode == 1) y->at(i) = f->at(i) - s;
    else if (mode == 2) y->at(i) = f->at(i) + s;
    else if (mode == 3) y->at(i) = x->at(i) + M->at(i) * (f->at(i) - s);
    else { y->at(i) = f->at(i) - s; xio->at(i) = M->at(i) * f->at(i); }
  }
}

struct Level {
  Dev A, P, R;
  Vec M, f, u, t;
};

static void vcycle(std::vector<Level>& L, const Vec& cinv, size_t l, const Vec& f, Vec& u) {
  Level& a = L[l];
  if (l + 1 == L.size()) {
    const int n = a.A.A.n;
    for (int i = 0; i < n; ++i) {
      double s = 0.;
      for (int j = 0; j < n; ++j) s += cinv.at((size_t)i * n + j) * f.at(j);
      u.at(i) = s;
    }
    return;
  }
  Level& c = L[l + 1];
  apply(a.A, 4, nullptr, &f, &a.t, &a.M, &u);
  apply(a.R, 0, &a.t, nullptr, &c.f, nullptr, nullptr);
  vcycle(L, cinv, l + 1, c.f, c.u);
  apply(a.P, 2, &c.u, &u, &a.t, nullptr, nullptr);
  apply(a.A, 3, &a.t, &f, &u, &a.M, nullptr);
}

int main() {
  try {
    const int g = 160;  // 25 600 rows: the finest A and P take the ELL form
    const CSR A0 = square(q1_stiffness(g));
    AMGHost H;
    build_sa_amg(A0, 768, H);
    std::vector<Level> L(H.levels.size());
    int n_ell = 0;
    for (size_t l = 0; l < H.levels.size(); ++l) {
      L[l].A = make_dev(H.levels[l].A);
      n_ell += L[l].A.w > 0;
      if (l + 1 < H.levels.size()) {
        L[l].P = make_dev(H.levels[l].P);
        L[l].R = make_dev(H.levels[l].R);
        n_ell += (L[l].P.w > 0) + (L[l].R.w > 0);
        if (L[l].P.A.n != H.levels[l].A.n || L[l].P.A.m != H.levels[l + 1].A.n ||
            L[l].R.A.n != H.levels[l + 1].A.n || L[l].R.A.m != H.levels[l].A.n)
          throw std::runtime_error("transfer operator shapes");
        L[l].M = H.levels[l].Minv;
      }
      const size_t n = H.levels[l].A.n;
      L[l].f.assign(n, 0.); L[l].u.assign(n, 0.); L[l].t.assign(n, 0.);
      std::printf("level %zu: n %d nnz %zu ell w/lanes A %d/%d P %d/%d R %d/%d\n", l, H.levels[l].A.n,
                  H.levels[l].A.val.size(), L[l].A.w, L[l].A.lanes, L[l].P.w, L[l].P.lanes, L[l].R.w, L[l].R.lanes);
    }
    if (n_ell == 0 || L[0].A.lanes != 4) throw std::runtime_error("the finest operator did not take the 4-lane ELL form");
    // ELL product == CSR product on every ELL matrix
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> U(-1., 1.);
    for (auto& lv : L)
      for (Dev* d : {&lv.A, &lv.P, &lv.R}) {
        if (!d->w) continue;
        Vec x(d->A.m), y1(d->A.n), y2(d->A.n);
        for (auto& v : x) v = U(rng);
        apply(*d, 0, &x, nullptr, &y1, nullptr, nullptr);
        Dev c = *d;
        c.w = 0;
        apply(c, 0, &x, nullptr, &y2, nullptr, nullptr);
        for (size_t i = 0; i < y1.size(); ++i)
          if (std::fabs(y1[i] - y2[i]) > 1e-13 * (1. + std::fabs(y2[i]))) throw std::runtime_error("ELL != CSR");
      }
    // the V-cycle is a fixed linear, symmetric operator
    const size_t n = A0.n;
    Vec f1(n), f2(n), f3(n), u1(n), u2(n), u3(n);
    for (size_t i = 0; i < n; ++i) { f1[i] = U(rng); f2[i] = U(rng); f3[i] = 0.3 * f1[i] - 1.7 * f2[i]; }
    vcycle(L, H.coarse_inv, 0, f1, u1);
    vcycle(L, H.coarse_inv, 0, f2, u2);
    vcycle(L, H.coarse_inv, 0, f3, u3);
    double lin = 0., nrm = 0., a = 0., b = 0., sa = 0.;
    for (size_t i = 0; i < n; ++i) {
      const double e = u3[i] - (0.3 * u1[i] - 1.7 * u2[i]);
      lin += e * e; nrm += u3[i] * u3[i];
      a += f2[i] * u1[i]; b += f1[i] * u2[i]; sa += std::fabs(f2[i] * u1[i]);
    }
    std::printf("linearity %.3e symmetry %.3e\n", std::sqrt(lin / nrm), std::fabs(a - b) / sa);
    if (std::sqrt(lin / nrm) > 1e-12) throw std::runtime_error("V-cycle is not linear");
    if (std::fabs(a - b) > 1e-11 * sa) throw std::runtime_error("V-cycle is not symmetric");
    std::printf("OK\n");
    return 0;
  } catch (const std::exception& e) {
    std::printf("ERROR %s\n", e.what());
    return 1;
  }
}
