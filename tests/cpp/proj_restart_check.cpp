// Host check of the projection basis restart (proj_restart_basis, hoc_host.cpp): on random Gram
// matrices and coefficient columns, the combinations it returns are orthonormal in the Gram inner
// product, span the newest columns, leave out vectors that take no part, and drop a column that
// repeats another.  Prints OK on success.
#include <cmath>
#include <cstdio>
#include <random>
#include <stdexcept>
#include <vector>

#include "hoc_host.hpp"

using hoc::proj_restart_basis;

static int fails = 0;
static void check(bool ok, const char* what, double v = 0.) {
  if (!ok) { std::printf("FAIL %s (%.3e)\n", what, v); ++fails; }
}

struct Case {
  int m, keep, ldg;
  std::vector<double> g, gcols, w;
};

// G = I + eps S (S symmetric, entries in [-1, 1]); gcols recorded from G; w upper triangular with a
// unit diagonal (column q: the coefficients of solution q on vectors 0 .. q).
static Case make(int m, int keep, double eps, unsigned seed) {
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> u(-1., 1.);
  Case c{m, keep, m + 1, std::vector<double>((size_t)m * m), std::vector<double>((size_t)m * (m + 1), 0.),
         std::vector<double>((size_t)m * m, 0.)};
  for (int i = 0; i < m; ++i)
    for (int j = 0; j <= i; ++j) {
      const double v = (i == j ? 1. : 0.) + eps * u(rng);
      c.g[(size_t)i * m + j] = c.g[(size_t)j * m + i] = v;
    }
  for (int j = 0; j < m; ++j) {
    for (int i = 0; i <= j; ++i) c.gcols[(size_t)j * c.ldg + i] = c.g[(size_t)i * m + j];
    c.gcols[(size_t)j * c.ldg + j + 1] = 1.;
  }
  for (int q = 0; q < m; ++q) {
    for (int j = 0; j < q; ++j) c.w[(size_t)q * m + j] = u(rng);
    c.w[(size_t)q * m + q] = 1.;
  }
  return c;
}

static double gdot(const std::vector<double>& g, int m, const double* a, const double* b) {
  double d = 0.;
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) d += a[i] * g[(size_t)i * m + j] * b[j];
  return d;
}

// The combinations are G-orthonormal and every kept column (masked) lies in their span.
static int verify(const Case& cs, const std::vector<char>& ok, const char* name) {
  Case c = cs;
  std::vector<double> g0 = c.g, comb;
  const int kk = proj_restart_basis(c.m, c.keep, c.gcols.data(), c.ldg, c.g.data(), c.w.data(), comb);
  const int m = c.m;
  // the inner product the routine used: g0 symmetric, cleared outside ok
  std::vector<double> gm((size_t)m * m, 0.);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j)
      if (ok[i] && ok[j]) gm[(size_t)i * m + j] = g0[(size_t)i * m + j];
  double orth = 0.;
  for (int a = 0; a < kk; ++a)
    for (int b = 0; b < kk; ++b)
      orth = std::fmax(orth, std::fabs(gdot(gm, m, &comb[(size_t)a * m], &comb[(size_t)b * m]) - (a == b)));
  check(orth < 1e-12, name, orth);
  for (int a = 0; a < kk; ++a)
    for (int j = 0; j < m; ++j)
      if (!ok[j]) check(comb[(size_t)a * m + j] == 0., "left-out vector has a coefficient");
  double span = 0.;
  for (int col = m - c.keep; col < m; ++col) {
    std::vector<double> v(m);
    for (int j = 0; j < m; ++j) v[j] = ok[j] ? c.w[(size_t)col * m + j] : 0.;
    const double n0 = gdot(gm, m, v.data(), v.data());
    for (int a = 0; a < kk; ++a) {
      const double d = gdot(gm, m, &comb[(size_t)a * m], v.data());
      for (int j = 0; j < m; ++j) v[j] -= d * comb[(size_t)a * m + j];
    }
    span = std::fmax(span, std::sqrt(std::fabs(gdot(gm, m, v.data(), v.data())) / n0));
  }
  check(span < 1e-12, "kept columns outside the new span", span);
  return kk;
}

int main() {
  // 16 vectors, keep 8: all participate
  for (unsigned seed = 1; seed <= 5; ++seed) {
    Case c = make(16, 8, 0.02, seed);
    const int kk = verify(c, std::vector<char>(16, 1), "G-orthonormal combinations");
    check(kk == 8, "kk with independent columns", kk);
  }
  // vector 3 takes no part (its recorded G_jj is 0)
  {
    Case c = make(16, 8, 0.02, 11);
    c.gcols[(size_t)3 * c.ldg + 3] = 0.;
    std::vector<char> ok(16, 1);
    ok[3] = 0;
    const int kk = verify(c, ok, "orthonormal with a left-out vector");
    check(kk == 8, "kk with a left-out vector", kk);
  }
  // the newest column repeats the one before it: one fewer vector
  {
    Case c = make(16, 8, 0.02, 12);
    for (int j = 0; j < 16; ++j) c.w[(size_t)15 * 16 + j] = c.w[(size_t)14 * 16 + j];
    const int kk = verify(c, std::vector<char>(16, 1), "orthonormal with a repeated column");
    check(kk == 7, "kk with a repeated column", kk);
  }
  // a wider spread of Gram entries (cosines up to 0.2) and another size
  {
    Case c = make(12, 6, 0.2 / 12., 13);
    const int kk = verify(c, std::vector<char>(12, 1), "orthonormal, m = 12");
    check(kk == 6, "kk, m = 12", kk);
  }
  // nothing kept
  {
    Case c = make(16, 0, 0.02, 14);
    std::vector<double> comb;
    check(proj_restart_basis(16, 0, c.gcols.data(), c.ldg, c.g.data(), c.w.data(), comb) == 0 && comb.empty(),
          "keep = 0");
  }
  // sizes are checked
  {
    Case c = make(16, 8, 0.02, 15);
    std::vector<double> comb;
    bool threw = false;
    try {
      proj_restart_basis(16, 17, c.gcols.data(), c.ldg, c.g.data(), c.w.data(), comb);
    } catch (const std::invalid_argument&) {
      threw = true;
    }
    check(threw, "keep > m rejected");
  }
  std::printf(fails ? "FAILED %d\n" : "OK\n", fails);
  return fails ? 1 : 0;
}
