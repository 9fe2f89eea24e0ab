// hoc_host.cpp - host-side setup for the MI355X hot path (see hoc_host.hpp).
#include "hoc_host.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <unordered_map>

namespace hoc {

namespace {
uint64_t pair_key(int a, int b) {
  if (a > b) std::swap(a, b);
  return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
}

// Node ids of cell side s (south, east, north, west) in the element's CCW direction,
// tensor order along the side (src/Mesh.cpp:372-377).
void side_nodes(const Mesh& m, int el, int s, int* out) {
  const int L = m.lnod;
  const int32_t* en = &m.elem[(size_t)el * L * L];
  out[0] = en[s];
  out[L - 1] = en[(s + 1) % 4];
  for (int q = 0; q < L - 2; ++q) out[1 + q] = en[4 + s * (L - 2) + q];
}
}  // namespace

// Face connectivity with the reference's conventions (src/Mesh.cpp:342-439): an element
// side shared with another element points at the lowest-numbered other element and its
// first matching side; otherwise it must be one of the listed boundary edges, which then
// records this element as its owner.  Edges are matched by their end-node pair.
void connect(Mesh& m) {
  const int L = m.lnod, N = m.n_elem;
  std::unordered_map<uint64_t, std::vector<std::pair<int, int>>> sides;  // key -> (el, cell side)
  sides.reserve((size_t)N * 3);
  int tmp[4];
  for (int el = 0; el < N; ++el)
    for (int s = 0; s < 4; ++s) {
      side_nodes(m, el, s, tmp);
      sides[pair_key(tmp[0], tmp[L - 1])].push_back({el, s});
    }
  std::unordered_map<uint64_t, int> bkey;
  for (int e = 0; e < m.n_bedges; ++e) {
    uint64_t k = pair_key(m.edges[e][0], m.edges[e][1]);  // Gmsh line: two end nodes first
    if (!bkey.count(k)) bkey[k] = e;
  }
  m.nb.assign((size_t)N * 4, 0);
  m.nb_face.assign((size_t)N * 4, -1);
  m.be_elem.assign(m.n_bedges, -1);
  m.be_ijp.assign(m.n_bedges, -1);
  for (int el = 0; el < N; ++el)
    for (int s = 0; s < 4; ++s) {
      side_nodes(m, el, s, tmp);
      const uint64_t k = pair_key(tmp[0], tmp[L - 1]);
      const auto& lst = sides[k];
      int best = -1, bside = -1;
      for (const auto& pr : lst) {
        if (pr.first == el) continue;
        if (best < 0 || pr.first < best || (pr.first == best && pr.second < bside)) { best = pr.first; bside = pr.second; }
      }
      const int ijp = kF2I[s];
      if (best >= 0) {
        m.nb[(size_t)el * 4 + ijp] = best;
        m.nb_face[(size_t)el * 4 + ijp] = (int8_t)kF2I[bside];
      } else {
        auto it = bkey.find(k);
        if (it == bkey.end())
          throw std::runtime_error("mesh: element side without neighbour is not a listed boundary edge");
        const int e = it->second;
        m.nb[(size_t)el * 4 + ijp] = -1 - e;
        m.be_elem[e] = el;
        m.be_ijp[e] = (int8_t)ijp;
      }
    }
  for (int e = 0; e < m.n_bedges; ++e)
    if (m.be_elem[e] < 0) throw std::runtime_error("mesh: a listed boundary edge belongs to no element side");
  // boundary edges take their owner's orientation (src/Mesh.cpp:385)
  for (int e = 0; e < m.n_bedges; ++e) {
    side_nodes(m, m.be_elem[e], kI2F[m.be_ijp[e]], tmp);
    m.edges[e].assign(tmp, tmp + L);
  }
}

int color_elements(const Mesh& m, int distance, std::vector<int32_t>& color) {
  const int N = m.n_elem;
  color.assign(N, -1);
  int ncol = 0;
  std::vector<int> mark(64, -1);
  for (int el = 0; el < N; ++el) {
    std::vector<int> forbid;
    auto add_nbrs = [&](int e, int depth) {
      for (int f = 0; f < 4; ++f) {
        int n = m.nb[(size_t)e * 4 + f];
        if (n < 0) continue;
        if (color[n] >= 0) forbid.push_back(color[n]);
        if (depth > 1)
          for (int f2 = 0; f2 < 4; ++f2) {
            int n2 = m.nb[(size_t)n * 4 + f2];
            if (n2 >= 0 && n2 != el && color[n2] >= 0) forbid.push_back(color[n2]);
          }
      }
    };
    add_nbrs(el, distance);
    int c = 0;
    std::sort(forbid.begin(), forbid.end());
    forbid.erase(std::unique(forbid.begin(), forbid.end()), forbid.end());
    for (int f : forbid) { if (f == c) ++c; else if (f > c) break; }
    color[el] = c;
    ncol = std::max(ncol, c + 1);
  }
  return ncol;
}

namespace {
double legendre(int n, double x, double* dp) {
  double p0 = 1., p1 = x;
  if (n == 0) { if (dp) *dp = 0.; return 1.; }
  for (int k = 2; k <= n; ++k) { double p2 = ((2. * k - 1.) * x * p1 - (k - 1.) * p0) / k; p0 = p1; p1 = p2; }
  if (dp) *dp = n * (x * p1 - p0) / (x * x - 1.);
  return p1;
}
}  // namespace

// Gauss-Legendre points (closed forms for K <= 5 as in src/HO_2D.cpp:488-521, Newton
// iteration beyond), Lagrange tables at the sps and at xi = -1/+1, g_DG Radau correction
// tables (src/HO_2D.cpp:642-671) and the geometry (equispaced, src/HO_2D.cpp:527-541) tables.
void build_tables(int K, int lnod, Tables& t) {
  if (K < 2 || K > kMaxK) throw std::runtime_error("K out of range");
  if (lnod < 2 || lnod > 4) throw std::runtime_error("lnod out of range");
  t.K = K; t.L = lnod;
  double* x = t.sps; double* w = t.wgt;
  if (K == 2) { x[1] = 1. / std::sqrt(3.); x[0] = -x[1]; w[0] = w[1] = 1.; }
  else if (K == 3) { x[2] = std::sqrt(0.6); x[1] = 0.; x[0] = -x[2]; w[0] = w[2] = 5. / 9.; w[1] = 8. / 9.; }
  else if (K == 4) {
    x[3] = std::sqrt((15. + 2. * std::sqrt(30.)) / 35.); x[2] = std::sqrt((15. - 2. * std::sqrt(30.)) / 35.);
    x[1] = -x[2]; x[0] = -x[3];
    w[0] = w[3] = (18. - std::sqrt(30.)) / 36.; w[1] = w[2] = (18. + std::sqrt(30.)) / 36.;
  } else if (K == 5) {
    x[4] = std::sqrt(5. + 2. * std::sqrt(70.) / 7.) / 3.; x[3] = std::sqrt(5. - 2. * std::sqrt(70.) / 7.) / 3.;
    x[2] = 0.; x[1] = -x[3]; x[0] = -x[4];
    w[0] = w[4] = (322. - 13. * std::sqrt(70.)) / 900.; w[1] = w[3] = (322. + 13. * std::sqrt(70.)) / 900.; w[2] = 128. / 225.;
  } else {
    for (int i = 0; i < (K + 1) / 2; ++i) {
      double z = std::cos(M_PI * (i + 0.75) / (K + 0.5)), dp = 0.;
      for (int it = 0; it < 100; ++it) {
        double p = legendre(K, z, &dp), dz = p / dp;
        z -= dz;
        if (std::fabs(dz) < 1e-16) break;
      }
      legendre(K, z, &dp);
      x[K - 1 - i] = z; x[i] = -z;
      w[K - 1 - i] = w[i] = 2. / ((1. - z * z) * dp * dp);
    }
    if (K % 2) x[K / 2] = 0.;
  }
  // Lagrange basis of the sps: values/derivatives at -1, +1 and derivative matrix
  auto lag = [&](const double* nodes, int n, int k, double xx) {
    double num = 1., den = 1.;
    for (int j = 0; j < n; ++j) if (j != k) { num *= xx - nodes[j]; den *= nodes[k] - nodes[j]; }
    return num / den;
  };
  auto dlag = [&](const double* nodes, int n, int k, double xx) {
    double s = 0., den = 1.;
    for (int j = 0; j < n; ++j) {
      if (j == k) continue;
      den *= nodes[k] - nodes[j];
      double p = 1.;
      for (int i = 0; i < n; ++i) if (i != k && i != j) p *= xx - nodes[i];
      s += p;
    }
    return s / den;
  };
  for (int k = 0; k < K; ++k) {
    t.bb[k][0] = lag(x, K, k, -1.); t.bb[k][1] = lag(x, K, k, 1.);
    t.bgb[k][0] = dlag(x, K, k, -1.); t.bgb[k][1] = dlag(x, K, k, 1.);
    for (int m = 0; m < K; ++m) t.D[k][m] = dlag(x, K, k, x[m]);
  }
  // g_DG = Radau of degree K: R(x) = (-1)^K/2 (P_K - P_{K-1}) is the left one
  const double c = 0.5 * ((K % 2 == 0) ? 1. : -1.), cd = 0.5 * K * ((K % 2 == 0) ? 1. : -1.);
  for (int i = 0; i < K; ++i) {
    const double s = x[i];
    t.rad[i][0] = c * (legendre(K, s, nullptr) - legendre(K - 1, s, nullptr));
    t.rad[i][1] = c * (legendre(K, -s, nullptr) - legendre(K - 1, -s, nullptr));
    t.grad[i][0] = cd * (legendre(K, s, nullptr) + legendre(K - 1, s, nullptr)) / (1. + s);
    t.grad[i][1] = -cd * (legendre(K, -s, nullptr) + legendre(K - 1, -s, nullptr)) / (1. - s);
  }
  t.gprime[0] = -0.5 * K * (K + 1.);
  t.gprime[1] = 0.5 * K * (K + 1.);
  const int L = lnod;
  for (int l = 0; l < L; ++l) t.gps[l] = -1. + 2. * l / (L - 1);
  if (L == 4) { t.gps[1] = -1. / 3.; t.gps[2] = 1. / 3.; }
  for (int l = 0; l < L; ++l) {
    t.gbb[l][0] = lag(t.gps, L, l, -1.); t.gbb[l][1] = lag(t.gps, L, l, 1.);
    t.gbgb[l][0] = dlag(t.gps, L, l, -1.); t.gbgb[l][1] = dlag(t.gps, L, l, 1.);
    for (int k = 0; k < K; ++k) { t.gsb[l][k] = lag(t.gps, L, l, x[k]); t.gsgb[l][k] = dlag(t.gps, L, l, x[k]); }
  }
}

namespace {
// Gmsh quad node slot of tensor position (i along xi, j along eta) (src/Mesh.cpp:34-91).
int t2g(int L, int i, int j) {
  static const int m2[2][2] = {{0, 1}, {3, 2}};
  static const int m3[3][3] = {{0, 4, 1}, {7, 8, 5}, {3, 6, 2}};
  static const int m4[4][4] = {{0, 4, 5, 1}, {11, 12, 13, 6}, {10, 15, 14, 7}, {3, 9, 8, 2}};
  return L == 2 ? m2[j][i] : (L == 3 ? m3[j][i] : m4[j][i]);
}
}  // namespace

// Volume metrics at the sps and face metrics at the flux points (src/HO_2D.cpp:1240-1388).
void build_geometry(const Mesh& m, const Tables& t, Geometry& g) {
  const int K = t.K, L = t.L, N = m.n_elem, K2 = K * K;
  auto rs = [&](std::vector<double>& v, size_t n) { v.assign(n, 0.); };
  rs(g.xxi, (size_t)N * K2); rs(g.xeta, (size_t)N * K2); rs(g.yxi, (size_t)N * K2); rs(g.yeta, (size_t)N * K2); rs(g.jac, (size_t)N * K2);
  for (auto* v : {&g.fxxi, &g.fxeta, &g.fyxi, &g.fyeta, &g.fjac, &g.fA, &g.fB, &g.fAn}) rs(*v, (size_t)N * 4 * K);
  double X[4][4], Y[4][4];
  for (int el = 0; el < N; ++el) {
    for (int j = 0; j < L; ++j)
      for (int i = 0; i < L; ++i) {
        int nd = m.elem[(size_t)el * L * L + t2g(L, i, j)];
        X[j][i] = m.xy[2 * (size_t)nd]; Y[j][i] = m.xy[2 * (size_t)nd + 1];
      }
    for (int j = 0; j < K; ++j)
      for (int i = 0; i < K; ++i) {
        double a = 0., b = 0., c = 0., d = 0.;
        for (int n = 0; n < L; ++n)
          for (int mm = 0; mm < L; ++mm) {
            a += t.gsgb[mm][i] * t.gsb[n][j] * X[n][mm];
            c += t.gsgb[mm][i] * t.gsb[n][j] * Y[n][mm];
            b += t.gsgb[n][j] * t.gsb[mm][i] * X[n][mm];
            d += t.gsgb[n][j] * t.gsb[mm][i] * Y[n][mm];
          }
        size_t v = (size_t)el * K2 + j * K + i;
        g.xxi[v] = a; g.xeta[v] = b; g.yxi[v] = c; g.yeta[v] = d;
        g.jac[v] = a * d - b * c;
      }
    for (int k = 0; k < K; ++k)
      for (int f = 0; f < 4; ++f) {
        double xx = 0., xe = 0., yx = 0., ye = 0.;
        for (int j = 0; j < L; ++j)
          for (int i = 0; i < L; ++i) {
            if (f < 2) {
              xx += t.gbgb[i][f] * t.gsb[j][k] * X[j][i]; yx += t.gbgb[i][f] * t.gsb[j][k] * Y[j][i];
              xe += t.gbb[i][f] * t.gsgb[j][k] * X[j][i]; ye += t.gbb[i][f] * t.gsgb[j][k] * Y[j][i];
            } else {
              xx += t.gbb[j][f - 2] * t.gsgb[i][k] * X[j][i]; yx += t.gbb[j][f - 2] * t.gsgb[i][k] * Y[j][i];
              xe += t.gbgb[j][f - 2] * t.gsb[i][k] * X[j][i]; ye += t.gbgb[j][f - 2] * t.gsb[i][k] * Y[j][i];
            }
          }
        size_t q = ((size_t)el * 4 + f) * K + k;
        g.fxxi[q] = xx; g.fxeta[q] = xe; g.fyxi[q] = yx; g.fyeta[q] = ye;
        g.fjac[q] = xx * ye - xe * yx;
        g.fA[q] = f < 2 ? xe * xe + ye * ye : xx * xx + yx * yx;
        g.fB[q] = xx * xe + yx * ye;
        g.fAn[q] = std::sqrt(g.fA[q]);
      }
  }
}

void solution_points(const Mesh& m, const Tables& t, std::vector<double>& xy) {
  const int K = t.K, L = t.L;
  xy.assign((size_t)m.n_elem * K * K * 2, 0.);
  for (int el = 0; el < m.n_elem; ++el)
    for (int j = 0; j < K; ++j)
      for (int k = 0; k < K; ++k) {
        double x = 0., y = 0.;
        for (int l = 0; l < L; ++l)
          for (int q = 0; q < L; ++q) {
            int nd = m.elem[(size_t)el * L * L + t2g(L, q, l)];
            x += t.gsb[l][j] * t.gsb[q][k] * m.xy[2 * (size_t)nd];
            y += t.gsb[l][j] * t.gsb[q][k] * m.xy[2 * (size_t)nd + 1];
          }
        size_t v = ((size_t)el * K * K + j * K + k) * 2;
        xy[v] = x; xy[v + 1] = y;
      }
}

// Hybrid boundaries (src/HO_2D.cpp:893-1134): switches by boundary name; with inlets, psi is
// marched around the closed loop of non-open boundary edges (wall edges take the running
// value, inlet edges a linear profile and a uniform vorticity from the neighbouring inflow
// differences, outlets advance the running value), which the reference does for Lnod = 2.
static void build_hybrid_bcs(const Mesh& m, const Tables& t, BCs& bc, const std::vector<std::vector<int>>& bedges,
                             const std::vector<double>& nvel_in, const std::vector<int8_t>& has_nvel) {
  const int K = t.K, E = m.n_bedges, NB = (int)m.bnames.size();
  bc.hybrid_kind.assign(E, 0);
  bc.inlet_nvel.assign(E, 0.);
  bool inlets = false;
  for (int b = 0; b < NB; ++b) {
    const std::string& nm = m.bnames[b];
    int kind;
    if (nm == "wall") kind = 0;
    else if (nm == "open") kind = 1;
    else if (nm == "inlet") kind = 2;
    else if (nm == "outlet") kind = 3;
    else throw std::runtime_error("hybrid boundaries must be named wall, open, inlet or outlet");
    inlets = inlets || kind == 2;
    for (int e : bedges[b]) {
      bc.hybrid_kind[e] = (uint8_t)kind;
      bc.sw_psi[e] = (kind == 1 || kind == 3) ? 1 : 0;
      bc.sw_diff[e] = (kind == 0 || kind == 3) ? 1 : 0;
      bc.no_slip_edge[e] = kind == 0;
      if (kind >= 2) {
        if (has_nvel.empty() || !has_nvel[e]) throw std::runtime_error("hybrid inlet/outlet edge without a normal velocity");
        bc.inlet_nvel[e] = nvel_in[e];
      }
    }
  }
  if (!inlets) return;
  if (m.lnod != 2) throw std::runtime_error("hybrid inlets need lnod = 2");
  // boundary loop through the non-open edges, start at the first wall/open edge in input order
  std::unordered_map<int32_t, int> starts;  // first node -> edge
  for (int e = 0; e < E; ++e)
    if (bc.hybrid_kind[e] != 1) starts[m.edges[e][0]] = e;
  int e0 = 0;
  while (e0 < E && bc.hybrid_kind[e0] >= 2) ++e0;
  if (e0 == E) throw std::runtime_error("hybrid boundary has no wall or open edge");
  std::vector<int> loop{e0};
  for (;;) {
    auto it = starts.find(m.edges[loop.back()][1]);
    if (it == starts.end()) throw std::runtime_error("hybrid boundary edges do not close a loop");
    if (it->second == loop[0]) break;
    loop.push_back(it->second);
    if ((int)loop.size() > E) throw std::runtime_error("hybrid boundary edges do not close a loop");
  }
  const int M = (int)loop.size();
  std::vector<double> len(M);
  for (int q = 0; q < M; ++q) {
    const auto& nd = m.edges[loop[q]];
    len[q] = std::hypot(m.xy[2 * (size_t)nd[0]] - m.xy[2 * (size_t)nd[1]],
                        m.xy[2 * (size_t)nd[0] + 1] - m.xy[2 * (size_t)nd[1] + 1]);
  }
  auto vin = [&](int q) { return bc.hybrid_kind[loop[q]] == 2 ? bc.inlet_nvel[loop[q]] : 0.; };
  double run = 0.;
  for (int q = 0; q < M; ++q) {
    const int e = loop[q];
    const double a = run;
    if (bc.hybrid_kind[e] == 2) {
      const double vn = bc.inlet_nvel[e];
      run -= len[q] * vn;
      const int qp = q == 0 ? M - 1 : q - 1, qn = q == M - 1 ? 0 : q + 1;
      const double w = (vin(qn) - vn) / (len[qn] + len[q]) + (vn - vin(qp)) / (len[q] + len[qp]);
      for (int k = 0; k < K; ++k) {
        bc.psi[(size_t)e * K + k] = t.gsb[0][k] * a + t.gsb[1][k] * run;
        bc.diff[(size_t)e * K + k] = w;
      }
    } else if (bc.hybrid_kind[e] == 3) {
      run -= len[q] * bc.inlet_nvel[e];
    } else if (bc.hybrid_kind[e] == 0) {
      for (int k = 0; k < K; ++k) bc.psi[(size_t)e * K + k] = run;
    }
  }
}

void build_bcs(const Mesh& m, const Tables& t, int problem_type, BCs& bc, const std::vector<int32_t>& edge_inst,
               const std::vector<double>& edge_normvel, const std::vector<int8_t>& edge_has_nvel) {
  const int K = t.K, E = m.n_bedges, L = t.L;
  bc.sw_psi.assign(E, 0); bc.sw_diff.assign(E, 0); bc.no_slip_edge.assign(E, 0);
  for (auto* v : {&bc.psi, &bc.psi_ramp, &bc.diff, &bc.diff_ramp, &bc.nvel, &bc.nvel_ramp, &bc.pvel, &bc.pvel_ramp})
    v->assign((size_t)E * K, 0.);
  bc.edge_sign.assign(E, 0.);
  for (int e = 0; e < E; ++e) bc.edge_sign[e] = 2. * (m.be_ijp[e] % 2) - 1.;
  const int NB = (int)m.bnames.size();
  std::vector<std::vector<int>> bedges(NB);
  for (int e = 0; e < E; ++e) bedges[m.bedge_bdry[e]].push_back(e);
  auto edge_y = [&](int e, int k) {  // y at flux point k of the (reoriented) edge (src/HO_2D.cpp:752)
    double y = 0.;
    for (int j = 0; j < L; ++j) y += t.gsb[j][k] * m.xy[2 * (size_t)m.edges[e][j] + 1];
    return y;
  };
  // inlet extent (src/HO_2D.cpp:730-741), per ensemble instance (edge_inst empty: one)
  int ni = 1;
  for (int32_t v : edge_inst) ni = std::max(ni, v + 1);
  auto inst = [&](int e) { return edge_inst.empty() ? 0 : edge_inst[e]; };
  std::vector<double> ymin(ni, 0.), ymax(ni, 0.);
  for (int b = 0; b < NB; ++b) {
    if (m.bnames[b] != "inlet" || bedges[b].empty()) continue;
    std::vector<char> started(ni, 0);
    for (int e : bedges[b]) {
      const int i = inst(e);
      if (!started[i]) { ymin[i] = ymax[i] = m.xy[2 * (size_t)m.edges[e][0] + 1]; started[i] = 1; }
      for (int end : {0, L - 1}) {
        double y = m.xy[2 * (size_t)m.edges[e][end] + 1];
        ymin[i] = std::min(ymin[i], y); ymax[i] = std::max(ymax[i], y);
      }
    }
  }
  auto set = [&](std::vector<double>& v, std::vector<double>& r, int e, int k, double val, bool ramp) {
    v[(size_t)e * K + k] = val; r[(size_t)e * K + k] = ramp ? 1. : 0.;
  };
  if (problem_type == 1) {  // src/HO_2D.cpp:700-726
    for (int e = 0; e < E; ++e) { bc.sw_psi[e] = 0; bc.sw_diff[e] = 1; bc.no_slip_edge[e] = 1; }
    for (int b = 0; b < NB; ++b)
      if (m.bnames[b] == "top")
        for (int e : bedges[b]) for (int k = 0; k < K; ++k) set(bc.pvel, bc.pvel_ramp, e, k, -1., false);
  } else if (problem_type == 2) {  // src/HO_2D.cpp:728-807, ramps 1999-2046
    bc.time_dependent = true;
    for (int b = 0; b < NB; ++b) {
      const std::string& nm = m.bnames[b];
      for (int e : bedges[b]) {
        const double y_min = ymin[inst(e)], y_max = ymax[inst(e)], h2 = (y_max - y_min) * (y_max - y_min);
        if (nm == "inlet") {
          bc.sw_psi[e] = 0; bc.sw_diff[e] = 0;
          for (int k = 0; k < K; ++k) {
            double y = edge_y(e, k);
            set(bc.psi, bc.psi_ramp, e, k, (y - y_min) * (y - y_min) * (3. * y_max - y_min - 2. * y) / h2, true);
            set(bc.diff, bc.diff_ramp, e, k, 6. * (2. * y - y_min - y_max) / h2, true);
            set(bc.nvel, bc.nvel_ramp, e, k, 6. * (y - y_min) * (y - y_max) / h2 * bc.edge_sign[e], true);
          }
        } else if (nm == "outlet") {
          bc.sw_psi[e] = 1; bc.sw_diff[e] = 1;
        } else if (nm == "bottom") {
          bc.sw_psi[e] = 0; bc.sw_diff[e] = 1; bc.no_slip_edge[e] = 1;
        } else if (nm == "top") {
          bc.sw_psi[e] = 0; bc.sw_diff[e] = 1; bc.no_slip_edge[e] = 1;
          for (int k = 0; k < K; ++k) set(bc.psi, bc.psi_ramp, e, k, y_max - y_min, true);
        }
      }
    }
  } else if (problem_type == 3) {  // src/HO_2D.cpp:809-891, ramps 2048-2102
    bc.time_dependent = true;
    for (int b = 0; b < NB; ++b) {
      const std::string& nm = m.bnames[b];
      for (int e : bedges[b]) {
        const double y_min = ymin[inst(e)], y_max = ymax[inst(e)];
        if (nm == "inlet") {
          bc.sw_psi[e] = 0; bc.sw_diff[e] = 0;
          for (int k = 0; k < K; ++k) {
            set(bc.psi, bc.psi_ramp, e, k, edge_y(e, k) - y_min, true);
            set(bc.nvel, bc.nvel_ramp, e, k, -1. * bc.edge_sign[e], true);
          }
        } else if (nm == "outlet") {
          bc.sw_psi[e] = 1; bc.sw_diff[e] = 1;
        } else if (nm == "top") {
          bc.sw_psi[e] = 0; bc.sw_diff[e] = 1;
          for (int k = 0; k < K; ++k) { set(bc.psi, bc.psi_ramp, e, k, y_max - y_min, true); set(bc.pvel, bc.pvel_ramp, e, k, -1., true); }
        } else if (nm == "bottom") {
          bc.sw_psi[e] = 0; bc.sw_diff[e] = 1;
          for (int k = 0; k < K; ++k) set(bc.pvel, bc.pvel_ramp, e, k, 1., true);
        } else if (nm == "object") {
          bc.sw_psi[e] = 0; bc.sw_diff[e] = 1; bc.no_slip_edge[e] = 1;
          for (int k = 0; k < K; ++k) set(bc.psi, bc.psi_ramp, e, k, 0.5 * (y_max - y_min), true);
        }
      }
    }
  } else if (problem_type == 11) {  // hybrid, src/HO_2D.cpp:893-1134
    build_hybrid_bcs(m, t, bc, bedges, edge_normvel, edge_has_nvel);
  } else {
    throw std::runtime_error("problem_type must be 1 (cavity), 2 (backward-facing step), 3 (cylinder) or 11 (hybrid)");
  }
}

// ------------------------------------------------------------------------------------------
// Smoothed aggregation (Vanek-Mandel-Brezina) with the parameters AMGCL uses by default:
// strength eps = 0.08 halved per level, prolongator smoothing omega = 2/3 on the filtered
// matrix, spai0 relaxation, Galerkin coarse operators.
// ------------------------------------------------------------------------------------------
namespace {
CSR transpose(const CSR& A) {
  CSR T;
  T.n = A.m; T.m = A.n;
  T.ptr.assign(T.n + 1, 0);
  for (int c : A.col) T.ptr[c + 1]++;
  for (int i = 0; i < T.n; ++i) T.ptr[i + 1] += T.ptr[i];
  T.col.resize(A.col.size()); T.val.resize(A.val.size());
  std::vector<int> pos(T.ptr.begin(), T.ptr.end() - 1);
  for (int i = 0; i < A.n; ++i)
    for (int j = A.ptr[i]; j < A.ptr[i + 1]; ++j) { int p = pos[A.col[j]]++; T.col[p] = i; T.val[p] = A.val[j]; }
  return T;
}
CSR matmul(const CSR& A, const CSR& B) {
  CSR C;
  C.n = A.n; C.m = B.m;
  C.ptr.assign(C.n + 1, 0);
  std::vector<int> mark(B.m, -1);
  std::vector<double> acc(B.m, 0.);
  std::vector<int> cols;
  for (int i = 0; i < A.n; ++i) {
    cols.clear();
    for (int ja = A.ptr[i]; ja < A.ptr[i + 1]; ++ja) {
      const int k = A.col[ja]; const double va = A.val[ja];
      for (int jb = B.ptr[k]; jb < B.ptr[k + 1]; ++jb) {
        const int c = B.col[jb];
        if (mark[c] != i) { mark[c] = i; acc[c] = 0.; cols.push_back(c); }
        acc[c] += va * B.val[jb];
      }
    }
    std::sort(cols.begin(), cols.end());
    for (int c : cols) { C.col.push_back(c); C.val.push_back(acc[c]); }
    C.ptr[i + 1] = (int)C.col.size();
  }
  return C;
}
}  // namespace

void csr_check(const CSR& A) {
  if (A.n < 0 || A.m < 0 || (int64_t)A.ptr.size() != (int64_t)A.n + 1 || A.ptr[0] != 0 ||
      (size_t)A.ptr[A.n] != A.col.size() || A.col.size() != A.val.size())
    throw std::invalid_argument("inconsistent CSR matrix");
  for (int i = 0; i < A.n; ++i)
    if (A.ptr[i + 1] < A.ptr[i]) throw std::invalid_argument("CSR row pointers decrease");
  for (int32_t c : A.col)
    if (c < 0 || c >= A.m) throw std::invalid_argument("CSR column index out of range");
}

bool ell_pack(const CSR& A, int min_rows, int max_width, int& w, int& lanes, std::vector<int32_t>& col,
              std::vector<double>& val) {
  w = 0;
  lanes = 1;
  for (int i = 0; i < A.n; ++i) w = std::max(w, A.ptr[i + 1] - A.ptr[i]);
  const double avg = A.n ? (double)A.val.size() / A.n : 0.;
  if (A.m <= 0 || A.n < min_rows || w > max_width || w > 1.25 * avg + 1.) { w = 0; return false; }
  lanes = w >= 16 ? 4 : 1;
  w = (w + lanes - 1) / lanes * lanes;
  col.assign((size_t)w * A.n, 0);
  val.assign((size_t)w * A.n, 0.);
  for (int i = 0; i < A.n; ++i) {
    const int pad = A.ptr[i + 1] > A.ptr[i] ? A.col[A.ptr[i]] : 0;
    for (int k = 0; k < w; ++k) {
      const int j = A.ptr[i] + k;
      const bool in = j < A.ptr[i + 1];
      const size_t q = ell_index(A.n, lanes, i, k);
      col[q] = in ? A.col[j] : pad;
      val[q] = in ? A.val[j] : 0.;
    }
  }
  return true;
}

void build_sa_amg(const CSR& A0, int coarse_enough, AMGHost& h) {
  h.levels.clear();
  CSR A = A0;
  double eps = 0.08;
  while (A.n > coarse_enough) {
    AMGLevelHost lv;
    const int n = A.n;
    std::vector<double> dia(n, 0.);
    lv.Minv.assign(n, 0.);
    for (int i = 0; i < n; ++i) {
      double s2 = 0.;
      for (int j = A.ptr[i]; j < A.ptr[i + 1]; ++j) {
        if (A.col[j] == i) dia[i] += A.val[j];
        s2 += A.val[j] * A.val[j];
      }
      lv.Minv[i] = s2 > 0. ? dia[i] / s2 : 0.;
    }
    std::vector<uint8_t> strong(A.col.size(), 0);
    for (int i = 0; i < n; ++i)
      for (int j = A.ptr[i]; j < A.ptr[i + 1]; ++j) {
        const int c = A.col[j];
        strong[j] = c != i && eps * eps * dia[i] * dia[c] < A.val[j] * A.val[j];
      }
    const int UNDEF = -1, REMOVED = -2;
    std::vector<int> agg(n, UNDEF);
    for (int i = 0; i < n; ++i) {
      bool any = false;
      for (int j = A.ptr[i]; j < A.ptr[i + 1] && !any; ++j) any = strong[j];
      if (!any) agg[i] = REMOVED;
    }
    int na = 0;
    std::vector<int> nbrs;
    for (int i = 0; i < n; ++i) {
      if (agg[i] != UNDEF) continue;
      const int id = na++;
      agg[i] = id;
      nbrs.clear();
      for (int j = A.ptr[i]; j < A.ptr[i + 1]; ++j)
        if (strong[j] && agg[A.col[j]] != REMOVED) { agg[A.col[j]] = id; nbrs.push_back(A.col[j]); }
      for (int c : nbrs)
        for (int j = A.ptr[c]; j < A.ptr[c + 1]; ++j)
          if (strong[j] && agg[A.col[j]] == UNDEF) agg[A.col[j]] = id;
    }
    if (na == 0) break;
    std::vector<int> cnt(na, 0), rn(na, -1);
    for (int i = 0; i < n; ++i) if (agg[i] >= 0) cnt[agg[i]]++;
    int nc = 0;
    for (int a = 0; a < na; ++a) if (cnt[a]) rn[a] = nc++;
    for (int i = 0; i < n; ++i) if (agg[i] >= 0) agg[i] = rn[agg[i]];
    if (nc >= n) break;
    eps *= 0.5;
    const double omega = 2. / 3.;
    CSR P;
    P.n = n; P.m = nc;
    P.ptr.assign(n + 1, 0);
    std::vector<int> mark(nc, -1);
    std::vector<double> acc(nc, 0.);
    std::vector<int> cols;
    for (int i = 0; i < n; ++i) {
      double df = 0.;
      for (int j = A.ptr[i]; j < A.ptr[i + 1]; ++j) if (A.col[j] == i || !strong[j]) df += A.val[j];
      const double s = df != 0. ? -omega / df : 0.;
      cols.clear();
      for (int j = A.ptr[i]; j < A.ptr[i + 1]; ++j) {
        const int c = A.col[j];
        if (c != i && !strong[j]) continue;
        const int g = agg[c];
        if (g < 0) continue;
        const double v = c == i ? 1. - omega : s * A.val[j];
        if (mark[g] != i) { mark[g] = i; acc[g] = 0.; cols.push_back(g); }
        acc[g] += v;
      }
      std::sort(cols.begin(), cols.end());
      for (int g : cols) { P.col.push_back(g); P.val.push_back(acc[g]); }
      P.ptr[i + 1] = (int)P.col.size();
    }
    lv.A = A;
    lv.P = P;
    lv.R = transpose(P);
    CSR Ac = matmul(lv.R, matmul(A, P));
    h.levels.push_back(std::move(lv));
    A = std::move(Ac);
  }
  AMGLevelHost last;
  last.A = A;
  h.levels.push_back(last);
  // dense inverse of the coarsest matrix (Gauss-Jordan, partial pivoting)
  const int n = A.n;
  std::vector<double> a((size_t)n * n, 0.), inv((size_t)n * n, 0.);
  for (int i = 0; i < n; ++i) {
    inv[(size_t)i * n + i] = 1.;
    for (int j = A.ptr[i]; j < A.ptr[i + 1]; ++j) a[(size_t)i * n + A.col[j]] += A.val[j];
  }
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r) if (std::fabs(a[(size_t)r * n + c]) > std::fabs(a[(size_t)p * n + c])) p = r;
    if (a[(size_t)p * n + c] == 0.) continue;  // singular direction: leave it out of the inverse
    if (p != c)
      for (int j = 0; j < n; ++j) { std::swap(a[(size_t)p * n + j], a[(size_t)c * n + j]); std::swap(inv[(size_t)p * n + j], inv[(size_t)c * n + j]); }
    const double d = 1. / a[(size_t)c * n + c];
    for (int j = 0; j < n; ++j) { a[(size_t)c * n + j] *= d; inv[(size_t)c * n + j] *= d; }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const double f = a[(size_t)r * n + c];
      if (f == 0.) continue;
      for (int j = 0; j < n; ++j) { a[(size_t)r * n + j] -= f * a[(size_t)c * n + j]; inv[(size_t)r * n + j] -= f * inv[(size_t)c * n + j]; }
    }
  }
  h.coarse_inv = std::move(inv);
}

// ---------------------------------------------------------------------------------------
// Gmsh 4.1 ASCII input (src/Mesh.cpp:93-293)
// ---------------------------------------------------------------------------------------
namespace {
struct Lines {
  std::vector<std::string> v;
  size_t find(const std::string& key) const {
    for (size_t i = 0; i < v.size(); ++i)
      if (v[i].compare(0, key.size(), key) == 0) return i;
    throw std::runtime_error("gmsh: section " + key + " not found");
  }
};
std::vector<long long> ints(const std::string& s) {
  std::vector<long long> out;
  const char* p = s.c_str();
  char* e = nullptr;
  for (;;) {
    while (*p == ' ' || *p == '\t' || *p == '\r') ++p;
    if (!*p) break;
    const long long x = std::strtoll(p, &e, 10);
    if (e == p) break;
    out.push_back(x);
    p = e;
    while (*p && *p != ' ' && *p != '\t' && *p != '\r') ++p;  // skip a fractional part
  }
  return out;
}
}  // namespace

void read_gmsh41(const std::string& path, int occ, Mesh& m) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("gmsh: cannot open " + path);
  Lines L;
  for (std::string ln; std::getline(in, ln);) L.v.push_back(ln);
  auto at = [&](size_t i) -> const std::string& {
    if (i >= L.v.size()) throw std::runtime_error("gmsh: unexpected end of file");
    return L.v[i];
  };
  // physical names of dimension 1: boundary names in file order
  size_t i = L.find("$PhysicalNames") + 1;
  const long long nphys = ints(at(i++)).at(0);
  std::vector<long long> phys_tags;
  m.bnames.clear();
  for (long long k = 0; k < nphys; ++k) {
    const std::string& ln = at(i++);
    const auto v = ints(ln);
    if (v.size() < 2 || v[0] != 1) continue;
    const size_t q0 = ln.find('"'), q1 = ln.rfind('"');
    if (q0 == std::string::npos || q1 <= q0) throw std::runtime_error("gmsh: bad physical name");
    phys_tags.push_back(v[1]);
    m.bnames.push_back(ln.substr(q0 + 1, q1 - q0 - 1));
  }
  // entities: curve tag -> its first physical tag (0 = none)
  i = L.find("$Entities") + 1;
  const auto cnt = ints(at(i++));
  if (cnt.size() < 4) throw std::runtime_error("gmsh: bad $Entities header");
  i += (size_t)cnt[0];
  std::unordered_map<long long, long long> curve_phys;
  for (long long k = 0; k < cnt[1]; ++k) {
    // curve: tag minX minY minZ maxX maxY maxZ numPhysical phys... numPoints points...
    std::istringstream ss(at(i++));
    long long tag, nph;
    double bb[6];
    ss >> tag;
    for (double& b : bb) ss >> b;
    ss >> nph;
    long long ph = 0;
    if (nph > 0) ss >> ph;
    curve_phys[tag] = ph;
  }
  // nodes (4.1: blocks of tags then coordinates); indices follow file order
  i = L.find("$Nodes") + 1;
  const auto nh = ints(at(i++));
  if (nh.size() < 4) throw std::runtime_error("gmsh: bad $Nodes header");
  if (nh[0] < 0 || nh[1] < 0 || nh[1] > INT32_MAX / 2) throw std::runtime_error("gmsh: bad $Nodes counts");
  std::unordered_map<long long, int32_t> tag2idx;
  m.xy.assign(2 * (size_t)nh[1], 0.);
  int32_t c = 0;
  for (long long b = 0; b < nh[0]; ++b) {
    const auto bh = ints(at(i++));
    if (bh.size() < 4) throw std::runtime_error("gmsh: bad node block");
    const long long nb = bh[3];
    if (nb < 0 || c + nb > nh[1]) throw std::runtime_error("gmsh: node block exceeds the $Nodes count");
    for (long long k = 0; k < nb; ++k) tag2idx[ints(at(i + k)).at(0)] = c + (int32_t)k;
    i += nb;
    for (long long k = 0; k < nb; ++k) {
      std::istringstream ss(at(i + k));
      ss >> m.xy[2 * (size_t)(c + k)] >> m.xy[2 * (size_t)(c + k) + 1];
    }
    i += nb;
    c += (int32_t)nb;
  }
  if (c != nh[1]) throw std::runtime_error("gmsh: node count mismatch");
  m.n_nodes = c;
  auto idx = [&](long long tag) {
    auto it = tag2idx.find(tag);
    if (it == tag2idx.end()) throw std::runtime_error("gmsh: element references an unknown node");
    return it->second;
  };
  // elements: line blocks (boundary edges of tagged curves) and quad blocks
  i = L.find("$Elements") + 1;
  const long long nblocks = ints(at(i++)).at(0);
  int lnod = 0, face_nodes = 0;
  m.elem.clear(); m.edges.clear(); m.bedge_bdry.clear();
  for (long long b = 0; b < nblocks; ++b) {
    const auto bh = ints(at(i++));
    if (bh.size() < 4) throw std::runtime_error("gmsh: bad element block");
    const long long dim = bh[0], etag = bh[1], etype = bh[2], nb = bh[3];
    if (nb < 0) throw std::runtime_error("gmsh: negative element count");
    if (dim == 1) {
      const int en = etype == 1 ? 2 : etype == 8 ? 3 : etype == 26 ? 4 : 0;
      if (!en) throw std::runtime_error("gmsh: unsupported line element type " + std::to_string(etype));
      lnod = en;
      const long long ph = curve_phys.count(etag) ? curve_phys[etag] : 0;
      if (ph != 0) {
        const auto pit = std::find(phys_tags.begin(), phys_tags.end(), ph);
        if (pit == phys_tags.end()) throw std::runtime_error("gmsh: curve with an unnamed physical tag");
        const int32_t bid = (int32_t)(pit - phys_tags.begin());
        for (long long k = 0; k < nb; ++k) {
          const auto v = ints(at(i + k));
          if ((int)v.size() < en + 1) throw std::runtime_error("gmsh: short line element");
          std::vector<int32_t> e(en);
          for (int q = 0; q < en; ++q) e[q] = idx(v[1 + q]);
          m.edges.push_back(e);
          m.bedge_bdry.push_back(bid);
        }
      }
    } else if (dim == 2) {
      const int fn = etype == 3 ? 4 : etype == 10 ? 9 : etype == 36 ? 16 : 0;
      if (!fn) throw std::runtime_error("gmsh: unsupported face element type " + std::to_string(etype));
      if (face_nodes && fn != face_nodes) throw std::runtime_error("gmsh: mixed quad orders");
      face_nodes = fn;
      for (long long k = 0; k < nb; ++k) {
        const auto v = ints(at(i + k));
        if ((int)v.size() < fn + 1) throw std::runtime_error("gmsh: short face element");
        for (int q = 0; q < fn; ++q) m.elem.push_back(idx(v[1 + q]));
      }
    }
    i += (size_t)nb;
  }
  if (!face_nodes) throw std::runtime_error("gmsh: no quadrilateral elements");
  const int Lq = face_nodes == 4 ? 2 : face_nodes == 9 ? 3 : 4;
  if (lnod && lnod != Lq) throw std::runtime_error("gmsh: line and quad orders differ");
  m.lnod = Lq;
  m.n_elem = (int)(m.elem.size() / face_nodes);
  m.n_bedges = (int)m.edges.size();
  if (occ) {  // reflect about the element diagonal: (i, j) -> (L-1-j, L-1-i)
    std::vector<int32_t> e(face_nodes);
    for (int el = 0; el < m.n_elem; ++el) {
      int32_t* p = m.elem.data() + (size_t)el * face_nodes;
      for (int jo = 0; jo < Lq; ++jo)
        for (int io = 0; io < Lq; ++io) e[t2g(Lq, Lq - 1 - jo, Lq - 1 - io)] = p[t2g(Lq, io, jo)];
      std::copy(e.begin(), e.end(), p);
    }
  }
}

// ---------------------------------------------------------------------------------------
// Legacy-VTK output of solution-point fields
// ---------------------------------------------------------------------------------------
int gmsh_slot(int lnod, int i, int j) { return t2g(lnod, i, j); }

void locate_sample_points(const Mesh& m, const Tables& t, int n, const double* xy, std::vector<int32_t>& elem,
                          std::vector<double>& local_xy, std::vector<double>& shape) {
  const int L = m.lnod, L2 = L * L, N = m.n_elem;
  constexpr double kSlope = 0.3819660112501051;  // fixed ray slope in [-1, 1]
  elem.assign(n, -1);
  local_xy.assign(2 * (size_t)n, 0.);
  shape.assign((size_t)n * L2, 0.);
  auto X = [&](int nd) { return m.xy[2 * (size_t)nd]; };
  auto Y = [&](int nd) { return m.xy[2 * (size_t)nd + 1]; };
  // the nodes of side s in order: corner s, the side's interior nodes, corner s+1 (Gmsh order)
  auto side_node = [&](int el, int s, int q) {
    const int32_t* en = &m.elem[(size_t)el * L2];
    if (q == 0) return en[s];
    if (q == L - 1) return en[(s + 1) % 4];
    return en[4 + s * (L - 2) + (q - 1)];
  };
  std::vector<double> mnx(N), mxx(N), mny(N), mxy(N);
  for (int el = 0; el < N; ++el) {  // bounding boxes of the element sides (src/HO_2D.cpp:342-354)
    const int n0 = m.elem[(size_t)el * L2];
    mnx[el] = mxx[el] = X(n0);
    mny[el] = mxy[el] = Y(n0);
    for (int s = 0; s < 4; ++s)
      for (int q = 0; q < L; ++q) {
        const int nd = side_node(el, s, q);
        mnx[el] = std::min(mnx[el], X(nd)); mxx[el] = std::max(mxx[el], X(nd));
        mny[el] = std::min(mny[el], Y(nd)); mxy[el] = std::max(mxy[el], Y(nd));
      }
  }
  // ray from p with slope kSlope against the segment a-b (are_intersecting, src/MathHelpers.hpp:99-114)
  auto hit = [&](double px, double py, int a, int b) {
    const double dx = X(b) - X(a), dy = Y(b) - Y(a);
    const double sp = (kSlope * (X(a) - px) + py - Y(a)) / (dy - kSlope * dx);
    const double tp = X(a) - px + sp * dx;
    return (tp > -1.e-8 && sp > -1.e-8 && sp < 1. + 1.e-8) ? 1 : 0;
  };
  std::vector<double> bx(L), by(L), gx(L), gy(L);
  for (int p = 0; p < n; ++p) {
    const double px = xy[2 * p], py = xy[2 * p + 1];
    int found = -1;
    for (int el = 0; el < N && found < 0; ++el) {  // src/HO_2D.cpp:357-378
      if (px < mnx[el] || px > mxx[el] || py < mny[el] || py > mxy[el]) continue;
      int hits = 0;
      for (int s = 0; s < 4; ++s)
        for (int q = 0; q + 1 < L; ++q) hits += hit(px, py, side_node(el, s, q), side_node(el, s, q + 1));
      if (hits & 1) found = el;
    }
    if (found < 0) {  // nearest mesh node (src/HO_2D.cpp:380-394)
      const int n0 = m.elem[0];
      double dist = std::sqrt((px - X(n0)) * (px - X(n0)) + (py - Y(n0)) * (py - Y(n0))) + 1.;
      for (int el = 0; el < N; ++el)
        for (int i = 0; i < L2; ++i) {
          const int nd = m.elem[(size_t)el * L2 + i];
          const double d = std::sqrt((px - X(nd)) * (px - X(nd)) + (py - Y(nd)) * (py - Y(nd)));
          if (dist > d) { dist = d; found = el; }
        }
    }
    elem[p] = found;
    // local coordinates (src/HO_2D.cpp:405-472)
    double xsi = 0.08, eta = 0.08, dxi = 0., deta = 0.;
    int it = 0;
    do {
      for (int k = 0; k < L; ++k) {
        gx[k] = gy[k] = 0.;
        bx[k] = by[k] = 1.;
        double den = 1.;
        for (int j = 0; j < L; ++j) {
          if (j == k) continue;
          den *= t.gps[k] - t.gps[j];
          bx[k] *= xsi - t.gps[j];
          by[k] *= eta - t.gps[j];
          double nxi = 1., net = 1.;
          for (int i = 0; i < L; ++i) {
            if (i == k || i == j) continue;
            nxi *= xsi - t.gps[i];
            net *= eta - t.gps[i];
          }
          gx[k] += nxi;
          gy[k] += net;
        }
        gx[k] /= den; bx[k] /= den; gy[k] /= den; by[k] /= den;
      }
      double f = px, g = py, fx = 0., fe = 0., gxi = 0., ge = 0.;
      for (int j = 0; j < L; ++j)
        for (int i = 0; i < L; ++i) {
          const int nd = m.elem[(size_t)found * L2 + t2g(L, i, j)];
          f -= bx[i] * by[j] * X(nd);
          g -= bx[i] * by[j] * Y(nd);
          fx -= by[j] * gx[i] * X(nd);
          gxi -= by[j] * gx[i] * Y(nd);
          fe -= gy[j] * bx[i] * X(nd);
          ge -= gy[j] * bx[i] * Y(nd);
        }
      dxi = -1. / (fx * ge - fe * gxi) * (f * ge - g * fe);
      deta = -1. / (fx * ge - fe * gxi) * (g * fx - f * gxi);
      xsi += dxi;
      eta += deta;
    } while ((std::fabs(dxi) > 0.001 || std::fabs(deta) > 0.001) && ++it < 100);
    if (!std::isfinite(xsi) || !std::isfinite(eta))
      throw std::runtime_error("sample point: local coordinates did not converge");
    local_xy[2 * p] = xsi;
    local_xy[2 * p + 1] = eta;
    for (int j = 0; j < L; ++j)
      for (int i = 0; i < L; ++i) shape[(size_t)p * L2 + t2g(L, i, j)] = bx[i] * by[j];
  }
}

void write_vtk_nodes(const std::string& path, const Mesh& m, const double* nodes) {
  std::ofstream out(path);
  if (!out) throw std::runtime_error("vtk: cannot open " + path);
  const int L2 = m.lnod * m.lnod;
  out << "# vtk DataFile Version 3.0\n2D Unstructured Grid of Quads\nASCII\n\nDATASET UNSTRUCTURED_GRID\n";
  out << "POINTS " << m.n_nodes << " double\n";
  for (int nd = 0; nd < m.n_nodes; ++nd) out << m.xy[2 * (size_t)nd] << ' ' << m.xy[2 * (size_t)nd + 1] << " 0.0\n";
  out << "CELLS " << m.n_elem << ' ' << (size_t)m.n_elem * (L2 + 1) << '\n';
  for (int el = 0; el < m.n_elem; ++el) {
    out << L2;
    for (int l = 0; l < L2; ++l) out << ' ' << m.elem[(size_t)el * L2 + l];
    out << '\n';
  }
  out << "CELL_TYPES " << m.n_elem << '\n';
  for (int el = 0; el < m.n_elem; ++el) out << "70\n";
  out << "POINT_DATA " << m.n_nodes << '\n';
  out << "SCALARS vorticity double 1 \nLOOKUP_TABLE default \n";
  for (int nd = 0; nd < m.n_nodes; ++nd) out << nodes[4 * (size_t)nd] << '\n';
  out << "SCALARS psi double 1 \nLOOKUP_TABLE default \n";
  for (int nd = 0; nd < m.n_nodes; ++nd) out << nodes[4 * (size_t)nd + 1] << '\n';
  out << "VECTORS velocity double \n";
  for (int nd = 0; nd < m.n_nodes; ++nd) out << nodes[4 * (size_t)nd + 2] << ' ' << nodes[4 * (size_t)nd + 3] << " 0.0\n";
  if (!out) throw std::runtime_error("vtk: write failed for " + path);
}

void write_samples(const std::string& path, int n, const double* xy, const double* vals) {
  std::ofstream out(path);
  if (!out) throw std::runtime_error("samples: cannot open " + path);
  for (int p = 0; p < n; ++p)
    out << xy[2 * p] << ' ' << xy[2 * p + 1] << ' ' << vals[3 * p] << ' ' << vals[3 * p + 1] << ' ' << vals[3 * p + 2]
        << '\n';
  if (!out) throw std::runtime_error("samples: write failed for " + path);
}

void write_vtk_sps(const std::string& path, int K, int n_elem, const double* pts, int n_fields,
                   const char* const* names, const double* const* fields) {
  std::ofstream out(path);
  if (!out) throw std::runtime_error("vtk: cannot open " + path);
  const size_t D = (size_t)n_elem * K * K;
  const size_t cells = (size_t)n_elem * (K - 1) * (K - 1);
  out << "# vtk DataFile Version 3.0\nHO-CXX solution points\nASCII\nDATASET UNSTRUCTURED_GRID\n";
  out << "POINTS " << D << " double\n";
  out.precision(17);
  for (size_t v = 0; v < D; ++v) out << pts[2 * v] << ' ' << pts[2 * v + 1] << " 0\n";
  out << "CELLS " << cells << ' ' << cells * 5 << '\n';
  for (int el = 0; el < n_elem; ++el)
    for (int j = 0; j + 1 < K; ++j)
      for (int i = 0; i + 1 < K; ++i) {
        const size_t b = (size_t)el * K * K;
        out << "4 " << b + j * K + i << ' ' << b + j * K + i + 1 << ' ' << b + (j + 1) * K + i + 1 << ' '
            << b + (j + 1) * K + i << '\n';
      }
  out << "CELL_TYPES " << cells << '\n';
  for (size_t q = 0; q < cells; ++q) out << "9\n";
  out << "POINT_DATA " << D << '\n';
  for (int f = 0; f < n_fields; ++f) {
    out << "SCALARS " << names[f] << " double 1\nLOOKUP_TABLE default\n";
    for (size_t v = 0; v < D; ++v) out << fields[f][v] << '\n';
  }
  if (!out) throw std::runtime_error("vtk: write failed for " + path);
}

int proj_restart_basis(int m, int keep, const double* gcols, int ldg, double* g, const double* w,
                       std::vector<double>& c) {
  if (m <= 0 || keep < 0 || keep > m || ldg < m + 1) throw std::invalid_argument("proj_restart_basis: sizes");
  std::vector<char> ok(m);
  for (int j = 0; j < m; ++j) {
    const double gjj = gcols[(size_t)j * ldg + j];
    ok[j] = gjj > 1e-24 * std::fabs(gcols[(size_t)j * ldg + j + 1]) && gjj > 0.;
  }
  for (int j = 0; j < m; ++j)
    for (int i = 0; i <= j; ++i) {
      const double v = (ok[i] && ok[j]) ? 0.5 * (g[(size_t)i * m + j] + g[(size_t)j * m + i]) : 0.;
      g[(size_t)i * m + j] = g[(size_t)j * m + i] = v;
    }
  auto gdot = [&](const double* u, const double* v) {
    double d = 0.;
    for (int i = 0; i < m; ++i) {
      double gi = 0.;
      for (int j = 0; j < m; ++j) gi += g[(size_t)i * m + j] * v[j];
      d += u[i] * gi;
    }
    return d;
  };
  c.clear();
  int kk = 0;
  std::vector<double> v(m);
  for (int col = m - keep; col < m; ++col) {
    for (int j = 0; j < m; ++j) v[j] = ok[j] ? w[(size_t)col * m + j] : 0.;
    const double n0 = gdot(v.data(), v.data());
    for (int pass = 0; pass < 2; ++pass)
      for (int q = 0; q < kk; ++q) {
        const double d = gdot(&c[(size_t)q * m], v.data());
        for (int j = 0; j < m; ++j) v[j] -= d * c[(size_t)q * m + j];
      }
    const double nv = gdot(v.data(), v.data());
    if (!(nv > 1e-24 * n0) || !(nv > 0.)) continue;  // in the span of those already kept
    const double f = 1. / std::sqrt(nv);
    for (int j = 0; j < m; ++j) c.push_back(v[j] * f);
    ++kk;
  }
  return kk;
}

}  // namespace hoc
