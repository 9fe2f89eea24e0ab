// hoc_host.hpp - host-side setup of the HO-CXX hot path on MI355X.
//
// Everything the GPU stage kernels read is produced here once per mesh:
// connectivity in face-local form, 1-D FR tables, volume/face metrics, the
// per-boundary-edge BC programme of problem types 1-3, element colourings for
// operator probing, and an SA-AMG hierarchy for the element-level Poisson matrix.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace hoc {

constexpr int kMaxK = 8;

// Faces are numbered by "ijp" as in the reference (src/HO_2D.cpp:2303-2336):
// 0 = west (xi=-1), 1 = east (xi=+1), 2 = south (eta=-1), 3 = north (eta=+1).
// Cell sides (south, east, north, west) = (0,1,2,3) (src/Mesh.hpp:114-119).
constexpr int kI2F[4] = {3, 1, 0, 2};   // ijp -> cell side   (src/MathHelpers.hpp:23)
constexpr int kF2I[4] = {2, 1, 3, 0};   // cell side -> ijp   (src/MathHelpers.hpp:24)

struct Mesh {
  int lnod = 2;
  int n_nodes = 0, n_elem = 0, n_bedges = 0;
  std::vector<double> xy;                 // interleaved
  std::vector<int32_t> elem;              // n_elem * lnod^2 (Gmsh order)
  std::vector<std::vector<int32_t>> edges;  // first n_bedges are the listed boundary edges
  std::vector<int32_t> bedge_bdry;
  std::vector<std::string> bnames;
  // connectivity by ijp face (filled by connect())
  std::vector<int32_t> nb;      // n_elem*4: neighbour element, or -1-edge for a boundary face
  std::vector<int8_t> nb_face;  // n_elem*4: neighbour's ijp of the shared face (or -1)
  std::vector<int32_t> be_elem; // per boundary edge: owner element
  std::vector<int8_t> be_ijp;   // per boundary edge: owner's face ijp
};

// Throws std::runtime_error on a non-conforming mesh.
void connect(Mesh& m);

// Gmsh 4.1 ASCII reader (Mesh::read_msh_file, src/Mesh.cpp:93-293): physical names of
// dimension 1 become the boundaries; only curves with a physical tag contribute boundary
// edges, in file order; 2-D elements keep file order; occ != 0 reflects every element's
// nodes into counter-clockwise order (src/Mesh.cpp:281-293).  Fills lnod, n_nodes, xy,
// n_elem, elem, n_bedges, edges, bedge_bdry, bnames.  Throws std::runtime_error.
void read_gmsh41(const std::string& path, int occ, Mesh& m);

// Legacy-VTK unstructured grid of the solution points (one quad cell per sub-square of
// each element's K x K point lattice) with point fields (src/HO_2D.cpp:3369-3692 writes
// the same fields as raw VTK).  fields: n_fields arrays of n_elem*K*K values.
void write_vtk_sps(const std::string& path, int K, int n_elem, const double* pts_xy, int n_fields,
                   const char* const* names, const double* const* fields);

// Greedy element colourings: distance 1 (face neighbours differ) and distance 2
// (additionally no element has two face neighbours of one colour).
int color_elements(const Mesh& m, int distance, std::vector<int32_t>& color);

struct Tables {
  int K = 0, L = 0;
  double sps[kMaxK], wgt[kMaxK], gps[4];
  double bb[kMaxK][2];        // Lagrange basis at xi=-1 / +1
  double bgb[kMaxK][2];       // its derivative at xi=-1 / +1
  double D[kMaxK][kMaxK];     // D[m][k] = l_m'(sps_k)
  double rad[kMaxK][2];       // left/right Radau (g_DG) correction values at sps
  double grad[kMaxK][2];      // their derivatives
  double gprime[2];           // g'(-1), g'(+1)
  double gsb[4][kMaxK];       // geometry basis at sps  [l][k]
  double gsgb[4][kMaxK];      // its derivative         [l][k]
  double gbb[4][2], gbgb[4][2];
};
void build_tables(int K, int lnod, Tables& t);

// Sample points (read_process_sample_points, src/HO_2D.cpp:319-474) for n points xy: the
// containing element (bounding box, then a ray-crossing test of the element's sides; the nearest
// mesh node's element when none contains the point), its local coordinates by Newton's method
// from (0.08, 0.08) with the reference's stopping test (both steps <= 1e-3; at most 100
// iterations here), and the geometry shape functions, in Gmsh node order, of the last iterate
// before the final update, as the reference keeps them.  The reference draws the ray's slope
// with std::rand(); here it is fixed, so the result does not depend on the run.
// elem: n, local_xy: 2n, shape: n * lnod^2.
void locate_sample_points(const Mesh& m, const Tables& t, int n, const double* xy, std::vector<int32_t>& elem,
                          std::vector<double>& local_xy, std::vector<double>& shape);
// Gmsh node slot of tensor position (i along xi, j along eta) for lnod = 2..4
// (src/Mesh.cpp:34-91).
int gmsh_slot(int lnod, int i, int j);
// Node-averaged output files in the reference's formats: the unstructured grid of the mesh
// nodes with vorticity, psi and velocity (save_smooth_vtk, src/HO_2D.cpp:3534-3589) from
// nodes[n_nodes][4] = vorticity, psi, u, v; and the sample-point lines "x y vorticity u v"
// (src/HO_2D.cpp:3593-3612) from vals[n][3].
void write_vtk_nodes(const std::string& path, const Mesh& m, const double* nodes);
void write_samples(const std::string& path, int n, const double* xy, const double* vals);


struct Geometry {
  // volume, [el][j][i]
  std::vector<double> xxi, xeta, yxi, yeta, jac;
  // faces, [el][ijp][k]
  std::vector<double> fxxi, fxeta, fyxi, fyeta, fjac, fA, fB, fAn;
};
void build_geometry(const Mesh& m, const Tables& t, Geometry& g);
// Solution-point coordinates (src/HO_2D.cpp:3946-3980), interleaved.
void solution_points(const Mesh& m, const Tables& t, std::vector<double>& xy);

// Boundary-condition programme of src/HO_2D.cpp:675-1238 and 1979-2217 for problem
// types 1-3.  Per-edge arrays of K values (edge flux-point order).  The values valid at
// time t are base * (ramp ? 2|sin(pi t/40)| : 1), with the outward->local sign folded in
// where update_BCs applies it.
struct BCs {
  std::vector<uint8_t> sw_psi, sw_diff;        // 0 Dirichlet, 1 Neumann
  std::vector<uint8_t> no_slip_edge;           // edge belongs to a no-slip boundary
  std::vector<double> psi, psi_ramp;           // BC_Poisson base, ramp flag (0/1)
  std::vector<double> diff, diff_ramp;         // BC_diffusion base (Dirichlet ones), ramp
  std::vector<double> nvel, nvel_ramp;         // BC_normal_vel base (sign folded), ramp
  std::vector<double> pvel, pvel_ramp;         // BC_parl_vel base, ramp
  std::vector<double> edge_sign;               // 2*(ijp%2)-1 of the owner face
  bool time_dependent = false;
  // problem_type 11 (hybrid): kind of each edge (0 wall, 1 open, 2 inlet, 3 outlet) and
  // the normal velocity an inlet edge carries
  std::vector<uint8_t> hybrid_kind;
  std::vector<double> inlet_nvel;
};
// edge_inst: ensemble instance of each boundary edge (empty: one instance); inlet extents
// are taken per instance.
// edge_normvel: normal velocity carried by each boundary edge (hybrid inlets/outlets;
// empty when none was given).
void build_bcs(const Mesh& m, const Tables& t, int problem_type, BCs& bc,
               const std::vector<int32_t>& edge_inst = {}, const std::vector<double>& edge_normvel = {},
               const std::vector<int8_t>& edge_has_nvel = {});

// Scalar CSR and smoothed-aggregation AMG hierarchy (host build, device apply).
struct CSR {
  int n = 0, m = 0;
  std::vector<int32_t> ptr, col;
  std::vector<double> val;
};
struct AMGLevelHost {
  CSR A, P, R;
  std::vector<double> Minv;   // spai0 weights
};
struct AMGHost {
  std::vector<AMGLevelHost> levels;  // last level: coarsest (A only)
  std::vector<double> coarse_inv;    // dense inverse of the coarsest matrix (row-major)
};
// Restart combination of a projection basis of m vectors X_j.  gcols: the basis's recorded Gram
// columns (column j at gcols + j ldg: G_ij for i < j, G_jj, then in row j + 1 the squared norm of
// the solution vector j came from; a vector with G_jj <= 1e-24 of that norm takes no part).  g:
// the measured Gram matrix (m x m, row-major), symmetrised in place with the non-participating
// rows and columns cleared.  w: column-major, column q the coefficients in the basis of the q-th
// appended solution.  The newest keep columns (m - keep .. m - 1) are orthonormalised in the g
// inner product by modified Gram-Schmidt applied twice, oldest first; a column whose remaining
// norm falls to 1e-24 of its own is left out.  c receives the combinations, kk rows of m (new
// vector k = sum_j c[k m + j] X_j); returns kk.
int proj_restart_basis(int m, int keep, const double* gcols, int ldg, double* g, const double* w,
                       std::vector<double>& c);
void build_sa_amg(const CSR& A, int coarse_enough, AMGHost& h);

// Throws std::invalid_argument unless A is a well-formed n x m CSR matrix (row pointers from
// 0 to nnz, non-decreasing; every column index in [0, m)).
void csr_check(const CSR& A);
// ELL packing for the device: w slots per row (a multiple of lanes), slot k of row i at
// ell_index(n, lanes, i, k): groups of `lanes` consecutive slots are stored row-major within
// a slot-group plane, so the `lanes` threads of a row, and consecutive rows, read consecutive
// words.  lanes = 4 for rows of 16 or more slots, else 1.  Padding slots repeat a column of the same
// row (column 0 for an empty row) with value 0, so every stored index is below A.m.  Returns
// false, packing nothing, when the rows are fewer than min_rows, longer than max_width, or too
// uneven (max > 1.25 avg + 1) for ELL to pay.
inline size_t ell_index(int n, int lanes, int i, int k) {
  return ((size_t)(k / lanes) * n + i) * lanes + k % lanes;
}
bool ell_pack(const CSR& A, int min_rows, int max_width, int& w, int& lanes, std::vector<int32_t>& col,
              std::vector<double>& val);

}  // namespace hoc
