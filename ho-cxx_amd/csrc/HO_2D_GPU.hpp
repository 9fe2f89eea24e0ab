// HO_2D_GPU.hpp - header-only C++ facade with the reference's HO_2D method names
// (src/HO_2D.hpp:220-290) over the C ABI of include/hoc.h.
//
// Existing callers of the reference class switch by type: the mesh/boundary loaders,
// process_mesh_input, the solution getters/setters, the hybrid driver's open-boundary
// setters, solveto_d and getallvorts_d keep their names, argument meaning and ownership
// (caller-owned buffers, copied).  Errors that the reference handles with exit()/assert
// throw hoc_gpu_error instead; methods the reference declares void still return nothing.
// Problem types 1-3 and 11 (hybrid) are supported.
#pragma once

#include <cmath>
#include <cstdint>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/hoc.h"

class hoc_gpu_error : public std::runtime_error {
 public:
  hoc_gpu_error(int code, const std::string& msg) : std::runtime_error(msg), code(code) {}
  int code;
};

class HO_2D_GPU {
 public:
  HO_2D_GPU() { hoc_default_config(&cfg_); }
  ~HO_2D_GPU() { clean_up(); }
  HO_2D_GPU(const HO_2D_GPU&) = delete;             // the reference's shallow copy is a double-free hazard
  HO_2D_GPU& operator=(const HO_2D_GPU&) = delete;

  // ---- configuration (src/HO_2D.cpp:3698-3721; input.dat fields :213-317) ----
  void set_defaults() { hoc_default_config(&cfg_); }
  void enable_hybrid() { hybrid_ = true; }
  void set_elemorder(const int32_t k) { cfg_.K = k; }
  void set_problem(int problem_type, double Re, double dt, int time_integration_type = 4, int advection_type = 1) {
    cfg_.problem_type = problem_type; cfg_.Re = Re; cfg_.dt = dt;
    cfg_.time_integration_type = time_integration_type; cfg_.advection_type = advection_type;
  }
  void set_linear_solver(double tol, int maxiter, int precision = HOC_FP64) {
    cfg_.tol = tol; cfg_.maxiter = maxiter; cfg_.precision = precision;
  }
  hoc_config& config() { return cfg_; }

  // ---- mesh input (src/HO_2D.cpp:3823-3898) ----
  // _iorder: geometry order (Lnod-1); _nnodes: 2*number of nodes; _nelems: nodes*elements;
  // _nwall/_nopen: Lnod*number of edges, edge node lists in Gmsh line order.
  void load_mesh_arrays_d(const int32_t _iorder, const int32_t _nnodes, const double* _xy, const int32_t _nelems,
                          const int32_t* _idxelems, const int32_t _nwall, const int32_t* _idxwall,
                          const int32_t _nopen, const int32_t* _idxopen) {
    lnod_ = _iorder + 1;
    xy_.assign(_xy, _xy + _nnodes);
    elem_.assign(_idxelems, _idxelems + _nelems);
    bedge_.clear(); bbdry_.clear(); names_.clear(); nval_.clear(); vals_.clear();
    add_boundary("wall", _nwall, _idxwall);
    add_boundary("open", _nopen, _idxopen);
  }
  // inlets/outlets with 1 (normal) or 2 (tangential, normal) velocity values per edge
  // (src/HO_2D.cpp:3882-3896, 3779-3793)
  void load_inout_arrays_d(const int32_t _ninval, const double* _velin, const int32_t _nin, const int32_t* _idxin,
                           const int32_t _noutval, const double* _velout, const int32_t _nout,
                           const int32_t* _idxout) {
    add_boundary("inlet", _nin, _idxin, _ninval, _velin);
    add_boundary("outlet", _nout, _idxout, _noutval, _velout);
  }
  // Named boundaries as read from a Gmsh file (src/Mesh.cpp:121-253), with optional values.
  void add_boundary(const std::string& name, int32_t n_idx, const int32_t* idx, int32_t n_val = 0,
                    const double* val = nullptr) {
    if (n_idx <= 0) return;
    const int b = (int)names_.size();
    const int32_t ne = n_idx / lnod_;
    const int per = (n_val > 0 && val) ? n_val / ne : 0;
    if (per < 0 || per > 2 || (per > 0 && per * ne != n_val))
      throw hoc_gpu_error(HOC_EINVAL, "boundary values must be 1 or 2 per edge");
    names_.push_back(name);
    for (int32_t i = 0; i < n_idx; ++i) bedge_.push_back(idx[i]);
    for (int32_t e = 0; e < ne; ++e) {
      bbdry_.push_back(b);
      nval_.push_back(per);
      vals_.push_back(per ? val[e * per] : 0.);
      vals_.push_back(per == 2 ? val[e * per + 1] : 0.);
    }
  }
  void process_mesh_input() {
    clean_up();
    n_samples_ = 0;
    check(hoc_create(&h_, &cfg_), nullptr);
    std::string joined;
    for (size_t i = 0; i < names_.size(); ++i) joined += (i ? "\n" : "") + names_[i];
    const int32_t L2 = lnod_ * lnod_;
    check(hoc_load_mesh(h_, lnod_, (int32_t)(xy_.size() / 2), xy_.data(), (int32_t)(elem_.size() / L2), elem_.data(),
                        (int32_t)bbdry_.size(), bedge_.data(), bbdry_.data(), (int32_t)names_.size(), joined.c_str()),
          h_);
    for (int32_t v : nval_)
      if (v) {
        check(hoc_set_boundary_values(h_, (int32_t)nval_.size(), nval_.data(), vals_.data()), h_);
        break;
      }
    check(hoc_setup(h_), h_);
  }

  // ---- getters (src/HO_2D.cpp:3942-3992, 4238-4259) ----
  int32_t getsolnptlen() { return (int32_t)(2 * dofs()); }
  void getsolnpts_d(const int32_t n, double* xy) { check(hoc_get_solution_points(h_, n, xy), h_); }
  void getsolnareas_d(const int32_t n, double* a) { check(hoc_get_solution_areas(h_, n, a), h_); }
  void getallvorts_d(const int32_t n, double* w) { check(hoc_get_vorticity(h_, n, w), h_); }
  void get_hoquad_weights_d(const int32_t n, double* out) {  // 0.25 w_j w_k (src/HO_2D.cpp:4252-4259)
    const int K = cfg_.K;
    if (n != K * K) throw hoc_gpu_error(HOC_EINVAL, "get_hoquad_weights_d: length must be K*K");
    const double* w = gauss_weights(K);
    for (int j = 0; j < K; ++j)
      for (int k = 0; k < K; ++k) out[j * K + k] = 0.25 * w[j] * w[k];
  }
  void getstreamfunction_d(const int32_t n, double* p) { check(hoc_get_streamfunction(h_, n, p), h_); }
  void getvelocity_d(const int32_t n, double* uv) { check(hoc_get_velocity(h_, n, uv), h_); }

  // ---- hybrid driver (src/HO_2D.cpp:3994-4184) ----
  int32_t getopenptlen() { return (int32_t)(2 * (h_ ? hoc_open_point_count(h_) : 0)); }
  void getopenpts_d(const int32_t n, double* xy) { check(hoc_get_open_points(h_, n, xy), h_); }
  void setopenvels_d(const int32_t n, double* uv) { check(hoc_set_open_velocities(h_, n, uv), h_); }
  void setopenvort_d(const int32_t n, double* w) { check(hoc_set_open_vorticity(h_, n, w), h_); }
  // Stores the driver's vorticity target; it enters the solution through the weighted blend of
  // problem_type 11 only, as in the reference.
  void setsolnvort_d(const int32_t n, double* w) {
    if (cfg_.problem_type == 11) check(hoc_set_vorticity_target(h_, n, w), h_);
    else if (n != getsolnptlen() / 2) throw hoc_gpu_error(HOC_EINVAL, "setsolnvort_d: length mismatch");
  }
  void setptogweights_d(const int32_t n, double* w) {
    if (cfg_.problem_type == 11) check(hoc_set_p2g_weights(h_, n, w), h_);
  }
  // trigger_write (src/HO_2D.cpp:4262-4271): the solution-point file and the node-averaged file
  // (save_vorticity_vtk / save_smooth_vtk), named as the reference names them, from the state of
  // the last stage; with sample points set, their values as well (src/HO_2D.cpp:3593-3612).
  void trigger_write(const int32_t indx, const int32_t subidx = -1) {
    const std::string mid = std::to_string(cfg_.problem_type) + "_K" + std::to_string(cfg_.K) +
                            (subidx > -1 ? "_" + pad(subidx, 2) : "") + "_" + pad(indx, 5);
    check(hoc_write_vtk(h_, ("vorticity" + mid + ".vtk").c_str()), h_);
    check(hoc_write_smooth_vtk(h_, 0, ("problem" + mid + ".vtk").c_str()), h_);
    if (n_samples_ > 0) {
      const std::string sp = "sampling_points" + std::to_string(cfg_.problem_type) + "_K" + std::to_string(cfg_.K) +
                             "_L" + std::to_string(lnod_) + "_" + pad(indx, 5) + ".dat";
      check(hoc_write_samples(h_, 0, sp.c_str()), h_);
    }
  }
  // read_process_sample_points (src/HO_2D.cpp:319-474) on a file of "x y" pairs; a last pair
  // within 1e-6 of the previous one is dropped, as the reference does.  Call after
  // process_mesh_input.
  void read_process_sample_points(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw hoc_gpu_error(HOC_EINVAL, "cannot open sample points file " + path);
    std::vector<double> xy;
    double a, b;
    while (in >> a >> b) { xy.push_back(a); xy.push_back(b); }
    const size_t n = xy.size() / 2;
    if (n >= 2 && std::hypot(xy[2 * n - 2] - xy[2 * n - 4], xy[2 * n - 1] - xy[2 * n - 3]) < 1e-6) {
      xy.resize(xy.size() - 2);
    }
    n_samples_ = (int32_t)(xy.size() / 2);
    check(hoc_set_sample_points(h_, n_samples_, xy.data()), h_);
  }
  // The vorticity field itself (initial conditions; no counterpart in the reference).
  void set_vorticity(const int32_t n, const double* w) { check(hoc_set_vorticity(h_, n, w), h_); }

  // ---- marching ----
  void solveto_d(const double outer_dt, const int32_t nstep, const int32_t integ, const double Re,
                 const int32_t /*region*/ = 0) {
    check(hoc_solveto(h_, outer_dt, nstep, integ, Re), h_);
  }
  // solve_advection_diffusion (src/HO_2D.cpp:1465-1531), nsteps times
  char solve_advection_diffusion(int32_t nsteps = 1) {
    check(hoc_step(h_, nsteps), h_);
    return 0;
  }
  double current_time() const { return h_ ? hoc_get_time(h_) : 0.; }

  // ---- wall velocity and gradients (no counterpart in the reference) ----
  // BC_parl_vel at the boundary flux points [E][K] (src/HO_2D.cpp:717-725 sets -1 along the lid).
  std::vector<double> wall_velocity() const {
    std::vector<double> v((size_t)hoc_boundary_point_count(h_));
    check(hoc_get_wall_velocity(h_, (int64_t)v.size(), v.data()), h_);
    return v;
  }
  void set_wall_velocity(const std::vector<double>& v) {
    check(hoc_set_wall_velocity(h_, (int64_t)v.size(), v.data()), h_);
  }
  // nsteps RK4 steps keeping checkpoints every `every` steps, then the gradients of an objective
  // with respect to the initial vorticity and the wall velocity from its gradient with respect
  // to the final vorticity (sizes: number of solution points, boundary flux points).
  void adjoint_forward(int32_t nsteps, int32_t every) { check(hoc_adjoint_forward(h_, nsteps, every), h_); }
  void adjoint_backward(const std::vector<double>& wbar_final, std::vector<double>& wbar_initial,
                        std::vector<double>& wall_velocity_bar) {
    wbar_initial.resize(wbar_final.size());
    wall_velocity_bar.resize((size_t)hoc_boundary_point_count(h_));
    check(hoc_adjoint_backward(h_, (int64_t)wbar_final.size(), wbar_final.data(), wbar_initial.data(),
                               (int64_t)wall_velocity_bar.size(), wall_velocity_bar.data()),
          h_);
  }

  void clean_up() {
    if (h_) hoc_destroy(h_);
    h_ = nullptr;
  }
  hoc_t handle() const { return h_; }

 private:
  static std::string pad(int32_t v, int width) {  // std::setw(width) << std::setfill('0') << v
    std::string t = std::to_string(v);
    return t.size() < (size_t)width ? std::string(width - t.size(), '0') + t : t;
  }
  int64_t dofs() const { return (int64_t)(elem_.size() / (lnod_ * lnod_)) * cfg_.K * cfg_.K; }
  static void check(int rc, hoc_t h) {
    if (rc != HOC_OK) throw hoc_gpu_error(rc, hoc_last_error(h));
  }
  static const double* gauss_weights(int K) {  // src/HO_2D.cpp:488-521
    static double w[7][6];
    static bool init = false;
    if (!init) {
      const double s30 = 5.477225575051661, s70 = 8.366600265340756;
      double t[7][6] = {{0}, {2.}, {1., 1.}, {5. / 9., 8. / 9., 5. / 9.},
                        {(18. - s30) / 36., (18. + s30) / 36., (18. + s30) / 36., (18. - s30) / 36.},
                        {(322. - 13. * s70) / 900., (322. + 13. * s70) / 900., 128. / 225., (322. + 13. * s70) / 900.,
                         (322. - 13. * s70) / 900.},
                        {0.17132449237917036, 0.36076157304813861, 0.46791393457269104, 0.46791393457269104,
                         0.36076157304813861, 0.17132449237917036}};
      for (int a = 0; a < 7; ++a)
        for (int b = 0; b < 6; ++b) w[a][b] = t[a][b];
      init = true;
    }
    if (K < 1 || K > 6) throw hoc_gpu_error(HOC_EINVAL, "K out of range");
    return w[K];
  }

  hoc_config cfg_{};
  hoc_t h_ = nullptr;
  bool hybrid_ = false;
  int32_t lnod_ = 2;
  int32_t n_samples_ = 0;
  std::vector<double> xy_;
  std::vector<int32_t> elem_, bedge_, bbdry_, nval_;
  std::vector<double> vals_;
  std::vector<std::string> names_;
};
