/*
 * ho_oracle.h - C ABI of the FP64 CPU oracle for the HO-CXX time-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/, the
 * smoke() entry of __graft_entry__.py and bench.py's cpu_baseline leg may
 * load it.  The product path (ho-cxx_amd/) never links or calls it.
 *
 * It restates, in plain C++ (OpenMP only in the Krylov/AMG linear algebra), the reference solver's
 * arithmetic (reference: src/HO_2D.cpp, src/Mesh.cpp, src/MathHelpers.hpp)
 * and the published algorithm of the third-party Poisson solver the
 * reference calls (AMGCL 1.4.0: smoothed aggregation + spai0 + BiCGSTAB,
 * src/HO_2D.hpp:53-54, src/HO_2D.cpp:3066-3072).
 */
#ifndef HO_ORACLE_H
#define HO_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ho_oracle ho_oracle;

/* Build the oracle for one mesh/problem.
 *   K, Lnod       : solution points / geometry nodes per direction
 *   xy            : n_nodes interleaved (x,y)
 *   elem_nodes    : n_elem * Lnod*Lnod node ids in Gmsh order (CCW corners first)
 *   bedge_nodes   : n_bedges * Lnod node ids (Gmsh line order: 2 ends, then interior)
 *   bedge_bdry    : boundary id of each boundary edge, 0..n_bdry-1
 *   bdry_names    : n_bdry names separated by '\n'
 *   Re            : Reynolds number as in input.dat (<=1e-3 disables viscosity)
 *   tol, maxiter  : BiCGSTAB relative tolerance / iteration cap
 * Returns NULL on error (message via ho_oracle_last_error(NULL)). */
ho_oracle* ho_oracle_create(int K, int Lnod, int n_nodes, const double* xy,
                            int n_elem, const int32_t* elem_nodes,
                            int n_bedges, const int32_t* bedge_nodes,
                            const int32_t* bedge_bdry, int n_bdry,
                            const char* bdry_names, int problem_type,
                            double Re, double dt, int time_integration_type,
                            int advection_type, double tol, int maxiter);
/* As ho_oracle_create, with the values a boundary carries (arrays_to_boundary,
 * src/HO_2D.cpp:3720-3820): edge_nval[e] = 0, 1 (normal velocity in edge_vals[2e]) or 2
 * (tangential in edge_vals[2e], normal in edge_vals[2e+1]).  Needed by problem_type 11
 * (hybrid: boundaries named wall, open, inlet, outlet). */
ho_oracle* ho_oracle_create_ex(int K, int Lnod, int n_nodes, const double* xy,
                               int n_elem, const int32_t* elem_nodes,
                               int n_bedges, const int32_t* bedge_nodes,
                               const int32_t* bedge_bdry, int n_bdry,
                               const char* bdry_names, int problem_type,
                               double Re, double dt, int time_integration_type,
                               int advection_type, double tol, int maxiter,
                               const int32_t* edge_nval, const double* edge_vals);
void ho_oracle_destroy(ho_oracle*);

/* Hybrid driver surface (src/HO_2D.cpp:3994-4233), problem_type 11.  Open-boundary arrays
 * follow the order of the "open" boundary's edges. */
int ho_oracle_open_points(ho_oracle*, double* xy);          /* returns N_open*K; xy may be NULL */
int ho_oracle_set_open_vels(ho_oracle*, const double* xyvel);   /* setopenvels_d  */
int ho_oracle_set_open_vort(ho_oracle*, const double* w);       /* setopenvort_d  */
int ho_oracle_set_soln_vort(ho_oracle*, const double* w);       /* setsolnvort_d  */
int ho_oracle_set_p2g_weights(ho_oracle*, const double* w);     /* setptogweights_d */
int ho_oracle_solveto(ho_oracle*, double outer_dt, int nstep, int integ, double Re);  /* solveto_d */
const char* ho_oracle_last_error(ho_oracle*);
/* Threads of the OpenMP-parallel linear algebra (SpMV, vector updates, dot products of
 * the AMG-preconditioned Krylov solve); returns the previous count.  Thread count only
 * changes the summation order of dot products. */
int ho_oracle_set_threads(int n);

int ho_oracle_dofs(ho_oracle*);           /* N_el*K*K            */
int ho_oracle_n_bedges(ho_oracle*);

int ho_oracle_set_vorticity(ho_oracle*, const double* w);
int ho_oracle_get_vorticity(ho_oracle*, double* w);
int ho_oracle_get_streamfunction(ho_oracle*, double* psi);
int ho_oracle_set_streamfunction(ho_oracle*, const double* psi);
int ho_oracle_get_velocity(ho_oracle*, double* uv);   /* 2*dofs, (u,v) interleaved */
double ho_oracle_time(ho_oracle*);

/* nsteps calls of solve_advection_diffusion (src/HO_2D.cpp:1465), advancing time by dt each. */
int ho_oracle_step(ho_oracle*, int nsteps);
/* One stage: update_BCs(t) ... k = dt*(adv+diff) (src/HO_2D.cpp:1806-1839) on input w. */
int ho_oracle_stage(ho_oracle*, const double* w, double t, double* k_out);
/* Poisson only: update_BCs(t) then solve_Poisson(w) (src/HO_2D.cpp:1841). psi warm-started. */
int ho_oracle_poisson(ho_oracle*, const double* w, double t, double* psi_out);
/* Last Poisson solve statistics. */
int ho_oracle_last_solve(ho_oracle*, int* iters, double* rel_residual);
/* Poisson solves so far and their summed Krylov iterations. */
int ho_oracle_solve_counts(ho_oracle*, int64_t* n_solves, int64_t* total_iters);

/* Laplacian in CSR (after the |v|>1e-14 filter of src/HO_2D.cpp:3029,3043). */
int64_t ho_oracle_nnz(ho_oracle*);
int ho_oracle_get_csr(ho_oracle*, int32_t* rowptr, int32_t* col, double* val);
/* boundary_source [N_eb][K*K][K] (src/HO_2D.cpp:1631-1698). */
int ho_oracle_get_boundary_source(ho_oracle*, double* out);

/* Tables for pin tests.  name in {"sps","wgt","sps_boundary_basis","sps_boundary_grad_basis",
 * "sps_sps_grad_basis","sps_radau","sps_grad_radau","gps_sps_basis","gps_sps_grad_basis",
 * "vol_jac","G","GB","face_jac","face_Acoef","face_Bcoef","face_Anorm","dxdxi","dxdeta","dydxi",
 * "dydeta","BC_Poisson","BC_normal_vel","BC_parl_vel","BC_diffusion","BC_vorticity",
 * "velocity_jump","nbr","nbr_side","bnd_index","be_elem","be_side","RHS_advective","RHS_diffusive",
 * "sps_xy","areas"}.
 * Returns the element count, copies up to cap doubles (integers are returned as doubles). */
int64_t ho_oracle_table(ho_oracle*, const char* name, double* out, int64_t cap);

/* Apply the FR diffusion operator (calc_RHS_diffusion without Re^-1/J scaling) with the
 * ψ boundary types and zero boundary values: the matrix-free form of the Laplacian. */
int ho_oracle_apply_laplacian_mf(ho_oracle*, const double* x, double* y);

/* Test hooks: overwrite a BC array ("BC_Poisson", "BC_diffusion", "BC_normal_vel",
 * "BC_parl_vel"; N_eb*K values), apply the FR Laplacian with given psi BC values, solve
 * the Poisson system with the current BC values (no update_BCs), velocity from the
 * current stream function. */
int ho_oracle_set_bc(ho_oracle*, const char* name, const double* vals);
int ho_oracle_apply_laplacian_bc(ho_oracle*, const double* x, const double* bcvals, double* y);
int ho_oracle_solve_poisson_bc(ho_oracle*, const double* w, double* psi);
int ho_oracle_velocity_from_psi(ho_oracle*, double* uv);

/* Node-averaged output (save_smooth_vtk, src/HO_2D.cpp:3369-3531): out[n_nodes][4] = vorticity,
 * stream function, u, v at the mesh nodes. */
int ho_oracle_n_nodes(ho_oracle*);
int ho_oracle_smooth_fields(ho_oracle*, double* out);
/* Sample points (read_process_sample_points, src/HO_2D.cpp:319-474, with a fixed ray slope):
 * containing element, local coordinates, and vorticity/u/v interpolated from the node-averaged
 * fields (src/HO_2D.cpp:3593-3612), out[n][3]. */
int ho_oracle_set_sample_points(ho_oracle*, int n, const double* xy);
int ho_oracle_sample_location(ho_oracle*, int32_t* elem, double* local_xy);
int ho_oracle_sample_values(ho_oracle*, double* out);
/* update_BCs(t), solve_Poisson(vorticity), velocity: what the reference's main loop runs
 * before each dump (src/HO_2D.cpp:1411-1414). */
int ho_oracle_refresh(ho_oracle*);

/* Test hooks for the pins: set the clock; run update_BCs(t) alone (src/HO_2D.cpp:1979-2217) so
 * the boundary tables can be read; evaluate the advective residual (update_advection_BC +
 * calc_RHS_advection or calc_RHS_advection_continuous, src/HO_2D.cpp:2219-2645) of w with a
 * prescribed velocity uv (2*dofs, interleaved), boundary vorticity BC_vorticity and boundary
 * normal velocity BC_normal_vel (both N_eb*K, edge flux-point order, the latter with the
 * element-side sign update_BCs applies). */
int ho_oracle_set_time(ho_oracle*, double t);
int ho_oracle_update_bcs(ho_oracle*, double t);
int ho_oracle_advection(ho_oracle*, const double* w, const double* uv, const double* bc_vort,
                        const double* bc_nvel, double* out);

#ifdef __cplusplus
}
#endif
#endif
