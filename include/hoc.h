/*
 * hoc.h - C ABI of the MI355X (gfx950) time-step hot path for HO-CXX.
 *
 * Drop-in for the HO_2D setup/advance surface of the reference
 * (src/HO_2D.hpp:220-290).  One handle = one mesh/problem on one device and
 * one HIP stream.  Host buffers are owned by the caller and copied; device
 * state is owned by the handle.  Every call is synchronous unless its name ends
 * in _device.  Status codes: 0 = OK, negative = error (message via
 * hoc_last_error).  No call exits, asserts or prints.
 *
 * Layouts follow the reference: fields at solution points are
 * [element][j (eta)][i (xi)] = el*K*K + j*K + i (src/HO_2D.cpp:4162, 4245);
 * velocities and points are interleaved (x,y) (src/HO_2D.cpp:3972-3974).
 */
#ifndef HOC_H
#define HOC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hoc_solver* hoc_t;

enum {
  HOC_OK = 0,
  HOC_EINVAL = -1,      /* bad argument / size mismatch / unsupported option */
  HOC_ESTATE = -2,      /* call out of order (e.g. step before setup)        */
  HOC_ENOCONV = -3,     /* Poisson solve did not reach the tolerance          */
  HOC_ENAN = -4,        /* NaN/Inf detected in a field                        */
  HOC_EDEVICE = -5,     /* HIP runtime error / no device / allocation failed  */
  HOC_EMESH = -6        /* mesh connectivity error                            */
};

enum { HOC_FP64 = 0, HOC_MIXED = 1 };
enum { HOC_KRYLOV_AUTO = 0, HOC_KRYLOV_BICGSTAB = 1, HOC_KRYLOV_PCG = 2 };

/* Parameters of input.dat (src/HO_2D.cpp:213-317) and of the linear solver
 * (src/HO_2D.cpp:3066-3069). */
typedef struct {
  int32_t K;                      /* Knod: solution points per direction (P+1), 2..6    */
  int32_t problem_type;           /* 1 cavity, 2 backward-facing step, 3 cylinder,
                                     11 hybrid (driver-supplied open-boundary values)    */
  double Re;                      /* Reynolds number (<= 1e-3 disables viscosity, :255) */
  double dt;                      /* time step                                           */
  int32_t time_integration_type;  /* 1 Euler, 2 RK2, 4 RK4                               */
  int32_t advection_type;         /* 1 discontinuous mass flux, 2 continuous             */
  double tol;                     /* Krylov relative residual tolerance                  */
  int32_t maxiter;                /* Krylov iteration cap                                */
  int32_t precision;              /* HOC_FP64, or HOC_MIXED: the FP64 Krylov method with
                                     the V-cycle preconditioner in FP32                  */
  int32_t krylov;                 /* HOC_KRYLOV_*; AUTO = PCG when W*A is symmetric     */
  int32_t device;                 /* HIP device ordinal                                  */
  int32_t fail_on_noconv;         /* nonzero: return HOC_ENOCONV when tol is not met     */
  double mixed_inner_tol;         /* unused (kept for layout); must lie in [0, 1)       */
  int32_t cheb_degree;            /* fine-level Chebyshev smoother degree, pre and post
                                     (0 = 3, or 4 with HOC_MIXED and PCG; at most 8)      */
  int32_t psi_guess;              /* initial guess of the stage Poisson solves: 1 = the
                                     previous solution (the reference's warm start),
                                     2 = linear extrapolation in the stage time from the
                                     solutions at the last two distinct stage times,
                                     3 = with RK4, the first two stages of a step from the
                                     previous step's stage solutions (psi is affine in
                                     omega), otherwise as 2;  4 = as 3, each predicted
                                     stage increment corrected by the same stage
                                     difference one step earlier;  5 = mode 4's guess (with
                                     Euler and RK2 the previous solution) plus its best
                                     correction in the span of 12 to 24 recent stage
                                     solutions: in the As-norm for a symmetric operator,
                                     least-squares residual otherwise  (0 = 5)            */
  double cheb_lower;              /* smoothing interval [cheb_lower, 1.1] x lambda_max(D^-1 As)
                                     (0 = 1/60)                                          */
  int32_t launch_mode;            /* Krylov iterations: 0 = chosen at hoc_setup by timing
                                     both, 1 = kernels queued one by one, 2 = replays of a
                                     captured hipGraph of 4 iterations (same iterates)   */
} hoc_config;

typedef struct {
  int64_t n_elem, n_dofs, n_bedges;
  int32_t n_levels;               /* Poisson hierarchy levels (fine level included)      */
  int32_t krylov;                 /* method in use                                       */
  int32_t last_iters;             /* iterations of the last Poisson solve                */
  double last_rel_residual;       /* ||b - A psi|| / ||b|| of the last solve             */
  int64_t total_iters;            /* summed over all solves                              */
  int64_t n_solves;
  double setup_seconds;
  double fine_op_bytes;           /* algorithmic bytes of one fine-operator application  */
  int32_t last_refinements;       /* 0 (the mixed solve is one FP64 Krylov solve)       */
  int32_t precision;              /* HOC_FP64 or HOC_MIXED                               */
  int64_t n_nodes;                /* mesh nodes (the node-averaged output's length)      */
  int32_t launch_graphs;          /* 1 when Krylov iterations replay a captured hipGraph */
  double launch_tune_ratio;       /* graph / direct time of the same iterations measured
                                     at hoc_setup (0 when launch_mode was set)           */
} hoc_stats;

/* Fill cfg with the reference defaults (src/HO_2D.cpp:3698-3713, 3066-3069). */
void hoc_default_config(hoc_config* cfg);

/* HO_2D() + set_defaults (src/HO_2D.cpp:3698). */
int hoc_create(hoc_t* h, const hoc_config* cfg);

/* = load_mesh_arrays_d / load_inout_arrays_d (src/HO_2D.cpp:3823-3898) and
 * Mesh::read_msh_file's output (src/Mesh.cpp:93-340):
 *   lnod         geometry nodes per edge (2..4)
 *   xy           n_nodes interleaved (x,y)
 *   elem_nodes   n_elem * lnod^2 node ids, Gmsh order (CCW corners, edge, interior)
 *   bedge_nodes  n_bedges * lnod node ids, Gmsh line order; listed boundary edges
 *                are numbered 0..n_bedges-1 in this order (BC arrays use it)
 *   bedge_bdry   boundary id of each edge; bdry_names holds n_bdry names joined by '\n'
 */
int hoc_load_mesh(hoc_t h, int32_t lnod, int32_t n_nodes, const double* xy,
                  int32_t n_elem, const int32_t* elem_nodes,
                  int32_t n_bedges, const int32_t* bedge_nodes,
                  const int32_t* bedge_bdry, int32_t n_bdry, const char* bdry_names);

/* = Mesh::read_msh_file (src/Mesh.cpp:93-293): a Gmsh 4.1 ASCII file; occ != 0 for files
 * written by the OCC kernel (element nodes reflected into CCW order, :281-293; the
 * OCC_format field of input.dat, src/HO_2D.cpp:244-246).  Boundaries are the physical
 * curves, named by $PhysicalNames.  Returns HOC_EMESH for unreadable or unsupported files. */
int hoc_load_gmsh(hoc_t h, const char* path, int32_t occ);

/* Ensembles (no counterpart in the reference, which runs one problem per process): the
 * loaded mesh is the union of n_inst independent meshes, instance i owning the elements
 * [elem_offsets[i], elem_offsets[i+1]) (instances may not share faces).  Every instance
 * advances with its own Re and dt (and its own clock for time-dependent boundary values)
 * in the same kernel launches.  With PCG every instance iterates with its own step lengths and
 * stops on its own ||b_i - A psi_i|| <= tol * ||b_i|| (instances that touch
 * at a corner share one iteration and stop when every instance meets that test); with
 * BiCGSTAB the instances share one iteration and stop on the joint bound
 * ||b - A psi|| <= tol * min_i ||b_i||, which implies it.  Call after hoc_load_mesh /
 * hoc_load_gmsh and before hoc_setup; hoc_solveto is rejected for an ensemble (it sets one dt
 * and Re). */
int hoc_set_instances(hoc_t h, int32_t n_inst, const int32_t* elem_offsets, const double* Re, const double* dt);
int hoc_get_instance_times(hoc_t h, int32_t n_inst, double* t);

/* Values carried by the boundary edges, as arrays_to_boundary receives them with
 * load_inout_arrays_d (src/HO_2D.cpp:3720-3820, 3882-3896): nval[e] = 0, 1 (normal velocity
 * in vals[2e]) or 2 (tangential in vals[2e], normal in vals[2e+1]).  Hybrid inlet and outlet
 * edges need their normal velocity.  Call before hoc_setup. */
int hoc_set_boundary_values(hoc_t h, int32_t n_bedges, const int32_t* nval, const double* vals);

/* = process_mesh_input (src/HO_2D.cpp:3901-3918): connectivity, bases, metrics,
 * IC/BC (setup_IC_BC_SRC), Poisson operator and hierarchy, device upload. */
int hoc_setup(hoc_t h);

int hoc_set_vorticity(hoc_t h, int64_t len, const double* w);     /* the vorticity field itself */
int hoc_get_vorticity(hoc_t h, int64_t len, double* w);           /* = getallvorts_d        */
int hoc_get_streamfunction(hoc_t h, int64_t len, double* psi);
int hoc_set_streamfunction(hoc_t h, int64_t len, const double* psi);
int hoc_get_velocity(hoc_t h, int64_t len, double* uv);           /* len = 2*dofs          */
int hoc_get_solution_points(hoc_t h, int64_t len, double* xy);   /* = getsolnpts_d        */
int hoc_get_solution_areas(hoc_t h, int64_t len, double* a);     /* = getsolnareas_d      */
/* Legacy-VTK file of the solution points (one quad cell per sub-square of each element's
 * K x K point lattice) with vorticity, streamfunction, u and v (the fields that
 * save_vorticity_vtk / save_smooth_vtk write, src/HO_2D.cpp:3369-3692). */
int hoc_write_vtk(hoc_t h, const char* path);
/* Node-averaged fields (save_smooth_vtk, src/HO_2D.cpp:3369-3531): vorticity, psi and velocity
 * extrapolated from each element's solution points to its geometry nodes, with the boundary
 * vorticity and the boundary velocity (normal and tangential values turned Cartesian) on boundary
 * faces, averaged over the elements sharing each mesh node.  out: [n_nodes][4] = vorticity,
 * psi, u, v (n_nodes = the mesh's node count).  refresh != 0 first sets the boundary values at
 * the current time and recomputes psi and the velocity for the current vorticity, as the
 * reference's main loop does before each dump (src/HO_2D.cpp:1411-1414); refresh == 0 uses the
 * state left by the last stage, as trigger_write does (src/HO_2D.cpp:4262-4271). */
int hoc_get_node_fields(hoc_t h, int32_t refresh, int32_t n_nodes, double* out);
/* The same fields as a legacy-VTK file of the mesh nodes in the reference's layout
 * (src/HO_2D.cpp:3534-3589: POINTS, CELLS of lnod^2 Gmsh-ordered nodes, cell type 70,
 * SCALARS vorticity, SCALARS psi, VECTORS velocity). */
int hoc_write_smooth_vtk(hoc_t h, int32_t refresh, const char* path);
/* Sample points (read_process_sample_points, src/HO_2D.cpp:319-474, for points passed as an
 * array instead of a file): each point's containing element (element bounding box, then a
 * ray-crossing test of its sides; the nearest mesh node's element when no element contains it)
 * and its local coordinates by Newton's method, with the reference's stopping test.  The
 * reference draws the ray's slope at random; here it is fixed.  n = 0 clears the points. */
int hoc_set_sample_points(hoc_t h, int32_t n, const double* xy);
/* Containing element (n) and local coordinates (2n, xi then eta) of each sample point. */
int hoc_get_sample_location(hoc_t h, int32_t n, int32_t* elem, double* local_xy);
/* Vorticity, u and v at the sample points, interpolated with the geometry shape functions from
 * the node-averaged fields (src/HO_2D.cpp:3593-3612).  out: [n][3]; refresh as above. */
int hoc_get_sample_values(hoc_t h, int32_t refresh, int32_t n, double* out);
/* The same values as a text file of lines "x y vorticity u v" (src/HO_2D.cpp:3593-3612). */
int hoc_write_samples(hoc_t h, int32_t refresh, const char* path);
/* The clock (current_time, src/HO_2D.cpp:1406-1418) that time-dependent boundary values follow
 * (the inflow ramps of problem types 2 and 3, src/HO_2D.cpp:2014, 2062).  Setting it also sets
 * every ensemble instance's clock. */
double hoc_get_time(hoc_t h);
int hoc_set_time(hoc_t h, double t);

/* ---- hybrid driver surface (problem_type 11, src/HO_2D.cpp:3994-4233) ----------------
 * Boundaries named wall / open / inlet / outlet.  Open-boundary arrays run over the
 * "open" boundary's edges in the order they were loaded, K points per edge.  Each setter
 * moves the previous "end" values to "start"; every stage blends start and end with the
 * fraction of the current solveto window (hoc_solveto opens a window [t_end, t_end + outer_dt]). */
int64_t hoc_open_point_count(hoc_t h);                                  /* = getopenptlen / 2 */
int hoc_get_open_points(hoc_t h, int64_t len, double* xy);             /* = getopenpts_d, len = 2*count */
int hoc_set_open_velocities(hoc_t h, int64_t len, const double* uv);   /* = setopenvels_d, len = 2*count */
int hoc_set_open_vorticity(hoc_t h, int64_t len, const double* w);     /* = setopenvort_d, len = count */
int hoc_set_vorticity_target(hoc_t h, int64_t len, const double* w);   /* = setsolnvort_d, len = dofs */
int hoc_set_p2g_weights(hoc_t h, int64_t len, const double* w);        /* = setptogweights_d, len = dofs */

/* nsteps x solve_advection_diffusion (src/HO_2D.cpp:1465-1531), time += dt each. */
int hoc_step(hoc_t h, int32_t nsteps);

/* = solveto_d (src/HO_2D.cpp:4194-4233): dt = outer_dt/nsub, integ, Re; the window
 * [t_end, t_end + outer_dt] starts at the previous window's end; then nsub steps. */
int hoc_solveto(hoc_t h, double outer_dt, int32_t nsub, int32_t integ, double Re);

/* Operator-level entry points for parity tests. */
/* update_BCs(t) + solve_Poisson(omega) (src/HO_2D.cpp:1841-1977); psi is warm-started
 * from the handle's current stream function and stored back into it. */
int hoc_poisson(hoc_t h, const double* omega, double t, double* psi);
/* One stage Euler_time_integrate(omega -> k, coeff=0 at time t) (src/HO_2D.cpp:1806-1839). */
int hoc_stage_residual(hoc_t h, const double* omega, double t, double* k);
/* y = A x with the Poisson matrix of form_Laplace_operator_matrix (src/HO_2D.cpp:1533-1804),
 * applied matrix-free. */
int hoc_apply_laplacian(hoc_t h, const double* x, double* y);
/* No counterpart in the reference: hoc_apply_laplacian for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * x and y are device buffers of n_dofs doubles (distinct); the call allocates nothing: its
 * scratch (the face traces, n_elem * 8 * K doubles) is a library-wide device array of 64 MiB per
 * GPU, held until the process ends, and larger meshes are rejected with HOC_EINVAL.  The work is
 * queued on `stream` (a hipStream_t; NULL = the handle's stream) and the call returns without
 * waiting; the caller synchronizes that stream before reading y.  Every call on one GPU shares
 * the array, so calls that overlap on one GPU (from any handles or streams) race.  Needs
 * hoc_setup on the handle (which takes the array's address on its device). */
int hoc_apply_laplacian_device(hoc_t h, const double* x, double* y, void* stream);

/* Boundary flux points and the tangential wall velocity.  Boundary arrays are [E][K] in the
 * edge flux-point order of the BC_* arrays (src/HO_2D.cpp:2229).  The wall velocity is
 * BC_parl_vel as setup_IC_BC_SRC sets it (src/HO_2D.cpp:717-725: -1 along the lid of the
 * cavity); setting it gives the walls any tangential velocity profile, e.g. a non-uniform lid.
 * hoc_boundary_point_count returns E*K. */
int64_t hoc_boundary_point_count(hoc_t h);
int hoc_get_boundary_points(hoc_t h, int64_t len, double* xy);          /* len = 2 E K */
int hoc_get_wall_velocity(hoc_t h, int64_t len, double* v);             /* len = E K   */
int hoc_set_wall_velocity(hoc_t h, int64_t len, const double* v);

/* Discrete adjoint of the RK4 trajectory (no counterpart in the reference; for gradient-based
 * callers).  Supported: FP64, one instance, problem_type 1 walls with Dirichlet stream
 * function, advection_type 1, time_integration_type 4, affine elements (symmetric Poisson
 * operator); otherwise HOC_EINVAL.
 * hoc_adjoint_forward advances nsteps RK4 steps like hoc_step, keeping on the device what the
 * reverse sweep needs: checkpoint_every >= 1: the vorticity every checkpoint_every steps (the sweep
 * recomputes each segment once); 0: every step's vorticity and stage stream functions when they
 * fit in half the free device memory (no recompute), else binomial checkpointing with the
 * vorticity snapshots that fit; < 0: binomial checkpointing with -checkpoint_every snapshots (the
 * schedule of fewest step advances, each step then recorded once).  hoc_adjoint_backward takes
 * the gradient of an objective with respect to the final vorticity and returns its gradient with
 * respect to the vorticity at the start of that run and (optional, len_wall = E K) to the wall
 * velocity BC_parl_vel; the upwind switches of
 * the advective flux (|dw| > 1e-4) are taken as fixed at their recorded state.  The handle's
 * vorticity, stream function, velocity and time are left as hoc_adjoint_forward left them.
 * Poisson solves of the sweep count in hoc_stats. */
int hoc_adjoint_forward(hoc_t h, int32_t nsteps, int32_t checkpoint_every);
int hoc_adjoint_backward(hoc_t h, int64_t len, const double* wbar_final, double* wbar_initial, int64_t len_wall,
                         double* wall_velocity_bar);
/* Transposed tangent of hoc_stage_residual's map omega -> k at time t: omega_bar =
 * (dk/domega)^T lam and (optional) wall_velocity_bar = (dk/dBC_parl_vel)^T lam. */
int hoc_stage_residual_vjp(hoc_t h, const double* omega, double t, const double* lam, double* omega_bar,
                           double* wall_velocity_bar);

int hoc_stats_get(hoc_t h, hoc_stats* s);

/* Live timing of the fine-level Poisson operator kernel (the dominant kernel): with
 * profiling on, launches are bracketed by HIP events recorded on the handle's own stream, and
 * only launches that did the work are counted
 * (every launch for on == 1, every on-th launch for on > 1: each event pair costs a little
 * device time, so sampling keeps a timed run's throughput intact).  One entry per fused
 * epilogue variant: entries 0-4 FP64, 5-9 FP32. */
#define HOC_PROF_VARIANTS 10
typedef struct {
  int64_t launches[HOC_PROF_VARIANTS];
  double total_ms[HOC_PROF_VARIANTS];
  double bytes_per_launch[HOC_PROF_VARIANTS];  /* algorithmic HBM bytes per launch (average) */
  char names[HOC_PROF_VARIANTS][24];
  int64_t noop_launches;   /* sampled launches left out: predicated off (the iteration after
                              convergence), they returned at once without doing the work */
} hoc_kernel_profile;
int hoc_profile_enable(hoc_t h, int32_t on);   /* on: 0 off, 1 every launch, n > 1 every n-th;
                                                  on != 0 also clears the counters */
int hoc_profile_get(hoc_t h, hoc_kernel_profile* p);

const char* hoc_last_error(hoc_t h);   /* h may be NULL for hoc_create failures */
void hoc_destroy(hoc_t h);

/* ---- CSR linear-solver plug-in ------------------------------------------------------
 * Replaces the reference's BaseLinearSolver interface (old/cxx/linsolvers/linsolver.hpp:42-109:
 * setMaxIterations / setRelativeTolerance / setAbsoluteTolerance, build(num_rows, rowptr,
 * colidx, values), solve(f, u), getNumIterations, getResidual) for the AMGCL solver the
 * reference's Poisson step uses (src/HO_2D.cpp:3011-3086): smoothed-aggregation AMG with
 * spai0 relaxation as the right preconditioner of BiCGSTAB, hierarchy built on the host,
 * iterations on the GPU.  Stops when ||f - A u|| <= max(reltol ||f||, abstol); defaults
 * maxiters 200, reltol 0, abstol 1e-10 (the interface's own defaults). */
typedef struct hoc_linsolver* hoc_ls_t;
int hoc_linsolver_create(hoc_ls_t* ls, int32_t device);
int hoc_linsolver_set_tolerances(hoc_ls_t ls, int32_t maxiters, double reltol, double abstol);
/* CSR with 0-based indices; the matrix is copied. */
int hoc_linsolver_build(hoc_ls_t ls, int32_t num_rows, const int32_t* rowptr, const int32_t* colidx,
                        const double* values);
/* u: initial guess in, solution out (also on HOC_ENOCONV, which reports a solve that ended
 * at maxiters or on breakdown above the tolerance; the reference reports it as Failure). */
int hoc_linsolver_solve(hoc_ls_t ls, const double* f, double* u);
int32_t hoc_linsolver_num_iterations(hoc_ls_t ls);
double hoc_linsolver_residual(hoc_ls_t ls);      /* ||f - A u|| / ||f|| of the last solve */
int32_t hoc_linsolver_levels(hoc_ls_t ls);       /* AMG levels, the system matrix included */
const char* hoc_linsolver_last_error(hoc_ls_t ls);
void hoc_linsolver_destroy(hoc_ls_t ls);

#ifdef __cplusplus
}
#endif
#endif
