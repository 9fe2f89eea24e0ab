"""Python host side of the MI355X HO-CXX hot path: a ctypes binding of libhoc.so.

The product path is the HIP library built from ho-cxx_amd/csrc (C ABI in
include/hoc.h).  This module loads the in-tree libhoc.so and fails loudly when
it is missing or when no GPU is present; there is no CPU fallback.

``HO2D`` mirrors the reference's HO_2D setup/advance surface
(src/HO_2D.hpp:220-290): load a mesh (``load_mesh`` ~ load_mesh_arrays_d /
read_msh_file), ``setup`` (~ process_mesh_input), ``step`` (~
solve_advection_diffusion), ``solveto`` (~ solveto_d), ``get_vorticity``
(~ getallvorts_d), ``set_vorticity`` (~ setsolnvort_d).
"""
from __future__ import annotations

import ctypes as C
import os
import numpy as np

from . import meshes  # noqa: F401

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "libhoc.so")

HOC_OK, HOC_EINVAL, HOC_ESTATE, HOC_ENOCONV, HOC_ENAN, HOC_EDEVICE, HOC_EMESH = 0, -1, -2, -3, -4, -5, -6
HOC_FP64, HOC_MIXED = 0, 1
KRYLOV_AUTO, KRYLOV_BICGSTAB, KRYLOV_PCG = 0, 1, 2

EXPORTED = [
    "hoc_default_config", "hoc_create", "hoc_load_mesh", "hoc_load_gmsh", "hoc_write_vtk", "hoc_setup",
    "hoc_get_node_fields", "hoc_write_smooth_vtk", "hoc_set_sample_points", "hoc_get_sample_location",
    "hoc_get_sample_values", "hoc_write_samples",
    "hoc_set_instances", "hoc_get_instance_times", "hoc_set_boundary_values", "hoc_open_point_count",
    "hoc_get_open_points", "hoc_set_open_velocities", "hoc_set_open_vorticity", "hoc_set_vorticity_target",
    "hoc_set_p2g_weights", "hoc_set_vorticity", "hoc_get_vorticity",
    "hoc_get_streamfunction", "hoc_set_streamfunction", "hoc_get_velocity", "hoc_get_solution_points",
    "hoc_get_solution_areas", "hoc_get_time", "hoc_set_time", "hoc_step", "hoc_solveto", "hoc_poisson",
    "hoc_stage_residual", "hoc_apply_laplacian", "hoc_apply_laplacian_device", "hoc_boundary_point_count", "hoc_get_boundary_points",
    "hoc_get_wall_velocity", "hoc_set_wall_velocity", "hoc_adjoint_forward", "hoc_adjoint_backward",
    "hoc_stage_residual_vjp", "hoc_stats_get", "hoc_profile_enable", "hoc_profile_get",
    "hoc_last_error", "hoc_destroy",
    "hoc_linsolver_create", "hoc_linsolver_set_tolerances", "hoc_linsolver_build", "hoc_linsolver_solve",
    "hoc_linsolver_num_iterations", "hoc_linsolver_residual", "hoc_linsolver_levels", "hoc_linsolver_last_error",
    "hoc_linsolver_destroy",
]
PROF_VARIANTS = 10


class hoc_config(C.Structure):
    _fields_ = [("K", C.c_int32), ("problem_type", C.c_int32), ("Re", C.c_double), ("dt", C.c_double),
                ("time_integration_type", C.c_int32), ("advection_type", C.c_int32), ("tol", C.c_double),
                ("maxiter", C.c_int32), ("precision", C.c_int32), ("krylov", C.c_int32), ("device", C.c_int32),
                ("fail_on_noconv", C.c_int32), ("mixed_inner_tol", C.c_double), ("cheb_degree", C.c_int32),
                ("psi_guess", C.c_int32), ("cheb_lower", C.c_double), ("launch_mode", C.c_int32)]


class hoc_stats(C.Structure):
    _fields_ = [("n_elem", C.c_int64), ("n_dofs", C.c_int64), ("n_bedges", C.c_int64), ("n_levels", C.c_int32),
                ("krylov", C.c_int32), ("last_iters", C.c_int32), ("last_rel_residual", C.c_double),
                ("total_iters", C.c_int64), ("n_solves", C.c_int64), ("setup_seconds", C.c_double),
                ("fine_op_bytes", C.c_double), ("last_refinements", C.c_int32), ("precision", C.c_int32),
                ("n_nodes", C.c_int64), ("launch_graphs", C.c_int32), ("launch_tune_ratio", C.c_double)]


class hoc_kernel_profile(C.Structure):
    _fields_ = [("launches", C.c_int64 * PROF_VARIANTS), ("total_ms", C.c_double * PROF_VARIANTS),
                ("bytes_per_launch", C.c_double * PROF_VARIANTS), ("names", (C.c_char * 24) * PROF_VARIANTS),
                ("noop_launches", C.c_int64)]


class HocError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"hoc error {code}: {msg}")
        self.code = code


_LIB = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def lib():
    """Load libhoc.so (raises when it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() or make -C ho-cxx_amd")
        L = C.CDLL(LIB_PATH)
        L.hoc_default_config.argtypes = [C.POINTER(hoc_config)]
        L.hoc_default_config.restype = None
        L.hoc_create.argtypes = [C.POINTER(C.c_void_p), C.POINTER(hoc_config)]
        L.hoc_load_mesh.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _dp, C.c_int32, _ip, C.c_int32, _ip, _ip,
                                    C.c_int32, C.c_char_p]
        L.hoc_setup.argtypes = [C.c_void_p]
        L.hoc_load_gmsh.argtypes = [C.c_void_p, C.c_char_p, C.c_int32]
        L.hoc_set_instances.argtypes = [C.c_void_p, C.c_int32, _ip, _dp, _dp]
        L.hoc_set_boundary_values.argtypes = [C.c_void_p, C.c_int32, _ip, _dp]
        L.hoc_open_point_count.argtypes = [C.c_void_p]
        L.hoc_open_point_count.restype = C.c_int64
        for nm in ("hoc_get_open_points", "hoc_set_open_velocities", "hoc_set_open_vorticity",
                   "hoc_set_vorticity_target", "hoc_set_p2g_weights"):
            getattr(L, nm).argtypes = [C.c_void_p, C.c_int64, _dp]
        L.hoc_get_instance_times.argtypes = [C.c_void_p, C.c_int32, _dp]
        L.hoc_write_vtk.argtypes = [C.c_void_p, C.c_char_p]
        L.hoc_get_node_fields.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _dp]
        L.hoc_write_smooth_vtk.argtypes = [C.c_void_p, C.c_int32, C.c_char_p]
        L.hoc_set_sample_points.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
        L.hoc_get_sample_location.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]
        L.hoc_get_sample_values.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p]
        L.hoc_write_samples.argtypes = [C.c_void_p, C.c_int32, C.c_char_p]
        for nm in ("hoc_set_vorticity", "hoc_get_vorticity", "hoc_get_streamfunction", "hoc_set_streamfunction",
                   "hoc_get_velocity", "hoc_get_solution_points", "hoc_get_solution_areas"):
            getattr(L, nm).argtypes = [C.c_void_p, C.c_int64, _dp]
        L.hoc_get_time.argtypes = [C.c_void_p]
        L.hoc_get_time.restype = C.c_double
        L.hoc_set_time.argtypes = [C.c_void_p, C.c_double]
        L.hoc_step.argtypes = [C.c_void_p, C.c_int32]
        L.hoc_solveto.argtypes = [C.c_void_p, C.c_double, C.c_int32, C.c_int32, C.c_double]
        L.hoc_poisson.argtypes = [C.c_void_p, _dp, C.c_double, _dp]
        L.hoc_stage_residual.argtypes = [C.c_void_p, _dp, C.c_double, _dp]
        L.hoc_apply_laplacian.argtypes = [C.c_void_p, _dp, _dp]
        L.hoc_apply_laplacian_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.hoc_boundary_point_count.argtypes = [C.c_void_p]
        L.hoc_boundary_point_count.restype = C.c_int64
        for nm in ("hoc_get_boundary_points", "hoc_get_wall_velocity", "hoc_set_wall_velocity"):
            getattr(L, nm).argtypes = [C.c_void_p, C.c_int64, C.c_void_p]
        L.hoc_adjoint_forward.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
        L.hoc_adjoint_backward.argtypes = [C.c_void_p, C.c_int64, _dp, _dp, C.c_int64, C.c_void_p]
        L.hoc_stage_residual_vjp.argtypes = [C.c_void_p, _dp, C.c_double, _dp, _dp, C.c_void_p]
        L.hoc_stats_get.argtypes = [C.c_void_p, C.POINTER(hoc_stats)]
        L.hoc_profile_enable.argtypes = [C.c_void_p, C.c_int32]
        L.hoc_profile_get.argtypes = [C.c_void_p, C.POINTER(hoc_kernel_profile)]
        L.hoc_last_error.argtypes = [C.c_void_p]
        L.hoc_last_error.restype = C.c_char_p
        L.hoc_destroy.argtypes = [C.c_void_p]
        L.hoc_destroy.restype = None
        L.hoc_linsolver_create.argtypes = [C.POINTER(C.c_void_p), C.c_int32]
        L.hoc_linsolver_set_tolerances.argtypes = [C.c_void_p, C.c_int32, C.c_double, C.c_double]
        L.hoc_linsolver_build.argtypes = [C.c_void_p, C.c_int32, _ip, _ip, _dp]
        L.hoc_linsolver_solve.argtypes = [C.c_void_p, _dp, _dp]
        for nm, rt in (("hoc_linsolver_num_iterations", C.c_int32), ("hoc_linsolver_residual", C.c_double),
                       ("hoc_linsolver_levels", C.c_int32), ("hoc_linsolver_last_error", C.c_char_p)):
            getattr(L, nm).argtypes = [C.c_void_p]
            getattr(L, nm).restype = rt
        L.hoc_linsolver_destroy.argtypes = [C.c_void_p]
        L.hoc_linsolver_destroy.restype = None
        _LIB = L
    return _LIB


def default_config(**kw) -> hoc_config:
    c = hoc_config()
    lib().hoc_default_config(C.byref(c))
    for k, v in kw.items():
        if not hasattr(c, k):
            raise KeyError(k)
        setattr(c, k, v)
    return c


class HO2D:
    """One problem on one GPU; mirrors the reference HO_2D call surface."""

    def __init__(self, K=3, problem_type=1, Re=100.0, dt=1e-3, time_integration_type=4, advection_type=1,
                 tol=1e-15, maxiter=150, precision=HOC_FP64, krylov=KRYLOV_AUTO, device=0, fail_on_noconv=False,
                 mixed_inner_tol=0.0, cheb_degree=0, cheb_lower=0.0, psi_guess=0, launch_mode=0):
        self.cfg = default_config(K=K, problem_type=problem_type, Re=Re, dt=dt,
                                  time_integration_type=time_integration_type, advection_type=advection_type,
                                  tol=tol, maxiter=maxiter, precision=precision, krylov=krylov, device=device,
                                  fail_on_noconv=int(fail_on_noconv), mixed_inner_tol=mixed_inner_tol,
                                  cheb_degree=cheb_degree, cheb_lower=cheb_lower, psi_guess=psi_guess,
                                  launch_mode=launch_mode)
        self.h = C.c_void_p()
        rc = lib().hoc_create(C.byref(self.h), C.byref(self.cfg))
        if rc != HOC_OK:
            raise HocError(rc, lib().hoc_last_error(None).decode())
        self.K = K
        self.D = 0

    def _chk(self, rc):
        if rc != HOC_OK:
            raise HocError(rc, lib().hoc_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().hoc_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load_mesh(self, m):
        self._chk(lib().hoc_load_mesh(self.h, m.lnod, m.n_nodes, np.ascontiguousarray(m.xy, dtype=np.float64).ravel(),
                                      m.n_elem, np.ascontiguousarray(m.elem, dtype=np.int32).ravel(),
                                      m.n_bedges, np.ascontiguousarray(m.bedges, dtype=np.int32).ravel(),
                                      np.ascontiguousarray(m.bedge_bdry, dtype=np.int32), len(m.bnames),
                                      "\n".join(m.bnames).encode()))
        self.D = m.n_elem * self.K * self.K

    def load_gmsh(self, path, occ=0):
        """Read a Gmsh 4.1 ASCII mesh in the library (= Mesh::read_msh_file)."""
        self._chk(lib().hoc_load_gmsh(self.h, os.fsencode(path), int(occ)))

    def set_instances(self, elem_offsets, Re, dt):
        """Ensemble of independent meshes (see meshes.concat): per-instance Re and dt."""
        off = np.ascontiguousarray(elem_offsets, dtype=np.int32)
        self._chk(lib().hoc_set_instances(self.h, len(off) - 1, off, np.ascontiguousarray(Re, dtype=np.float64),
                                          np.ascontiguousarray(dt, dtype=np.float64)))
        self.n_inst = len(off) - 1

    # ---- hybrid driver surface (problem_type 11) ----
    def set_boundary_values(self, nval, vals):
        """Per boundary edge: nval 0, 1 (normal velocity in vals[e, 0]) or 2 (tangential, normal)."""
        nv = np.ascontiguousarray(nval, dtype=np.int32)
        vv = np.ascontiguousarray(np.reshape(vals, (len(nv), 2)), dtype=np.float64).ravel()
        self._chk(lib().hoc_set_boundary_values(self.h, len(nv), nv, vv))

    def open_points(self):
        n = lib().hoc_open_point_count(self.h)
        return self._get("hoc_get_open_points", 2 * n).reshape(-1, 2)

    def _put(self, fn, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        self._chk(getattr(lib(), fn)(self.h, a.size, a))

    def set_open_velocities(self, uv):
        self._put("hoc_set_open_velocities", uv)

    def set_open_vorticity(self, w):
        self._put("hoc_set_open_vorticity", w)

    def set_vorticity_target(self, w):
        self._put("hoc_set_vorticity_target", w)

    def set_p2g_weights(self, w):
        self._put("hoc_set_p2g_weights", w)

    def instance_times(self):
        n = getattr(self, "n_inst", 1)
        out = np.zeros(n)
        self._chk(lib().hoc_get_instance_times(self.h, n, out))
        return out

    def write_vtk(self, path):
        self._chk(lib().hoc_write_vtk(self.h, os.fsencode(path)))

    # ---- node-averaged output and sample points (save_smooth_vtk, read_process_sample_points)
    def node_fields(self, refresh=False, n_nodes=None):
        """[n_nodes, 4] = vorticity, psi, u, v averaged at the mesh nodes.  refresh: recompute psi
        and the velocity for the current vorticity first (as before a dump in the reference's
        main loop)."""
        n = int(self.stats()["n_nodes"] if n_nodes is None else n_nodes)
        out = np.zeros(4 * n)
        self._chk(lib().hoc_get_node_fields(self.h, int(bool(refresh)), n, out))
        return out.reshape(n, 4)

    def write_smooth_vtk(self, path, refresh=False):
        self._chk(lib().hoc_write_smooth_vtk(self.h, int(bool(refresh)), os.fsencode(path)))

    def set_sample_points(self, xy):
        xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1, 2)
        self._nsp = len(xy)
        self._chk(lib().hoc_set_sample_points(self.h, self._nsp, xy.ctypes.data if self._nsp else None))

    def load_sample_points(self, path):
        """Reads "x y" pairs as read_process_sample_points does (src/HO_2D.cpp:323-336): a last pair
        within 1e-6 of the previous one is dropped."""
        v = np.array(open(path).read().split(), dtype=np.float64)
        xy = v[: len(v) // 2 * 2].reshape(-1, 2)
        if len(xy) >= 2 and np.hypot(*(xy[-1] - xy[-2])) < 1e-6:
            xy = xy[:-1]
        self.set_sample_points(xy)
        return xy

    def sample_location(self):
        n = getattr(self, "_nsp", 0)
        el = np.zeros(n, dtype=np.int32)
        loc = np.zeros(2 * n)
        self._chk(lib().hoc_get_sample_location(self.h, n, el.ctypes.data if n else None,
                                                loc.ctypes.data if n else None))
        return el, loc.reshape(-1, 2)

    def sample_values(self, refresh=False):
        """[n, 3] = vorticity, u, v at the sample points."""
        n = getattr(self, "_nsp", 0)
        out = np.zeros(3 * n)
        self._chk(lib().hoc_get_sample_values(self.h, int(bool(refresh)), n, out.ctypes.data if n else None))
        return out.reshape(-1, 3)

    def write_samples(self, path, refresh=False):
        self._chk(lib().hoc_write_samples(self.h, int(bool(refresh)), os.fsencode(path)))

    def setup(self):
        self._chk(lib().hoc_setup(self.h))
        self.D = int(self.stats()["n_dofs"])

    def _get(self, fn, n):
        out = np.zeros(n)
        self._chk(getattr(lib(), fn)(self.h, n, out))
        return out

    def get_vorticity(self):
        return self._get("hoc_get_vorticity", self.D)

    def set_vorticity(self, w):
        w = np.ascontiguousarray(w, dtype=np.float64)
        self._chk(lib().hoc_set_vorticity(self.h, w.size, w))

    def get_streamfunction(self):
        return self._get("hoc_get_streamfunction", self.D)

    def set_streamfunction(self, p):
        p = np.ascontiguousarray(p, dtype=np.float64)
        self._chk(lib().hoc_set_streamfunction(self.h, p.size, p))

    def get_velocity(self):
        return self._get("hoc_get_velocity", 2 * self.D).reshape(-1, 2)

    def solution_points(self):
        return self._get("hoc_get_solution_points", 2 * self.D).reshape(-1, 2)

    def solution_areas(self):
        return self._get("hoc_get_solution_areas", self.D)

    @property
    def time(self):
        return lib().hoc_get_time(self.h)

    def set_time(self, t):
        """Set the clock (every instance's clock for an ensemble)."""
        self._chk(lib().hoc_set_time(self.h, float(t)))

    def step(self, n=1):
        self._chk(lib().hoc_step(self.h, n))

    def solveto(self, outer_dt, nsub, integ, Re):
        self._chk(lib().hoc_solveto(self.h, outer_dt, nsub, integ, Re))

    def poisson(self, w, t=0.0):
        out = np.zeros(self.D)
        self._chk(lib().hoc_poisson(self.h, np.ascontiguousarray(w, dtype=np.float64), t, out))
        return out

    def stage(self, w, t=0.0):
        out = np.zeros(self.D)
        self._chk(lib().hoc_stage_residual(self.h, np.ascontiguousarray(w, dtype=np.float64), t, out))
        return out

    # ---- wall velocity and the discrete adjoint ------------------------------------------------
    def boundary_points(self):
        """[E*K, 2] boundary flux points in the edge flux-point order of the boundary arrays."""
        n = int(lib().hoc_boundary_point_count(self.h))
        out = np.zeros(2 * n)
        self._chk(lib().hoc_get_boundary_points(self.h, 2 * n, out.ctypes.data if n else None))
        return out.reshape(-1, 2)

    def wall_velocity(self):
        """BC_parl_vel at the boundary flux points [E*K] (-1 along the lid of the cavity)."""
        n = int(lib().hoc_boundary_point_count(self.h))
        out = np.zeros(n)
        self._chk(lib().hoc_get_wall_velocity(self.h, n, out.ctypes.data if n else None))
        return out

    def set_wall_velocity(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        self._chk(lib().hoc_set_wall_velocity(self.h, v.size, v.ctypes.data if v.size else None))

    def adjoint_forward(self, nsteps, checkpoint_every):
        """Advance nsteps RK4 steps keeping checkpoints for adjoint_backward."""
        self._chk(lib().hoc_adjoint_forward(self.h, int(nsteps), int(checkpoint_every)))

    def adjoint_backward(self, wbar_final):
        """Gradient of an objective w.r.t. the vorticity and the wall velocity at the start of the
        last adjoint_forward run, from its gradient w.r.t. the final vorticity."""
        wb = np.ascontiguousarray(wbar_final, dtype=np.float64)
        out = np.zeros(self.D)
        n = int(lib().hoc_boundary_point_count(self.h))
        pv = np.zeros(n)
        self._chk(lib().hoc_adjoint_backward(self.h, wb.size, wb, out, n, pv.ctypes.data if n else None))
        return out, pv

    def stage_vjp(self, w, lam, t=0.0):
        """(dk/domega)^T lam and (dk/dwall_velocity)^T lam of stage(w, t)."""
        out = np.zeros(self.D)
        n = int(lib().hoc_boundary_point_count(self.h))
        pv = np.zeros(n)
        self._chk(lib().hoc_stage_residual_vjp(self.h, np.ascontiguousarray(w, dtype=np.float64), t,
                                               np.ascontiguousarray(lam, dtype=np.float64), out,
                                               pv.ctypes.data if n else None))
        return out, pv

    def apply_laplacian(self, x):
        out = np.zeros(self.D)
        self._chk(lib().hoc_apply_laplacian(self.h, np.ascontiguousarray(x, dtype=np.float64), out))
        return out

    def apply_laplacian_device(self, x_ptr, y_ptr, stream=0):
        """y = A x for device buffers (raw addresses of D doubles each); allocates nothing and
        returns without waiting on `stream` (0 = the handle's stream)."""
        self._chk(lib().hoc_apply_laplacian_device(self.h, C.c_void_p(x_ptr), C.c_void_p(y_ptr),
                                                   C.c_void_p(stream or None)))

    def stats(self):
        s = hoc_stats()
        self._chk(lib().hoc_stats_get(self.h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in s._fields_}

    def profile(self, on=True, every=1):
        """Turn live per-launch timing of the fine-level operator kernel on/off; every > 1 times
        every every-th launch only."""
        self._chk(lib().hoc_profile_enable(self.h, max(1, int(every)) if on else 0))

    def profile_read(self):
        p = hoc_kernel_profile()
        self._chk(lib().hoc_profile_get(self.h, C.byref(p)))
        out = {}
        for v in range(PROF_VARIANTS):
            out[p.names[v].value.decode()] = {"launches": p.launches[v], "total_ms": p.total_ms[v],
                                              "bytes_per_launch": p.bytes_per_launch[v]}
        self.profile_noop_launches = int(p.noop_launches)
        return out


class LinearSolver:
    """GPU plug-in with the shape of the reference's BaseLinearSolver
    (old/cxx/linsolvers/linsolver.hpp:42-109): build(num_rows, rowptr, colidx, values),
    then solve(f, u) with u the initial guess; SA-AMG (spai0) preconditioned BiCGSTAB as the
    reference's AMGCL configuration (src/HO_2D.cpp:3011-3086)."""

    def __init__(self, device=0, maxiters=200, reltol=0.0, abstol=1e-10):
        self.h = C.c_void_p()
        rc = lib().hoc_linsolver_create(C.byref(self.h), device)
        if rc != HOC_OK:
            raise HocError(rc, lib().hoc_linsolver_last_error(None).decode())
        self.set_tolerances(maxiters, reltol, abstol)
        self.n = 0

    def _chk(self, rc):
        if rc != HOC_OK:
            raise HocError(rc, lib().hoc_linsolver_last_error(self.h).decode())

    def set_tolerances(self, maxiters, reltol, abstol):
        self._chk(lib().hoc_linsolver_set_tolerances(self.h, maxiters, reltol, abstol))

    def build(self, num_rows, rowptr, colidx, values):
        self._chk(lib().hoc_linsolver_build(self.h, num_rows, np.ascontiguousarray(rowptr, dtype=np.int32),
                                            np.ascontiguousarray(colidx, dtype=np.int32),
                                            np.ascontiguousarray(values, dtype=np.float64)))
        self.n = num_rows

    def solve(self, f, u=None):
        """Returns the solution; raises HocError(HOC_ENOCONV) when the tolerance was not met."""
        out = np.zeros(self.n) if u is None else np.array(u, dtype=np.float64, copy=True)
        self._chk(lib().hoc_linsolver_solve(self.h, np.ascontiguousarray(f, dtype=np.float64), out))
        return out

    @property
    def num_iterations(self):
        return lib().hoc_linsolver_num_iterations(self.h)

    @property
    def residual(self):
        return lib().hoc_linsolver_residual(self.h)

    @property
    def levels(self):
        return lib().hoc_linsolver_levels(self.h)

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().hoc_linsolver_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
