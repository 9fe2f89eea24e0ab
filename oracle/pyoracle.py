"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (run make -C oracle)")
        L = C.CDLL(path)
        L.ho_oracle_create.restype = C.c_void_p
        L.ho_oracle_create.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.c_int, _ip, C.c_int, _ip, _ip, C.c_int,
                                       C.c_char_p, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                                       C.c_double, C.c_int]
        L.ho_oracle_destroy.argtypes = [C.c_void_p]
        L.ho_oracle_last_error.restype = C.c_char_p
        L.ho_oracle_last_error.argtypes = [C.c_void_p]
        for nm in ("ho_oracle_dofs", "ho_oracle_n_bedges"):
            getattr(L, nm).argtypes = [C.c_void_p]
            getattr(L, nm).restype = C.c_int
        for nm in ("ho_oracle_set_vorticity", "ho_oracle_get_vorticity", "ho_oracle_get_streamfunction",
                   "ho_oracle_set_streamfunction", "ho_oracle_get_velocity", "ho_oracle_get_boundary_source"):
            getattr(L, nm).argtypes = [C.c_void_p, _dp]
        L.ho_oracle_time.argtypes = [C.c_void_p]
        L.ho_oracle_time.restype = C.c_double
        L.ho_oracle_step.argtypes = [C.c_void_p, C.c_int]
        L.ho_oracle_stage.argtypes = [C.c_void_p, _dp, C.c_double, _dp]
        L.ho_oracle_poisson.argtypes = [C.c_void_p, _dp, C.c_double, _dp]
        L.ho_oracle_last_solve.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.ho_oracle_solve_counts.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.ho_oracle_nnz.argtypes = [C.c_void_p]
        L.ho_oracle_nnz.restype = C.c_int64
        L.ho_oracle_get_csr.argtypes = [C.c_void_p, _ip, _ip, _dp]
        L.ho_oracle_table.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p, C.c_int64]
        L.ho_oracle_table.restype = C.c_int64
        L.ho_oracle_apply_laplacian_mf.argtypes = [C.c_void_p, _dp, _dp]
        L.ho_oracle_set_bc.argtypes = [C.c_void_p, C.c_char_p, _dp]
        L.ho_oracle_apply_laplacian_bc.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.ho_oracle_solve_poisson_bc.argtypes = [C.c_void_p, _dp, _dp]
        L.ho_oracle_velocity_from_psi.argtypes = [C.c_void_p, _dp]
        L.ho_oracle_create_ex.restype = C.c_void_p
        L.ho_oracle_create_ex.argtypes = L.ho_oracle_create.argtypes + [_ip, _dp]
        L.ho_oracle_open_points.argtypes = [C.c_void_p, C.c_void_p]
        for nm in ("ho_oracle_set_open_vels", "ho_oracle_set_open_vort", "ho_oracle_set_soln_vort",
                   "ho_oracle_set_p2g_weights"):
            getattr(L, nm).argtypes = [C.c_void_p, _dp]
        L.ho_oracle_solveto.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_double]
        L.ho_oracle_n_nodes.argtypes = [C.c_void_p]
        L.ho_oracle_n_nodes.restype = C.c_int
        L.ho_oracle_smooth_fields.argtypes = [C.c_void_p, _dp]
        L.ho_oracle_set_sample_points.argtypes = [C.c_void_p, C.c_int, _dp]
        L.ho_oracle_sample_location.argtypes = [C.c_void_p, _ip, _dp]
        L.ho_oracle_sample_values.argtypes = [C.c_void_p, _dp]
        L.ho_oracle_refresh.argtypes = [C.c_void_p]
        L.ho_oracle_set_time.argtypes = [C.c_void_p, C.c_double]
        L.ho_oracle_update_bcs.argtypes = [C.c_void_p, C.c_double]
        L.ho_oracle_advection.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp]
        L.ho_oracle_set_threads.argtypes = [C.c_int]
        L.ho_oracle_set_threads.restype = C.c_int
        L.ho_oracle_set_threads(1)   # deterministic single-thread checker unless asked otherwise
        _LIB = L
    return _LIB


def set_threads(n):
    """Threads of the oracle's OpenMP linear algebra; returns the previous count."""
    return lib().ho_oracle_set_threads(int(n))


class Oracle:
    """FP64 CPU restatement of the HO_2D hot path for one mesh/problem."""

    def __init__(self, mesh, K: int, problem_type: int = 1, Re: float = 100.0, dt: float = 1e-3,
                 time_integration_type: int = 4, advection_type: int = 1, tol: float = 1e-15,
                 maxiter: int = 150, edge_nval=None, edge_vals=None):
        """edge_nval / edge_vals: values carried by the boundary edges (hybrid inlets and
        outlets): per edge 0, 1 (normal velocity) or 2 (tangential, normal) values in
        edge_vals[e, 0:2]."""
        L = lib()
        self.mesh, self.K = mesh, K
        names = "\n".join(mesh.bnames).encode()
        args = (K, mesh.lnod, mesh.n_nodes, np.ascontiguousarray(mesh.xy, dtype=np.float64).ravel(),
                mesh.n_elem, np.ascontiguousarray(mesh.elem, dtype=np.int32).ravel(),
                mesh.n_bedges, np.ascontiguousarray(mesh.bedges, dtype=np.int32).ravel(),
                np.ascontiguousarray(mesh.bedge_bdry, dtype=np.int32),
                len(mesh.bnames), names, problem_type, Re, dt, time_integration_type, advection_type, tol, maxiter)
        if edge_nval is None:
            self.h = L.ho_oracle_create(*args)
        else:
            nv = np.ascontiguousarray(edge_nval, dtype=np.int32)
            vv = np.ascontiguousarray(np.reshape(edge_vals, (mesh.n_bedges, 2)), dtype=np.float64).ravel()
            self.h = L.ho_oracle_create_ex(*args, nv, vv)
        if not self.h:
            raise RuntimeError("oracle create failed: " + L.ho_oracle_last_error(None).decode())
        self.D = L.ho_oracle_dofs(self.h)
        self.n_bedges = L.ho_oracle_n_bedges(self.h)

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.ho_oracle_destroy(self.h)
            self.h = None

    def _chk(self, rc):
        if rc != 0:
            raise RuntimeError(lib().ho_oracle_last_error(self.h).decode())

    # hybrid driver surface (problem_type 11)
    def open_points(self):
        n = lib().ho_oracle_open_points(self.h, None)
        out = np.zeros(2 * n)
        lib().ho_oracle_open_points(self.h, out.ctypes.data)
        return out.reshape(-1, 2)

    def set_open_vels(self, xyvel):
        self._chk(lib().ho_oracle_set_open_vels(self.h, np.ascontiguousarray(xyvel, dtype=np.float64).ravel()))

    def set_open_vort(self, w):
        self._chk(lib().ho_oracle_set_open_vort(self.h, np.ascontiguousarray(w, dtype=np.float64)))

    def set_soln_vort(self, w):
        self._chk(lib().ho_oracle_set_soln_vort(self.h, np.ascontiguousarray(w, dtype=np.float64)))

    def set_p2g_weights(self, w):
        self._chk(lib().ho_oracle_set_p2g_weights(self.h, np.ascontiguousarray(w, dtype=np.float64)))

    def solveto(self, outer_dt, nstep, integ, Re):
        self._chk(lib().ho_oracle_solveto(self.h, outer_dt, nstep, integ, Re))

    def vorticity(self):
        out = np.zeros(self.D); lib().ho_oracle_get_vorticity(self.h, out); return out

    def set_vorticity(self, w):
        lib().ho_oracle_set_vorticity(self.h, np.ascontiguousarray(w, dtype=np.float64))

    def streamfunction(self):
        out = np.zeros(self.D); lib().ho_oracle_get_streamfunction(self.h, out); return out

    def set_streamfunction(self, p):
        lib().ho_oracle_set_streamfunction(self.h, np.ascontiguousarray(p, dtype=np.float64))

    def velocity(self):
        out = np.zeros(2 * self.D); lib().ho_oracle_get_velocity(self.h, out); return out.reshape(-1, 2)

    def time(self):
        return lib().ho_oracle_time(self.h)

    def set_time(self, t):
        self._chk(lib().ho_oracle_set_time(self.h, float(t)))

    def update_bcs(self, t):
        """update_BCs(t) alone, so the boundary tables can be read."""
        self._chk(lib().ho_oracle_update_bcs(self.h, float(t)))

    def advection(self, w, uv, bc_vort, bc_nvel):
        """Advective residual of w (the configured advection type) with a prescribed velocity uv
        [dofs, 2] and boundary values BC_vorticity / BC_normal_vel ([n_bedges, K], edge order)."""
        out = np.zeros(self.D)
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64).ravel()
        self._chk(lib().ho_oracle_advection(self.h, f(w), f(uv), f(bc_vort), f(bc_nvel), out))
        return out

    def step(self, n=1):
        self._chk(lib().ho_oracle_step(self.h, n))

    def stage(self, w, t=0.0):
        out = np.zeros(self.D)
        self._chk(lib().ho_oracle_stage(self.h, np.ascontiguousarray(w, dtype=np.float64), t, out))
        return out

    def poisson(self, w, t=0.0):
        out = np.zeros(self.D)
        self._chk(lib().ho_oracle_poisson(self.h, np.ascontiguousarray(w, dtype=np.float64), t, out))
        return out

    def last_solve(self):
        it, rel = C.c_int(), C.c_double()
        lib().ho_oracle_last_solve(self.h, C.byref(it), C.byref(rel))
        return it.value, rel.value

    def solve_counts(self):
        """(Poisson solves so far, their summed Krylov iterations)."""
        n, it = C.c_int64(), C.c_int64()
        lib().ho_oracle_solve_counts(self.h, C.byref(n), C.byref(it))
        return n.value, it.value

    def csr(self):
        L = lib()
        nnz = L.ho_oracle_nnz(self.h)
        ptr = np.zeros(self.D + 1, dtype=np.int32)
        col = np.zeros(nnz, dtype=np.int32)
        val = np.zeros(nnz)
        L.ho_oracle_get_csr(self.h, ptr, col, val)
        return ptr, col, val

    def boundary_source(self):
        out = np.zeros(self.n_bedges * self.K ** 3)
        lib().ho_oracle_get_boundary_source(self.h, out)
        return out.reshape(self.n_bedges, self.K * self.K, self.K)

    def table(self, name):
        L = lib()
        n = L.ho_oracle_table(self.h, name.encode(), None, 0)
        if n < 0:
            raise KeyError(name)
        out = np.zeros(max(n, 1))
        L.ho_oracle_table(self.h, name.encode(), out.ctypes.data, n)
        return out[:n]

    def laplacian_mf(self, x):
        y = np.zeros(self.D)
        self._chk(lib().ho_oracle_apply_laplacian_mf(self.h, np.ascontiguousarray(x, dtype=np.float64), y))
        return y

    def set_bc(self, name, vals):
        vals = np.ascontiguousarray(vals, dtype=np.float64).ravel()
        if vals.size != self.n_bedges * self.K:
            raise ValueError("need n_bedges*K values")
        self._chk(lib().ho_oracle_set_bc(self.h, name.encode(), vals))

    def laplacian_bc(self, x, bcvals):
        y = np.zeros(self.D)
        self._chk(lib().ho_oracle_apply_laplacian_bc(self.h, np.ascontiguousarray(x, dtype=np.float64),
                                                     np.ascontiguousarray(bcvals, dtype=np.float64).ravel(), y))
        return y

    def solve_poisson_bc(self, w):
        out = np.zeros(self.D)
        self._chk(lib().ho_oracle_solve_poisson_bc(self.h, np.ascontiguousarray(w, dtype=np.float64), out))
        return out

    def velocity_from_psi(self, psi):
        self.set_streamfunction(psi)
        out = np.zeros(2 * self.D)
        self._chk(lib().ho_oracle_velocity_from_psi(self.h, out))
        return out.reshape(-1, 2)

    def refresh(self):
        """Boundary values, stream function and velocity for the current vorticity (what the
        reference's main loop runs before each dump)."""
        self._chk(lib().ho_oracle_refresh(self.h))

    def smooth_fields(self):
        """Node-averaged [n_nodes, 4] = vorticity, stream function, u, v."""
        n = lib().ho_oracle_n_nodes(self.h)
        out = np.zeros(4 * n)
        self._chk(lib().ho_oracle_smooth_fields(self.h, out))
        return out.reshape(n, 4)

    def set_sample_points(self, xy):
        xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1, 2)
        self._nsp = len(xy)
        self._chk(lib().ho_oracle_set_sample_points(self.h, self._nsp, xy.ravel()))

    def sample_location(self):
        el = np.zeros(self._nsp, dtype=np.int32)
        loc = np.zeros(2 * self._nsp)
        self._chk(lib().ho_oracle_sample_location(self.h, el, loc))
        return el, loc.reshape(-1, 2)

    def sample_values(self):
        out = np.zeros(3 * self._nsp)
        self._chk(lib().ho_oracle_sample_values(self.h, out))
        return out.reshape(-1, 3)
