"""The no-allocation operator apply (hoc_apply_laplacian_device) against hoc_apply_laplacian
through the C ABI: the same kernels on device buffers, with the face traces in the library's
scratch array.  Two calls in a row on one GPU, the second on a smaller mesh (the array keeps
the first call's contents), the second on a stream of the caller's; a mesh too large for the
array and aliased buffers are rejected.  Device buffers come from the HIP runtime itself."""
import ctypes as C
import os

import numpy as np
import pytest

import hoc
from hoc import meshes

pytestmark = pytest.mark.gpu


def _hip():
    for path in ("libamdhip64.so", os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so")):
        try:
            L = C.CDLL(path)
            break
        except OSError:
            L = None
    if L is None:
        pytest.skip("HIP runtime library not found")
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipFree.argtypes = [C.c_void_p]
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    L.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    L.hipStreamSynchronize.argtypes = [C.c_void_p]
    L.hipStreamDestroy.argtypes = [C.c_void_p]
    return L


class DevBuf:
    def __init__(self, hip, n):
        self.hip, self.n, self.p = hip, n, C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), n * 8) == 0

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == self.n
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, self.n * 8, 1) == 0   # host to device

    def get(self):
        out = np.empty(self.n)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, self.n * 8, 2) == 0   # device to host
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _handle(m, K, problem_type=1):
    g = hoc.HO2D(K=K, problem_type=problem_type, Re=100.0, dt=1e-3)
    g.load_mesh(m)
    g.setup()
    return g


def test_noalloc_apply_matches_host_apply(gpu_available):
    hip = _hip()
    rng = np.random.default_rng(11)
    st = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(st)) == 0
    cases = [(meshes.cavity(24, 2, seed=3, rotate=True), 5, 1, None),    # the handle's stream
             (meshes.backward_facing_step(12, 4, 2), 4, 2, st)]        # a caller's stream, smaller mesh
    sizes = []
    for m, K, pt, stream in cases:
        g = _handle(m, K, pt)
        x = rng.standard_normal(g.D)
        ref = g.apply_laplacian(x)
        xd, yd = DevBuf(hip, g.D), DevBuf(hip, g.D)
        xd.put(x)
        yd.put(np.full(g.D, np.nan))
        g.apply_laplacian_device(xd.p.value, yd.p.value, stream.value if stream is not None else 0)
        assert hip.hipDeviceSynchronize() == 0
        got = yd.get()
        print("no-alloc apply: mesh", m.n_elem, "elements, K", K, "max |diff|", np.abs(got - ref).max())
        assert np.array_equal(got, ref)
        sizes.append(g.D)
        xd.free(); yd.free()
        g.close()
    hip.hipStreamDestroy(st)
    assert sizes[1] < sizes[0]


def test_noalloc_apply_size_limit(gpu_available):
    hip = _hip()
    # 160000 elements x 8 x 6 x 8 bytes = 58.6 MiB of traces: fits the 64 MiB array
    g = _handle(meshes.cavity(400, 2), 6)
    x = np.random.default_rng(1).standard_normal(g.D)
    xd, yd = DevBuf(hip, g.D), DevBuf(hip, g.D)
    xd.put(x)
    g.apply_laplacian_device(xd.p.value, yd.p.value, 0)
    assert hip.hipDeviceSynchronize() == 0
    assert np.array_equal(yd.get(), g.apply_laplacian(x))
    with pytest.raises(hoc.HocError) as e:             # aliased buffers
        g.apply_laplacian_device(xd.p.value, xd.p.value, 0)
    assert e.value.code == hoc.HOC_EINVAL
    xd.free(); yd.free()
    g.close()
    # 176400 x 8 x 6 x 8 bytes = 64.6 MiB: rejected before any work
    g = _handle(meshes.cavity(420, 2), 6)
    xd, yd = DevBuf(hip, g.D), DevBuf(hip, g.D)
    with pytest.raises(hoc.HocError) as e:
        g.apply_laplacian_device(xd.p.value, yd.p.value, 0)
    assert e.value.code == hoc.HOC_EINVAL
    xd.free(); yd.free()
    g.close()
