"""The drop-in boundary: libhoc.so loads, exports every entry point include/hoc.h
declares, and fails loudly (a status code, no crash, no CPU fallback) where no GPU
is present.  CPU only."""
import ctypes as C
import os
import re

import pytest

import hoc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "hoc.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(hoc_[a-z0-9_]+)\s*\(", src)))


def test_header_lists_the_binding():
    assert set(declared_functions()) == set(hoc.EXPORTED)


def test_library_exports_every_declared_symbol():
    L = C.CDLL(hoc.LIB_PATH)
    missing = [f for f in declared_functions() if not hasattr(L, f)]
    assert not missing, missing


def test_config_defaults():
    c = hoc.default_config()
    assert c.K == 3 and c.time_integration_type == 2 and c.advection_type == 2   # set_defaults
    assert c.tol == 1e-15 and c.maxiter == 150                                   # :3067-3069


def test_struct_layouts_match_header(tmp_path):
    """The ctypes mirrors of hoc_config / hoc_stats agree with the C compiler's layout."""
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{os.path.join(ROOT, "include", "hoc.h")}"',
             "int main(void) {"]
    for st, cls in (("hoc_config", hoc.hoc_config), ("hoc_stats", hoc.hoc_stats)):
        lines.append(f'printf("%zu\\n", sizeof({st}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("%zu\\n", offsetof({st}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = []
    for cls in (hoc.hoc_config, hoc.hoc_stats):
        want.append(C.sizeof(cls))
        want += [getattr(cls, f).offset for f, _ in cls._fields_]
    assert got == want


def test_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(hoc.HocError) as e:
        hoc.HO2D(K=4)
    assert e.value.code == hoc.HOC_EDEVICE


def test_rejects_bad_config():
    for kw in ({"K": 9}, {"time_integration_type": 3}, {"advection_type": 0}, {"problem_type": 11}):
        with pytest.raises(hoc.HocError) as e:
            hoc.HO2D(**kw)
        assert e.value.code in (hoc.HOC_EINVAL, hoc.HOC_EDEVICE)
