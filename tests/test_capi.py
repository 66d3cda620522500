"""The C-ABI library loads and exports every function include/isoext_amd.h declares."""

import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def declared():
    text = (ROOT / "include/isoext_amd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(isoext_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(ROOT / "isoext_amd/libisoext_amd.so"))
    names = declared()
    assert len(names) >= 5
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_error_without_gpu():
    lib = ctypes.CDLL(str(ROOT / "isoext_amd/libisoext_amd.so"))
    lib.isoext_version.restype = ctypes.c_int
    assert lib.isoext_version() == 2
    lib.isoext_last_error.restype = ctypes.c_char_p
    assert lib.isoext_last_error() == b""


def test_non_finite_aabb_rejected_before_any_device_work():
    """A NaN or infinite box is an argument error, returned before any HIP
    call (so this runs without a GPU)."""
    import math
    lib = ctypes.CDLL(str(ROOT / "isoext_amd/libisoext_amd.so"))
    lib.isoext_last_error.restype = ctypes.c_char_p

    class Grid(ctypes.Structure):
        _fields_ = [("values", ctypes.c_void_p), ("sparse_cell_ids", ctypes.c_void_p), ("n_sparse", ctypes.c_int64),
                    ("shape", ctypes.c_int64 * 3), ("aabb_min", ctypes.c_float * 3), ("aabb_max", ctypes.c_float * 3)]
    for bad in (math.nan, math.inf, -math.inf):
        g = Grid(None, None, 16, (ctypes.c_int64 * 3)(8, 8, 8), (ctypes.c_float * 3)(-1, bad, -1),
                 (ctypes.c_float * 3)(1, 1, 1))
        n = ctypes.c_int64(7)
        assert lib.isoext_mc_sparse_count(ctypes.byref(g), ctypes.c_float(0.0), None, ctypes.byref(n), None) == -1
        assert n.value == 0 and b"finite" in lib.isoext_last_error()
        counts = (ctypes.c_int64 * 6)()
        rc = lib.isoext_mc_dense(None, g.shape, g.aabb_min, g.aabb_max, ctypes.c_float(0.0), 0, 0, None,
                                 ctypes.c_int64(0), ctypes.c_int64(0), None, ctypes.c_int64(0), None, ctypes.c_int64(0),
                                 counts, None)
        assert rc == -1 and b"finite" in lib.isoext_last_error()


def test_python_grid_rejects_non_finite_aabb():
    import pytest
    torch = pytest.importorskip("torch")
    from isoext_amd.grid import _vec3
    with pytest.raises(ValueError):
        _vec3([0.0, float("nan"), 0.0], "aabb_min")
    assert torch is not None


def test_plan_entry_points_check_arguments_without_gpu():
    """isoext_mc_plan_*: null handles and bad arguments are argument errors
    returned before any HIP call; destroying a null plan is a no-op."""
    import math
    lib = ctypes.CDLL(str(ROOT / "isoext_amd/libisoext_amd.so"))
    lib.isoext_last_error.restype = ctypes.c_char_p
    vp, i64 = ctypes.c_void_p, ctypes.c_int64
    I3, F3, I6 = i64 * 3, ctypes.c_float * 3, i64 * 6
    lib.isoext_mc_plan_create.restype = ctypes.c_int
    lib.isoext_mc_plan_create.argtypes = [vp, I3, F3, F3, ctypes.c_float, ctypes.c_int, ctypes.c_int, vp, i64, i64,
                                          vp, i64, vp, i64, ctypes.POINTER(vp)]
    lib.isoext_mc_plan_run.restype = ctypes.c_int
    lib.isoext_mc_plan_run.argtypes = [vp, I6, vp]
    lib.isoext_mc_plan_destroy.restype = None
    lib.isoext_mc_plan_destroy.argtypes = [vp]
    shape, amin, amax = I3(8, 8, 8), F3(-1, -1, -1), F3(1, 1, 1)
    dummy = vp(256)   # never dereferenced: every call below fails its checks first

    def create(values, method=0, lo=amin, out=True):
        h = vp(12345)
        rc = lib.isoext_mc_plan_create(values, shape, lo, amax, 0.0, method, 0, dummy, 1 << 20, 65536, dummy,
                                       1024, dummy, 1024, ctypes.byref(h) if out else None)
        return rc, h

    rc, _ = create(dummy, out=False)
    assert rc == -1 and b"plan" in lib.isoext_last_error()
    rc, h = create(None)
    assert rc == -1 and h.value is None and b"values" in lib.isoext_last_error()
    rc, h = create(dummy, method=7)
    assert rc == -1 and h.value is None and b"method" in lib.isoext_last_error()
    rc, h = create(dummy, lo=F3(-1, math.nan, -1))
    assert rc == -1 and h.value is None
    assert lib.isoext_mc_plan_run(None, I6(), None) == -1
    lib.isoext_mc_plan_destroy(None)


def test_plan_wrapper_takes_device_tensors_only():
    import pytest
    import torch
    from isoext_amd.mc import MarchingCubesPlan
    with pytest.raises(TypeError):
        MarchingCubesPlan(torch.zeros((4, 4, 4)))
