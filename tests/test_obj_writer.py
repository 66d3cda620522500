"""OBJ writer: the C++ formatter against Python's own float repr (the
reference writes f"v {x} {y} {z}" of float32 values widened to Python
floats, src/isoext/utils.py:42-63), and whole files byte for byte."""

import ctypes

import numpy as np
import pytest
import torch


def _fmt(L, x):
    buf = ctypes.create_string_buffer(64)
    n = L.isoext_format_float(float(x), buf, 64)
    assert n > 0
    return buf.value.decode()


def test_float_repr_matches_python():
    from isoext_amd import _lib
    L = _lib.lib()
    rng = np.random.default_rng(0)
    bits = rng.integers(0, 2 ** 32, 50000, dtype=np.uint64).astype(np.uint32)
    vals = [float(v) for v in bits.view(np.float32)]                       # every binade, nan and inf included
    vals += [float(np.float32(x)) for x in rng.uniform(-1, 1, 20000)]       # mesh-like coordinates
    vals += [0.0, -0.0, 1.0, -1.0, 0.5, 1e16, 1e15, 1e-4, 1e-5, 9999999999999998.0, 123456789012345678.0,
             1e100, 1e-300, 5e-324, float("inf"), float("-inf"), float("nan"), 2.0 ** 60, 0.1]
    vals += [10.0 ** k for k in range(-30, 30)] + [float(np.float32(10.0 ** k)) for k in range(-38, 38)]
    bad = [(x, _fmt(L, x), repr(x)) for x in vals if _fmt(L, x) != repr(x)]
    assert not bad, bad[:5]


def _python_obj(v, f):
    lines = [f"v {a} {b} {c}\n" for a, b, c in v.tolist()]
    lines += [f"f {a} {b} {c}\n" for a, b, c in (f + 1).tolist()]
    return "".join(lines).encode()


@pytest.mark.parametrize("nv,nf", [(1, 1), (1000, 1999), (70000, 140001)])
def test_write_obj_bytes(tmp_path, nv, nf):
    from isoext_amd.utils import write_obj
    rng = np.random.default_rng(nv)
    v = torch.from_numpy(rng.uniform(-1, 1, (nv, 3)).astype(np.float32))
    v[0, 0] = 0.0
    v[0, 1] = -0.0
    f = torch.from_numpy(rng.integers(0, nv, (nf, 3)).astype(np.int32))
    p = tmp_path / "m.obj"
    write_obj(str(p), v, f, threads=4)
    assert p.read_bytes() == _python_obj(v, f)


def test_write_obj_empty(tmp_path):
    from isoext_amd.utils import write_obj
    p = tmp_path / "e.obj"
    p.write_text("old")
    write_obj(str(p), None, None)
    assert p.read_bytes() == b""
    write_obj(str(p), torch.zeros((0, 3)), torch.zeros((0, 3), dtype=torch.int32))
    assert p.read_bytes() == b""


def test_write_obj_bad_path():
    from isoext_amd.utils import write_obj
    with pytest.raises(OSError):
        write_obj("/nonexistent_dir/x.obj", torch.zeros((1, 3)), torch.zeros((1, 3), dtype=torch.int32))
