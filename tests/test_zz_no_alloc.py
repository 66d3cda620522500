"""isoext_mc_dense_noalloc, called through the C ABI: its meshes equal
isoext_mc_dense's (same default tile configuration) element for element, on
two calls in a row (the second on a smaller grid, so stale contents of the
fixed scratch would show), on fields whose weld fix-up takes the
single-workgroup and the multi-kernel paths; grids beyond its fixed scratch
are refused before any device work."""

import ctypes

import numpy as np
import pytest

from conftest import multi_sphere_field, sphere_field

pytestmark = pytest.mark.gpu

c_i64, c_f, vp = ctypes.c_int64, ctypes.c_float, ctypes.c_void_p
I64x3, F32x3, I64x6 = c_i64 * 3, c_f * 3, c_i64 * 6


def _lib():
    from isoext_amd import _lib
    L = _lib.lib()
    L.isoext_mc_dense_noalloc.restype = ctypes.c_int
    L.isoext_mc_dense_noalloc.argtypes = [vp, I64x3, F32x3, F32x3, c_f, ctypes.c_int, vp, c_i64, vp, c_i64, I64x6,
                                          vp]
    L.isoext_mc_dense_noalloc_limit.restype = c_i64
    L.isoext_mc_dense_noalloc_limit.argtypes = []
    return L


def _reference(L, vals, level, method):
    """isoext_mc_dense with the default configuration and a fresh workspace."""
    import torch
    shape = I64x3(*vals.shape)
    cfg = L.isoext_mc_auto_config(shape)
    cap = 1 << 16
    wsb = L.isoext_mc_dense_workspace_bytes(shape, cfg, cap)
    ws = torch.zeros(wsb, dtype=torch.uint8, device=vals.device)
    counts = I64x6()
    cap_v, cap_f = 1 << 22, 1 << 23
    for _ in range(4):
        v = torch.empty((cap_v, 3), dtype=torch.float32, device=vals.device)
        f = torch.empty((cap_f, 3), dtype=torch.int32, device=vals.device)
        rc = L.isoext_mc_dense(vals.data_ptr(), shape, F32x3(-1, -1, -1), F32x3(1, 1, 1), level, method, cfg,
                               ws.data_ptr(), wsb, cap, v.data_ptr(), cap_v, f.data_ptr(), cap_f, counts,
                               torch.cuda.current_stream().cuda_stream)
        if rc == -3:
            cap_v, cap_f = counts[0] + 1, counts[1] + 1
            continue
        if rc == -4:   # special-vertex lists: grow to the reported need
            cap = int(counts[2])
            wsb = L.isoext_mc_dense_workspace_bytes(shape, cfg, cap)
            ws = torch.zeros(wsb, dtype=torch.uint8, device=vals.device)
            continue
        assert rc == 0, L.isoext_last_error()
        return v[:counts[0]].clone(), f[:counts[1]].clone()
    raise AssertionError("sizing did not converge")


def _noalloc(L, vals, level, method):
    import torch
    shape = I64x3(*vals.shape)
    counts = I64x6()
    cap_v, cap_f = 1 << 22, 1 << 23
    for _ in range(3):
        v = torch.empty((cap_v, 3), dtype=torch.float32, device=vals.device)
        f = torch.empty((cap_f, 3), dtype=torch.int32, device=vals.device)
        rc = L.isoext_mc_dense_noalloc(vals.data_ptr(), shape, F32x3(-1, -1, -1), F32x3(1, 1, 1), level, method,
                                       v.data_ptr(), cap_v, f.data_ptr(), cap_f, counts,
                                       torch.cuda.current_stream().cuda_stream)
        if rc == -3:
            cap_v, cap_f = counts[0] + 1, counts[1] + 1
            continue
        assert rc == 0, L.isoext_last_error()
        return v[:counts[0]], f[:counts[1]], int(counts[2])
    raise AssertionError("sizing did not converge")


def _check(L, arr, level=0.0, method=0):
    import torch
    vals = torch.from_numpy(np.ascontiguousarray(arr, np.float32)).cuda()
    rv, rf = _reference(L, vals, level, method)
    nv, nf, need = _noalloc(L, vals, level, method)
    print(f"noalloc {tuple(arr.shape)} level {level}: {rv.shape[0]} vertices, special need {need}", flush=True)
    assert rv.shape == nv.shape and rf.shape == nf.shape, (rv.shape, nv.shape, rf.shape, nf.shape)
    assert torch.equal(rv, nv) and torch.equal(rf, nf)
    return rv.shape[0], need


def test_noalloc_equals_mc_dense_two_calls(gpu):
    L = _lib()
    big = multi_sphere_field((300, 280, 320), 24, seed=31)
    small = multi_sphere_field((150, 130, 170), 9, seed=32)
    assert _check(L, big)[0] > 10000
    assert _check(L, small, 0.0, 1)[0] > 1000


def test_noalloc_weld_paths(gpu):
    """A small quantised field puts about a hundred corners exactly at the
    level (the single-workgroup fix-up); an occupancy field at level 1.0
    snaps every vertex to a corner, several thousand weld candidates (the
    multi-kernel fix-up, its tables in the fixed scratch)."""
    L = _lib()
    q = (np.round(sphere_field((10, 10, 10), 0.6) * 8) / 8).astype(np.float32)
    assert _check(L, q, 0.0)[1] > 0
    occ = (sphere_field((60, 56, 64), 0.7) < 0).astype(np.float32)
    assert _check(L, occ, 1.0)[1] > 2048
    _check(L, multi_sphere_field((64, 64, 64), 5, seed=33), 0.0)


def test_noalloc_refuses_grids_beyond_its_scratch(gpu):
    L = _lib()
    counts = I64x6()
    rc = L.isoext_mc_dense_noalloc(None, I64x3(700, 700, 700), F32x3(-1, -1, -1), F32x3(1, 1, 1), 0.0, 0, None, 0,
                                   None, 0, counts, None)
    assert rc == -5 and b"allocation-free" in L.isoext_last_error()
    assert L.isoext_mc_dense_noalloc_limit() > 0
