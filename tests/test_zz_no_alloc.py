"""The allocation-free surface entry point (cam_egrad_noalloc) against the
allocating one (cam_egrad_h3 / _ch4h / _oh3 / _brh2 / _o3), through the C
ABI: two calls in a row per surface, the second with fewer beads, bit for
bit; sizes beyond its limit are rejected."""
import ctypes

import numpy as np
import pytest

import caracal_amd as cam
from caracal_amd import _abi, workloads

pytestmark = pytest.mark.gpu

LIMIT = 16384
CASES = [(cam.CAM_SURFACE_H3, 3, workloads.H3_TS, "cam_egrad_h3"),
         (cam.CAM_SURFACE_CH4H, 6, workloads.CH4H_TS, "cam_egrad_ch4h"),
         (cam.CAM_SURFACE_OH3, 4, workloads.OH3_TS, "cam_egrad_oh3"),
         (cam.CAM_SURFACE_BRH2, 3, workloads.BRH2_TS, "cam_egrad_brh2"),
         (cam.CAM_SURFACE_O3, 3, workloads.O3_TS, "cam_egrad_o3")]


def _call(fn, *args):
    rc = fn(*args)
    assert rc == 0, _abi.lib().cam_last_error()


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@pytest.mark.parametrize("surface,natoms,ts,name", CASES)
def test_noalloc_equals_allocating_entry_point(surface, natoms, ts, name):
    lib = _abi.lib()
    rng = np.random.default_rng(surface)
    for nb in (LIMIT, 777):
        print(f"surface {surface}: {nb} beads", flush=True)
        q = np.ascontiguousarray(ts[None] + rng.normal(0.0, 0.05, (nb, natoms, 3)))
        v0, g0 = np.zeros(nb), np.zeros_like(q)
        v1, g1 = np.full(nb, np.nan), np.full_like(q, np.nan)
        _call(getattr(lib, name), _p(q), natoms, nb, _p(v0), _p(g0))
        _call(lib.cam_egrad_noalloc, surface, _p(q), natoms, nb, _p(v1), _p(g1))
        assert np.array_equal(v0, v1) and np.array_equal(g0, g1)


def test_noalloc_rejects_sizes_beyond_its_limit():
    lib = _abi.lib()
    q = np.zeros((LIMIT + 1, 3, 3))
    v, g = np.zeros(LIMIT + 1), np.zeros_like(q)
    assert lib.cam_egrad_noalloc(cam.CAM_SURFACE_H3, _p(q), 3, LIMIT + 1, _p(v), _p(g)) == 1
    assert b"16384" in lib.cam_last_error()
    assert lib.cam_egrad_noalloc(99, _p(q), 3, 1, _p(v), _p(g)) == 1
