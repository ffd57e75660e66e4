"""The transform library on the CPU: it loads, exports every symbol
include/speedy_sht.h declares, builds the same bases as the oracle (nodes,
weights, Fourier and Legendre tables), and rejects bad arguments before any
GPU work."""
import ctypes
import os
import re

import numpy as np
import pytest

from jcm_amd import transforms
from oracle import sht_oracle as so

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_declared_symbols():
    text = open(os.path.join(ROOT, "include", "speedy_sht.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = sorted(set(re.findall(r"\b(speedy_\w+)\s*\(", text)))
    assert "speedy_dynamics_to_physics" in names and "speedy_sht_to_modal" in names
    lib = transforms.load()
    for n in names:
        assert hasattr(lib, n), n


def test_struct_layouts():
    assert ctypes.sizeof(transforms._DynamicsSpec) == 48
    assert ctypes.sizeof(transforms._Spectral) == 48
    assert ctypes.sizeof(transforms._Nodal) == 56


@pytest.mark.parametrize("trunc", [21, 31, 85])
def test_host_basis_matches_the_oracle(trunc):
    n_m, n_l, nlon, nlat = transforms.TRUNCATIONS[trunc]
    x, w, f, p = transforms.host_basis(n_m, n_l, nlon, nlat)
    g = so.Grid.truncation(trunc)
    np.testing.assert_allclose(x, g.sin_lat, rtol=0, atol=2e-15)
    np.testing.assert_allclose(w * 2.0 * np.pi / nlon, g.weights, rtol=1e-10)
    np.testing.assert_allclose(f, g.f.astype(np.float32), rtol=0, atol=2e-7)
    np.testing.assert_allclose(p, g.p.astype(np.float32), rtol=2e-6, atol=1e-6)


def test_truncation_table_matches_the_oracle():
    assert transforms.TRUNCATIONS == so.TRUNCATIONS
    assert transforms.GRAVITY == so.GRAVITY and transforms.IDEAL_GAS_CONSTANT == so.IDEAL_GAS_CONSTANT


def test_bad_grids_are_rejected_on_the_host():
    lib = transforms.load()
    x, w = np.zeros(8), np.zeros(8)
    dp = ctypes.POINTER(ctypes.c_double)
    # n_m > n_l, 2 n_m - 1 > nlon, n_l > nlat
    for args in ((10, 9, 32, 16), (17, 17, 32, 16), (4, 9, 32, 8)):
        rc = lib.speedy_sht_host_basis(*args, x.ctypes.data_as(dp), w.ctypes.data_as(dp), None, None)
        assert rc == 1, args
        assert b"bad grid" in lib.speedy_sht_last_error()
    assert lib.speedy_sht_workspace_bytes(None, 4) == 0
    h = ctypes.c_void_p()
    assert lib.speedy_sht_create(4, 9, 32, 8, 1.0, ctypes.byref(h)) == 1 and not h.value
