"""gapx_min_distance_noalloc: FGET_BOND in the library's fixed per-GPU
scratch array (no allocation), against gapx_min_distance (which allocates its
workspace) bit for bit, on calls in a row: 1500 atoms (the three-launch
binning), then 120 atoms (the one-launch binning), then 1500 again; inputs
beyond the scratch size are refused before any GPU work."""
import ctypes

import numpy as np
import pytest

import gapx
from gapx import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if gapx.device_count() == 0:
        pytest.skip("no GPU visible")


def noalloc(lat, pos, rcut, flags=0):
    L = gapx.lib()
    f = L.gapx_min_distance_noalloc
    f.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_void_p]
    lat = np.ascontiguousarray(lat, np.float64) if not flags else np.asfortranarray(lat, np.float64)
    pos = np.ascontiguousarray(pos, np.float64) if not flags else np.asfortranarray(pos, np.float64)
    out = np.zeros(1)
    rc = f(len(pos), lat.ctypes.data, pos.ctypes.data, float(rcut), int(flags), out.ctypes.data)
    if rc:
        raise gapx.GapxError(rc, L.gapx_last_error().decode())
    return float(out[0])


def test_noalloc_equals_allocating_entry_on_calls_in_a_row():
    big = synth.structure(1500, 1500 * 11.0, (1, 8, 22), np.random.default_rng(701))
    small = synth.structure(120, 120 * 11.0, (1, 8, 22), np.random.default_rng(702))
    for s in (big, small, big):
        want = gapx.min_distance(s["lat"], s["pos"], 6.0)
        got = noalloc(s["lat"], s["pos"], 6.0)
        print(len(s["pos"]), got, want)
        assert got == want
    # Fortran layout (FGET_BOND's LAT(3,3), POS(NA,3))
    assert noalloc(small["lat"], small["pos"], 6.0, flags=gapx.LAYOUT_F) == gapx.min_distance(
        small["lat"], small["pos"], 6.0)


def test_noalloc_refuses_inputs_beyond_its_scratch():
    rng = np.random.default_rng(703)
    lat = np.eye(3) * 120.0
    pos = rng.random((150000, 3)) * 120.0
    with pytest.raises(gapx.GapxError) as e:
        noalloc(lat, pos, 6.0)
    assert e.value.code == 1 and "scratch" in str(e.value)
