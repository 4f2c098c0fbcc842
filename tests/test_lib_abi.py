"""CPU checks of the HIP library boundary: it loads, exports every entry point declared in
include/pond.h, and its host-evaluated lattice tables equal the oracle's bit for bit.
No GPU is touched here."""
import ctypes
import os
import re

import numpy as np
import pytest

import pyoracle as po
import lbmpond_amd
from lbmpond_amd import pond as pondmod

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    inc = os.path.join(ROOT, "include")
    for fn in sorted(os.listdir(inc)):
        if fn.endswith(".h"):
            txt = open(os.path.join(inc, fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names += re.findall(r"^[A-Za-z_][\w \*]*?\b(pond_\w+)\s*\(", txt, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(pondmod.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 17
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the Python mirror binds all of them
    assert set(names) <= set(pondmod.SIGNATURES), set(names) - set(pondmod.SIGNATURES)


@pytest.mark.parametrize("lat", ["D1Q5", "D3Q27", "D3Q125"])
def test_lattice_tables_match_oracle_bitwise(lat):
    e, ef, w, T, D = lbmpond_amd.lattice_tables(lat)
    eo, efo, wo, To, Do = po.lattice(lat)
    assert np.array_equal(e, eo) and np.array_equal(ef, efo) and np.array_equal(w, wo)
    assert (T, D) == (To, Do)


def test_params_default_matches_reference_constants():
    L = lbmpond_amd.lib()
    p = pondmod.PondParams()
    L.pond_params_default(ctypes.byref(p), 1)
    assert p.dtau == 1.0 / 0.9 and p.tol_abs == 1e-12 and p.tol_rel == 1e-10 and p.max_iter == 100
    assert [L.pond_qn(i) for i in range(4)] == [5, 27, 125, 0]


def test_invalid_params_are_rejected_without_touching_the_gpu():
    L = lbmpond_amd.lib()
    for kw in ({"lattice": 7}, {"nx": 0}, {"interp": 9}, {"max_iter": 0}):
        p = pondmod.PondParams()
        L.pond_params_default(ctypes.byref(p), 1)
        p.nx = p.ny = p.nz = 4
        for k, v in kw.items():
            setattr(p, k, v)
        h = ctypes.c_void_p()
        st = L.pond_create(ctypes.byref(p), ctypes.byref(h))
        assert st == 1, kw
        assert L.pond_last_error(None)


def test_split_grid_limit_is_checked_before_allocation():
    """The split step addresses two FP32 population quad planes (16 bytes per cell each) through
    a 32-bit buffer offset: 520^3 cells are more than it can address."""
    L = lbmpond_amd.lib()
    p = pondmod.PondParams()
    L.pond_params_default(ctypes.byref(p), 1)
    p.precision = 1
    p.nx = p.ny = p.nz = 520
    h = ctypes.c_void_p()
    assert L.pond_create(ctypes.byref(p), ctypes.byref(h)) == 1
    assert b"POND_SPLIT D3Q27 supports up to" in L.pond_last_error(None)


def test_arr_reader_rejects_bad_files(tmp_path):
    bad = tmp_path / "bad.arr"
    bad.write_bytes(b"\x00\x00\x00\x00\x02\x00\x00\x00")
    with pytest.raises(lbmpond_amd.PondError):
        lbmpond_amd.read_arr(str(bad))


def test_arr_reader_reads_reference_layout(tmp_path):
    """.arr layout of drop.cu:13-31: int comment length, comment, int dim=3, int sizeof(double),
    int Nx, Ny, Nz, then doubles z-outer / x-fastest."""
    nx, ny, nz = 3, 2, 4
    data = np.arange(nx * ny * nz, dtype=np.float64) * 0.5
    path = tmp_path / "Rho_0000007.arr"
    with open(path, "wb") as fp:
        np.array([8], np.int32).tofile(fp)
        np.array([0, 0], np.int32).tofile(fp)
        np.array([3, 8, nx, ny, nz], np.int32).tofile(fp)
        data.tofile(fp)
    arr = lbmpond_amd.read_arr(str(path))
    assert arr.shape == (nz, ny, nx)
    assert np.array_equal(arr.reshape(-1), data)


def test_amr_is_limited_to_fp64_d1q5_lines():
    """the 1D cell activation (amr.cu) is accepted for POND_FP64 D1Q5 lines only, and the
    removal pass stages the line in LDS, so its length is bounded"""
    L = lbmpond_amd.lib()
    cases = [(1, 0, (64, 4, 4)), (0, 0, (64, 2, 1)), (0, 0, (40000, 1, 1)), (0, 0, (64, 1, 1))]
    for lat, prec, shape in cases:
        p = pondmod.PondParams()
        L.pond_params_default(ctypes.byref(p), lat)
        p.precision = prec
        p.nx, p.ny, p.nz = shape
        p.amr = 1 if shape != (64, 1, 1) else 2
        h = ctypes.c_void_p()
        assert L.pond_create(ctypes.byref(p), ctypes.byref(h)) == 1, (lat, shape)
        assert b"amr" in L.pond_last_error(None)
