"""ctypes loader of the CPU oracle (oracle/_build/libpond_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU-baseline leg of
bench.py, never by the lbmpond_amd package.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# POND_ORACLE_LIB names another build of the oracle (oracle/mutation_check.py runs the pin
# tests against deliberately broken builds that way)
LIB_PATH = os.environ.get("POND_ORACLE_LIB") or os.path.join(_HERE, "_build", "libpond_oracle.so")
LATTICES = {"D1Q5": 0, "D3Q27": 1, "D3Q125": 2}
INTERPS = {"weno5": 0, "lagr4": 1, "weno5c": 2, "weno5g": 3}
QN = {"D1Q5": 5, "D3Q27": 27, "D3Q125": 125}

_D = ctypes.POINTER(ctypes.c_double)
_I32 = ctypes.POINTER(ctypes.c_int32)


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("lattice", ctypes.c_int),
        ("interp", ctypes.c_int),
        ("nx", ctypes.c_long),
        ("ny", ctypes.c_long),
        ("nz", ctypes.c_long),
        ("dtau", ctypes.c_double),
        ("tol_abs", ctypes.c_double),
        ("tol_rel", ctypes.c_double),
        ("max_iter", ctypes.c_int),
        ("mirror", ctypes.c_int),
        ("amr", ctypes.c_int),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: run make -C oracle")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_qn.restype = ctypes.c_int
        L.orc_qn.argtypes = [ctypes.c_int]
        L.orc_lattice.restype = None
        L.orc_lattice.argtypes = [ctypes.c_int, _I32, _D, _D, _D]
        L.orc_reverse.restype = None
        L.orc_reverse.argtypes = [ctypes.c_int, _I32]
        L.orc_transfer.restype = ctypes.c_double
        L.orc_transfer.argtypes = [ctypes.c_int, _D, _D, _D, ctypes.c_int]
        L.orc_lagrpol.restype = ctypes.c_double
        L.orc_lagrpol.argtypes = [ctypes.c_int] * 4 + [ctypes.c_double] * 3
        L.orc_moments.restype = None
        L.orc_moments.argtypes = [ctypes.c_int, _D, _D, _D]
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.POINTER(OrcParams)]
        L.orc_destroy.restype = None
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_set_state.restype = None
        L.orc_set_state.argtypes = [ctypes.c_void_p] + [_D] * 5
        L.orc_set_populations.restype = None
        L.orc_set_populations.argtypes = [ctypes.c_void_p] + [_D] * 6
        L.orc_step.restype = None
        L.orc_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_get_fields.restype = None
        L.orc_get_fields.argtypes = [ctypes.c_void_p] + [_D] * 5 + [_I32]
        L.orc_get_populations.restype = None
        L.orc_get_populations.argtypes = [ctypes.c_void_p, _D]
        L.orc_interp.restype = ctypes.c_double
        L.orc_interp.argtypes = [ctypes.c_void_p, ctypes.c_int, _D, _D]
        L.orc_get_active.restype = None
        L.orc_get_active.argtypes = [ctypes.c_void_p, _I32]
        L.orc_set_active.restype = None
        L.orc_set_active.argtypes = [ctypes.c_void_p, _I32]
        L.orc_totals.restype = None
        L.orc_totals.argtypes = [ctypes.c_void_p, _D]
        L.orc_threads.restype = ctypes.c_int
        L.orc_threads.argtypes = []
        L.orc_mlp_nparam.restype = ctypes.c_int
        L.orc_mlp_nparam.argtypes = []
        _F = ctypes.POINTER(ctypes.c_float)
        L.orc_mlp_forward.restype = None
        L.orc_mlp_forward.argtypes = [_F, ctypes.c_int64, _F, _F, ctypes.c_int]
        L.orc_mlp_grad.restype = None
        L.orc_mlp_grad.argtypes = [_F, ctypes.c_int64, _F, _F, _D, _D, ctypes.c_int]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_D)


def _c(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))


def lattice(name):
    q = QN[name]
    e = np.zeros((q, 3), np.int32)
    ef = np.zeros((q, 3))
    w = np.zeros(q)
    td = np.zeros(2)
    lib().orc_lattice(LATTICES[name], e.ctypes.data_as(_I32), _p(ef), _p(w), _p(td))
    return e, ef, w, float(td[0]), float(td[1])


def reverse(name):
    """reverseX/Y/Z of the lattice as int[3][Qn]"""
    r = np.zeros((3, QN[name]), np.int32)
    lib().orc_reverse(LATTICES[name], r.ctypes.data_as(_I32))
    return r


def transfer(name, f, uT0, base, inew):
    f, uT0, base = _c(f), _c(uT0), _c(base)
    return lib().orc_transfer(LATTICES[name], _p(f), _p(uT0), _p(base), int(inew))


def lagrpol(n, ix, iy, iz, sx, sy, sz):
    return lib().orc_lagrpol(n, ix, iy, iz, sx, sy, sz)


def moments(name, f, gauge):
    out = np.zeros(5)
    lib().orc_moments(LATTICES[name], _p(_c(f)), _p(_c(gauge)), _p(out))
    return out


def threads():
    return int(lib().orc_threads())


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def mlp_nparam():
    return int(lib().orc_mlp_nparam())


def mlp_forward(params, X, bf16=True):
    """network output for normalized features X[B][135] (oracle/mlp_oracle.c)"""
    P, X = _f32(params), _f32(X)
    Y = np.zeros((X.shape[0], 4), np.float32)
    F = ctypes.POINTER(ctypes.c_float)
    lib().orc_mlp_forward(P.ctypes.data_as(F), X.shape[0], X.ctypes.data_as(F), Y.ctypes.data_as(F), int(bf16))
    return Y


def mlp_grad(params, X, T, bf16=True):
    P, X, T = _f32(params), _f32(X), _f32(T)
    g = np.zeros(P.size)
    loss = np.zeros(1)
    F = ctypes.POINTER(ctypes.c_float)
    lib().orc_mlp_grad(P.ctypes.data_as(F), X.shape[0], X.ctypes.data_as(F), T.ctypes.data_as(F), _p(g), _p(loss),
                       int(bf16))
    return g, float(loss[0])


class Oracle:
    """CPU restatement of the reference step on one grid."""

    def __init__(self, lattice, shape, interp="weno5", dtau=1.0 / 0.9, tol_abs=1e-12, tol_rel=1e-10, max_iter=100,
                 mirror=0, amr=0):
        """mirror: bit d set = mirror boundaries on axis d (Lagrange-4 and D3Q125 only, as in the
        reference, whose WENO5 interpolation has no mirror path).  amr: 1 = the 1D cell activation
        and removal passes (amr.cu) after every step."""
        p = OrcParams(LATTICES[lattice], INTERPS[interp], int(shape[0]), int(shape[1]), int(shape[2]), float(dtau),
                      float(tol_abs), float(tol_rel), int(max_iter), int(mirror), int(amr))
        self._h = lib().orc_create(ctypes.byref(p))
        if not self._h:
            raise MemoryError("orc_create failed")
        self.lattice = lattice
        self.qn = QN[lattice]
        self.shape = tuple(int(s) for s in shape)
        self.n = self.shape[0] * self.shape[1] * self.shape[2]

    def close(self):
        if getattr(self, "_h", None):
            lib().orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, rho, ux, uy, uz, T):
        a = [_c(v) for v in (rho, ux, uy, uz, T)]
        lib().orc_set_state(self._h, *[_p(v) for v in a])

    def set_populations(self, f, rho, ux, uy, uz, T):
        a = [_c(v) for v in (f, rho, ux, uy, uz, T)]
        lib().orc_set_populations(self._h, *[_p(v) for v in a])

    def step(self, n=1):
        lib().orc_step(self._h, int(n))

    def active(self):
        """Cinfo::set of every cell (1 = stepped)"""
        a = np.zeros(self.n, np.int32)
        lib().orc_get_active(self._h, a.ctypes.data_as(_I32))
        return a

    def set_active(self, a):
        a = np.ascontiguousarray(a, dtype=np.int32)
        lib().orc_set_active(self._h, a.ctypes.data_as(_I32))

    def fields(self):
        out = {k: np.zeros(self.n) for k in ("rho", "ux", "uy", "uz", "T")}
        out["niter"] = np.zeros(self.n, np.int32)
        lib().orc_get_fields(self._h, _p(out["rho"]), _p(out["ux"]), _p(out["uy"]), _p(out["uz"]), _p(out["T"]),
                             out["niter"].ctypes.data_as(_I32))
        return out

    def populations(self):
        f = np.zeros(self.n * self.qn)
        lib().orc_get_populations(self._h, _p(f))
        return f.reshape(self.qn, self.n)

    def interp(self, i, crd, gauge):
        return lib().orc_interp(self._h, int(i), _p(_c(crd)), _p(_c(gauge)))

    def totals(self):
        out = np.zeros(4)
        lib().orc_totals(self._h, _p(out))
        return dict(mass=out[0], momentum_x=out[1], energy=out[2], k=out[3])
