"""Python host mirror of the PonD step API, over the C ABI in include/pond.h.

Two layers:

* ``Pond`` -- a thin object wrapper of ``pond_create / pond_set_state / pond_step /
  pond_get_fields / pond_diagnostics / pond_drop / pond_destroy``.
* the reference-named facade ``LBMParams``, ``init``, ``calcStep``, ``drop``,
  ``diagnostics`` with the meaning of the reference's host driver
  (lbm.cu:134-168, init.cu:109-128, update.cu:86-142, drop.cu:2-34, diagn.cu:24-30).

All compute runs in the HIP library ``libpond.so`` built from ``csrc/``.  There is no
CPU fallback: if the library is missing or no GPU is present, calls raise.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpond.so")

LATTICES = {"D1Q5": 0, "D3Q27": 1, "D3Q125": 2}
INTERPS = {"weno5": 0, "lagr4": 1, "weno5c": 2, "weno5g": 3}
PRECISIONS = {"fp64": 0, "split": 1}
QN = {"D1Q5": 5, "D3Q27": 27, "D3Q125": 125}

POND_OK = 0
_STATUS = {0: "OK", 1: "INVALID", 2: "HIP", 3: "NOMEM", 4: "IO", 5: "STATE"}


class PondError(RuntimeError):
    """A pond_* call returned an error status (the reference would CHECK_ERROR -> exit)."""

    def __init__(self, status: int, msg: str):
        super().__init__(f"pond error {_STATUS.get(status, status)}: {msg}")
        self.status = status


class PondParams(ctypes.Structure):
    _fields_ = [
        ("lattice", ctypes.c_int),
        ("interp", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("nx", ctypes.c_int64),
        ("ny", ctypes.c_int64),
        ("nz", ctypes.c_int64),
        ("dtau", ctypes.c_double),
        ("tol_abs", ctypes.c_double),
        ("tol_rel", ctypes.c_double),
        ("max_iter", ctypes.c_int),
        ("use_mlp", ctypes.c_int),
        ("split_tol_rel", ctypes.c_double),
        ("split_tol_abs", ctypes.c_double),
        ("tile_patch", ctypes.c_int),
        ("refill_chunk", ctypes.c_int),
        ("pass1_iters", ctypes.c_int),
        ("global_loads", ctypes.c_int),
        ("learned_trust", ctypes.c_double),
        ("mirror", ctypes.c_int),
        ("amr", ctypes.c_int),
    ]


class PondDiag(ctypes.Structure):
    _fields_ = [
        ("mass", ctypes.c_double),
        ("momentum_x", ctypes.c_double),
        ("energy", ctypes.c_double),
        ("k", ctypes.c_double),
        ("niter_max", ctypes.c_int32),
        ("niter_mean", ctypes.c_double),
        ("step", ctypes.c_int64),
        ("max_departure", ctypes.c_double),
    ]


class PondTiming(ctypes.Structure):
    _fields_ = [
        ("last_ms", ctypes.c_double),
        ("total_ms", ctypes.c_double),
        ("steps", ctypes.c_int64),
        ("last_mlp_ms", ctypes.c_double),
    ]


_lib = None

# exported symbol -> (restype, argtypes)
_D = ctypes.POINTER(ctypes.c_double)
_I32 = ctypes.POINTER(ctypes.c_int32)
_I64 = ctypes.POINTER(ctypes.c_int64)
_F = ctypes.POINTER(ctypes.c_float)
_CTX = ctypes.c_void_p
SIGNATURES = {
    "pond_params_default": (None, [ctypes.POINTER(PondParams), ctypes.c_int]),
    "pond_create": (ctypes.c_int, [ctypes.POINTER(PondParams), ctypes.POINTER(ctypes.c_void_p)]),
    "pond_destroy": (None, [_CTX]),
    "pond_last_error": (ctypes.c_char_p, [_CTX]),
    "pond_qn": (ctypes.c_int, [ctypes.c_int]),
    "pond_lattice_tables": (ctypes.c_int, [ctypes.c_int, _I32, _D, _D, _D]),
    "pond_set_state": (ctypes.c_int, [_CTX, _D, _D, _D, _D, _D]),
    "pond_set_populations": (ctypes.c_int, [_CTX, _D, _D, _D, _D, _D, _D]),
    "pond_copy_state": (ctypes.c_int, [_CTX, _CTX]),
    "pond_get_active": (ctypes.c_int, [_CTX, _I32]),
    "pond_set_active": (ctypes.c_int, [_CTX, _I32]),
    "pond_step": (ctypes.c_int, [_CTX, ctypes.c_int]),
    "pond_sync": (ctypes.c_int, [_CTX]),
    "pond_get_fields": (ctypes.c_int, [_CTX, _D, _D, _D, _D, _D, _I32]),
    "pond_get_populations": (ctypes.c_int, [_CTX, _D]),
    "pond_diagnostics": (ctypes.c_int, [_CTX, ctypes.POINTER(PondDiag)]),
    "pond_timing_info": (ctypes.c_int, [_CTX, ctypes.POINTER(PondTiming)]),
    "pond_drop": (ctypes.c_int, [_CTX, ctypes.c_char_p]),
    "pond_checkpoint_write": (ctypes.c_int, [_CTX, ctypes.c_char_p]),
    "pond_checkpoint_read": (ctypes.c_int, [_CTX, ctypes.c_char_p]),
    "pond_read_arr": (ctypes.c_int, [ctypes.c_char_p, _I64, _D]),
    "pond_device_bytes": (ctypes.c_int64, [_CTX]),
    "pond_mlp_num_params": (ctypes.c_int64, []),
    "pond_mlp_init": (ctypes.c_int, [_CTX, ctypes.c_uint64]),
    "pond_mlp_collect": (ctypes.c_int, [_CTX, ctypes.c_int64, ctypes.c_uint64]),
    "pond_mlp_dataset_size": (ctypes.c_int, [_CTX, _I64]),
    "pond_mlp_train": (ctypes.c_int, [_CTX, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_uint64, _F]),
    "pond_mlp_get_params": (ctypes.c_int, [_CTX, _F, _F]),
    "pond_mlp_set_params": (ctypes.c_int, [_CTX, _F, _F]),
    "pond_mlp_forward": (ctypes.c_int, [_CTX, _F, ctypes.c_int64, _F]),
    "pond_mlp_forward_noalloc": (ctypes.c_int, [_CTX, _F, ctypes.c_int64, _F]),
    "pond_mlp_predict": (ctypes.c_int, [_CTX, _F]),
    "pond_mlp_grad": (ctypes.c_int, [_CTX, _F, _F, ctypes.c_int64, _F, _F]),
}


def lib():
    """Load libpond.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                              "(make -C lbmpond_amd/csrc)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _dptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(_D)


def _c64(a, n):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))
    if a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


def lattice_tables(lattice: str):
    """(e int[Qn,3], ef double[Qn,3], w double[Qn], TLat, Dim) as evaluated by the library."""
    L = lib()
    q = QN[lattice]
    e = np.zeros((q, 3), np.int32)
    ef = np.zeros((q, 3))
    w = np.zeros(q)
    td = np.zeros(2)
    st = L.pond_lattice_tables(LATTICES[lattice], e.ctypes.data_as(_I32), _dptr(ef), _dptr(w), _dptr(td))
    if st != POND_OK:
        raise PondError(st, "pond_lattice_tables")
    return e, ef, w, float(td[0]), float(td[1])


def read_arr(path: str) -> np.ndarray:
    """Read a .arr field file (drop.cu:2-34 format) -> array of shape (Nz, Ny, Nx)."""
    L = lib()
    dims = np.zeros(3, np.int64)
    st = L.pond_read_arr(path.encode(), dims.ctypes.data_as(_I64), None)
    if st != POND_OK:
        raise PondError(st, f"pond_read_arr({path})")
    out = np.zeros(int(dims.prod()))
    st = L.pond_read_arr(path.encode(), dims.ctypes.data_as(_I64), _dptr(out))
    if st != POND_OK:
        raise PondError(st, f"pond_read_arr({path})")
    return out.reshape(int(dims[2]), int(dims[1]), int(dims[0]))


class Pond:
    """One simulation context (device state, stream, timers)."""

    def __init__(self, lattice="D3Q27", shape=(16, 16, 16), interp="weno5", precision="fp64", dtau=None,
                 tol_abs=None, tol_rel=None, max_iter=None, use_mlp=False, split_tol_rel=None,
                 split_tol_abs=None, tile_patch=None, refill_chunk=None, pass1_iters=None, global_loads=None,
                 learned_trust=None, mirror=None, amr=None):
        """Parameters left at None keep pond_params_default's value.  tile_patch, refill_chunk,
        pass1_iters and global_loads set the D3Q27 kernels' work distribution (no effect on
        results); learned_trust the learned start's trust radius; mirror the mirror-boundary axes,
        bit d for axis d; amr the 1D cell activation after every step (include/pond.h)."""
        L = lib()
        p = PondParams()
        L.pond_params_default(ctypes.byref(p), LATTICES[lattice])
        p.interp = INTERPS[interp]
        p.precision = PRECISIONS[precision]
        p.nx, p.ny, p.nz = (int(s) for s in shape)
        if dtau is not None:
            p.dtau = float(dtau)
        if tol_abs is not None:
            p.tol_abs = float(tol_abs)
        if tol_rel is not None:
            p.tol_rel = float(tol_rel)
        if max_iter is not None:
            p.max_iter = int(max_iter)
        if split_tol_rel is not None:
            p.split_tol_rel = float(split_tol_rel)
        if split_tol_abs is not None:
            p.split_tol_abs = float(split_tol_abs)
        for name, val in (("tile_patch", tile_patch), ("refill_chunk", refill_chunk), ("pass1_iters", pass1_iters),
                          ("global_loads", global_loads)):
            if val is not None:
                setattr(p, name, int(val))
        if learned_trust is not None:
            p.learned_trust = float(learned_trust)
        if mirror is not None:
            p.mirror = int(mirror)
        if amr is not None:
            p.amr = int(amr)
        p.use_mlp = 1 if use_mlp else 0
        h = ctypes.c_void_p()
        st = L.pond_create(ctypes.byref(p), ctypes.byref(h))
        if st != POND_OK:
            raise PondError(st, L.pond_last_error(None).decode())
        self._h = h
        self.params = p
        self.lattice = lattice
        self.qn = QN[lattice]
        self.shape = (int(p.nx), int(p.ny), int(p.nz))
        self.n = self.shape[0] * self.shape[1] * self.shape[2]

    # -- helpers
    def _check(self, st, what):
        if st != POND_OK:
            raise PondError(st, f"{what}: {lib().pond_last_error(self._h).decode()}")

    def close(self):
        if getattr(self, "_h", None):
            lib().pond_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- state
    def set_state(self, rho, ux, uy, uz, T):
        """Equilibrium f_i = w_i*rho in each cell's own gauge (Cell::set_base)."""
        a = [_c64(v, self.n) for v in (rho, ux, uy, uz, T)]
        self._check(lib().pond_set_state(self._h, *[_dptr(v) for v in a]), "pond_set_state")

    def set_populations(self, f, rho, ux, uy, uz, T):
        fa = _c64(f, self.n * self.qn)
        a = [_c64(v, self.n) for v in (rho, ux, uy, uz, T)]
        self._check(lib().pond_set_populations(self._h, _dptr(fa), *[_dptr(v) for v in a]),
                    "pond_set_populations")

    def copy_state_from(self, other):
        """the current state of another context of the same lattice and grid (pond_copy_state),
        converted to this context's precision on the device"""
        self._check(lib().pond_copy_state(self._h, other._h), "pond_copy_state")

    def active(self):
        """Cinfo::set of every cell (1 = stepped, 0 = switched off)"""
        a = np.zeros(self.n, np.int32)
        self._check(lib().pond_get_active(self._h, a.ctypes.data_as(_I32)), "pond_get_active")
        return a

    def set_active(self, a):
        a = np.ascontiguousarray(a, dtype=np.int32).reshape(-1)
        if a.size != self.n:
            raise ValueError(f"set_active: {a.size} flags for {self.n} cells")
        self._check(lib().pond_set_active(self._h, a.ctypes.data_as(_I32)), "pond_set_active")

    def step(self, n=1):
        self._check(lib().pond_step(self._h, int(n)), "pond_step")

    def sync(self):
        self._check(lib().pond_sync(self._h), "pond_sync")

    def fields(self):
        out = {k: np.zeros(self.n) for k in ("rho", "ux", "uy", "uz", "T")}
        out["niter"] = np.zeros(self.n, np.int32)
        self._check(lib().pond_get_fields(self._h, _dptr(out["rho"]), _dptr(out["ux"]), _dptr(out["uy"]),
                                          _dptr(out["uz"]), _dptr(out["T"]),
                                          out["niter"].ctypes.data_as(_I32)), "pond_get_fields")
        return out

    def niter(self):
        """gauge iterations of every cell in the last step (pond_get_fields with only niter)"""
        it = np.zeros(self.n, np.int32)
        self._check(lib().pond_get_fields(self._h, None, None, None, None, None, it.ctypes.data_as(_I32)),
                    "pond_get_fields")
        return it

    def populations(self):
        f = np.zeros(self.n * self.qn)
        self._check(lib().pond_get_populations(self._h, _dptr(f)), "pond_get_populations")
        return f.reshape(self.qn, self.n)

    def diagnostics(self):
        d = PondDiag()
        self._check(lib().pond_diagnostics(self._h, ctypes.byref(d)), "pond_diagnostics")
        return {k: getattr(d, k) for k, _ in PondDiag._fields_}

    def timing(self):
        t = PondTiming()
        self._check(lib().pond_timing_info(self._h, ctypes.byref(t)), "pond_timing_info")
        return {k: getattr(t, k) for k, _ in PondTiming._fields_}

    def drop(self, directory):
        self._check(lib().pond_drop(self._h, str(directory).encode()), "pond_drop")

    def checkpoint(self, path):
        """the full state (populations included) for an exact restart (pond_checkpoint_write)"""
        self._check(lib().pond_checkpoint_write(self._h, str(path).encode()), "pond_checkpoint_write")

    def restore(self, path):
        """state and step count from a checkpoint of the same lattice, precision and grid"""
        self._check(lib().pond_checkpoint_read(self._h, str(path).encode()), "pond_checkpoint_read")

    def device_bytes(self):
        return int(lib().pond_device_bytes(self._h))

    # -- learned gauge initializer
    def mlp_init(self, seed=0):
        self._check(lib().pond_mlp_init(self._h, int(seed)), "pond_mlp_init")

    def mlp_collect(self, nsamples, seed=0):
        self._check(lib().pond_mlp_collect(self._h, int(nsamples), int(seed)), "pond_mlp_collect")

    def mlp_dataset_size(self):
        n = ctypes.c_int64(0)
        self._check(lib().pond_mlp_dataset_size(self._h, ctypes.byref(n)), "pond_mlp_dataset_size")
        return int(n.value)

    def mlp_train(self, epochs=4, batch=4096, lr=2e-3, seed=0):
        losses = np.zeros(int(epochs), np.float32)
        self._check(lib().pond_mlp_train(self._h, int(epochs), int(batch), float(lr), int(seed),
                                         losses.ctypes.data_as(_F)), "pond_mlp_train")
        return losses

    def mlp_params(self):
        n = int(lib().pond_mlp_num_params())
        p = np.zeros(n, np.float32)
        sc = np.zeros(9, np.float32)
        self._check(lib().pond_mlp_get_params(self._h, p.ctypes.data_as(_F), sc.ctypes.data_as(_F)),
                    "pond_mlp_get_params")
        return p, sc

    def mlp_set_params(self, params, scales=None):
        p = np.ascontiguousarray(params, np.float32)
        sc = None if scales is None else np.ascontiguousarray(scales, np.float32)
        self._check(lib().pond_mlp_set_params(self._h, p.ctypes.data_as(_F),
                                              None if sc is None else sc.ctypes.data_as(_F)), "pond_mlp_set_params")

    def mlp_forward(self, X):
        X = np.ascontiguousarray(X, np.float32)
        B = X.shape[0]
        Y = np.zeros((B, 4), np.float32)
        self._check(lib().pond_mlp_forward(self._h, X.ctypes.data_as(_F), B, Y.ctypes.data_as(_F)),
                    "pond_mlp_forward")
        return Y

    def mlp_forward_noalloc(self, X):
        """mlp_forward through the entry point that allocates no device memory (at most 16384 rows)"""
        X = np.ascontiguousarray(X, np.float32)
        Y = np.zeros((X.shape[0], 4), np.float32)
        self._check(lib().pond_mlp_forward_noalloc(self._h, X.ctypes.data_as(_F), X.shape[0], Y.ctypes.data_as(_F)),
                    "pond_mlp_forward_noalloc")
        return Y

    def mlp_predict(self):
        """predicted start of every cell from the current state: (4, n) float32"""
        out = np.zeros(4 * self.n, np.float32)
        self._check(lib().pond_mlp_predict(self._h, out.ctypes.data_as(_F)), "pond_mlp_predict")
        return out.reshape(4, self.n)

    def mlp_grad(self, X, T):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        g = np.zeros(int(lib().pond_mlp_num_params()), np.float32)
        loss = ctypes.c_float(0)
        self._check(lib().pond_mlp_grad(self._h, X.ctypes.data_as(_F), T.ctypes.data_as(_F), X.shape[0],
                                        g.ctypes.data_as(_F), ctypes.byref(loss)), "pond_mlp_grad")
        return g, float(loss.value)


# ------------------------------------------------------------------ reference-named facade


@dataclass
class LBMParams:
    """Host parameter record with the reference's field names (structs.cuh:96-149)."""
    Nx: int = 128
    Ny: int = 1
    Nz: int = 1
    dtau: float = 1.0 / 0.9          # pc.dtau (PhysPars::set)
    lattice: str = "D1Q5"            # consts.h:6-12 selection, at run time
    interp: str = "weno5"            # steps.cu:451/458 selection, at run time
    precision: str = "fp64"
    max_iter: int = 100
    iStep: int = 0
    curtime: float = 0.0
    dt: float = 1.0
    extra: dict = field(default_factory=dict)


_current = {"pond": None, "pars": None}


def init(pars: LBMParams, rho, ux, uy, uz, T):
    """init() + fill(): allocate the state and load an initial condition (init.cu:109-128)."""
    if _current["pond"] is not None:
        _current["pond"].close()
    p = Pond(pars.lattice, (pars.Nx, pars.Ny, pars.Nz), pars.interp, pars.precision, dtau=pars.dtau,
             max_iter=pars.max_iter, **pars.extra)
    p.set_state(rho, ux, uy, uz, T)
    pars.iStep = 0
    pars.curtime = 0.0
    _current["pond"], _current["pars"] = p, pars
    return p


def calcStep():
    """One PonD step (update.cu:86-142 without the I/O it times)."""
    p, pars = _current["pond"], _current["pars"]
    if p is None:
        raise PondError(5, "calcStep before init")
    p.step(1)
    pars.iStep += 1
    pars.curtime += pars.dt


def drop(directory):
    """Write Rho/Vx/Vy/Vz/T_%07d.arr of the current state (drop.cu:2-34)."""
    _current["pond"].drop(directory)


def diagnostics():
    """mass, momentum_x, energy (calcTotMoments), k (diagn), NiterMax, mean iterations."""
    return _current["pond"].diagnostics()
