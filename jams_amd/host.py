"""ctypes binding of the JAMS-compatible host layer (``include/jams_host.h``):
config parsing, lattice, neighbour list, CSR builder, plugins and run loop.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .engine import LIB_DIR, Engine, EngineError, load_library

HOST_LIB = LIB_DIR / "libjamshost.so"

ABI = {
    "jams_sim_create": (C.c_int, [C.POINTER(C.c_char_p), C.c_int, C.c_char_p, C.c_char_p, C.c_int,
                                  C.POINTER(C.c_void_p)]),
    "jams_sim_destroy": (C.c_int, [C.c_void_p]),
    "jams_sim_last_error": (C.c_char_p, [C.c_void_p]),
    "jams_sim_num_spins": (C.c_int64, [C.c_void_p]),
    "jams_sim_max_steps": (C.c_int64, [C.c_void_p]),
    "jams_sim_engine": (C.c_void_p, [C.c_void_p]),
    "jams_sim_run": (C.c_int, [C.c_void_p, C.c_int64]),
    "jams_sim_sites": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "jams_sim_exchange_counts": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int64),
                                           C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "jams_sim_exchange_pairs": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "jams_sim_exchange_csr": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "jams_sim_exchange_values": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                           C.POINTER(C.c_int64)]),
    "jams_sim_iteration": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
    "jams_sim_write_checkpoint": (C.c_int, [C.c_void_p, C.c_char_p]),
    "jams_sim_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                  C.POINTER(C.c_uint64)]),
}

_lib = None


def load_host_library() -> C.CDLL:
    global _lib
    if _lib is None:
        load_library()  # the engine must load first (libjamshost links it)
        if not HOST_LIB.exists():
            raise RuntimeError(f"jams_amd host library not built: {HOST_LIB}")
        lib = C.CDLL(str(HOST_LIB))
        for name, (res, args) in ABI.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Simulation:
    """A JAMS config set up by the host layer.  device < 0: set-up only."""

    def __init__(self, sources, device: int = -1, output_dir: str = ".", name: str = "jams"):
        self.lib = load_host_library()
        if isinstance(sources, (str, bytes)):
            sources = [sources]
        enc = [s.encode() if isinstance(s, str) else s for s in sources]
        arr = (C.c_char_p * len(enc))(*enc)
        h = C.c_void_p()
        rc = self.lib.jams_sim_create(arr, len(enc), output_dir.encode(), name.encode(), int(device),
                                      C.byref(h))
        if rc != 0:
            raise RuntimeError(self.lib.jams_sim_last_error(None).decode())
        self.h = h
        self.n = self.lib.jams_sim_num_spins(h)

    def close(self):
        if getattr(self, "h", None):
            self.lib.jams_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != 0:
            raise RuntimeError(self.lib.jams_sim_last_error(self.h).decode())

    @property
    def max_steps(self) -> int:
        return self.lib.jams_sim_max_steps(self.h)

    def engine_handle(self):
        return self.lib.jams_sim_engine(self.h)

    def engine(self) -> "Engine":
        """A non-owning Engine view of the simulation's context."""
        h = self.engine_handle()
        if not h:
            raise RuntimeError("set-up-only simulation has no engine")
        e = Engine.__new__(Engine)
        e.lib = load_library()
        e.h = C.c_void_p(h)
        e.n = self.n
        e._keep = [self]
        e.close = lambda: None
        return e

    def run(self, limit: int = -1):
        self._chk(self.lib.jams_sim_run(self.h, int(limit)))

    def sites(self):
        mus, gyro, alpha = (np.empty(self.n) for _ in range(3))
        s = np.empty((self.n, 3))
        self._chk(self.lib.jams_sim_sites(self.h, _p(mus), _p(gyro), _p(alpha), _p(s)))
        return mus, gyro, alpha, s

    def exchange_counts(self, index=0):
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        self._chk(self.lib.jams_sim_exchange_counts(self.h, index, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def exchange_pairs(self, index=0):
        _, npairs, _ = self.exchange_counts(index)
        pi = np.empty(npairs, np.int32)
        pj = np.empty(npairs, np.int32)
        J = np.empty((npairs, 9), np.float64)
        self._chk(self.lib.jams_sim_exchange_pairs(self.h, index, _p(pi), _p(pj), _p(J)))
        return pi, pj, J

    def exchange_list(self, index=0):
        """(pi, pj, value index per pair, distinct tensors nvalues x 9): the
        neighbour list exactly as the engine receives it."""
        _, npairs, _ = self.exchange_counts(index)
        pi = np.empty(npairs, np.int32)
        pj = np.empty(npairs, np.int32)
        self._chk(self.lib.jams_sim_exchange_pairs(self.h, index, _p(pi), _p(pj), None))
        nv = C.c_int64()
        self._chk(self.lib.jams_sim_exchange_values(self.h, index, None, None, C.byref(nv)))
        vidx = np.empty(npairs, np.int32)
        vals = np.empty((nv.value, 9), np.float64)
        self._chk(self.lib.jams_sim_exchange_values(self.h, index, _p(vidx), _p(vals), C.byref(nv)))
        return pi, pj, vidx, vals

    def exchange_csr(self, index=0):
        _, _, nnz = self.exchange_counts(index)
        rowptr = np.empty(3 * self.n + 1, np.int64)
        col = np.empty(nnz, np.int32)
        val = np.empty(nnz, np.float64)
        self._chk(self.lib.jams_sim_exchange_csr(self.h, index, _p(rowptr), _p(col), _p(val)))
        return rowptr, col, val

    def iteration(self):
        """(iteration, time in ps) of the run."""
        it, t = C.c_int64(), C.c_double()
        self._chk(self.lib.jams_sim_iteration(self.h, C.byref(it), C.byref(t)))
        return it.value, t.value

    def write_checkpoint(self, path: str):
        self._chk(self.lib.jams_sim_write_checkpoint(self.h, str(path).encode()))

    def params(self):
        dt, T, seed = C.c_double(), C.c_double(), C.c_uint64()
        self._chk(self.lib.jams_sim_params(self.h, C.byref(dt), C.byref(T), C.byref(seed)))
        return dt.value, T.value, seed.value
