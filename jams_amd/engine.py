"""ctypes binding of the gfx950 Heun-LLG engine (``include/jams_hip.h``).

The engine library is built in-tree by ``jams_amd.build`` into
``jams_amd/lib/libjamship.so``.  There is no fallback: if the library is
missing or no HIP device is present, creating an :class:`Engine` raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_DIR = Path(__file__).resolve().parent / "lib"
ENGINE_LIB = LIB_DIR / "libjamship.so"

JH_OK = 0
PREC_FP64 = 0
PREC_FP32_FIELD = 1
INTEGRATOR_HEUN = 0
INTEGRATOR_RK4 = 1
XFMT_AUTO, XFMT_TYPED, XFMT_BLOCK, XFMT_PATTERN, XFMT_PATTERN_GLOBAL, XFMT_STENCIL = 0, 1, 2, 3, 4, 5
XFMT_LINE = 6
STEP_LAUNCH_TWO, STEP_LAUNCH_PERSISTENT = 0, 1
TERM_EXCHANGE, TERM_ZEEMAN, TERM_UNIAXIAL, TERM_DIPOLE = 0, 1, 2, 3

# every entry point declared in include/jams_hip.h
ABI = {
    "jh_version": (C.c_int, []),
    "jh_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "jh_destroy": (C.c_int, [C.c_void_p]),
    "jh_last_error": (C.c_char_p, [C.c_void_p]),
    "jh_set_sites": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]),
    "jh_set_lattice": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    "jh_set_spins": (C.c_int, [C.c_void_p, C.c_void_p]),
    "jh_get_spins": (C.c_int, [C.c_void_p, C.c_void_p]),
    "jh_set_time_step": (C.c_int, [C.c_void_p, C.c_double]),
    "jh_add_exchange_csr": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.POINTER(C.c_int)]),
    "jh_add_exchange_list": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_int64, C.c_void_p, C.POINTER(C.c_int)]),
    "jh_add_zeeman": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.POINTER(C.c_int)]),
    "jh_add_uniaxial": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                  C.POINTER(C.c_int)]),
    "jh_add_dipole_fft": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                    C.c_void_p, C.c_double, C.c_double, C.POINTER(C.c_int)]),
    "jh_num_terms": (C.c_int, [C.c_void_p]),
    "jh_term_kind": (C.c_int, [C.c_void_p, C.c_int]),
    "jh_set_thermostat": (C.c_int, [C.c_void_p, C.c_double, C.c_uint64, C.c_int]),
    "jh_set_noise": (C.c_int, [C.c_void_p, C.c_void_p]),
    "jh_get_noise": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p]),
    "jh_set_precision": (C.c_int, [C.c_void_p, C.c_int]),
    "jh_set_integrator": (C.c_int, [C.c_void_p, C.c_int]),
    "jh_set_exchange_format": (C.c_int, [C.c_void_p, C.c_int]),
    "jh_step": (C.c_int, [C.c_void_p, C.c_int64]),
    "jh_get_iteration": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
    "jh_set_iteration": (C.c_int, [C.c_void_p, C.c_int64]),
    "jh_compute_term_field": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_void_p]),
    "jh_compute_fields": (C.c_int, [C.c_void_p, C.c_double, C.c_void_p]),
    "jh_term_energy": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.POINTER(C.c_double)]),
    "jh_reduce_magnetisation": (C.c_int, [C.c_void_p, C.c_void_p]),
    "jh_reduce_magnetisation_noalloc": (C.c_int, [C.c_void_p, C.c_void_p]),
    "jh_time_steps": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_double)]),
    "jh_time_phase": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "jh_phase_bytes": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double),
                                 C.POINTER(C.c_double)]),
    "jh_set_replicas": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "jh_reduce_magnetisation_replicas": (C.c_int, [C.c_void_p, C.c_void_p]),
    "jh_set_graph_steps": (C.c_int, [C.c_void_p, C.c_int]),
    "jh_set_block_size": (C.c_int, [C.c_void_p, C.c_int]),
    "jh_set_step_launch": (C.c_int, [C.c_void_p, C.c_int]),
    "jh_autotune": (C.c_int, [C.c_void_p, C.c_int, C.c_char_p, C.c_int64]),
    "jh_time_term": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "jh_term_bytes": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double)]),
    "jh_describe": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int64]),
    "jh_synchronize": (C.c_int, [C.c_void_p]),
}

_lib = None


def load_library(path: Path | None = None) -> C.CDLL:
    """Load libjamship.so and declare its signatures.  Raises if absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else ENGINE_LIB
    if not p.exists():
        raise RuntimeError(f"jams_amd engine library not built: {p} (run __graft_entry__.build())")
    lib = C.CDLL(str(p))
    for name, (res, args) in ABI.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


class EngineError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"jams_hip error {code}: {msg}")
        self.code = code


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a, shape=None) -> np.ndarray:
    x = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        x = x.reshape(shape)
    return x


class Engine:
    """One context on one GPU (see include/jams_hip.h for the semantics)."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        rc = self.lib.jh_create(device, C.byref(h))
        if rc != JH_OK:
            raise EngineError(rc, self.lib.jh_last_error(None).decode())
        self.h = h
        self.n = 0
        self._keep = []

    def close(self):
        if getattr(self, "h", None):
            self.lib.jh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc: int):
        if rc != JH_OK:
            raise EngineError(rc, self.lib.jh_last_error(self.h).decode())

    # ---- set-up -------------------------------------------------------
    def set_sites(self, mus, gyro, alpha):
        mus, gyro, alpha = _f64(mus), _f64(gyro), _f64(alpha)
        self.n = int(mus.size)
        self._chk(self.lib.jh_set_sites(self.h, self.n, _ptr(mus), _ptr(gyro), _ptr(alpha)))

    def set_lattice(self, size, periodic, nbasis):
        """Lattice shape: spin ((x Ly + y) Lz + z) nbasis + b (optional; it
        lets a crystal's exchange run as a stencil over LDS tiles)."""
        sz = np.ascontiguousarray(size, dtype=np.int32).reshape(3)
        pe = np.ascontiguousarray([1 if p else 0 for p in periodic], dtype=np.int32)
        self._chk(self.lib.jh_set_lattice(self.h, _ptr(sz), _ptr(pe), int(nbasis)))

    def set_spins(self, s):
        s = _f64(s, (self.n, 3))
        self._chk(self.lib.jh_set_spins(self.h, _ptr(s)))

    def get_spins(self) -> np.ndarray:
        out = np.empty((self.n, 3), np.float64)
        self._chk(self.lib.jh_get_spins(self.h, _ptr(out)))
        return out

    def set_time_step(self, dt_ps: float):
        self._chk(self.lib.jh_set_time_step(self.h, float(dt_ps)))

    def add_exchange_csr(self, rowptr, col, val) -> int:
        rowptr = np.ascontiguousarray(rowptr, np.int64)
        col = np.ascontiguousarray(col, np.int32)
        val = _f64(val)
        tid = C.c_int(-1)
        self._chk(self.lib.jh_add_exchange_csr(self.h, rowptr.size - 1, val.size, _ptr(rowptr),
                                               _ptr(col), _ptr(val), C.byref(tid)))
        return tid.value

    def add_exchange_list(self, pi, pj, vidx, values9) -> int:
        pi = np.ascontiguousarray(pi, np.int32)
        pj = np.ascontiguousarray(pj, np.int32)
        vidx = np.ascontiguousarray(vidx, np.int32)
        values9 = _f64(values9).reshape(-1, 9)
        tid = C.c_int(-1)
        self._chk(self.lib.jh_add_exchange_list(self.h, pi.size, _ptr(pi), _ptr(pj), _ptr(vidx),
                                                values9.shape[0], _ptr(values9), C.byref(tid)))
        return tid.value

    def add_zeeman(self, dc, ac=None, omega=None) -> int:
        dc = _f64(dc, (self.n, 3))
        ac = None if ac is None else _f64(ac, (self.n, 3))
        omega = None if omega is None else _f64(omega)
        tid = C.c_int(-1)
        self._chk(self.lib.jh_add_zeeman(self.h, _ptr(dc), _ptr(ac), _ptr(omega), C.byref(tid)))
        return tid.value

    def add_uniaxial(self, K, axis, power=2) -> int:
        K = _f64(K)
        axis = _f64(axis, (self.n, 3))
        tid = C.c_int(-1)
        self._chk(self.lib.jh_add_uniaxial(self.h, _ptr(K), _ptr(axis), int(power), C.byref(tid)))
        return tid.value

    def add_dipole_fft(self, size, periodic, nbasis, cell, basis_frac, a_m, r_cutoff) -> int:
        size = np.ascontiguousarray(size, np.int32)
        periodic = np.ascontiguousarray(periodic, np.int32)
        cell = _f64(cell).reshape(9)
        basis_frac = _f64(basis_frac).reshape(-1, 3)
        tid = C.c_int(-1)
        self._chk(self.lib.jh_add_dipole_fft(self.h, _ptr(size), _ptr(periodic), int(nbasis),
                                             _ptr(cell), _ptr(basis_frac), float(a_m),
                                             float(r_cutoff), C.byref(tid)))
        return tid.value

    def set_thermostat(self, temperature: float, seed: int, gilbert: bool = False):
        self._chk(self.lib.jh_set_thermostat(self.h, float(temperature), int(seed) & (2**64 - 1),
                                             int(bool(gilbert))))

    def set_noise(self, w):
        w = _f64(w, (self.n, 3))
        self._chk(self.lib.jh_set_noise(self.h, _ptr(w)))

    def get_noise(self, step: int) -> np.ndarray:
        out = np.empty((self.n, 3), np.float64)
        self._chk(self.lib.jh_get_noise(self.h, int(step), _ptr(out)))
        return out

    def set_precision(self, mode: int):
        self._chk(self.lib.jh_set_precision(self.h, int(mode)))

    def set_integrator(self, integrator: int):
        """JH_INTEGRATOR_HEUN (0, llg-heun-gpu) or JH_INTEGRATOR_RK4 (1, llg-rk4-gpu)."""
        self._chk(self.lib.jh_set_integrator(self.h, int(integrator)))

    def set_exchange_format(self, fmt: int):
        self._chk(self.lib.jh_set_exchange_format(self.h, int(fmt)))

    # ---- running ------------------------------------------------------
    def step(self, nsteps: int = 1):
        self._chk(self.lib.jh_step(self.h, int(nsteps)))

    def iteration(self):
        it = C.c_int64()
        t = C.c_double()
        self._chk(self.lib.jh_get_iteration(self.h, C.byref(it), C.byref(t)))
        return it.value, t.value

    def set_iteration(self, it: int):
        self._chk(self.lib.jh_set_iteration(self.h, int(it)))

    def term_field(self, term: int, t: float = 0.0) -> np.ndarray:
        out = np.empty((self.n, 3), np.float64)
        self._chk(self.lib.jh_compute_term_field(self.h, int(term), float(t), _ptr(out)))
        return out

    def fields(self, t: float = 0.0) -> np.ndarray:
        out = np.empty((self.n, 3), np.float64)
        self._chk(self.lib.jh_compute_fields(self.h, float(t), _ptr(out)))
        return out

    def term_energy(self, term: int, t: float = 0.0) -> float:
        e = C.c_double()
        self._chk(self.lib.jh_term_energy(self.h, int(term), float(t), C.byref(e)))
        return e.value

    def magnetisation(self) -> np.ndarray:
        out = np.empty(4, np.float64)
        self._chk(self.lib.jh_reduce_magnetisation(self.h, _ptr(out)))
        return out

    def time_steps(self, nsteps: int) -> float:
        ms = C.c_double()
        self._chk(self.lib.jh_time_steps(self.h, int(nsteps), C.byref(ms)))
        return ms.value

    def time_phase(self, phase: int, reps: int) -> float:
        ms = C.c_double()
        self._chk(self.lib.jh_time_phase(self.h, int(phase), int(reps), C.byref(ms)))
        return ms.value

    def phase_bytes(self, phase: int):
        a, b = C.c_double(), C.c_double()
        self._chk(self.lib.jh_phase_bytes(self.h, int(phase), C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_replicas(self, offsets, temperatures, seeds):
        """Declare replicas: spins [offsets[r], offsets[r+1]) with their own
        temperature and noise seed."""
        off = np.ascontiguousarray(offsets, np.int64)
        T = np.ascontiguousarray(temperatures, np.float64)
        sd = np.ascontiguousarray(np.asarray(seeds, np.uint64))
        if T.size != off.size - 1 or sd.size != off.size - 1:
            raise ValueError("one temperature and one seed per replica")
        self._chk(self.lib.jh_set_replicas(self.h, int(T.size), _ptr(off), _ptr(T), _ptr(sd)))
        self.nrep = int(T.size)

    def replica_magnetisation(self) -> np.ndarray:
        out = np.empty((self.nrep, 4), np.float64)
        self._chk(self.lib.jh_reduce_magnetisation_replicas(self.h, _ptr(out)))
        return out

    def set_graph_steps(self, steps: int):
        self._chk(self.lib.jh_set_graph_steps(self.h, int(steps)))

    def set_block_size(self, threads: int):
        self._chk(self.lib.jh_set_block_size(self.h, int(threads)))

    def set_step_launch(self, mode: int):
        """STEP_LAUNCH_TWO (predictor and corrector launches) or
        STEP_LAUNCH_PERSISTENT (one launch per step with a grid barrier)."""
        self._chk(self.lib.jh_set_step_launch(self.h, int(mode)))

    def autotune(self, steps: int = 20) -> str:
        buf = C.create_string_buffer(4096)
        self._chk(self.lib.jh_autotune(self.h, int(steps), buf, 4096))
        return buf.value.decode()

    def num_terms(self) -> int:
        return int(self.lib.jh_num_terms(self.h))

    def term_kind(self, term: int) -> int:
        return int(self.lib.jh_term_kind(self.h, int(term)))

    def time_term(self, term: int, reps: int) -> float:
        ms = C.c_double(0.0)
        self._chk(self.lib.jh_time_term(self.h, int(term), int(reps), C.byref(ms)))
        return ms.value

    def term_bytes(self, term: int) -> float:
        b = C.c_double(0.0)
        self._chk(self.lib.jh_term_bytes(self.h, int(term), C.byref(b)))
        return b.value

    def describe(self) -> str:
        buf = C.create_string_buffer(1024)
        self._chk(self.lib.jh_describe(self.h, buf, 1024))
        return buf.value.decode()

    def synchronize(self):
        self._chk(self.lib.jh_synchronize(self.h))
