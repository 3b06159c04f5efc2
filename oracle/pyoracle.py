"""ctypes front end of the test oracle (liboracle.so built from oracle/*.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import math
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"

TERM_EXCHANGE, TERM_ZEEMAN, TERM_UNIAXIAL, TERM_FIELD, TERM_DIPOLE = 0, 1, 2, 3, 4

_lib = None


def build() -> Path:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-B", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = C.CDLL(str(LIB_PATH))
        d = C.c_double
        _lib.or_gyromagnetic_ratio.restype = d
        _lib.or_vacuum_permeability.restype = d
        _lib.or_u01.restype = d
        _lib.or_noise_log.restype = d
        _lib.or_noise_log.argtypes = [d]
        _lib.or_setup_uniform.restype = d
        _lib.or_setup_uniform.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        _lib.or_kahan_sum.restype = d
        _lib.or_term_energy.restype = d
        _lib.or_neighbour_list.restype = C.c_int64
        _lib.or_neighbour_bruteforce.restype = C.c_int64
        _lib.or_exchange_csr.restype = C.c_int64
        _lib.or_coo_to_csr.restype = C.c_int64
        _lib.or_annulus_pairs.restype = C.c_int64
        _lib.or_site_index.restype = C.c_int64
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


GYRO = None
MU0 = None
BOHR = 0.0578838181
KB = 0.0861733326


def gyromagnetic_ratio() -> float:
    return lib().or_gyromagnetic_ratio()


def vacuum_permeability() -> float:
    return lib().or_vacuum_permeability()


# ---- RNG ------------------------------------------------------------------
def philox(ctr, key) -> list[int]:
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().or_philox4x32_10(c, k, o)
    return list(o)


def normals(seed: int, step: int, n: int) -> np.ndarray:
    out = np.empty((n, 3), np.float64)
    lib().or_normals_fill(C.c_uint64(seed), C.c_uint64(step), C.c_int64(n), _p(out))
    return out


def random_spins(seed: int, n: int) -> np.ndarray:
    out = np.empty((n, 3), np.float64)
    lib().or_random_spins(C.c_uint64(seed), C.c_int64(n), _p(out))
    return out


def noise_log(x: float) -> float:
    return lib().or_noise_log(float(x))


def noise_sincos2pi(u: float):
    c, s = C.c_double(), C.c_double()
    lib().or_noise_sincos2pi(C.c_double(u), C.byref(c), C.byref(s))
    return c.value, s.value


# ---- lattice / neighbours / CSR ------------------------------------------
class Lattice(C.Structure):
    _fields_ = [("cell", C.c_double * 9), ("nbasis", C.c_int), ("basis_frac", C.c_void_p),
                ("basis_material", C.c_void_p), ("size", C.c_int * 3), ("periodic", C.c_int * 3)]


class Template(C.Structure):
    _fields_ = [("type_i", C.c_int), ("type_j", C.c_int), ("r", C.c_double * 3),
                ("J", C.c_double * 9)]


class LatticeSpec:
    """A lattice: unit cell (rows of the matrix whose columns are a1 a2 a3),
    fractional basis positions, material ids, size, periodicity."""

    def __init__(self, cell, basis_frac, basis_material, size, periodic):
        self.cell = np.ascontiguousarray(cell, np.float64).reshape(3, 3)
        self.basis_frac = np.ascontiguousarray(basis_frac, np.float64).reshape(-1, 3)
        self.basis_material = np.ascontiguousarray(basis_material, np.int32)
        self.size = tuple(int(x) for x in size)
        self.periodic = tuple(int(bool(x)) for x in periodic)
        self.nbasis = self.basis_frac.shape[0]
        self.n = self.nbasis * self.size[0] * self.size[1] * self.size[2]

    def c(self) -> Lattice:
        L = Lattice()
        L.cell[:] = list(self.cell.reshape(9))
        L.nbasis = self.nbasis
        L.basis_frac = self.basis_frac.ctypes.data
        L.basis_material = self.basis_material.ctypes.data
        L.size[:] = list(self.size)
        L.periodic[:] = list(self.periodic)
        return L

    def positions(self) -> np.ndarray:
        """Cartesian positions (lattice constants) in site-index order."""
        i, j, k = np.meshgrid(*[np.arange(s) for s in self.size], indexing="ij")
        cells = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.float64)
        frac = (cells[:, None, :] + self.basis_frac[None, :, :]).reshape(-1, 3)
        return frac @ self.cell.T


def templates(rows) -> tuple:
    """rows: iterable of (type_i, type_j, r[3], J scalar or 9 values)."""
    rows = list(rows)
    arr = (Template * max(len(rows), 1))()
    for q, (ti, tj, r, J) in enumerate(rows):
        arr[q].type_i = ti
        arr[q].type_j = tj
        arr[q].r[:] = list(map(float, r))
        Jv = np.asarray(J, np.float64)
        if Jv.size == 1:
            Jv = float(Jv) * np.eye(3)
        arr[q].J[:] = list(Jv.reshape(9))
    return arr, len(rows)


def neighbour_list(lat: LatticeSpec, rows, tol=1e-4, radius_cutoff=100.0, max_out=None):
    T, nt = templates(rows)
    if max_out is None:
        max_out = lat.n * max(nt, 1) + 1
    pi = np.empty(max_out, np.int32)
    pj = np.empty(max_out, np.int32)
    pJ = np.empty((max_out, 9), np.float64)
    Lc = lat.c()
    m = lib().or_neighbour_list(C.byref(Lc), T, nt, C.c_double(tol), C.c_double(radius_cutoff),
                                _p(pi), _p(pj), _p(pJ), C.c_int64(max_out))
    if m < 0:
        raise ValueError(f"or_neighbour_list failed with {m}")
    return pi[:m].copy(), pj[:m].copy(), pJ[:m].copy()


def neighbour_bruteforce(lat: LatticeSpec, rows, tol=1e-4, max_out=None):
    T, nt = templates(rows)
    if max_out is None:
        max_out = lat.n * max(nt, 1) + 1
    pi = np.empty(max_out, np.int32)
    pj = np.empty(max_out, np.int32)
    pJ = np.empty((max_out, 9), np.float64)
    Lc = lat.c()
    m = lib().or_neighbour_bruteforce(C.byref(Lc), T, nt, C.c_double(tol), _p(pi), _p(pj),
                                      _p(pJ), C.c_int64(max_out))
    if m < 0:
        raise ValueError(f"or_neighbour_bruteforce failed with {m}")
    return pi[:m].copy(), pj[:m].copy(), pJ[:m].copy()


def annulus_pairs(lat: LatticeSpec, inner: float, outer: float, eps: float):
    """or_annulus_pairs: (i, j, rij (m x 3), d) sorted by (i, j, d)."""
    cap = 64 * lat.n + 1024
    while True:
        pi = np.empty(cap, np.int32)
        pj = np.empty(cap, np.int32)
        rij = np.empty((cap, 3), np.float64)
        d = np.empty(cap, np.float64)
        Lc = lat.c()
        m = lib().or_annulus_pairs(C.byref(Lc), C.c_double(inner), C.c_double(outer),
                                   C.c_double(eps), _p(pi), _p(pj), _p(rij), _p(d),
                                   C.c_int64(cap))
        if m >= 0:
            return pi[:m].copy(), pj[:m].copy(), rij[:m].copy(), d[:m].copy()
        cap *= 4


def _material(lat: LatticeSpec, site: int) -> int:
    return int(lat.basis_material[site % lat.nbasis])


def neartree_pairs(lat: LatticeSpec, rules, width: float, energy_cutoff: float):
    """hamiltonian/exchange_neartree.cc:58-139: rules (type_a, type_b, radius,
    J) in lattice constants and engine energy units; J I between sites of the
    two types in the shell (radius - w/2, radius + w/2] (tolerance w/10).
    Returns [(i, j, J)], sorted; raises on a pair found twice."""
    ntypes = int(lat.basis_material.max()) + 1
    shells = [[] for _ in range(ntypes)]
    for ta, tb, radius, J in rules:
        shells[ta].append((tb, radius, J))
        if ta != tb:
            shells[tb].append((ta, radius, J))
    eps = width / 10.0
    cand = {}
    for ta, tb, radius, J in rules:
        if radius not in cand:
            pi, pj, _, _ = annulus_pairs(lat, radius - 0.5 * width, radius + 0.5 * width, eps)
            by_i = {}
            for i, j in zip(pi.tolist(), pj.tolist()):
                by_i.setdefault(i, []).append(j)
            cand[radius] = by_i
    out = []
    for i in range(lat.n):
        seen = set()
        for tb, radius, J in shells[_material(lat, i)]:
            for j in cand[radius].get(i, []):
                if j == i or _material(lat, j) != tb:
                    continue
                if j in seen:
                    raise ValueError(f"multiple interactions between spins {i} and {j}")
                seen.add(j)
                if abs(J) > energy_cutoff:
                    out.append((i, j, J))
    return sorted(out)


def functional_pairs(lat: LatticeSpec, table, eps: float = 1e-4):
    """hamiltonian/exchange_functional.cc:203-250: table {(type_i, type_j):
    (r_cutoff, f)} with f(rij) the functional; pairs within the largest
    cutoff (annulus (eps, max]) whose own cutoff passes r - rc < eps max(r, rc).
    Returns ([(i, j, J)] with J != 0, the reference's counter)."""
    rmax = max(rc for rc, _ in table.values())
    pi, pj, rij, _ = annulus_pairs(lat, eps, rmax, eps)
    out, seen, counter = [], set(), 0
    for i, j, r3 in zip(pi.tolist(), pj.tolist(), rij.tolist()):
        if i == j:
            continue
        key = (_material(lat, i), _material(lat, j))
        if key not in table:
            continue
        rc, f = table[key]
        r = math.sqrt(r3[0] * r3[0] + r3[1] * r3[1] + r3[2] * r3[2])
        if not (r - rc) < max(abs(r), abs(rc)) * eps:
            continue
        if (i, j) in seen:
            raise ValueError(f"multiple interactions between spins {i} and {j}")
        seen.add((i, j))
        J = f(r3)
        counter += 1
        if J != 0.0:
            out.append((i, j, J))
    return sorted(out), counter


def pairs_csr(n: int, pairs):
    """CSR of scalar pairs J I (insert_interaction_scalar / _tensor, then the
    builder of or_exchange_csr)."""
    pi = np.array([p[0] for p in pairs], np.int32)
    pj = np.array([p[1] for p in pairs], np.int32)
    pJ = np.array([(np.eye(3) * p[2]).reshape(9) for p in pairs], np.float64).reshape(-1, 9)
    return exchange_csr(n, pi, pj, pJ)


def exchange_csr(n: int, pi, pj, pJ, scale=1.0, cutoff=0.0):
    pi = np.ascontiguousarray(pi, np.int32)
    pj = np.ascontiguousarray(pj, np.int32)
    pJ = np.ascontiguousarray(pJ, np.float64).reshape(-1, 9)
    cap = pi.size * 9 + 1
    rowptr = np.empty(3 * n + 1, np.int64)
    col = np.empty(cap, np.int32)
    val = np.empty(cap, np.float64)
    nnz = lib().or_exchange_csr(C.c_int64(n), C.c_int64(pi.size), _p(pi), _p(pj), _p(pJ),
                                C.c_double(scale), C.c_double(cutoff), _p(rowptr), _p(col),
                                _p(val), C.c_int64(cap))
    if nnz < 0:
        raise ValueError("or_exchange_csr failed")
    return rowptr, col[:nnz].copy(), val[:nnz].copy()


def coo_to_csr(nrows: int, row, col, val):
    row = np.ascontiguousarray(row, np.int64).copy()
    col = np.ascontiguousarray(col, np.int64).copy()
    val = np.ascontiguousarray(val, np.float64).copy()
    rowptr = np.empty(nrows + 1, np.int64)
    oc = np.empty(max(val.size, 1), np.int32)
    ov = np.empty(max(val.size, 1), np.float64)
    nnz = lib().or_coo_to_csr(C.c_int64(nrows), C.c_int64(val.size), _p(row), _p(col), _p(val),
                              _p(rowptr), _p(oc), _p(ov))
    return rowptr, oc[:nnz].copy(), ov[:nnz].copy()


def csr_is_symmetric(rowptr, col, val) -> bool:
    return bool(lib().or_csr_is_symmetric(C.c_int64(rowptr.size - 1), _p(rowptr), _p(col),
                                          _p(val)))


def csrmv(rowptr, col, val, x, nthreads=1) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float64).reshape(-1)
    y = np.empty(rowptr.size - 1, np.float64)
    lib().or_csrmv(C.c_int64(rowptr.size - 1), _p(rowptr), _p(col), _p(val), _p(x), _p(y),
                   C.c_int(nthreads))
    return y


# ---- system + Heun ----------------------------------------------------------
class System(C.Structure):
    _fields_ = [("n", C.c_int64), ("mus", C.c_void_p), ("gyro", C.c_void_p),
                ("alpha", C.c_void_p), ("rowptr", C.c_void_p), ("col", C.c_void_p),
                ("val", C.c_void_p), ("zee_dc", C.c_void_p), ("zee_ac", C.c_void_p),
                ("zee_omega", C.c_void_p), ("ani_K", C.c_void_p), ("ani_axis", C.c_void_p),
                ("ani_power", C.c_int), ("ext_field", C.c_void_p), ("nterms", C.c_int),
                ("kind", C.c_int * 8), ("nthreads", C.c_int),
                ("dip_lat", C.c_void_p), ("dip_pos", C.c_void_p), ("dip_rcut", C.c_double),
                ("dip_a", C.c_double)]


class OracleSystem:
    """Arrays of one spin system, in the reference's AoS layout, plus the
    ordered list of Hamiltonian terms."""

    def __init__(self, mus, gyro, alpha, nthreads=1):
        self.mus = np.ascontiguousarray(mus, np.float64)
        self.gyro = np.ascontiguousarray(gyro, np.float64)
        self.alpha = np.ascontiguousarray(alpha, np.float64)
        self.n = self.mus.size
        self.kinds = []
        self.rowptr = self.col = self.val = None
        self.zee_dc = self.zee_ac = self.zee_omega = None
        self.ani_K = self.ani_axis = None
        self.ani_power = 2
        self.ext = None
        self.dip = None
        self.nthreads = nthreads
        self._work = np.empty(15 * self.n, np.float64)

    def add_exchange(self, rowptr, col, val):
        self.rowptr = np.ascontiguousarray(rowptr, np.int64)
        self.col = np.ascontiguousarray(col, np.int32)
        self.val = np.ascontiguousarray(val, np.float64)
        self.kinds.append(TERM_EXCHANGE)
        return len(self.kinds) - 1

    def add_zeeman(self, dc, ac=None, omega=None):
        self.zee_dc = np.ascontiguousarray(dc, np.float64).reshape(self.n, 3)
        if ac is not None:
            self.zee_ac = np.ascontiguousarray(ac, np.float64).reshape(self.n, 3)
            self.zee_omega = np.ascontiguousarray(omega, np.float64)
        self.kinds.append(TERM_ZEEMAN)
        return len(self.kinds) - 1

    def add_uniaxial(self, K, axis, power=2):
        self.ani_K = np.ascontiguousarray(K, np.float64)
        self.ani_axis = np.ascontiguousarray(axis, np.float64).reshape(self.n, 3)
        self.ani_power = int(power)
        self.kinds.append(TERM_UNIAXIAL)
        return len(self.kinds) - 1

    def add_field(self, h):
        self.ext = np.ascontiguousarray(h, np.float64).reshape(self.n, 3)
        self.kinds.append(TERM_FIELD)
        return len(self.kinds) - 1

    def add_dipole(self, lat: "LatticeSpec", r_cutoff: float, a_m: float):
        """Dipole term by brute force (dipole_bruteforce.cc:75-120)."""
        Lc = lat.c()
        pos = np.ascontiguousarray(lat.positions(), np.float64)
        self.dip = (lat, Lc, pos, float(r_cutoff), float(a_m))
        self.kinds.append(TERM_DIPOLE)
        return len(self.kinds) - 1

    def c(self) -> System:
        S = System()
        S.n = self.n
        S.mus, S.gyro, S.alpha = _p(self.mus), _p(self.gyro), _p(self.alpha)
        S.rowptr, S.col, S.val = _p(self.rowptr), _p(self.col), _p(self.val)
        S.zee_dc, S.zee_ac, S.zee_omega = _p(self.zee_dc), _p(self.zee_ac), _p(self.zee_omega)
        S.ani_K, S.ani_axis, S.ani_power = _p(self.ani_K), _p(self.ani_axis), self.ani_power
        S.ext_field = _p(self.ext)
        S.nterms = len(self.kinds)
        S.kind[:len(self.kinds)] = self.kinds
        S.nthreads = self.nthreads
        if self.dip is not None:
            _, Lc, pos, rc, a = self.dip
            S.dip_lat = C.cast(C.pointer(Lc), C.c_void_p)
            S.dip_pos = _p(pos)
            S.dip_rcut = rc
            S.dip_a = a
        return S

    def term_field(self, term, s, t=0.0) -> np.ndarray:
        s = np.ascontiguousarray(s, np.float64).reshape(self.n, 3)
        h = np.empty((self.n, 3), np.float64)
        S = self.c()
        lib().or_term_field(C.byref(S), C.c_int(term), _p(s), C.c_double(t), _p(h))
        return h

    def fields(self, s, t=0.0) -> np.ndarray:
        s = np.ascontiguousarray(s, np.float64).reshape(self.n, 3)
        h = np.empty((self.n, 3), np.float64)
        scratch = np.empty((self.n, 3), np.float64)
        S = self.c()
        lib().or_compute_fields(C.byref(S), _p(s), C.c_double(t), _p(h), _p(scratch))
        return h

    def heun_step(self, s, dt, t0=0.0, w=None):
        """Advance s (N x 3, modified in place) by one Heun step."""
        assert s.flags.c_contiguous and s.dtype == np.float64
        S = self.c()
        wp = None if w is None else np.ascontiguousarray(w, np.float64)
        lib().or_heun_step(C.byref(S), _p(s), _p(wp), C.c_double(dt), C.c_double(t0),
                           _p(self._work))

    def rk4_step(self, s, dt, t0=0.0, w=None):
        """Advance s (N x 3, modified in place) by one RK4 step (llg-rk4-gpu)."""
        assert s.flags.c_contiguous and s.dtype == np.float64
        S = self.c()
        wp = None if w is None else np.ascontiguousarray(w, np.float64)
        lib().or_rk4_step(C.byref(S), _p(s), _p(wp), C.c_double(dt), C.c_double(t0),
                          _p(self._work))

    def term_energy(self, term, s, t=0.0) -> float:
        s = np.ascontiguousarray(s, np.float64).reshape(self.n, 3)
        S = self.c()
        return lib().or_term_energy(C.byref(S), C.c_int(term), _p(s), C.c_double(t))


def thermostat_sigma(mus, gyro, alpha, dt, gilbert=False) -> np.ndarray:
    mus, gyro, alpha = (np.ascontiguousarray(a, np.float64) for a in (mus, gyro, alpha))
    out = np.empty(mus.size, np.float64)
    lib().or_thermostat_sigma(C.c_int64(mus.size), _p(mus), _p(gyro), _p(alpha), C.c_double(dt),
                              C.c_int(int(gilbert)), _p(out))
    return out


def scale_noise(xi, sigma, T) -> np.ndarray:
    xi = np.ascontiguousarray(xi, np.float64)
    sigma = np.ascontiguousarray(sigma, np.float64)
    out = np.empty_like(xi)
    lib().or_scale_noise(C.c_int64(sigma.size), _p(xi), _p(sigma), C.c_double(T), _p(out))
    return out


def sum_moments(s, mus) -> np.ndarray:
    s = np.ascontiguousarray(s, np.float64)
    mus = np.ascontiguousarray(mus, np.float64)
    out = np.empty(3, np.float64)
    lib().or_sum_moments(C.c_int64(mus.size), _p(s), _p(mus), _p(out))
    return out


def kahan_sum(x) -> float:
    x = np.ascontiguousarray(x, np.float64).reshape(-1)
    return lib().or_kahan_sum(C.c_int64(x.size), _p(x))


def kahan_sum3(x) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
    out = np.empty(3, np.float64)
    lib().or_kahan_sum3(C.c_int64(x.shape[0]), _p(x), _p(out))
    return out


def dipole_fft_direct(lat: LatticeSpec, mus, s, r_cutoff, a_m, one_sided_open=False,
                      nthreads=1) -> np.ndarray:
    """The FFT dipole's convolution as a direct offset sum (oracle_dipole_fft.c)."""
    mus = np.ascontiguousarray(mus, np.float64)
    s = np.ascontiguousarray(s, np.float64).reshape(-1, 3)
    if mus.size != lat.n or s.shape[0] != lat.n:
        raise ValueError("mus and s must hold one entry per lattice site")
    h = np.empty_like(s)
    Lc = lat.c()
    lib().or_dipole_fft_direct(C.byref(Lc), _p(mus), _p(s), C.c_double(r_cutoff),
                               C.c_double(a_m), C.c_int(int(one_sided_open)), _p(h),
                               C.c_int(nthreads))
    return h


def dipole_bruteforce(lat: LatticeSpec, mus, s, r_cutoff, a_m, nthreads=1) -> np.ndarray:
    pos = np.ascontiguousarray(lat.positions(), np.float64)
    mus = np.ascontiguousarray(mus, np.float64)
    s = np.ascontiguousarray(s, np.float64).reshape(-1, 3)
    if mus.size != lat.n or s.shape[0] != lat.n:
        raise ValueError("mus and s must hold one entry per lattice site")
    h = np.empty_like(s)
    Lc = lat.c()
    lib().or_dipole_bruteforce(C.byref(Lc), C.c_int64(mus.size), _p(pos), _p(mus), _p(s),
                               C.c_double(r_cutoff), C.c_double(a_m), _p(h), C.c_int(nthreads))
    return h


def dipole_bruteforce_sites(lat: LatticeSpec, mus, s, r_cutoff, a_m, sites,
                            nthreads=1) -> np.ndarray:
    """or_dipole_bruteforce at the listed sites only (nsites x 3)."""
    pos = np.ascontiguousarray(lat.positions(), np.float64)
    mus = np.ascontiguousarray(mus, np.float64)
    s = np.ascontiguousarray(s, np.float64).reshape(-1, 3)
    sites = np.ascontiguousarray(sites, np.int64)
    if mus.size != lat.n or s.shape[0] != lat.n:
        raise ValueError("mus and s must hold one entry per lattice site")
    if sites.size and (sites.min() < 0 or sites.max() >= lat.n):
        raise ValueError("site index out of range")
    h = np.empty((sites.size, 3), np.float64)
    Lc = lat.c()
    lib().or_dipole_bruteforce_sites(C.byref(Lc), C.c_int64(mus.size), _p(pos), _p(mus), _p(s),
                                     C.c_double(r_cutoff), C.c_double(a_m), C.c_int64(sites.size),
                                     _p(sites), _p(h), C.c_int(nthreads))
    return h
