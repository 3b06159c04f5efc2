"""The benchmark workloads as JAMS config text (libconfig syntax), built the
way a JAMS user would write them, plus their seeded synthetic ingredients.

  C1  simple-cubic 16^3 periodic, NN J = 1 meV, B = (0,0,1) T, alpha 0.1,
      dt 1 fs, T = 0, random spins (seed 1), 1000 steps
  C2  bcc 64^3 x 2, J1 = 3 meV (8), J2 = 1.5 meV (6), K1 = 0.05 meV z,
      alpha 0.05, dt 0.5 fs, T = 300 K, spins +z, 10^4 steps
  C3  bcc 128^3 x 2, five shells (58 neighbours), J_n = 3 exp(-r/a) meV with
      DMI 0.1 J_n d (seeded unit d per bond pair), T = 100 K, dt 0.5 fs
  C4  256 independent replicas, sizes drawn from a seed: simple-cubic L^3
      or bcc (L/2)^3 x 2 with L in 8..48, NN J = 1 meV, T_r = 1000 r/255 K,
      B_r = (0, 0, 2 r/255) T, alpha 0.1, dt 1 fs, 5000 steps
  C5  simple-cubic film 256 x 256 x 8 open, NN J = 2 meV + dipole-fft,
      mu_s = 2.2 muB, a = 0.3 nm, alpha 0.02, dt 1 fs, T = 0, 2000 steps

Values the workload descriptions leave open are fixed here and recorded:
mu_s = 2.0 muB where unstated, alpha = 0.05 for C3, the DMI convention
s_i.[D]x s_j = D.(s_i x s_j) with d(-r) = -d(r) so that J(-r) = J(r)^T.
"""
from __future__ import annotations

import itertools
import math

import numpy as np


def _vec(v) -> str:
    return "[" + ", ".join(repr(float(x)) for x in v) + "]"


def _common(materials: str, unitcell: str, lattice: str, hamiltonians: str, t_step: float,
            t_max: float, temperature: float, seed: int, precision: str = "fp64",
            extra: str = "") -> str:
    return f"""
materials = ( {materials} );
unitcell : {{ {unitcell} }};
lattice : {{ {lattice} }};
hamiltonians = ( {hamiltonians} );
solver : {{ module = "llg-heun-gpu"; t_step = {t_step!r}; t_max = {t_max!r}; precision = "{precision}"; }};
physics : {{ temperature = {temperature!r}; }};
sim : {{ seed = {seed}; }};
{extra}
"""


def sc_neighbour_rows(J: float):
    return [("A", "A", v, J) for v in ([1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0],
                                        [0, 0, 1], [0, 0, -1])]


def bcc_two_shell_rows(J1: float, J2: float):
    rows = [("A", "A", [sx * 0.5, sy * 0.5, sz * 0.5], J1)
            for sx, sy, sz in itertools.product((1, -1), repeat=3)]
    return rows + sc_neighbour_rows(J2)


def bcc_shell_vectors():
    """The first five bcc shells (8, 6, 12, 24, 8 = 58 vectors), lattice
    constants, at r/a = sqrt(3)/2, 1, sqrt(2), sqrt(11)/2, sqrt(3)."""
    vecs = [(0.5 * a, 0.5 * b, 0.5 * c) for a, b, c in itertools.product((1, -1), repeat=3)]
    for ax in range(3):
        for sg in (1.0, -1.0):
            v = [0.0, 0.0, 0.0]
            v[ax] = sg
            vecs.append(tuple(v))
    for a, b in ((0, 1), (0, 2), (1, 2)):
        for sa, sb in itertools.product((1.0, -1.0), repeat=2):
            v = [0.0, 0.0, 0.0]
            v[a], v[b] = sa, sb
            vecs.append(tuple(v))
    for ax in range(3):
        others = [q for q in range(3) if q != ax]
        for s0, s1, s2 in itertools.product((1, -1), repeat=3):
            v = [0.0, 0.0, 0.0]
            v[ax] = 1.5 * s0
            v[others[0]] = 0.5 * s1
            v[others[1]] = 0.5 * s2
            vecs.append(tuple(v))
    vecs += [(float(a), float(b), float(c)) for a, b, c in itertools.product((1, -1), repeat=3)]
    assert len(vecs) == 58 and len(set(vecs)) == 58
    return vecs


def dmi_matrix(D):
    """[D]x with s_i . [D]x s_j = D . (s_i x s_j)."""
    dx, dy, dz = D
    return np.array([[0.0, dz, -dy], [-dz, 0.0, dx], [dy, -dx, 0.0]])


def bcc_five_shell_dmi_rows(seed: int = 7, J0: float = 3.0, dmi: float = 0.1):
    """J_n = J0 exp(-r/a) meV plus D = dmi J_n d, d a seeded random unit
    vector per bond pair with d(-r) = -d(r), so J(-r) = J(r)^T exactly."""
    rng = np.random.default_rng(seed)
    chosen = {}
    rows = []
    for v in bcc_shell_vectors():
        neg = tuple(-x for x in v)
        if neg in chosen:
            d = -chosen[neg]
        else:
            z = rng.uniform(-1.0, 1.0)
            phi = rng.uniform(0.0, 2.0 * math.pi)
            r = math.sqrt(1.0 - z * z)
            d = np.array([r * math.cos(phi), r * math.sin(phi), z])
            chosen[v] = d
        Jn = J0 * math.exp(-math.sqrt(sum(x * x for x in v)))
        T = Jn * np.eye(3) + dmi_matrix(dmi * Jn * d)
        rows.append(("A", "A", list(v), T.reshape(9)))
    return rows


def interactions_text(rows) -> str:
    out = []
    for ti, tj, r, J in rows:
        Jv = np.asarray(J, np.float64).reshape(-1)
        jt = repr(float(Jv[0])) if Jv.size == 1 else _vec(Jv)
        out.append(f'("{ti}", "{tj}", {_vec(r)}, {jt})')
    return ",\n    ".join(out)


def sc_config(L=(16, 16, 16), periodic=(True, True, True), J=1.0, moment=2.0, alpha=0.1,
              field=(0.0, 0.0, 1.0), t_step=1.0e-15, steps=1000, temperature=0.0, seed=1,
              spin="\"random\"", a=0.3e-9, dipole_rcut=None, precision="fp64", monitors=""):
    per = ", ".join("true" if p else "false" for p in periodic)
    ham = [f'{{ module = "exchange"; energy_units = "meV"; symops = false; '
           f'interactions = ( {interactions_text(sc_neighbour_rows(J))} ); }}']
    if field is not None:
        ham.append(f'{{ module = "applied-field"; field = {_vec(field)}; }}')
    if dipole_rcut is not None:
        ham.append(f'{{ module = "dipole-fft"; r_cutoff = {dipole_rcut!r}; }}')
    mat = f'{{ name = "A"; moment = {moment!r}; alpha = {alpha!r}; gyro = 1.0; spin = {spin}; }}'
    uc = (f"parameter = {a!r}; basis = ([1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]); "
          f'positions = ( ("A", [0.0, 0.0, 0.0]) ); symops = false;')
    lat = f"size = [{L[0]}, {L[1]}, {L[2]}]; periodic = [{per}];"
    return _common(mat, uc, lat, ",\n  ".join(ham), t_step, t_step * steps, temperature, seed,
                   precision, monitors)


def bcc_config(L=64, rows=None, moment=2.0, alpha=0.05, K=0.05, t_step=0.5e-15, steps=10000,
               temperature=300.0, seed=1, spin="[0.0, 0.0, 1.0]", precision="fp64", monitors=""):
    if rows is None:
        rows = bcc_two_shell_rows(3.0, 1.5)
    ham = [f'{{ module = "exchange"; energy_units = "meV"; symops = false; '
           f'interactions = ( {interactions_text(rows)} ); }}']
    if K is not None:
        ham.append(f'{{ module = "uniaxial"; order = "K1"; energy_units = "meV"; '
                   f'anisotropies = ( ("A", [0.0, 0.0, 1.0], {K!r}) ); }}')
    mat = f'{{ name = "A"; moment = {moment!r}; alpha = {alpha!r}; gyro = 1.0; spin = {spin}; }}'
    uc = ('parameter = 0.3e-9; basis = ([1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]); '
          'positions = ( ("A", [0.0, 0.0, 0.0]), ("A", [0.5, 0.5, 0.5]) ); symops = false;')
    lat = f"size = [{L}, {L}, {L}]; periodic = [true, true, true];"
    return _common(mat, uc, lat, ",\n  ".join(ham), t_step, t_step * steps, temperature, seed,
                   precision, monitors)


def c1(L: int = 16, **kw) -> str:
    return sc_config(L=(L, L, L), **kw)


def c2(L: int = 64, **kw) -> str:
    return bcc_config(L=L, **kw)


def c3(L: int = 128, precision="fp64", **kw) -> str:
    kw.setdefault("temperature", 100.0)
    kw.setdefault("spin", '"random"')
    kw.setdefault("seed", 3)
    return bcc_config(L=L, rows=bcc_five_shell_dmi_rows(), K=None, precision=precision, **kw)


def c5(Lxy: int = 256, Lz: int = 8, precision="fp64", steps=2000, **kw) -> str:
    # r_cutoff = the longest diagonal of the supercell: every pair interacts
    rcut = math.sqrt(Lxy * Lxy * 2 + Lz * Lz)
    return sc_config(L=(Lxy, Lxy, Lz), periodic=(False, False, False), J=2.0, moment=2.2,
                     alpha=0.02, field=None, t_step=1e-15, steps=steps, temperature=0.0, seed=5,
                     dipole_rcut=rcut, precision=precision, **kw)


def bcc_nn_rows(J: float):
    return [("A", "A", [sx * 0.5, sy * 0.5, sz * 0.5], J)
            for sx, sy, sz in itertools.product((1, -1), repeat=3)]


def c4(nrep: int = 256, seed: int = 4, precision: str = "fp64", steps: int = 5000):
    """The replica batch: a list of (config text, temperature K, noise seed,
    applied field T, description).  The lattice kind and L are drawn from
    numpy's default_rng(seed); replica r has T = 1000 r/(nrep-1) K and
    B = (0, 0, 2 r/(nrep-1)) T; its spins start random from its own seed."""
    rng = np.random.default_rng(seed)
    out = []
    for r in range(nrep):
        kind = int(rng.integers(0, 2))
        L = int(rng.integers(8, 49))
        frac = r / (nrep - 1) if nrep > 1 else 0.0
        T, B = 1000.0 * frac, (0.0, 0.0, 2.0 * frac)
        rseed = seed * 100000 + r
        if kind == 0:
            text = sc_config(L=(L, L, L), J=1.0, moment=2.0, alpha=0.1, field=None, t_step=1e-15,
                             steps=steps, temperature=T, seed=rseed, precision=precision)
            desc = f"sc {L}^3"
        else:
            text = bcc_config(L=L // 2, rows=bcc_nn_rows(1.0), moment=2.0, alpha=0.1, K=None,
                              t_step=1e-15, steps=steps, temperature=T, seed=rseed,
                              spin='"random"', precision=precision)
            desc = f"bcc {L // 2}^3 x 2"
        out.append((text, T, rseed, B, desc))
    return out
