"""The FP32-field precision form against the FP64 path (the oracle for it):
FP32 storage of spins and tensors in the exchange matvec, exact products
accumulated in FP64; Heun state, renormalisation and noise in FP64."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests import systems as S

pytestmark = pytest.mark.gpu


def angle(a, b):
    c = np.clip(np.sum(a * b, axis=1) / (np.linalg.norm(a, axis=1) * np.linalg.norm(b, axis=1)), -1, 1)
    return np.arccos(c)


def open_film():
    """sc 12 x 10 x 4 film, all boundaries open: rows of 3 to 6 neighbours, so
    most slices mix tensor sequences (the non-uniform path)."""
    lat = po.LatticeSpec(np.eye(3), [[0, 0, 0]], [0], (12, 10, 4), (0, 0, 0))
    return S.build(lat, S.sc_nn_rows(2.0), 2.2, 0.02, ["exchange"], dt=1e-3, spins="random", seed=4)


def open_bcc_dmi():
    """bcc 5^3 x 2 with the five-shell DMI template, open in z only."""
    return S.build(S.bcc_lattice(5, (1, 1, 0)), S.bcc_five_shell_dmi_rows(), 2.0, 0.05,
                   ["exchange"], dt=S.dt_ps(0.5e-15), spins="random", seed=3)


def bcc_aniso_dmi(L=6, with_csr=True):
    """bcc L^3 x 2, the five-shell template with each J tensor's zz entry
    scaled by 0.8: not isotropic + DMI, so no pair programs (the plain
    wave-uniform path where row patterns apply)."""
    rows = []
    for ti, tj, r, J in S.bcc_five_shell_dmi_rows():
        J = np.array(J, np.float64).reshape(3, 3).copy()
        J[2, 2] *= 0.8
        rows.append((ti, tj, r, J))
    return S.build(S.bcc_lattice(L), rows, 2.0, 0.05, ["exchange"], dt=S.dt_ps(0.5e-15),
                   spins="random", seed=3, with_csr=with_csr)


SYSTEMS = {
    "bcc_dmi_periodic": lambda: S.c3(6, with_csr=True),
    "bcc_aniso_dmi": bcc_aniso_dmi,
    "bcc_two_shell_diag": lambda: S.c2(6),
    "sc_open_film": open_film,
    "bcc_dmi_open_z": open_bcc_dmi,
}


@pytest.mark.parametrize("name", list(SYSTEMS))
@pytest.mark.parametrize("fmt", [0, 1, 2])
def test_fp32_field_error_bound(engine_factory, fmt, name):
    """|h32 - h64| <= 2^-22 sum_j |J_ij| |s_j| per component: the inputs are
    rounded to FP32 (2^-24 each), products and sums are exact to FP64 (in
    whatever order the form sums a row)."""
    b = SYSTEMS[name]()
    s = po.random_spins(21, b.n)
    e64 = b.engine(engine_factory, prec=0, fmt=fmt)
    e32 = b.engine(engine_factory, prec=1, fmt=fmt)
    e64.set_spins(s)
    e32.set_spins(s)
    h64, h32 = e64.term_field(0), e32.term_field(0)
    rowptr, col, val = b.csr
    bound = po.csrmv(rowptr, col, np.abs(val), np.abs(s).reshape(-1)).reshape(-1, 3) * 2.0**-22
    assert np.all(np.abs(h32 - h64) <= bound)
    assert np.abs(h32 - h64).max() > 0  # it is the FP32 path


def test_fp32_field_trajectory_angle_c3_small(engine_factory):
    """bcc 8^3 x 2 with five shells + DMI at 100 K, same noise stream:
    per-spin angle between the FP32-field and FP64 trajectories after 2000
    steps stays below 1e-5 rad."""
    b = S.c3(8, with_csr=True)
    es = []
    for prec in (0, 1):
        e = b.engine(engine_factory, prec=prec)
        e.set_thermostat(b.temperature, 5)
        e.step(2000)
        es.append(e.get_spins())
    ang = angle(es[0], es[1])
    dm = np.linalg.norm(es[0].mean(axis=0) - es[1].mean(axis=0))
    print(f"FP32-field vs FP64 after 2000 steps: max angle {ang.max():.3e} rad, "
          f"rms {np.sqrt((ang ** 2).mean()):.3e}, |dm| {dm:.3e}")
    assert ang.max() < 1e-5


def test_fp32_uniform_slices_are_used(engine_factory):
    """Periodic crystals large enough for row patterns (bcc 16^3 x 2 with five
    shells, bcc 8^3 x 2 with two) get the wave-uniform tensor variant in every
    slice, with pair programs where the tensors are isotropic + DMI and
    without where they are isotropic alone or anisotropic; an open film, whose every slice mixes surface
    and interior rows, keeps the per-lane variant; FP64 keeps the reference's
    row order."""
    e = S.c3(16).engine(engine_factory, prec=1, use_list=True)
    d = e.describe()
    assert "exchange=pattern-uniform-pairs-dmi" in d and "uniform_slices=128/128" in d, d
    assert "paired_slices=128 " in d, d
    e = S.c2(8).engine(engine_factory, prec=1)
    d = e.describe()
    assert "exchange=pattern-uniform-diag" in d and "uniform_slices=16/16" in d, d
    e = bcc_aniso_dmi(16, with_csr=False).engine(engine_factory, prec=1, use_list=True)
    d = e.describe()
    assert "exchange=pattern-uniform-full" in d and "paired_slices=0 " in d, d
    e = open_film().engine(engine_factory, prec=1)
    d = e.describe()
    assert "exchange=pattern-diag " in d, d
    e = S.c3(16).engine(engine_factory, prec=0, use_list=True)
    assert "exchange=pattern-full " in e.describe()
