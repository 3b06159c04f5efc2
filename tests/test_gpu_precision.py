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


@pytest.mark.parametrize("fmt", [0, 1, 2])
def test_fp32_field_error_bound(engine_factory, fmt):
    """|h32 - h64| <= 2^-22 sum_j |J_ij| |s_j| per component: the inputs are
    rounded to FP32 (2^-24 each), products and sums are exact to FP64."""
    b = S.c3(6, with_csr=True)
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
