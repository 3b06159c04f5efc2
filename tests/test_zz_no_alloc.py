"""The allocation-free magnetisation reduction (jh_reduce_magnetisation_noalloc)
through the C ABI against jh_reduce_magnetisation, the entry point it is a
variant of: the same kernels, with their partial sums in the library's static
device array instead of the context's buffers.

Two calls in a row on two contexts, the second with fewer spins, so the second
reads an array that still holds the first call's partial sums beyond its own
range; then the limit and the state checks.
"""
import ctypes as C

import numpy as np
import pytest

from jams_amd import engine as E
from oracle import pyoracle as po
from tests import systems as S

pytestmark = pytest.mark.gpu

JH_ERR_INVALID = 1  # include/jams_hip.h


def _noalloc(e):
    out = np.full(4, np.nan)
    rc = e.lib.jh_reduce_magnetisation_noalloc(e.h, out.ctypes.data_as(C.c_void_p))
    return rc, out


def _alloc(e):
    out = np.full(4, np.nan)
    rc = e.lib.jh_reduce_magnetisation(e.h, out.ctypes.data_as(C.c_void_p))
    assert rc == E.JH_OK, e.lib.jh_last_error(e.h)
    return out


def test_noalloc_magnetisation_matches_allocating_entry_point(engine_factory):
    # bcc 24^3 x 2 (27,648 spins, 108 workgroups of partial sums), then
    # sc 10^3 (1,000 spins, 4 workgroups), with random moments and spins
    results = []
    for k, b in enumerate((S.c2(24), S.c1(10))):
        rng = np.random.default_rng(40 + k)
        b.mus = b.mus * rng.uniform(0.5, 2.0, b.n)
        b.s0 = po.random_spins(7 + k, b.n)
        e = b.engine(engine_factory)
        e.step(3)  # lays the context out (the no-alloc call refuses otherwise)
        ref = _alloc(e)
        rc, got = _noalloc(e)
        assert rc == E.JH_OK, e.lib.jh_last_error(e.h)
        print(f"n={b.n}: alloc {ref.tolist()} noalloc {got.tolist()}", flush=True)
        assert np.array_equal(got, ref)
        # plain numpy check of the same sums (another order: a tolerance)
        s = e.get_spins()
        want = np.concatenate([(b.mus[:, None] * s).sum(0), [b.mus.sum()]])
        assert np.allclose(got, want, rtol=1e-12, atol=1e-12 * b.mus.sum())
        results.append(got)
    assert results[1][3] != results[0][3]


def test_noalloc_rejects_unprepared_context_and_oversized_input(engine_factory):
    b = S.c1(6)
    e = b.engine(engine_factory)
    e.set_spins(b.s0)
    # set-up changed since the last layout: refused, nothing launched
    rc, out = _noalloc(e)
    assert rc != E.JH_OK
    assert "jh_step" in e.lib.jh_last_error(e.h).decode()
    assert np.all(np.isnan(out))
    e.step(1)
    rc, out = _noalloc(e)
    assert rc == E.JH_OK
    assert np.array_equal(out, _alloc(e))
    # more spins than the static array covers: refused before any launch
    big = engine_factory()
    n = (1 << 24) + 1
    ones = np.ones(n)
    big.set_sites(ones, ones, 0.1 * ones)
    rc, out = _noalloc(big)
    assert rc == JH_ERR_INVALID
    assert "scratch limit of 16777216" in big.lib.jh_last_error(big.h).decode()
    assert np.all(np.isnan(out))
