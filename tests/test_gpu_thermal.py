"""Thermal statistics of the GPU Heun step with its in-kernel Philox noise
(thermostats/cuda_thermostat_classical.cc:22-56, solvers/cpu_llg_heun.cc:53-81).

The FP64 trajectories equal the oracle's bit for bit when both draw the same
stream (tests/test_gpu_heun.py); these tests check the physics the stream
produces:

- non-interacting spins in a field follow the Langevin function
  <s_z> = coth(x) - 1/x, x = mus B / (kB T) (a closed form, no reference
  code involved);
- an interacting ferromagnet's time-averaged |m| at T > 0 agrees with the
  oracle's, run on an independent noise stream, within three combined
  standard errors (blocking analysis).
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from tests import systems as S

pytestmark = pytest.mark.gpu


def langevin(x):
    return 1.0 / math.tanh(x) - 1.0 / x


def block_mean_err(samples, nblocks=16):
    """Mean and standard error from nblocks contiguous blocks."""
    a = np.asarray(samples, np.float64)
    a = a[: len(a) // nblocks * nblocks].reshape(nblocks, -1).mean(axis=1)
    return a.mean(), a.std(ddof=1) / math.sqrt(nblocks)


def test_langevin_function_paramagnet_gpu(engine_factory):
    """32768 non-interacting spins, B = 10 T along z, T = 10 K, alpha = 1, the
    Gilbert prefactor form (gyro / (1 + alpha^2), sigma with the same
    denominator), dt = 1 fs: <s_z> over steps 5000-10000 within 0.01 of the
    Langevin function (the oracle's own run of this test holds the same bar)."""
    n, B, T, alpha, dt = 32768, 10.0, 10.0, 1.0, 1e-3
    mu = 2.0 * S.BOHR
    mus = np.full(n, mu)
    gyro = np.full(n, S.gyro_e() / (1 + alpha * alpha))
    e = engine_factory()
    e.set_sites(mus, gyro, np.full(n, alpha))
    e.add_zeeman(mus[:, None] * np.array([[0.0, 0.0, B]]))
    e.set_time_step(dt)
    e.set_spins(np.tile([0.0, 0.0, 1.0], (n, 1)))
    e.set_thermostat(T, 77, gilbert=True)
    e.step(5000)
    acc = []
    for _ in range(500):
        e.step(10)
        m = e.magnetisation()
        acc.append(m[2] / m[3])
    got, err = block_mean_err(acc)
    want = langevin(mu * B / (S.KB * T))
    print(f"<s_z> = {got:.5f} +- {err:.1e}, Langevin {want:.5f}")
    assert abs(got - want) < 0.01


def test_mean_m_matches_oracle_on_independent_stream(engine_factory):
    """sc 8^3 periodic, J = 1 meV (6 NN), mu_s = 2 muB, alpha = 0.5, dt = 1 fs,
    T = 8 K (about half the classical T_c of 16.7 K), from +z: |m| sampled
    every 10 steps over steps 2000-10000 on the GPU (seed 11) and in the
    oracle (seed 12) agree within three combined standard errors."""
    lat = po.LatticeSpec(S.SC_CELL, [[0, 0, 0]], [0], (8, 8, 8), (1, 1, 1))
    b = S.build(lat, S.sc_nn_rows(1.0), 2.0, 0.5, ["exchange"], dt=S.dt_ps(1.0e-15), spins="+z",
                temperature=8.0)
    sigma = po.thermostat_sigma(b.mus, b.gyro, b.alpha, b.dt)

    e = b.engine(engine_factory)
    e.set_thermostat(b.temperature, 11)
    e.step(2000)
    g = []
    for _ in range(800):
        e.step(10)
        m = e.magnetisation()
        g.append(np.linalg.norm(m[:3]) / m[3])

    o = b.oracle(nthreads=8)
    s = b.s0.copy()
    c = []
    for k in range(10000):
        o.heun_step(s, b.dt, k * b.dt, po.scale_noise(po.normals(12, k, b.n), sigma, b.temperature))
        if k >= 2000 and (k + 1) % 10 == 0:
            mv = (b.mus[:, None] * s).sum(axis=0) / b.mus.sum()
            c.append(np.linalg.norm(mv))

    mg, eg = block_mean_err(g)
    mc, ec = block_mean_err(c)
    print(f"<|m|> GPU {mg:.5f} +- {eg:.1e}, oracle {mc:.5f} +- {ec:.1e}")
    assert 0.3 < mg < 0.99
    assert abs(mg - mc) <= 3.0 * math.sqrt(eg * eg + ec * ec)


@pytest.mark.parametrize("T", [4.0, 8.0, 24.0, 40.0])
def test_mean_m_curve_across_tc(engine_factory, T):
    """<|m|>(T) of the sc 8^3 ferromagnet (J = 1 meV, classical T_c about
    16.7 K) at two temperatures below T_c and two above: the GPU's time
    average on one stream (seed 101) and the oracle's on another (seed 202),
    |m| every 10 steps over steps 2000-22000, agree within 4 combined
    standard errors from an 8-block analysis (the critical region itself,
    whose correlation times outlast such a run, is left out)."""
    lat = po.LatticeSpec(S.SC_CELL, [[0, 0, 0]], [0], (8, 8, 8), (1, 1, 1))
    b = S.build(lat, S.sc_nn_rows(1.0), 2.0, 0.5, ["exchange"], dt=S.dt_ps(1.0e-15), spins="+z",
                temperature=T)
    sigma = po.thermostat_sigma(b.mus, b.gyro, b.alpha, b.dt)
    e = b.engine(engine_factory)
    e.set_thermostat(T, 101)
    e.step(2000)
    g = []
    for _ in range(2000):
        e.step(10)
        m = e.magnetisation()
        g.append(np.linalg.norm(m[:3]) / m[3])
    o = b.oracle(nthreads=1)
    s = b.s0.copy()
    c = []
    for k in range(22000):
        o.heun_step(s, b.dt, k * b.dt, po.scale_noise(po.normals(202, k, b.n), sigma, T))
        if k >= 2000 and (k + 1) % 10 == 0:
            c.append(np.linalg.norm((b.mus[:, None] * s).sum(axis=0)) / b.mus.sum())
    mg, eg = block_mean_err(g, 8)
    mc, ec = block_mean_err(c, 8)
    print(f"T = {T} K: <|m|> GPU {mg:.5f} +- {eg:.1e}, oracle {mc:.5f} +- {ec:.1e}")
    assert abs(mg - mc) <= 4.0 * math.sqrt(eg * eg + ec * ec)
    lo, hi = {4.0: (0.85, 0.95), 8.0: (0.70, 0.82), 24.0: (0.03, 0.15), 40.0: (0.02, 0.10)}[T]
    assert lo < mg < hi
