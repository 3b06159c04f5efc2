"""GPU parity of the RK4 integrator (llg-rk4-gpu, solvers/cuda_llg_rk4.cu over
solvers/cuda_rk4_base.cu) against the oracle's or_rk4_step, through the C ABI.

Bar: FP64 trajectories equal the oracle bit for bit where the arithmetic is
the same IEEE sequence (static fields, K1); with the ac Zeeman term (cos on
the device vs libm) within 1e-10.  The FP32-field form stays within the
precision form's angle bound of FP64.  Graph replays equal eager launches.
"""
import numpy as np
import pytest

from jams_amd import engine as E
from oracle import pyoracle as po
from tests import systems as S

pytestmark = pytest.mark.gpu


def rk4_oracle(b, nsteps, noise=None, o=None):
    o = o or b.oracle(nthreads=8)
    s = b.s0.copy()
    for k in range(nsteps):
        o.rk4_step(s, b.dt, k * b.dt, None if noise is None else noise(k))
    return s


def rk4_engine(b, make, prec=0):
    e = b.engine(make, prec=prec)
    e.set_integrator(E.INTEGRATOR_RK4)
    return e


def test_c1_rk4_trajectory_bitwise(engine_factory):
    """C1 lattice 16^3 (J = 1 meV, B = z, alpha = 0.1), T = 0, 500 RK4 steps."""
    b = S.c1()
    e = rk4_engine(b, engine_factory)
    e.step(500)
    got = e.get_spins()
    ref = rk4_oracle(b, 500)
    print(f"C1 RK4 500 steps: max |ds| = {np.abs(got - ref).max():.3e}")
    assert np.array_equal(got, ref)
    assert "integrator=rk4" in e.describe()


@pytest.mark.parametrize("fmt", [1, 2, 3])
def test_rk4_formats_bitwise(engine_factory, fmt):
    b = S.c1(8)
    e = b.engine(engine_factory, fmt=fmt)
    e.set_integrator(E.INTEGRATOR_RK4)
    e.step(40)
    assert np.array_equal(e.get_spins(), rk4_oracle(b, 40))


def test_c2_small_rk4_injected_noise_bitwise(engine_factory):
    """bcc 8^3 x 2, two shells + K1, T = 300 K, the oracle's noise fed in per
    step (one draw serves the four stages): 60 steps."""
    b = S.c2(8)
    sigma = po.thermostat_sigma(b.mus, b.gyro, b.alpha, b.dt)
    noise = lambda k: po.scale_noise(po.normals(7, k, b.n), sigma, b.temperature)
    e = rk4_engine(b, engine_factory)
    for k in range(60):
        e.set_noise(noise(k))
        e.step(1)
    got = e.get_spins()
    ref = rk4_oracle(b, 60, noise)
    print(f"C2 small RK4 injected noise: max |ds| = {np.abs(got - ref).max():.3e}")
    assert np.array_equal(got, ref)


def test_c2_small_rk4_generated_noise_bitwise(engine_factory):
    b = S.c2(8)
    e = rk4_engine(b, engine_factory)
    e.set_thermostat(b.temperature, 42)
    e.step(100)
    sigma = po.thermostat_sigma(b.mus, b.gyro, b.alpha, b.dt)
    noise = lambda k: po.scale_noise(po.normals(42, k, b.n), sigma, b.temperature)
    ref = rk4_oracle(b, 100, noise)
    assert np.array_equal(e.get_spins(), ref)


def test_rk4_ac_zeeman_and_k2_within_tolerance(engine_factory):
    """The general kernel (ac Zeeman cos(omega t) at the four stage times, K
    with power 4): within 1e-10 of the oracle over 200 steps (cos and pow are
    the device's and libm's, last-ulp parity unpinned)."""
    lat = po.LatticeSpec(np.eye(3), [[0, 0, 0]], [0], (6, 5, 4), (1, 1, 0))
    b = S.build(lat, S.sc_nn_rows(1.5), 2.0, 0.1, ["exchange"], dt=1e-3, spins="random", seed=9)
    n = b.n
    dc = b.mus[:, None] * np.array([[0.0, 0.0, 0.5]])
    ac = b.mus[:, None] * np.array([[2.0, 0.0, 0.0]])
    om = np.full(n, 40.0)
    K = np.full(n, 0.3)
    ax = np.tile([0.0, 0.6, 0.8], (n, 1))
    e = engine_factory()
    e.set_sites(b.mus, b.gyro, b.alpha)
    e.add_exchange_csr(*b.csr)
    e.add_zeeman(dc, ac, om)
    e.add_uniaxial(K, ax, 4)
    e.set_time_step(b.dt)
    e.set_spins(b.s0)
    e.set_integrator(E.INTEGRATOR_RK4)
    e.step(200)
    o = po.OracleSystem(b.mus, b.gyro, b.alpha, nthreads=8)
    o.add_exchange(*b.csr)
    o.add_zeeman(dc, ac, om)
    o.add_uniaxial(K, ax, 4)
    s = b.s0.copy()
    for k in range(200):
        o.rk4_step(s, b.dt, k * b.dt)
    diff = np.abs(e.get_spins() - s).max()
    print(f"RK4 ac Zeeman + K p=4, 200 steps: max |ds| = {diff:.2e}")
    assert diff <= 1e-10


def test_rk4_fp32_field_angle(engine_factory):
    """bcc 8^3 x 2 with five shells + DMI at 100 K, same Philox stream: the
    FP32-field RK4 trajectory stays within 1e-6 rad of FP64 after 500 steps."""
    b = S.c3(8, with_csr=True)
    runs = []
    for prec in (0, 1):
        e = rk4_engine(b, engine_factory, prec=prec)
        e.set_thermostat(b.temperature, 5)
        e.step(500)
        runs.append(e.get_spins())
    c = np.clip(np.sum(runs[0] * runs[1], axis=1), -1, 1)
    ang = np.arccos(c).max()
    print(f"RK4 FP32-field vs FP64 after 500 steps: max angle {ang:.3e} rad")
    assert 0 < ang <= 1e-6


def test_rk4_graph_replay_matches_eager(engine_factory):
    b = S.c2(8)
    eager = rk4_engine(b, engine_factory)
    graph = rk4_engine(b, engine_factory)
    for e in (eager, graph):
        e.set_thermostat(b.temperature, 11)
    graph.set_graph_steps(8)
    for k, n in enumerate((8, 16, 5, 24)):
        eager.step(n)
        graph.step(n)
        eager.fields()
        graph.fields()
        assert np.array_equal(eager.get_spins(), graph.get_spins()), k


def test_rk4_with_dipole_trajectory(engine_factory):
    """sc 8 x 6 x 3 open film, J = 2 meV + dipole, T = 0, 100 RK4 steps: the
    engine (FFT dipole at each stage) against the oracle (brute force)."""
    lat = po.LatticeSpec(np.eye(3), [[0, 0, 0]], [0], (8, 6, 3), (0, 0, 0))
    b = S.build(lat, S.sc_nn_rows(2.0), 2.2, 0.02, ["exchange"], dt=S.dt_ps(1e-15), spins="random",
                seed=5)
    o = b.oracle(nthreads=8)
    o.add_dipole(lat, 20.0, 0.3e-9)
    e = rk4_engine(b, engine_factory)
    e.add_dipole_fft(lat.size, lat.periodic, 1, lat.cell, lat.basis_frac, 0.3e-9, 20.0)
    e.set_spins(b.s0)
    e.step(100)
    ref = rk4_oracle(b, 100, o=o)
    diff = np.abs(e.get_spins() - ref).max()
    print(f"RK4 exchange + dipole, 100 steps: max |ds| = {diff:.2e}")
    assert diff <= 1e-9


def test_rk4_rejects_phase_timing(engine_factory):
    b = S.c1(4)
    e = rk4_engine(b, engine_factory)
    with pytest.raises(E.EngineError):
        e.time_phase(0, 1)
