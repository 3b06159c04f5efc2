"""GPU parity of the Heun-LLG step against the oracle (CPU restatement of
solvers/cpu_llg_heun.cc), called through the C ABI.

Bar: the FP64 path reproduces the oracle bit for bit (the brief's bar is a
trajectory difference <= 1e-10; every FP64 test asserts that bar and, where
the arithmetic is the same sequence of IEEE operations, exact equality).
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests import systems as S

pytestmark = pytest.mark.gpu

TOL_TRAJ = 1e-10


def run_oracle(b, nsteps, noise=None, nthreads=8):
    o = b.oracle(nthreads=nthreads)
    s = b.s0.copy()
    t = 0.0
    for k in range(nsteps):
        w = None if noise is None else noise(k)
        o.heun_step(s, b.dt, t, w)
        t = (k + 1) * b.dt
    return s


def test_c1_trajectory_1000_steps_bitwise(engine_factory):
    """C1: SC 16^3, J = 1 meV, B = z, alpha = 0.1, T = 0, 1000 steps."""
    b = S.c1()
    e = b.engine(engine_factory)
    e.step(1000)
    got = e.get_spins()
    ref = run_oracle(b, 1000)
    diff = np.abs(got - ref).max()
    print(f"C1 1000 steps: max |ds| = {diff:.3e}; bitwise equal: {np.array_equal(got, ref)}")
    assert diff <= TOL_TRAJ
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("fmt", [1, 2, 3])
def test_formats_are_bit_identical(engine_factory, fmt):
    """TYPED and BLOCK storage give the same bits (C1 lattice, 50 steps)."""
    b = S.c1(8)
    e = b.engine(engine_factory, fmt=fmt)
    e.step(50)
    ref = run_oracle(b, 50)
    assert np.array_equal(e.get_spins(), ref)


def test_exchange_list_input_matches_csr(engine_factory):
    b = S.c2(6)
    e1 = b.engine(engine_factory, use_list=False)
    e2 = b.engine(engine_factory, use_list=True)
    for e in (e1, e2):
        e.step(20)
    assert np.array_equal(e1.get_spins(), e2.get_spins())


def test_c2_small_injected_noise_bitwise(engine_factory):
    """bcc 8^3 x 2, two shells + K1, T = 300 K with the oracle's noise stream
    fed in through jh_set_noise: 100 steps."""
    b = S.c2(8)
    sigma = po.thermostat_sigma(b.mus, b.gyro, b.alpha, b.dt)
    noise = lambda k: po.scale_noise(po.normals(99, k, b.n), sigma, b.temperature)
    e = b.engine(engine_factory)
    for k in range(100):
        e.set_noise(noise(k))
        e.step(1)
    got = e.get_spins()
    ref = run_oracle(b, 100, noise)
    diff = np.abs(got - ref).max()
    print(f"C2 small injected noise: max |ds| = {diff:.3e}")
    assert diff <= TOL_TRAJ
    assert np.array_equal(got, ref)


def test_generated_noise_stream_bitwise(engine_factory):
    """The in-kernel Philox stream equals the oracle's bit for bit."""
    b = S.c2(4)
    e = b.engine(engine_factory)
    e.set_thermostat(b.temperature, 1234)
    sigma = po.thermostat_sigma(b.mus, b.gyro, b.alpha, b.dt)
    for step in (0, 1, 77, 2**33 + 5):
        got = e.get_noise(step)
        ref = po.scale_noise(po.normals(1234, step, b.n), sigma, b.temperature)
        assert np.array_equal(got, ref), step


def test_c2_small_generated_noise_trajectory_bitwise(engine_factory):
    b = S.c2(8)
    e = b.engine(engine_factory)
    e.set_thermostat(b.temperature, 42)
    e.step(200)
    sigma = po.thermostat_sigma(b.mus, b.gyro, b.alpha, b.dt)
    noise = lambda k: po.scale_noise(po.normals(42, k, b.n), sigma, b.temperature)
    ref = run_oracle(b, 200, noise)
    got = e.get_spins()
    print(f"C2 small generated noise: max |ds| = {np.abs(got - ref).max():.3e}")
    assert np.array_equal(got, ref)


def test_term_fields_and_sum(engine_factory):
    b = S.c2(6)
    o = b.oracle()
    e = b.engine(engine_factory)
    s = po.random_spins(5, b.n)
    e.set_spins(s)
    for t in range(len(b.order)):
        assert np.array_equal(e.term_field(t), o.term_field(t, s))
    assert np.array_equal(e.fields(), o.fields(s))


def test_energies_and_magnetisation(engine_factory):
    b = S.c2(6)
    o = b.oracle()
    e = b.engine(engine_factory)
    s = po.random_spins(6, b.n)
    e.set_spins(s)
    for t in range(len(b.order)):
        eg, er = e.term_energy(t), o.term_energy(t, s)
        assert abs(eg - er) <= 1e-12 * max(1.0, abs(er)), (t, eg, er)
    m = e.magnetisation()
    ref = po.kahan_sum3(b.mus[:, None] * s)
    assert np.allclose(m[:3], ref, rtol=0, atol=1e-12 * b.mus.sum())
    assert abs(m[3] - po.kahan_sum(b.mus)) <= 1e-12 * b.mus.sum()


@pytest.mark.parametrize("n_cells", [(1, 1, 1), (3, 1, 1), (5, 7, 3)])
def test_ragged_and_tiny_systems(engine_factory, n_cells):
    """Open boundaries (ragged rows), N not a multiple of 64, a single spin."""
    lat = po.LatticeSpec(np.eye(3), [[0, 0, 0]], [0], n_cells, (0, 0, 0))
    b = S.build(lat, S.sc_nn_rows(1.0), 2.0, 0.1, ["exchange", "zeeman"], zeeman_B=[0.3, 0, 1],
                dt=1e-3, spins="random", seed=9)
    if b.csr[0][-1] == 0:
        b.order = ["zeeman"]
    e = b.engine(engine_factory)
    e.step(30)
    assert np.array_equal(e.get_spins(), run_oracle(b, 30))


@pytest.mark.parametrize("fmt", [0, 1, 2])
def test_vacancies_zero_length_spins_bitwise(engine_factory, fmt):
    """Zero-length spins (vacancies), for which the reference switches to its
    zero-safe kernels (solvers/cuda_llg_heun.cu:42-52,
    cuda_llg_heun_kernel.cuh:116-211): bcc 8^3 x 2, two shells + K1, 10 % of
    the sites empty, T = 300 K with the generated stream, 200 steps.  The
    trajectory equals the oracle's bit for bit, the empty sites stay exactly
    zero (unit_vector leaves |v| <= epsilon alone, containers/vec3.h:276-283)
    and every other spin keeps unit length."""
    b = S.c2(8)
    rng = np.random.default_rng(8)
    empty = rng.choice(b.n, b.n // 10, replace=False)
    b.s0 = po.random_spins(8, b.n)
    b.s0[empty] = 0.0
    e = b.engine(engine_factory, fmt=fmt)
    e.set_thermostat(b.temperature, 55)
    e.step(200)
    got = e.get_spins()
    sigma = po.thermostat_sigma(b.mus, b.gyro, b.alpha, b.dt)
    noise = lambda k: po.scale_noise(po.normals(55, k, b.n), sigma, b.temperature)
    ref = run_oracle(b, 200, noise)
    assert np.array_equal(got, ref)
    assert np.all(got[empty] == 0.0)
    full = np.setdiff1d(np.arange(b.n), empty)
    assert np.abs(np.linalg.norm(got[full], axis=1) - 1.0).max() < 1e-15


def test_vacancies_fp32_field_form(engine_factory):
    """The FP32-field form on the same vacancy lattice: empty sites stay
    exactly zero and the others stay within 1e-6 rad of FP64 after 200
    steps on the same stream."""
    b = S.c2(8)
    rng = np.random.default_rng(8)
    empty = rng.choice(b.n, b.n // 10, replace=False)
    b.s0 = po.random_spins(8, b.n)
    b.s0[empty] = 0.0
    out = []
    for prec in (0, 1):
        e = b.engine(engine_factory, prec=prec)
        e.set_thermostat(b.temperature, 55)
        e.step(200)
        out.append(e.get_spins())
    assert np.all(out[1][empty] == 0.0)
    full = np.setdiff1d(np.arange(b.n), empty)
    a, c = out[0][full], out[1][full]
    ang = np.arctan2(np.linalg.norm(np.cross(a, c), axis=1), np.einsum("ij,ij->i", a, c))
    assert ang.max() < 1e-6
