"""Pins of the oracle (CPU restatement of the reference path) against
published known answers, closed forms, brute force and the reference's own
golden values.  Each test names the mistake it would catch."""
import math

import numpy as np
import pytest
from scipy import stats

from oracle import pyoracle as po
from tests import systems as S


# ---- RNG ---------------------------------------------------------------------
# Random123 known-answer vectors for philox4x32_10 (counter, key) -> output.
KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_known_answers(ctr, key, out):
    assert po.philox(ctr, key) == out


def test_noise_log_and_sincos_match_libm():
    rng = np.random.default_rng(3)
    for x in np.concatenate([rng.random(2000), [2.0**-53, 0.5, 1.0 - 2**-53, 0.7071067811865476]]):
        ref = math.log(x)
        assert abs(po.noise_log(x) - ref) <= 4 * np.spacing(abs(ref)) + 1e-300
    # reference values in 80-bit extended precision (2 pi u rounded in double
    # would itself be off by ~u ulp(2 pi))
    two_pi = np.longdouble("6.28318530717958647692528676655900577")
    for u in np.concatenate([rng.random(2000), [0.0, 0.125, 0.25, 0.5, 0.75, 0.999999]]):
        c, s = po.noise_sincos2pi(u)
        rc = float(np.cos(two_pi * np.longdouble(u)))
        rs = float(np.sin(two_pi * np.longdouble(u)))
        assert abs(c - rc) <= 3e-16 and abs(s - rs) <= 3e-16, u


def test_normal_stream_statistics():
    z = po.normals(2024, 17, 100000).reshape(-1)
    assert abs(z.mean()) < 5 * (1 / math.sqrt(z.size))
    assert abs(z.var() - 1.0) < 5 * math.sqrt(2 / z.size)
    assert stats.kstest(z, "norm").pvalue > 1e-3
    # independent streams for different steps / seeds
    z2 = po.normals(2024, 18, 100000).reshape(-1)
    assert abs(np.corrcoef(z, z2)[0, 1]) < 0.02
    assert not np.array_equal(po.normals(1, 0, 8), po.normals(2, 0, 8))


def test_random_spins_uniform_on_sphere():
    s = po.random_spins(1, 50000)
    assert np.allclose(np.linalg.norm(s, axis=1), 1.0, atol=4e-16)
    assert stats.kstest(s[:, 2], "uniform", args=(-1, 2)).pvalue > 1e-3
    phi = np.arctan2(s[:, 1], s[:, 0])
    assert stats.kstest(phi, "uniform", args=(-math.pi, 2 * math.pi)).pvalue > 1e-3


# ---- lattice, neighbour list, CSR ----------------------------------------------
def test_site_index_closed_form():
    lat = po.LatticeSpec(np.eye(3), [[0, 0, 0], [0.5, 0.5, 0.5]], [0, 0], (3, 4, 5), (1, 1, 1))
    L = lat.c()
    for (i, j, k, m) in [(0, 0, 0, 0), (2, 3, 4, 1), (1, 2, 3, 0)]:
        assert po.lib().or_site_index(po.C.byref(L), i, j, k, m) == ((i * 4 + j) * 5 + k) * 2 + m


CASES = [
    ("sc periodic", po.LatticeSpec(np.eye(3), [[0, 0, 0]], [0], (5, 6, 7), (1, 1, 1)), S.sc_nn_rows(1.0), 6),
    ("sc open", po.LatticeSpec(np.eye(3), [[0, 0, 0]], [0], (5, 4, 3), (0, 0, 0)), S.sc_nn_rows(1.0), None),
    ("sc mixed", po.LatticeSpec(np.eye(3), [[0, 0, 0]], [0], (6, 5, 3), (1, 0, 1)), S.sc_nn_rows(1.0), None),
    ("bcc two shells", S.bcc_lattice(5), S.bcc_two_shell_rows(3.0, 1.5), 14),
    ("bcc five shells dmi", S.bcc_lattice(6), S.bcc_five_shell_dmi_rows(), 58),
    ("bcc open", S.bcc_lattice(4, (0, 1, 0)), S.bcc_two_shell_rows(3.0, 1.5), None),
]


@pytest.mark.parametrize("name,lat,rows,per_site", CASES, ids=[c[0] for c in CASES])
def test_neighbour_list_matches_bruteforce(name, lat, rows, per_site):
    pi, pj, pJ = po.neighbour_list(lat, rows)
    bi, bj, bJ = po.neighbour_bruteforce(lat, rows)
    assert np.array_equal(pi, bi) and np.array_equal(pj, bj) and np.array_equal(pJ, bJ)
    if per_site is not None:
        assert pi.size == lat.n * per_site
        assert np.all(np.bincount(pi, minlength=lat.n) == per_site)


def test_open_film_pair_count_closed_form():
    """C5 film 256x256x8 open: 2 (255*256*8 + 256*255*8 + 256*256*7) pairs."""
    lat = po.LatticeSpec(np.eye(3), [[0, 0, 0]], [0], (256, 256, 8), (0, 0, 0))
    pi, pj, _ = po.neighbour_list(lat, S.sc_nn_rows(2.0))
    assert pi.size == 2 * (255 * 256 * 8 + 256 * 255 * 8 + 256 * 256 * 7) == 3006464


def test_duplicate_pairs_are_rejected():
    # a 2-cell periodic chain maps +x and -x onto the same neighbour
    lat = po.LatticeSpec(np.eye(3), [[0, 0, 0]], [0], (2, 3, 3), (1, 1, 1))
    with pytest.raises(ValueError):
        po.neighbour_list(lat, S.sc_nn_rows(1.0))


def dense_from_pairs(n, pi, pj, pJ):
    A = np.zeros((3 * n, 3 * n))
    for i, j, J in zip(pi, pj, pJ.reshape(-1, 3, 3)):
        A[3 * i:3 * i + 3, 3 * j:3 * j + 3] += J
    return A


@pytest.mark.parametrize("name,lat,rows,per_site", CASES[:5], ids=[c[0] for c in CASES[:5]])
def test_exchange_csr_structure_and_values(name, lat, rows, per_site):
    pi, pj, pJ = po.neighbour_list(lat, rows)
    rowptr, col, val = po.exchange_csr(lat.n, pi, pj, pJ)
    # nnz closed form: isotropic -> 3 per pair, full tensor -> 9 per pair
    nz_per_pair = (pJ != 0).sum(axis=1)
    assert rowptr[-1] == nz_per_pair.sum()
    # rows sorted, strictly ascending columns, exact symmetry
    for r in range(3 * lat.n):
        c = col[rowptr[r]:rowptr[r + 1]]
        assert np.all(np.diff(c) > 0)
    assert po.csr_is_symmetric(rowptr, col, val)
    # structure rule: row 3i+m holds 3j+n for neighbours j with J[m][n] != 0
    A = dense_from_pairs(lat.n, pi, pj, pJ)
    D = np.zeros_like(A)
    for r in range(3 * lat.n):
        D[r, col[rowptr[r]:rowptr[r + 1]]] = val[rowptr[r]:rowptr[r + 1]]
    assert np.array_equal(A, D)
    assert np.array_equal(np.nonzero(A), np.nonzero(D))


def test_builder_quirks_as_written():
    """merge() skips the element after an erase (triple duplicates survive as
    two entries); build_csr() starts from row_[1] so an empty row 0 followed
    by non-empty rows gives row 0 a phantom entry."""
    rowptr, col, val = po.coo_to_csr(3, [1, 1, 1, 2], [0, 0, 0, 1], [1.0, 2.0, 4.0, 8.0])
    assert list(col) == [0, 0, 1] and list(val) == [3.0, 4.0, 8.0]
    assert list(rowptr) == [0, 1, 2, 3]  # row 0 claims entry 0, which belongs to row 1
    rowptr, col, val = po.coo_to_csr(3, [0, 1, 1, 2], [1, 0, 0, 2], [1.0, 2.0, 4.0, 8.0])
    assert list(rowptr) == [0, 1, 2, 3] and list(val) == [1.0, 6.0, 8.0]


def test_csrmv_dense_brute_force_and_uniform_state():
    lat = S.bcc_lattice(6)
    rows = S.bcc_five_shell_dmi_rows()
    pi, pj, pJ = po.neighbour_list(lat, rows)
    rowptr, col, val = po.exchange_csr(lat.n, pi, pj, pJ)
    A = dense_from_pairs(lat.n, pi, pj, pJ)
    x = po.random_spins(11, lat.n).reshape(-1)
    y = po.csrmv(rowptr, col, val, x)
    assert np.allclose(y, A @ x, rtol=0, atol=1e-13)
    # uniform state: h_i = (sum_j J_ij) s (DMI cancels in the sum over +-r)
    s = np.tile([0.3, -0.4, math.sqrt(1 - 0.25)], lat.n)
    J0 = sum(np.asarray(J, float).reshape(3, 3) for (_, _, _, J) in rows)
    y = po.csrmv(rowptr, col, val, s).reshape(-1, 3)
    assert np.allclose(y, (J0 @ s[:3])[None, :], rtol=0, atol=1e-12)


def test_csrmv_plane_wave_eigenvector():
    L = 8
    lat = po.LatticeSpec(np.eye(3), [[0, 0, 0]], [0], (L, L, L), (1, 1, 1))
    J = 1.7
    pi, pj, pJ = po.neighbour_list(lat, S.sc_nn_rows(J))
    rowptr, col, val = po.exchange_csr(lat.n, pi, pj, pJ)
    r = lat.positions()
    k = 2 * math.pi * np.array([1, 2, 3]) / L
    eps = 0.1
    ph = r @ k
    s = np.stack([eps * np.cos(ph), eps * np.sin(ph), np.full(lat.n, math.sqrt(1 - eps * eps))], 1)
    h = po.csrmv(rowptr, col, val, s.reshape(-1)).reshape(-1, 3)
    Jk = J * 2 * np.cos(k).sum()
    assert np.allclose(h[:, 0], Jk * s[:, 0], atol=1e-13)
    assert np.allclose(h[:, 1], Jk * s[:, 1], atol=1e-13)
    assert np.allclose(h[:, 2], 6 * J * s[:, 2], atol=1e-13)


# ---- Heun LLG ----------------------------------------------------------------
def test_macrospin_precession_and_damping_closed_form():
    """Single spin in B z: tan(theta/2) = tan(theta0/2) exp(-alpha gamma B t),
    phi = phi0 + gamma B t (Landau-Lifshitz form)."""
    g = S.gyro_e()
    alpha, B, dt, nsteps = 0.1, 2.0, 1e-3, 2000
    mus = np.array([2.0 * S.BOHR])
    o = po.OracleSystem(mus, np.array([g]), np.array([alpha]))
    o.add_zeeman(mus[:, None] * np.array([[0.0, 0.0, B]]))
    th0, ph0 = 1.1, 0.3
    s = np.array([[math.sin(th0) * math.cos(ph0), math.sin(th0) * math.sin(ph0), math.cos(th0)]])
    for k in range(nsteps):
        o.heun_step(s, dt, k * dt)
    t = nsteps * dt
    th = 2 * math.atan(math.tan(th0 / 2) * math.exp(-alpha * g * B * t))
    ph = ph0 + g * B * t
    ref = np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
    assert np.abs(s[0] - ref).max() < 1e-8
    assert abs(np.linalg.norm(s[0]) - 1) < 4e-16


def test_unit_length_and_energy_descent_at_zero_temperature():
    b = S.c1(8)
    o = b.oracle()
    s = b.s0.copy()
    e_prev = sum(o.term_energy(t, s) for t in range(len(b.order)))
    for k in range(200):
        o.heun_step(s, b.dt, k * b.dt)
        if k % 20 == 19:
            e = sum(o.term_energy(t, s) for t in range(len(b.order)))
            assert e <= e_prev + 1e-9
            e_prev = e
    assert np.abs(np.linalg.norm(s, axis=1) - 1).max() < 1e-15


def test_thermostat_sigma_closed_form():
    mus = np.array([2.0 * S.BOHR, 1.5 * S.BOHR])
    g = np.array([S.gyro_e()] * 2)
    a = np.array([0.05, 0.2])
    dt = 5e-4
    sig = po.thermostat_sigma(mus, g, a, dt)
    assert np.allclose(sig, np.sqrt(2 * S.KB * a / (mus * g * dt)), rtol=1e-15)
    sig_g = po.thermostat_sigma(mus, g, a, dt, gilbert=True)
    assert np.allclose(sig_g, np.sqrt(2 * S.KB * a / (mus * g * dt * (1 + a * a))), rtol=1e-15)


def paramagnet_mean_sz(gilbert, n=2048, B=10.0, T=10.0, alpha=1.0, dt=1e-3, steps=10000):
    mu = 2.0 * S.BOHR
    mus = np.full(n, mu)
    g = np.full(n, S.gyro_e() / (1 + alpha * alpha) if gilbert else S.gyro_e())
    al = np.full(n, alpha)
    o = po.OracleSystem(mus, g, al)
    o.add_zeeman(mus[:, None] * np.array([[0.0, 0.0, B]]))
    sigma = po.thermostat_sigma(mus, g, al, dt, gilbert=gilbert)
    s = np.tile([0.0, 0.0, 1.0], (n, 1))
    acc = []
    for k in range(steps):
        w = po.scale_noise(po.normals(77, k, n), sigma, T)
        o.heun_step(s, dt, k * dt, w)
        if k >= 5000 and k % 10 == 0:
            acc.append(s[:, 2].mean())
    return np.mean(acc), mu * B / (S.KB * T)


def langevin(x):
    return 1 / math.tanh(x) - 1 / x


def test_langevin_function_paramagnet_gilbert_form():
    """Non-interacting spins in B with solver.gilbert_prefactor = true
    (gyro / (1+alpha^2), sigma with the same denominator):
    <s_z> = coth(x) - 1/x, x = mus B / (kB T)."""
    m, x = paramagnet_mean_sz(True)
    assert abs(m - langevin(x)) < 0.01, (m, langevin(x))


def test_langevin_function_paramagnet_ll_form_as_coded():
    """Without gilbert_prefactor the reference pairs the LL equation with
    sigma^2 = 2 kB alpha / (mus gyro dt): equilibrium is at (1 + alpha^2) T."""
    m, x = paramagnet_mean_sz(False)
    assert abs(m - langevin(x / 2.0)) < 0.01, (m, langevin(x / 2.0))


def test_kahan_sums_exact_on_integers():
    rng = np.random.default_rng(5)
    x = rng.integers(-1000, 1000, size=(10000, 3)).astype(np.float64)
    assert np.array_equal(po.kahan_sum3(x), x.sum(axis=0))
    assert po.kahan_sum(x[:, 0]) == x[:, 0].sum()
    mus = np.full(10000, 2.0)
    assert np.array_equal(po.sum_moments(x, mus), (2.0 * x).sum(axis=0))


# ---- dipole (reference known answers, test/hamiltonian/test_dipole.h) --------
DIP_PREFACTOR = None


def dipole_prefactor():
    mu0 = po.vacuum_permeability()
    return -((0.5 * mu0) / (4 * math.pi)) * (2 * S.BOHR) ** 2 / (0.3e-9) ** 3


@pytest.mark.parametrize("spin,expect", [((1, 0, 0), 4.808228), ((0, 1, 0), -2.404114),
                                         ((0, 0, 1), -2.404114)])
def test_dipole_bruteforce_1d_chain_known_answers(spin, expect):
    """1D FM chain of 2000 cells periodic along x (test_dipole.h:256-284)."""
    lat = po.LatticeSpec(np.eye(3), [[0, 0, 0]], [0], (2000, 1, 1), (1, 0, 0))
    n = lat.n
    mus = np.full(n, 2.0 * S.BOHR)
    s = np.tile(np.asarray(spin, float), (n, 1))
    h = po.dipole_bruteforce(lat, mus, s, 1000.0, 0.3e-9, nthreads=8)
    e = -0.5 * np.sum(s * h) / n
    assert abs(e - dipole_prefactor() * expect) < 1e-5


def test_dipole_bruteforce_1d_afm_known_answer():
    """1D AFM chain along x (test_dipole.h:371-404): x = -3.60617."""
    lat = po.LatticeSpec(np.eye(3), [[0, 0, 0]], [0], (2000, 1, 1), (1, 0, 0))
    n = lat.n
    mus = np.full(n, 2.0 * S.BOHR)
    s = np.zeros((n, 3))
    s[:, 0] = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    h = po.dipole_bruteforce(lat, mus, s, 1000.0, 0.3e-9, nthreads=8)
    e = -0.5 * np.sum(s * h) / n
    assert abs(e - dipole_prefactor() * -3.60617) < 1e-5


# ---- the FFT dipole's convolution (oracle_dipole_fft.c) ---------------------
DIRECT_CASES = [
    ("sc_open_6x5x4", po.LatticeSpec(np.eye(3), [[0, 0, 0]], [0], (6, 5, 4), (0, 0, 0)), 30.0),
    ("bcc_mixed_4x3x5", po.LatticeSpec(np.eye(3), [[0, 0, 0], [0.5, 0.5, 0.5]], [0, 0], (4, 3, 5),
                                       (1, 0, 1)), 1.9),
    ("sc_periodic_5x5x5", po.LatticeSpec(np.eye(3), [[0, 0, 0]], [0], (5, 5, 5), (1, 1, 1)), 2.0),
]


@pytest.mark.parametrize("name,lat,rc", DIRECT_CASES, ids=[c[0] for c in DIRECT_CASES])
def test_dipole_fft_direct_two_sided_equals_bruteforce(name, lat, rc):
    """The convolution with two-sided kernels on open dimensions is the
    brute-force sum (which the reference's known answers pin)."""
    mus = np.full(lat.n, 2.0 * S.BOHR)
    s = po.random_spins(3, lat.n)
    h_d = po.dipole_fft_direct(lat, mus, s, rc, 0.3e-9, one_sided_open=False, nthreads=8)
    h_b = po.dipole_bruteforce(lat, mus, s, rc, 0.3e-9, nthreads=8)
    assert np.abs(h_d - h_b).max() <= 1e-12 * np.abs(h_b).max()


@pytest.mark.parametrize("name,lat,rc", DIRECT_CASES, ids=[c[0] for c in DIRECT_CASES])
def test_dipole_bruteforce_sites_equals_full_sum(name, lat, rc):
    """The sampled-site form is the brute-force sum at those sites, bit for
    bit (the full sum is pinned by the reference's known answers)."""
    mus = np.full(lat.n, 2.0 * S.BOHR)
    s = po.random_spins(4, lat.n)
    full = po.dipole_bruteforce(lat, mus, s, rc, 0.3e-9, nthreads=8)
    sites = np.array([0, lat.n - 1, lat.n // 2, 3], np.int64)
    assert np.array_equal(po.dipole_bruteforce_sites(lat, mus, s, rc, 0.3e-9, sites, 4), full[sites])


def test_dipole_fft_direct_one_sided_as_written_differs_on_open_dimensions():
    """dipole_fft.cc:227-251 as written fills only n >= 0 on an open
    dimension: its field is not the dipole field there (not copied)."""
    lat = DIRECT_CASES[0][1]
    mus = np.full(lat.n, 2.0 * S.BOHR)
    s = po.random_spins(3, lat.n)
    h_one = po.dipole_fft_direct(lat, mus, s, 30.0, 0.3e-9, one_sided_open=True, nthreads=8)
    h_b = po.dipole_bruteforce(lat, mus, s, 30.0, 0.3e-9, nthreads=8)
    assert np.abs(h_one - h_b).max() > 0.1 * np.abs(h_b).max()
    # on periodic dimensions both forms coincide
    lat = DIRECT_CASES[2][1]
    mus = np.full(lat.n, 2.0 * S.BOHR)
    s = po.random_spins(4, lat.n)
    h1 = po.dipole_fft_direct(lat, mus, s, 2.0, 0.3e-9, one_sided_open=True)
    h2 = po.dipole_fft_direct(lat, mus, s, 2.0, 0.3e-9, one_sided_open=False)
    assert np.array_equal(h1, h2)


# ---- RK4 (llg-rk4-gpu) ------------------------------------------------------


def _macrospin(alpha, B, ac=None, omega=None):
    g = S.gyro_e()
    mus = np.array([2.0 * S.BOHR])
    o = po.OracleSystem(mus, np.array([g]), np.array([alpha]))
    if ac is None:
        o.add_zeeman(mus[:, None] * np.array([[0.0, 0.0, B]]))
    else:
        o.add_zeeman(mus[:, None] * np.array([[0.0, 0.0, B]]), mus[:, None] * np.asarray([ac]),
                     np.array([omega]))
    return o, g


def _rk4_run(o, s0, dt, nsteps):
    s = s0.copy()
    for k in range(nsteps):
        o.rk4_step(s, dt, k * dt)
    return s


def test_rk4_macrospin_closed_form_fourth_order():
    """Single spin in B z (closed form as for the Heun step): the RK4 error
    falls ~16x per halving of dt.  A wrong stage weight or stage time drops
    the order to 1 or 2 (ratio <= 4); Heun itself gives ~4."""
    alpha, B, T = 0.1, 10.0, 1.6
    o, g = _macrospin(alpha, B)
    th0, ph0 = 1.1, 0.3
    s0 = np.array([[math.sin(th0) * math.cos(ph0), math.sin(th0) * math.sin(ph0), math.cos(th0)]])
    th = 2 * math.atan(math.tan(th0 / 2) * math.exp(-alpha * g * B * T))
    ph = ph0 + g * B * T
    ref = np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
    errs = []
    for nsteps in (32, 64, 128):
        s = _rk4_run(o, s0, T / nsteps, nsteps)
        errs.append(np.abs(s[0] - ref).max())
        assert abs(np.linalg.norm(s[0]) - 1) < 4e-16
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    print(f"RK4 macrospin errors {errs}, ratios {r1:.2f} {r2:.2f}")
    assert errs[2] < 1e-7
    assert 12 < r1 < 20 and 12 < r2 < 20
    # the Heun step on the same problem is second order (contrast)
    s = s0.copy()
    for k in range(128):
        o.heun_step(s, T / 128, k * T / 128)
    assert np.abs(s[0] - ref).max() > 100 * errs[2]


def test_rk4_stage_times_with_an_ac_field():
    """Time-dependent field B(t) = B z + b cos(w t) x: fourth-order convergence
    to a run at dt/16 pins the stage times t0, t0 + dt/2, t0 + dt/2, t0 + dt."""
    o, _ = _macrospin(0.05, 4.0, ac=(3.0, 0.0, 0.0), omega=5.0)
    s0 = np.array([[0.0, 0.6, 0.8]])
    T = 1.28
    fine = _rk4_run(o, s0, T / 2048, 2048)
    errs = [np.abs(_rk4_run(o, s0, T / m, m)[0] - fine[0]).max() for m in (32, 64, 128)]
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    print(f"RK4 ac-field errors {errs}, ratios {r1:.2f} {r2:.2f}")
    assert 12 < r1 < 20 and 12 < r2 < 20


def test_rk4_unit_length_and_agreement_with_heun():
    """C1 lattice 6^3: |s| = 1 after every RK4 step; RK4 and Heun agree to
    O(dt^2) and the gap shrinks ~4x when dt halves."""
    b = S.c1(6)
    o = b.oracle()
    gaps = []
    for scale in (1, 2):
        dt = b.dt / scale
        s4, s2 = b.s0.copy(), b.s0.copy()
        for k in range(100 * scale):
            o.rk4_step(s4, dt, k * dt)
            o.heun_step(s2, dt, k * dt)
            assert np.abs(np.linalg.norm(s4, axis=1) - 1).max() < 1e-15
        gaps.append(np.abs(s4 - s2).max())
    print(f"RK4 vs Heun gaps {gaps}")
    assert gaps[0] < 1e-4
    assert 3.0 < gaps[0] / gaps[1] < 5.0


# ---- dipole stencil ------------------------------------------------------------
def test_dipole_stencil_matches_direct_sum():
    """The dipole tensor read off the direct offset sum with one spin at the
    centre of a small cube, applied as a stencil, is the reference of the GPU
    test on the 256 x 256 x 8 open film: it equals the direct sum on a
    12 x 10 x 4 open film (r_cutoff 2.5). Catches a dropped or misplaced
    offset, a wrong boundary (wrap instead of zero) and a wrong kernel size."""
    from tests.test_gpu_dipole import A, SC, offset_kernel, stencil_field
    rc, mu = 2.5, 2.0 * S.BOHR
    K = offset_kernel(rc, mu)
    assert K.shape == (5, 5, 5, 3, 3)
    lat = po.LatticeSpec(SC, [[0, 0, 0]], [0], (12, 10, 4), (0, 0, 0))
    s = po.random_spins(5, lat.n)
    ref = po.dipole_fft_direct(lat, np.full(lat.n, mu), s, rc, A, nthreads=4)
    scale = np.abs(ref).max()
    assert np.abs(stencil_field(lat.size, K, s) - ref).max() <= 1e-13 * scale
    dropped = K.copy()
    dropped[2, 2, 4] = 0.0
    assert np.abs(stencil_field(lat.size, dropped, s) - ref).max() > 1e-3 * scale
    wrapped = np.roll(s.reshape(12, 10, 4, 3), 1, axis=2).reshape(-1, 3)
    assert np.abs(stencil_field(lat.size, K, wrapped) - ref).max() > 1e-1 * scale


# ---- uniaxial anisotropy: field against the energy it derives from -----------
def _uniaxial_system(n, power, seed):
    rng = np.random.default_rng(seed)
    mus = np.full(n, 2.0 * S.BOHR)
    o = po.OracleSystem(mus, np.full(n, S.gyro_e()), np.full(n, 0.1))
    K = rng.uniform(0.02, 0.5, n)
    axis = rng.normal(size=(n, 3))
    axis /= np.linalg.norm(axis, axis=1)[:, None]
    o.add_uniaxial(K, axis, power)
    return o, K, axis


def _energy_gradient_field(o, s, eps=1e-6):
    """-dE/ds by central differences of or_term_energy, per spin (each site's
    energy depends on its own spin only)."""
    h = np.empty_like(s)
    for c in range(3):
        sp, sm = s.copy(), s.copy()
        sp[:, c] += eps
        sm[:, c] -= eps
        # per-site energies: the term energy of one site at a time is the sum
        # over a system whose other spins are perpendicular to their axes
        # (zero energy for even powers), so sum differences site by site
        ep = np.array([o.term_energy(0, _only(sp, i, o)) for i in range(s.shape[0])])
        em = np.array([o.term_energy(0, _only(sm, i, o)) for i in range(s.shape[0])])
        h[:, c] = -(ep - em) / (2 * eps)
    return h


def _only(s, i, o):
    """s with every spin but i turned perpendicular to its axis."""
    out = np.cross(o.ani_axis, np.array([0.0, 0.0, 1.0]))
    bad = np.linalg.norm(out, axis=1) < 1e-6
    out[bad] = np.cross(o.ani_axis[bad], np.array([1.0, 0.0, 0.0]))
    out[i] = s[i]
    return out


def _field_matches_gradient(h, g):
    return bool(np.all(np.abs(h - g) <= 1e-7 * (1.0 + np.abs(g))))


@pytest.mark.parametrize("power", [2, 4, 6])
def test_uniaxial_field_is_minus_energy_gradient(power):
    """uniaxial_anisotropy.cc:116-123 (E = -K (s.e)^p) and :155-163 (the field
    K p (s.e)^(p-1) e): the oracle's field equals -dE/ds of its own energy by
    central differences at random spins on tilted axes.  The same check
    rejects the plausible slips: the factor p dropped, the exponent p instead
    of p - 1, the sign flipped."""
    o, K, axis = _uniaxial_system(12, power, 40 + power)
    s = po.random_spins(41, 12)
    h = o.term_field(0, s)
    g = _energy_gradient_field(o, s)
    assert _field_matches_gradient(h, g)
    d = np.einsum("ij,ij->i", s, axis)
    wrong = {"no factor p": (K * d ** (power - 1))[:, None] * axis,
             "exponent p": (K * power * d ** power)[:, None] * axis,
             "sign": -h}
    for name, hw in wrong.items():
        assert not _field_matches_gradient(hw, g), name


@pytest.mark.parametrize("power", [2, 4, 6])
def test_uniaxial_precession_closed_form(power):
    """One spin, easy axis z, no damping, T = 0: s precesses about z at
    omega = gamma p K cos^(p-1)(theta) / mus with theta fixed (the field is
    along z with magnitude p K cos^(p-1) theta in meV, / mus in tesla)."""
    g = S.gyro_e()
    K = 0.8 if power == 2 else 2.0
    mu = 2.0 * S.BOHR
    o = po.OracleSystem(np.array([mu]), np.array([g]), np.array([0.0]))
    o.add_uniaxial(np.array([K]), np.array([[0.0, 0.0, 1.0]]), power)
    th0, ph0 = 0.6, 0.2
    s = np.array([[math.sin(th0) * math.cos(ph0), math.sin(th0) * math.sin(ph0), math.cos(th0)]])
    omega = g * power * K * math.cos(th0) ** (power - 1) / mu
    dt = 2e-4
    nsteps = int(round(2.5 / (omega * dt)))
    for k in range(nsteps):
        o.heun_step(s, dt, k * dt)
    ph = ph0 + omega * nsteps * dt
    ref = np.array([math.sin(th0) * math.cos(ph), math.sin(th0) * math.sin(ph), math.cos(th0)])
    err = np.abs(s[0] - ref).max()
    assert err < 1e-6, err
    # a field short of the factor p would precess p times slower
    ph_wrong = ph0 + omega / power * nsteps * dt
    wrong = np.array([math.sin(th0) * math.cos(ph_wrong), math.sin(th0) * math.sin(ph_wrong),
                      math.cos(th0)])
    assert np.abs(s[0] - wrong).max() > 1e-2
