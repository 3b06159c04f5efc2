"""GPU parity of the LINE exchange layout (a wavefront holds one whole z line
of one basis site, loads each neighbour line once and reads the neighbours
along z from the lanes that hold them; jh_set_lattice + JH_XFMT_LINE)
against the oracle.

FP64 must reproduce the reference's ordered CSR row sums bit for bit
(hamiltonian/sparse_interaction.cc:25-33, interface/sparse_blas.h:15-27):
the lines' programs keep the column order of rows away from the z ends, and
the rows whose order differs because a z neighbour wraps round the periodic
boundary are computed by the row-list launch from their own row patterns.
The FP32-field form must stay within 1e-5 rad of FP64 over 200 steps.

Lz = 64 puts one cell in each lane, Lz = 128 two (z = 2 lane + h); x and y
are 8 cells, so the oracle builds the CSR in seconds.
"""
import numpy as np
import pytest

from jams_amd import engine as E
from oracle import pyoracle as po
from tests import systems as S

pytestmark = pytest.mark.gpu


def bcc(size):
    return po.LatticeSpec(S.SC_CELL, [[0, 0, 0], [0.5, 0.5, 0.5]], [0, 0], size, (1, 1, 1))


def sc(size):
    return po.LatticeSpec(S.SC_CELL, [[0, 0, 0]], [0], size, (1, 1, 1))


def _dmi(lz, seed):
    # C3's template: five shells with DMI, full tensors, T = 100 K
    return S.build(bcc((8, 8, lz)), S.bcc_five_shell_dmi_rows(), 2.0, 0.05, ["exchange"],
                   dt=S.dt_ps(0.5e-15), spins="random", seed=seed, temperature=100.0)


def _two_shell(lz, seed):
    # C2's template: two isotropic shells (diagonal tensors) and K1
    return S.build(bcc((8, 8, lz)), S.bcc_two_shell_rows(3.0, 1.5), 2.0, 0.05,
                   ["exchange", "uniaxial"], K=0.05, dt=S.dt_ps(0.5e-15), spins="random",
                   seed=seed, temperature=300.0)


SYSTEMS = {
    "dmi_lz64": lambda: _dmi(64, 3),
    "dmi_lz128": lambda: _dmi(128, 4),
    "two_shell_lz64": lambda: _two_shell(64, 6),
    "two_shell_lz128": lambda: _two_shell(128, 7),
    # simple cubic, one basis site, Zeeman after the exchange
    "sc_nn_zeeman_lz128": lambda: S.build(sc((8, 8, 128)), S.sc_nn_rows(1.0), 2.0, 0.1,
                                          ["exchange", "zeeman"], zeeman_B=[0, 0, 1],
                                          dt=S.dt_ps(1.0e-15), spins="random", seed=8,
                                          temperature=50.0),
}


def noise_fn(b, seed):
    sigma = po.thermostat_sigma(b.mus, b.gyro, b.alpha, b.dt)
    return lambda k: po.scale_noise(po.normals(seed, k, b.n), sigma, b.temperature)


def line_engine(factory, b, prec=E.PREC_FP64):
    e = b.engine(factory, prec=prec, fmt=E.XFMT_LINE, lattice=True)
    d = e.describe()
    assert " line=" in d, d
    return e


def oracle_run(b, seed, steps):
    noise = noise_fn(b, seed)
    o = b.oracle(nthreads=8)
    s = b.s0.copy()
    for k in range(steps):
        o.heun_step(s, b.dt, k * b.dt, noise(k))
    return s


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_line_fp64_heun_bitwise(engine_factory, name):
    """60 FP64 Heun steps with the generated Philox stream equal
    or_heun_step bit for bit."""
    b = SYSTEMS[name]()
    e = line_engine(engine_factory, b)
    print(name, e.describe(), flush=True)
    seed = 41
    e.set_thermostat(b.temperature, seed)
    steps = 60
    e.step(steps)
    got = e.get_spins()
    s = oracle_run(b, seed, steps)
    print(f"{name}: max |ds| = {np.abs(got - s).max():.3e}", flush=True)
    assert np.array_equal(got, s)


def test_line_wrapped_rows_go_to_the_row_list(engine_factory):
    """The rows whose z neighbours wrap are the row list's: the C3 template
    reaches from two cells down to one up in z from the first basis site and
    from one down to two up from the second, so at most four z values of each
    of the 2 x 8 x 8 lines."""
    e = line_engine(engine_factory, SYSTEMS["dmi_lz128"]())
    d = e.describe()
    nwrap = int(d.split("wrapped_rows=")[1].split()[0])
    assert 0 < nwrap <= 2 * 8 * 8 * 4, d
    assert "line=2x64" in d, d


def test_line_graph_replay_matches_eager(engine_factory):
    b = SYSTEMS["dmi_lz128"]()
    e = line_engine(engine_factory, b)
    e.set_thermostat(b.temperature, 9)
    e.step(30)
    eager = e.get_spins()
    e.set_spins(b.s0)
    e.set_iteration(0)
    e.set_graph_steps(10)
    for rep in range(3):
        e.step(10)
        e.synchronize()
    got = e.get_spins()
    assert np.array_equal(got, eager)


@pytest.mark.parametrize("name", ["dmi_lz64", "dmi_lz128", "two_shell_lz128"])
def test_line_fp32_field_trajectory(engine_factory, name):
    """FP32-field form over the LINE layout: 200 steps stay within 1e-5 rad
    of the FP64 trajectory on the same noise stream."""
    b = SYSTEMS[name]()
    e64 = b.engine(engine_factory, prec=E.PREC_FP64)
    e32 = line_engine(engine_factory, b, prec=E.PREC_FP32_FIELD)
    for e in (e64, e32):
        e.set_thermostat(b.temperature, 13)
        e.step(200)
    s64, s32 = e64.get_spins(), e32.get_spins()
    cosang = np.clip(np.einsum("ij,ij->i", s64, s32), -1.0, 1.0)
    ang = np.arccos(cosang)
    print(f"{name}: FP32-field line vs FP64 after 200 steps: max angle {ang.max():.3e}")
    assert ang.max() < 1e-5


def test_line_fp32_pair_programs_are_used(engine_factory):
    """The five-shell DMI template pairs every bond (T, T^T) in the FP32-field
    LINE programs; the two-shell isotropic one too (diagonal tensors)."""
    for name in ("dmi_lz128", "two_shell_lz64"):
        e = line_engine(engine_factory, SYSTEMS[name](), prec=E.PREC_FP32_FIELD)
        assert " pairs=1 " in e.describe(), e.describe()
    e = line_engine(engine_factory, SYSTEMS["dmi_lz128"]())
    assert " pairs=0 " in e.describe(), e.describe()


def test_line_fp32_field_one_step(engine_factory):
    """One FP32-field step over the LINE layout from random spins against the
    oracle's FP64 step: FP32 spins in an FP64 sum move s by far less than
    1e-9."""
    b = SYSTEMS["dmi_lz128"]()
    e = line_engine(engine_factory, b, prec=E.PREC_FP32_FIELD)
    s = po.random_spins(19, b.n)
    e.set_spins(s)
    e.step(1)
    got = e.get_spins()
    o = b.oracle(nthreads=8)
    ref = s.copy()
    o.heun_step(ref, b.dt, 0.0)
    assert np.abs(got - ref).max() < 1e-9


def test_line_refuses_open_boundaries(engine_factory):
    """LINE needs every axis periodic (z neighbours come round the wave):
    asking for it on an open lattice is an error, not a silent fallback."""
    lat = po.LatticeSpec(S.SC_CELL, [[0, 0, 0], [0.5, 0.5, 0.5]], [0, 0], (8, 8, 64), (1, 1, 0))
    b = S.build(lat, S.bcc_two_shell_rows(3.0, 1.5), 2.0, 0.05, ["exchange"],
                dt=S.dt_ps(0.5e-15), spins="random", seed=2)
    e = b.engine(engine_factory, fmt=E.XFMT_LINE, lattice=True)
    with pytest.raises(Exception, match="LINE"):
        e.step(1)


def test_autotune_with_line_candidate_keeps_the_bits(engine_factory):
    """jh_autotune times LINE beside the other formats; whichever it keeps,
    FP64 trajectories stay the oracle's bit for bit."""
    b = SYSTEMS["dmi_lz128"]()
    e = b.engine(engine_factory, lattice=True)
    e.set_thermostat(b.temperature, 23)
    rep = e.autotune(3)
    print(rep, flush=True)
    assert "format=6 block=256" in rep, rep
    e.step(40)
    got = e.get_spins()
    assert np.array_equal(got, oracle_run(b, 23, 40))


def _general_engine(factory, b, fmt, prec, dipole):
    """The system's exchange plus the other terms the Heun kernels fuse:
    uniaxial anisotropy of power 4 about a tilted axis and an ac Zeeman field
    (both take the general path, GEN), optionally the dipole field (FFT),
    then the thermostat."""
    n = b.n
    e = factory()
    e.set_sites(b.mus, b.gyro, b.alpha)
    e.set_lattice(b.lat.size, b.lat.periodic, b.lat.nbasis)
    e.add_exchange_csr(*b.csr)
    axis = np.tile(np.array([0.36, 0.48, 0.8]), (n, 1))
    e.add_uniaxial(np.full(n, 0.05), axis, 4)
    dc = b.mus[:, None] * np.array([0.0, 0.2, 1.0])[None, :]
    ac = b.mus[:, None] * np.array([0.5, 0.0, 0.0])[None, :]
    e.add_zeeman(dc, ac, np.full(n, 2.0e3))
    if dipole:
        e.add_dipole_fft(b.lat.size, b.lat.periodic, b.lat.nbasis, b.lat.cell,
                         b.lat.basis_frac, 0.3e-9, 2.0)
    e.set_time_step(b.dt)
    e.set_precision(prec)
    e.set_exchange_format(fmt)
    e.set_spins(b.s0)
    e.set_thermostat(100.0, 61)
    return e


@pytest.mark.parametrize("name,dipole", [("dmi_lz64", False), ("sc_nn_zeeman_lz128", True)])
def test_line_general_terms_match_pattern(engine_factory, name, dipole):
    """With the general local terms (power-4 anisotropy, ac Zeeman: the GEN
    kernels), and on the one-basis lattice a dipole field (which keeps the
    reference spin order, so LINE with it needs one basis site), the LINE
    kernels give the gathering PATTERN kernels' FP64 trajectory bit for bit
    (both sum the same ordered rows and share the Heun arithmetic), and the
    FP32-field forms agree to 1e-7 per component after 30 steps (their FP32
    roundings differ: the LINE pair sums are exact, the gathering ones are
    FP32)."""
    b = SYSTEMS[name]()
    got = {}
    for prec in (E.PREC_FP64, E.PREC_FP32_FIELD):
        for fmt in (E.XFMT_LINE, E.XFMT_PATTERN):
            e = _general_engine(engine_factory, b, fmt, prec, dipole)
            d = e.describe()
            assert (" line=" in d) == (fmt == E.XFMT_LINE), d
            e.step(30)
            got[(prec, fmt)] = e.get_spins()
    assert np.array_equal(got[(E.PREC_FP64, E.XFMT_LINE)], got[(E.PREC_FP64, E.XFMT_PATTERN)])
    a, p = got[(E.PREC_FP32_FIELD, E.XFMT_LINE)], got[(E.PREC_FP32_FIELD, E.XFMT_PATTERN)]
    # |a - p| rather than an arccos, which cannot resolve angles below ~2e-8
    diff = np.abs(a - p).max()
    print(f"{name}: FP32-field LINE vs PATTERN, general terms, 30 steps: max |ds| {diff:.3e}")
    assert diff < 1e-7


def test_line_refuses_two_basis_sites_with_a_dipole(engine_factory):
    """The dipole term keeps the reference spin order, which LINE cannot use
    with two basis sites: asking for it is an error."""
    b = SYSTEMS["dmi_lz64"]()
    e = _general_engine(engine_factory, b, E.XFMT_LINE, E.PREC_FP64, True)
    with pytest.raises(Exception, match="LINE"):
        e.step(1)
