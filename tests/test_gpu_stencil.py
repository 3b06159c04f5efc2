"""GPU parity of the STENCIL exchange layout (each workgroup stages the spins
its rows read in LDS; jh_set_lattice + JH_XFMT_STENCIL) against the oracle.

FP64 must reproduce the reference's ordered CSR row sums bit for bit
(hamiltonian/sparse_interaction.cc:25-33, interface/sparse_blas.h:15-27),
including the rows whose order differs because a z neighbour wraps round a
periodic boundary (the row-list launch computes those from their own row
patterns) and the rows at open boundaries (their missing neighbours read zero
spins from the tile, which adds +0 to a sum that starts at +0).  The FP32-field
form must stay within its per-evaluation bound.

The lattices are 8 x 8 x 64 cells: z runs of 64 are the layout's unit, and
the oracle builds their CSR in seconds.
"""
import numpy as np
import pytest

from jams_amd import engine as E
from oracle import pyoracle as po
from tests import systems as S

pytestmark = pytest.mark.gpu

SIZE = (8, 8, 64)


def bcc(periodic):
    return po.LatticeSpec(S.SC_CELL, [[0, 0, 0], [0.5, 0.5, 0.5]], [0, 0], SIZE, periodic)


def sc(periodic):
    return po.LatticeSpec(S.SC_CELL, [[0, 0, 0]], [0], SIZE, periodic)


SYSTEMS = {
    # C3's template: five shells with DMI, full tensors, T = 100 K
    "dmi_periodic": lambda: S.build(bcc((1, 1, 1)), S.bcc_five_shell_dmi_rows(), 2.0, 0.05,
                                    ["exchange"], dt=S.dt_ps(0.5e-15), spins="random", seed=3,
                                    temperature=100.0),
    "dmi_open_z": lambda: S.build(bcc((1, 1, 0)), S.bcc_five_shell_dmi_rows(), 2.0, 0.05,
                                  ["exchange"], dt=S.dt_ps(0.5e-15), spins="random", seed=4,
                                  temperature=100.0),
    "dmi_open_x": lambda: S.build(bcc((0, 1, 1)), S.bcc_five_shell_dmi_rows(), 2.0, 0.05,
                                  ["exchange"], dt=S.dt_ps(0.5e-15), spins="random", seed=5,
                                  temperature=100.0),
    # C2's template: two isotropic shells (diagonal tensors) and K1
    "two_shell_periodic": lambda: S.build(bcc((1, 1, 1)), S.bcc_two_shell_rows(3.0, 1.5), 2.0,
                                          0.05, ["exchange", "uniaxial"], K=0.05,
                                          dt=S.dt_ps(0.5e-15), spins="random", seed=6,
                                          temperature=300.0),
    "two_shell_open": lambda: S.build(bcc((0, 0, 0)), S.bcc_two_shell_rows(3.0, 1.5), 2.0, 0.05,
                                      ["exchange", "uniaxial"], K=0.05, dt=S.dt_ps(0.5e-15),
                                      spins="random", seed=7, temperature=300.0),
    # simple cubic, one basis site, Zeeman after the exchange
    "sc_nn_zeeman": lambda: S.build(sc((1, 1, 1)), S.sc_nn_rows(1.0), 2.0, 0.1,
                                    ["exchange", "zeeman"], zeeman_B=[0, 0, 1],
                                    dt=S.dt_ps(1.0e-15), spins="random", seed=8,
                                    temperature=50.0),
}


def noise_fn(b, seed):
    sigma = po.thermostat_sigma(b.mus, b.gyro, b.alpha, b.dt)
    return lambda k: po.scale_noise(po.normals(seed, k, b.n), sigma, b.temperature)


def stencil_engine(factory, b, prec=E.PREC_FP64):
    e = b.engine(factory, prec=prec, fmt=E.XFMT_STENCIL, lattice=True)
    d = e.describe()
    assert "stencil=off" not in d and " stencil=" in d, d
    return e


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_stencil_fp64_heun_bitwise(engine_factory, name):
    """100 FP64 Heun steps with the generated Philox stream equal
    or_heun_step bit for bit."""
    b = SYSTEMS[name]()
    e = stencil_engine(engine_factory, b)
    print(name, e.describe(), flush=True)
    seed = 41
    e.set_thermostat(b.temperature, seed)
    steps = 100
    e.step(steps)
    got = e.get_spins()
    noise = noise_fn(b, seed)
    o = b.oracle(nthreads=8)
    s = b.s0.copy()
    for k in range(steps):
        o.heun_step(s, b.dt, k * b.dt, noise(k))
    print(f"{name}: max |ds| = {np.abs(got - s).max():.3e}", flush=True)
    assert np.array_equal(got, s)


def test_stencil_wrapped_rows_go_to_the_row_list(engine_factory):
    """Periodic z: the rows whose z neighbours wrap are the row list's (the
    C3 template reaches from two cells down to one up in z from the first
    basis site and from one down to two up from the second: at most four z
    values of each of the 2 x 8 x 8 lines); open z has none."""
    e = stencil_engine(engine_factory, SYSTEMS["dmi_periodic"]())
    d = e.describe()
    assert "wrapped_rows=" in d
    nwrap = int(d.split("wrapped_rows=")[1].split()[0])
    assert 0 < nwrap <= 2 * 8 * 8 * 4, d
    e2 = stencil_engine(engine_factory, SYSTEMS["dmi_open_z"]())
    assert "wrapped_rows=0" in e2.describe()


def test_stencil_graph_replay_matches_eager(engine_factory):
    b = SYSTEMS["dmi_periodic"]()
    e = stencil_engine(engine_factory, b)
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


@pytest.mark.parametrize("name", ["dmi_periodic", "dmi_open_x", "two_shell_open"])
def test_stencil_fp32_field_bound_and_trajectory(engine_factory, name):
    """FP32-field form over the stencil: one step's s* and s stay within the
    FP32 gather rounding of FP64; 200 steps stay within 1e-5 rad of FP64."""
    b = SYSTEMS[name]()
    e64 = b.engine(engine_factory, prec=E.PREC_FP64)
    e32 = stencil_engine(engine_factory, b, prec=E.PREC_FP32_FIELD)
    for e in (e64, e32):
        e.set_thermostat(b.temperature, 13)
        e.step(200)
    s64, s32 = e64.get_spins(), e32.get_spins()
    cosang = np.clip(np.einsum("ij,ij->i", s64, s32), -1.0, 1.0)
    ang = np.arccos(cosang)
    print(f"{name}: FP32-field stencil vs FP64 after 200 steps: max angle {ang.max():.3e}")
    assert ang.max() < 1e-5


def test_stencil_fp32_pattern_and_stencil_agree(engine_factory):
    """The FP32-field stencil's field per evaluation against the FP64 oracle,
    within 2^-22 sum_j |J_ij| |s_j| (its FP32 spins and FP64 FMA sums)."""
    b = SYSTEMS["dmi_periodic"]()
    e = stencil_engine(engine_factory, b, prec=E.PREC_FP32_FIELD)
    # one step's predictor field is not exposed; compare trajectories of one
    # step from random spins against the oracle's FP64 step instead
    s = po.random_spins(19, b.n)
    e.set_spins(s)
    e.step(1)
    got = e.get_spins()
    o = b.oracle(nthreads=8)
    ref = s.copy()
    o.heun_step(ref, b.dt, 0.0)
    assert np.abs(got - ref).max() < 1e-9


def test_autotune_with_stencil_candidates_keeps_the_bits(engine_factory):
    """jh_autotune times the STENCIL shapes beside the gather formats; whichever
    it keeps, FP64 trajectories stay the oracle's bit for bit."""
    b = SYSTEMS["dmi_periodic"]()
    e = b.engine(engine_factory, lattice=True)
    e.set_thermostat(b.temperature, 23)
    rep = e.autotune(3)
    print(rep, flush=True)
    assert "format=5 block=1024" in rep or "format=5 block=512" in rep, rep
    e.step(40)
    got = e.get_spins()
    noise = noise_fn(b, 23)
    o = b.oracle(nthreads=8)
    s = b.s0.copy()
    for k in range(40):
        o.heun_step(s, b.dt, k * b.dt, noise(k))
    assert np.array_equal(got, s)
