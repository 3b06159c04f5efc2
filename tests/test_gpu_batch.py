"""The batching form: independent replicas in one context (jh_set_replicas),
hipGraph replay (jh_set_graph_steps) and the launch-shape autotuner
(jh_autotune), through the C ABI.

Bar: a replica in the batch evolves bit for bit as the same system alone in
its own context (FP64; thermal replicas included, since the noise stream is
keyed by the replica's seed and local index); graph replays equal eager
launches bit for bit, replay after replay, with host work in between; the
autotuned launch shape gives the same bits.
"""
import numpy as np
import pytest

from jams_amd import configs
from jams_amd.batch import ReplicaBatch
from jams_amd import engine as E
from jams_amd.engine import Engine
from jams_amd.host import Simulation

pytestmark = pytest.mark.gpu


def small_replicas():
    def sc(L, T, B, seed):
        return (configs.sc_config(L=L, J=1.0, moment=2.0, alpha=0.1, field=None, t_step=1e-15,
                                  temperature=T, seed=seed), T, seed, (0.0, 0.0, B), f"sc {L}")

    def bcc(L, T, B, seed):
        return (configs.bcc_config(L=L, rows=configs.bcc_nn_rows(1.0), moment=2.0, alpha=0.1,
                                   K=None, t_step=1e-15, temperature=T, seed=seed,
                                   spin='"random"'), T, seed, (0.0, 0.0, B), f"bcc {L}")

    return [sc((5, 5, 5), 0.0, 0.5, 11), bcc(3, 300.0, 0.0, 12), sc((8, 8, 8), 1000.0, 2.0, 13),
            bcc(4, 50.0, 1.0, 14), sc((3, 4, 5), 10.0, 0.0, 15), bcc(2, 0.0, 0.0, 16)]


def alone(rep, steps):
    """The replica in a context of its own, thermostat from jh_set_thermostat."""
    text, T, seed, B = rep[:4]
    sim = Simulation(text, device=-1, output_dir="/tmp")
    m, g, a, s = sim.sites()
    dt, _, _ = sim.params()
    i, j, v, table = sim.exchange_list(0)
    e = Engine(0)
    e.set_sites(m, g, a)
    e.add_exchange_list(i, j, v, table)
    dc = m[:, None] * np.asarray(B, np.float64)[None, :]
    if np.any(dc != 0.0):
        e.add_zeeman(np.ascontiguousarray(dc))
    e.set_time_step(dt)
    e.set_spins(s)
    e.set_thermostat(T, seed)
    e.step(steps)
    out = e.get_spins()
    e.close()
    sim.close()
    return out


def test_replicas_evolve_as_alone():
    reps = small_replicas()
    b = ReplicaBatch(reps)
    b.engine.step(40)
    for r, rep in enumerate(reps):
        got = b.spins(r)
        ref = alone(rep, 40)
        print(f"replica {r} ({rep[4]}, T={rep[1]}): bitwise {np.array_equal(got, ref)}")
        assert np.array_equal(got, ref), r
    b.close()


def test_replica_magnetisation():
    """Per-replica sums of mus s and mus against numpy on the spins (every
    site here has moment 2 muB)."""
    from tests import systems as S
    b = ReplicaBatch(small_replicas())
    b.engine.step(10)
    m = b.engine.replica_magnetisation()
    s = b.spins()
    mu = 2.0 * S.BOHR
    for r in range(b.nrep):
        sl = slice(b.offsets[r], b.offsets[r + 1])
        n_r = b.offsets[r + 1] - b.offsets[r]
        assert abs(m[r, 3] - n_r * mu) <= 1e-12 * n_r * mu
        assert np.allclose(m[r, :3], mu * s[sl].sum(0), rtol=0, atol=1e-12 * n_r * mu)
    b.close()


def test_graph_replay_matches_eager_launches():
    reps = small_replicas()
    eager = ReplicaBatch(reps)
    graph = ReplicaBatch(reps)
    graph.engine.set_graph_steps(10)
    for k, n in enumerate((10, 10, 25, 3, 20)):
        eager.engine.step(n)
        graph.engine.step(n)
        a, g = eager.spins(), graph.spins()
        # host work between replays: a field evaluation and a reduction
        eager.engine.fields()
        graph.engine.fields()
        graph.engine.replica_magnetisation()
        print(f"after call {k} ({n} steps): bitwise {np.array_equal(a, g)}")
        assert np.array_equal(a, g), k
    assert "graph=10" in graph.engine.describe()
    eager.close()
    graph.close()


def test_autotuned_launch_shape_gives_the_same_bits():
    reps = small_replicas()
    ref = ReplicaBatch(reps)
    tuned = ReplicaBatch(reps)
    report = tuned.engine.autotune(5)
    print(report)
    assert "chosen" in report
    # the autotuner restores the state it started from
    assert np.array_equal(tuned.spins(), ref.spins())
    ref.engine.step(30)
    tuned.engine.step(30)
    assert np.array_equal(tuned.spins(), ref.spins())
    ref.close()
    tuned.close()


@pytest.mark.parametrize("bs", [64, 128, 512])
def test_block_sizes_give_the_same_bits(engine_factory, bs):
    from tests import systems as S
    b = S.c2(8)
    e1 = b.engine(engine_factory)
    e2 = b.engine(engine_factory)
    e2.set_block_size(bs)
    for e in (e1, e2):
        e.set_thermostat(300.0, 3)
        e.step(20)
    assert np.array_equal(e1.get_spins(), e2.get_spins())


# ---------------------------------------------------------------------------
# The 256-replica workload against the oracle (solvers/cpu_llg_heun.cc:45-148
# per replica; the batch itself has no reference counterpart)

def replica_oracle(b, r, desc):
    """The oracle for replica r of a batch of configs.c4(): its lattice and
    nearest-neighbour rows restated from the description, its sites, Zeeman
    field mus B (after the exchange, as the batch orders the terms) and its
    initial spins."""
    from oracle import pyoracle as po
    from tests import systems as S
    a, z = b.offsets[r], b.offsets[r + 1]
    kind, size = desc.split()[0], int(desc.split()[1].split("^")[0])
    if kind == "sc":
        lat = po.LatticeSpec(S.SC_CELL, [[0, 0, 0]], [0], (size,) * 3, (1, 1, 1))
        rows = S.sc_nn_rows(1.0)
    else:
        lat = S.bcc_lattice(size)
        rows = [(0, 0, [sx * 0.5, sy * 0.5, sz * 0.5], 1.0)
                for sx in (1, -1) for sy in (1, -1) for sz in (1, -1)]
    assert lat.n == z - a
    o = po.OracleSystem(b.mus[a:z], b.gyro[a:z], b.alpha[a:z], nthreads=8)
    o.add_exchange(*po.exchange_csr(lat.n, *po.neighbour_list(lat, rows)))
    o.add_zeeman(b.zeeman_dc[a:z])
    sigma = po.thermostat_sigma(b.mus[a:z], b.gyro[a:z], b.alpha[a:z], b.dt)
    T, seed = float(b.temperatures[r]), int(b.seeds[r])
    noise = (lambda k: po.scale_noise(po.normals(seed, k, lat.n), sigma, T)) if T > 0 else (lambda k: None)
    return o, b.s0[a:z].copy(), noise


def chosen_replicas(reps):
    sizes = [(int(d.split()[1].split("^")[0]), r) for r, (*_, d) in enumerate(reps)]
    sc = max((s, r) for (s, r), (*_, d) in zip(sizes, reps) if d.startswith("sc"))[1]
    bcc = max((s, r) for (s, r), (*_, d) in zip(sizes, reps) if d.startswith("bcc"))[1]
    return sorted({0, len(reps) - 1, sc, bcc, 97})


def test_c4_workload_replicas_match_oracle_bitwise():
    """configs.c4(): 256 seeded replicas (sc L^3 and bcc (L/2)^3 x 2, L in
    8..48, T 0-1000 K, B 0-2 T), 6,356,301 spins in one context, 100 FP64 Heun
    steps with the in-kernel stream.  The coldest, the hottest, the largest sc
    and bcc replicas and one in the middle equal or_heun_step on their own
    Philox streams (seed, local index), bit for bit."""
    reps = configs.c4()
    b = ReplicaBatch(reps)
    assert b.nrep == 256 and b.n == 6356301
    b.engine.step(100)
    got = b.spins()
    for r in chosen_replicas(reps):
        o, s, noise = replica_oracle(b, r, reps[r][4])
        for k in range(100):
            o.heun_step(s, b.dt, k * b.dt, noise(k))
        g = got[b.offsets[r]:b.offsets[r + 1]]
        print(f"replica {r} ({reps[r][4]}, T = {reps[r][1]:.1f} K, B = {reps[r][3][2]:.3f} T, "
              f"first spin {b.offsets[r]}): max |ds| = {np.abs(g - s).max():.3e}")
        assert np.array_equal(g, s), r
    b.close()


def test_c4_workload_fp32_field_against_fp64():
    """The FP32-field batch of the same workload after 100 steps on the same
    streams: every replica within 1e-6 rad of FP64 per spin; the exchange
    field of the chosen replicas within 2^-22 sum_j |J_ij| |s_j| of the
    oracle's FP64 field."""
    from oracle import pyoracle as po
    reps = configs.c4()
    finals, fields = {}, {}
    for prec in (E.PREC_FP64, E.PREC_FP32_FIELD):
        b = ReplicaBatch(reps, precision=prec)
        fields[prec] = b.engine.term_field(0)
        b.engine.step(100)
        finals[prec] = b.spins()
        offsets = b.offsets
        if prec == E.PREC_FP64:
            keep = b
        else:
            b.close()
    a64, a32 = finals[E.PREC_FP64], finals[E.PREC_FP32_FIELD]
    cos = np.einsum("ij,ij->i", a64, a32)
    ang = np.arctan2(np.linalg.norm(np.cross(a64, a32), axis=1), cos)
    worst = max(ang[offsets[r]:offsets[r + 1]].max() for r in range(len(reps)))
    print(f"FP32-field vs FP64 after 100 steps: max angle {worst:.3e} rad")
    assert worst < 1e-6
    for r in chosen_replicas(reps):
        o, s, _ = replica_oracle(keep, r, reps[r][4])
        sl = slice(offsets[r], offsets[r + 1])
        ref = o.term_field(0, s)
        assert np.array_equal(fields[E.PREC_FP64][sl], ref), r
        bound = po.csrmv(o.rowptr, o.col, np.abs(o.val), np.abs(s).reshape(-1)).reshape(-1, 3)
        assert np.all(np.abs(fields[E.PREC_FP32_FIELD][sl] - ref) <= bound * 2.0**-22), r
    keep.close()
