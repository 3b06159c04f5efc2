"""One persistent launch per Heun step (jh_set_step_launch with
JH_STEP_LAUNCH_PERSISTENT, the batching form's launch): the workgroups that
are resident at once take the predictor over every spin of every replica,
meet at a grid barrier and take the corrector, through the C ABI.

Bar: bit for bit the two-launch step (the same per-spin arithmetic in the
same order), in FP64 and in the FP32-field form, in plain launches and in
graph replays with host work between them; on the 256-replica workload, where
every workgroup takes several spin blocks per phase, also bit for bit the
oracle's replicas (solvers/cpu_llg_heun.cc:45-148 per replica); a context the
one launch cannot serve (a dipole term, whose transforms run between the
phases; the LINE layout) is refused, not run another way.
"""
import numpy as np
import pytest

from jams_amd import configs
from jams_amd import engine as E
from jams_amd.batch import ReplicaBatch
from oracle import pyoracle as po
from tests import systems as S
from tests.test_gpu_batch import chosen_replicas, replica_oracle, small_replicas

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [E.PREC_FP64, E.PREC_FP32_FIELD])
@pytest.mark.parametrize("graph", [0, 10])
def test_small_batch_persistent_equals_two_launches(prec, graph):
    reps = small_replicas()
    two = ReplicaBatch(reps, precision=prec)
    one = ReplicaBatch(reps, precision=prec)
    one.engine.set_step_launch(E.STEP_LAUNCH_PERSISTENT)
    one.engine.set_graph_steps(graph)
    for k, n in enumerate((10, 25, 3, 20)):
        two.engine.step(n)
        one.engine.step(n)
        a, b = two.spins(), one.spins()
        # host work between calls: a field evaluation and a reduction
        one.engine.fields()
        one.engine.replica_magnetisation()
        print(f"prec {prec} graph {graph} call {k} ({n} steps): bitwise {np.array_equal(a, b)}",
              flush=True)
        assert np.array_equal(a, b), k
    d = one.engine.describe()
    assert "launch=persistent" in d, d
    assert "launch=two" in two.engine.describe()
    two.close()
    one.close()


@pytest.mark.parametrize("bs", [128, 512])
def test_single_system_persistent_equals_two_launches(engine_factory, bs):
    """One system (the C2 template at L = 8, basis-major order, T = 300 K):
    the persistent launch at 128 and 512 threads gives the two-launch bits."""
    b = S.c2(8)
    e2 = b.engine(engine_factory)
    e1 = b.engine(engine_factory)
    e1.set_step_launch(E.STEP_LAUNCH_PERSISTENT)
    for e in (e1, e2):
        e.set_block_size(bs)
        e.set_thermostat(300.0, 5)
        e.step(40)
    assert np.array_equal(e1.get_spins(), e2.get_spins())


def test_c4_workload_persistent_matches_oracle_bitwise():
    """configs.c4() (6,356,301 spins: more spin blocks than resident
    workgroups) as one persistent launch per step in 20-step graph replays,
    60 FP64 steps: the coldest, the hottest, the largest sc and bcc replicas
    and one in the middle equal or_heun_step on their own streams, and the
    whole batch equals the two-launch batch, bit for bit."""
    reps = configs.c4()
    one = ReplicaBatch(reps)
    one.engine.set_step_launch(E.STEP_LAUNCH_PERSISTENT)
    one.engine.set_graph_steps(20)
    one.engine.step(60)
    got = one.spins()
    assert "launch=persistent" in one.engine.describe()
    for r in chosen_replicas(reps):
        o, s, noise = replica_oracle(one, r, reps[r][4])
        for k in range(60):
            o.heun_step(s, one.dt, k * one.dt, noise(k))
        g = got[one.offsets[r]:one.offsets[r + 1]]
        print(f"replica {r} ({reps[r][4]}): max |ds| = {np.abs(g - s).max():.3e}", flush=True)
        assert np.array_equal(g, s), r
    one.close()
    two = ReplicaBatch(reps)
    two.engine.step(60)
    assert np.array_equal(two.spins(), got)
    two.close()


def test_autotune_times_the_persistent_launch():
    reps = small_replicas()
    ref = ReplicaBatch(reps)
    tuned = ReplicaBatch(reps)
    report = tuned.engine.autotune(5)
    print(report)
    assert "launch=persistent" in report and "launch=two" in report
    assert report.splitlines()[-1].startswith("chosen ")
    assert np.array_equal(tuned.spins(), ref.spins())
    ref.engine.step(30)
    tuned.engine.step(30)
    assert np.array_equal(tuned.spins(), ref.spins())
    ref.close()
    tuned.close()


def test_persistent_refuses_dipole_and_line(engine_factory):
    lat = S.bcc_lattice(8)
    b = S.build(lat, S.bcc_two_shell_rows(3.0, 1.5), 2.0, 0.05, ["exchange"],
                dt=S.dt_ps(0.5e-15), spins="random", seed=2)
    e = b.engine(engine_factory)
    e.add_dipole_fft(lat.size, lat.periodic, lat.nbasis, lat.cell, lat.basis_frac, 0.3e-9, 2.0)
    e.set_step_launch(E.STEP_LAUNCH_PERSISTENT)
    with pytest.raises(Exception, match="persistent"):
        e.step(1)
    lat3 = po.LatticeSpec(S.SC_CELL, [[0, 0, 0], [0.5, 0.5, 0.5]], [0, 0], (8, 8, 64), (1, 1, 1))
    c3 = S.build(lat3, S.bcc_five_shell_dmi_rows(), 2.0, 0.05, ["exchange"],
                 dt=S.dt_ps(0.5e-15), spins="random", seed=3)
    e = c3.engine(engine_factory, fmt=E.XFMT_LINE, lattice=True)
    e.set_step_launch(E.STEP_LAUNCH_PERSISTENT)
    with pytest.raises(Exception, match="persistent"):
        e.step(1)
    # AUTO does not take LINE when the step is one launch
    e = c3.engine(engine_factory, lattice=True)
    e.set_step_launch(E.STEP_LAUNCH_PERSISTENT)
    e.step(2)
    d = e.describe()
    assert " line=" not in d and "launch=persistent" in d, d


def test_persistent_with_general_local_terms(engine_factory):
    """The general local terms (uniaxial anisotropy of power 4 about a tilted
    axis, an ac Zeeman field: the GEN instantiations) on the C2 template at
    L = 8, FP64 and FP32-field, T = 300 K: one persistent launch per step
    gives the two-launch bits."""
    b = S.c2(8)
    n = b.n
    axis = np.tile(np.array([0.36, 0.48, 0.8]), (n, 1))
    dc = b.mus[:, None] * np.array([0.0, 0.2, 1.0])[None, :]
    ac = b.mus[:, None] * np.array([0.5, 0.0, 0.0])[None, :]
    for prec in (E.PREC_FP64, E.PREC_FP32_FIELD):
        got = []
        for launch in (E.STEP_LAUNCH_TWO, E.STEP_LAUNCH_PERSISTENT):
            e = engine_factory()
            e.set_sites(b.mus, b.gyro, b.alpha)
            e.add_exchange_csr(*b.csr)
            e.add_uniaxial(np.full(n, 0.05), axis, 4)
            e.add_zeeman(dc, ac, np.full(n, 2.0e3))
            e.set_time_step(b.dt)
            e.set_precision(prec)
            e.set_spins(b.s0)
            e.set_thermostat(300.0, 17)
            e.set_step_launch(launch)
            e.step(30)
            got.append(e.get_spins())
        assert np.array_equal(got[0], got[1]), prec
