"""The drop-in path on the GPU: config text -> host layer (include/jams_host.h)
-> engine C ABI -> gfx950 kernels, against the oracle."""
import numpy as np
import pytest

from jams_amd import configs
from jams_amd.host import Simulation
from oracle import pyoracle as po
from tests import systems as S

pytestmark = pytest.mark.gpu


def normalised(s):
    length = np.sqrt((s[:, 0] * s[:, 0] + s[:, 1] * s[:, 1]) + s[:, 2] * s[:, 2])
    return np.ascontiguousarray(s / length[:, None])


def test_c1_config_run_bitwise(tmp_path):
    """C1 from its config, 1000 steps through the run loop with a
    magnetisation monitor every 100 steps."""
    mon = 'monitors = ( { module = "magnetisation"; output_steps = 100; } );'
    sim = Simulation(configs.c1(16, monitors=mon), device=0, output_dir=str(tmp_path), name="c1")
    sim.run()
    got = sim.engine().get_spins()
    b = S.c1(16)
    o = b.oracle(nthreads=8)
    s = normalised(b.s0)
    mags = []
    for k in range(1000):
        if k % 100 == 0:
            mags.append(po.kahan_sum3(b.mus[:, None] * s) / po.kahan_sum(b.mus))
        o.heun_step(s, b.dt, k * b.dt)
    mags.append(po.kahan_sum3(b.mus[:, None] * s) / po.kahan_sum(b.mus))
    assert np.array_equal(got, s)
    rows = [l.split("\t") for l in (tmp_path / "c1_mag.tsv").read_text().splitlines()[1:]]
    assert len(rows) == 11
    m = np.array([[float(x) for x in r[5:8]] for r in rows])
    assert np.allclose(m, np.array(mags), rtol=0, atol=1e-7)  # printed with 9 digits


def test_c2_small_config_thermal_run_matches_oracle_stream():
    sim = Simulation(configs.c2(8, steps=100), device=0, output_dir="/tmp", name="c2s")
    sim.run()
    got = sim.engine().get_spins()
    dt, T, seed = sim.params()
    b = S.c2(8)
    o = b.oracle(nthreads=8)
    sigma = po.thermostat_sigma(b.mus, b.gyro, b.alpha, b.dt)
    s = b.s0.copy()
    for k in range(100):
        o.heun_step(s, b.dt, k * b.dt, po.scale_noise(po.normals(seed, k, b.n), sigma, T))
    assert np.array_equal(got, s)


def test_energy_monitor_and_term_energies(tmp_path):
    mon = 'monitors = ( { module = "energy"; output_steps = 10; } );'
    sim = Simulation(configs.c2(6, steps=20, monitors=mon), device=0, output_dir=str(tmp_path),
                     name="e")
    sim.run()
    lines = (tmp_path / "e_eng.tsv").read_text().splitlines()
    assert lines[0].split("\t")[1:3] == ["exchange_E_meV", "uniaxial_E_meV"]
    assert len(lines) == 1 + 3
    # spins start along +z: E_ex = -1/2 N sum_j J_ij, E_ani = -N K
    e0 = [float(x) for x in lines[1].split("\t")[1:3]]
    n = 432
    assert abs(e0[0] - (-0.5 * n * (8 * 3.0 + 6 * 1.5))) < 1e-9
    assert abs(e0[1] - (-n * 0.05)) < 1e-12


def test_rk4_solver_module_from_config(tmp_path):
    """solver.module = "llg-rk4-gpu" (core/solver.cc:71) selects the RK4
    integrator: the C1 lattice at 8^3, 200 steps, equal to the oracle's RK4."""
    text = configs.c1(8, steps=200).replace('module = "llg-heun-gpu"', 'module = "llg-rk4-gpu"')
    sim = Simulation(text, device=0, output_dir=str(tmp_path), name="rk4")
    sim.run()
    e = sim.engine()
    assert "integrator=rk4" in e.describe()
    got = e.get_spins()
    b = S.c1(8)
    o = b.oracle(nthreads=8)
    s = normalised(b.s0)
    for k in range(200):
        o.rk4_step(s, b.dt, k * b.dt)
    assert np.array_equal(got, s)


@pytest.mark.parametrize("module", ["llg-heun-gpu", "llg-rk4-gpu"])
def test_checkpoint_resume_continues_bitwise(tmp_path, module):
    """bcc 6^3 x 2 at 300 K: 120 steps straight equal 50 steps, a checkpoint
    (the "checkpoint" monitor's file), and 70 more steps of a new simulation
    resumed from it: the spins and the thermostat stream continue exactly."""
    mon = 'monitors = ( { module = "checkpoint"; output_steps = 50; } );'
    text = configs.c2(6, steps=120, monitors=mon).replace('module = "llg-heun-gpu"',
                                                          f'module = "{module}"')
    whole = Simulation(text, device=0, output_dir=str(tmp_path), name="whole")
    whole.run()
    ref = whole.engine().get_spins()
    assert whole.iteration()[0] == 120
    whole.close()
    part = Simulation(text, device=0, output_dir=str(tmp_path), name="part")
    part.run(50)
    ck = tmp_path / "part_checkpoint.tsv"
    assert "# iteration 50" in ck.read_text()
    part.close()
    resumed = Simulation([text, f'sim : {{ resume = "{ck}"; }};'], device=0,
                         output_dir=str(tmp_path), name="resumed")
    assert resumed.iteration()[0] == 50
    resumed.run()
    assert resumed.iteration()[0] == 120
    assert np.array_equal(resumed.engine().get_spins(), ref)
    # the explicit writer gives the same file format
    resumed.write_checkpoint(str(tmp_path / "final.tsv"))
    assert "# iteration 120" in (tmp_path / "final.tsv").read_text()


def test_exchange_functional_config_run_bitwise(tmp_path):
    """A config whose exchange comes from exchange-functional (gaussian, fcc
    4^3 periodic, two shells within r_cutoff) runs 100 Heun steps on the GPU
    equal, bit for bit, to the oracle's Heun step on the oracle's CSR."""
    import math
    from tests.test_host import distance_config, spec
    ham = ('{ module = "exchange-functional"; energy_units = "meV"; interactions = ( '
           '("A", "A", "gaussian", 1.05, [2.0, 0.7, 0.3]) ); }')
    text = distance_config("fcc", (4, 4, 4), (1, 1, 1), ham).replace("spin = [0.0, 0.0, 1.0]",
                                                                      'spin = "random"')
    sim = Simulation(text, device=0, output_dir=str(tmp_path), name="fun")
    sim.run()
    got = sim.engine().get_spins()
    lat = spec("fcc", (4, 4, 4), (1, 1, 1))

    def f(r3):
        r = math.sqrt(r3[0] * r3[0] + r3[1] * r3[1] + r3[2] * r3[2])
        return 2.0 * math.exp(-((r - 0.7) * (r - 0.7)) / (2 * (0.3 * 0.3)))
    pairs, _ = po.functional_pairs(lat, {(0, 0): (1.05, f)})
    rowptr, col, val = po.pairs_csr(lat.n, pairs)
    mus, gyro, alpha, s0 = sim.sites()
    o = po.OracleSystem(mus, gyro, alpha, nthreads=8)
    o.add_exchange(rowptr, col, val)
    dt, _, _ = sim.params()
    s = np.ascontiguousarray(s0.copy())
    for k in range(sim.max_steps):
        o.heun_step(s, dt, k * dt)
    assert sim.max_steps == 100
    assert np.array_equal(got, s)
