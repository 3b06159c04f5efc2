"""Benchmark of the gfx950 Heun-LLG step (one JSON line on stdout).

Workload (headline): bcc 128^3 x 2 = 4,194,304 spins, full 3x3 exchange
tensors to the fifth shell (58 neighbours, J_n = 3.0 exp(-r/a) meV plus a DMI
part 0.1 J_n along seeded unit vectors), alpha = 0.05, dt = 0.5 fs, Langevin
thermostat at 100 K from the Philox stream, random initial spins from seed 3,
FP64; the same system in the FP32-field form beside it.  It is set up through
the JAMS-compatible host layer from config text, the way a JAMS user selects
the solver, and runs on one GPU.

A "step" is one full Heun step (predictor + corrector, both field
evaluations, in-kernel noise).  value = spins x steps / event time of the
step loop, inputs resident in HBM.  Before the warm-up, 0.3 s of untimed steps
let the GPU clock recover from the host-side work that precedes each timing
(set-up, spin copies).

The precision section runs the same system for --accuracy-steps steps in
FP64, in the FP32-field form and in FP64 from a state rotated by ~1e-12 rad,
from the ordered (+z) and from the random start, and states the per-spin
angle and |dm| differences and the <|m|> and exchange-energy statistics over
the last 10^3 steps.  Further legs: the bcc 64^3 x 2 two-shell system (C2),
RK4, the 256-replica batch, the dipolar film, and the CPU baseline (the
oracle's restatement of llg-heun-cpu at 1 thread and at every host thread).

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--headline-only]
                  [--no-cpu] [--no-c2] [--no-rk4] [--no-c4] [--no-c5]
                  [--accuracy-steps S] [--reps R]
  python bench.py --copy-calibration   # a plain 1 GiB copy kernel, for
                                       # scaling rocprofv3 FETCH/WRITE_SIZE
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "LLG spin-updates/s, FP64 vs FP32-field mixed at stated error; % HBM peak"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def measure(sim_text: str, steps: int, warmup: int, reps: int = 200):
    from jams_amd.host import Simulation

    t0 = time.perf_counter()
    sim = Simulation(sim_text, device=0, output_dir="/tmp")
    eng = sim.engine()
    setup_s = time.perf_counter() - t0
    log(f"  set-up {setup_s:.1f} s: {eng.describe()}")
    _settle(eng)
    eng.step(warmup)
    eng.synchronize()
    w0 = time.perf_counter()
    ms = eng.time_steps(steps)
    wall = time.perf_counter() - w0
    phases = []
    for ph in (0, 1):
        t_ms = eng.time_phase(ph, reps)
        fmt_b, csr_b = eng.phase_bytes(ph)
        phases.append({"phase": ("predictor", "corrector")[ph], "ms": t_ms, "bytes": fmt_b,
                       "gbs": fmt_b / (t_ms * 1e-3) / 1e9,
                       "csr_equiv_gbs": csr_b / (t_ms * 1e-3) / 1e9})
    n = sim.n
    out = {"n": n, "ms": ms, "wall_s": wall, "setup_s": setup_s, "phases": phases,
           "value": n * steps / (ms * 1e-3), "describe": eng.describe()}
    sim.close()
    return out


def precision_errors(a, b, mus):
    """Per-spin angle between two spin sets and |dm| of their magnetisations."""
    import numpy as np
    ang = _angles(a, b)
    m_of = lambda s: (mus[:, None] * s).sum(0) / mus.sum()
    return {"angle_rad_max": float(ang.max()), "angle_rad_rms": float(np.sqrt((ang ** 2).mean())),
            "dm_abs": float(np.linalg.norm(m_of(a) - m_of(b)))}


def c2_accuracy(L: int, steps: int, early: int = 100):
    """C2 run for `steps` steps (untimed) in FP64, in the FP32-field form and
    in FP64 from the start state rotated by ~1e-12 rad per spin, all on the
    same noise stream: the FP32-field and the perturbed run against FP64,
    after `early` steps and after `steps`.  At 300 K this lattice is
    paramagnetic and its stochastic trajectories are chaotic: the perturbed
    run shows how far the dynamics alone carries a tiny difference."""
    import numpy as np

    from jams_amd import configs
    from jams_amd.engine import PREC_FP32_FIELD, PREC_FP64
    from jams_amd.host import Simulation

    sim = Simulation(configs.c2(L), device=0, output_dir="/tmp")
    eng = sim.engine()
    s0 = eng.get_spins()
    mus = sim.sites()[0]
    d = np.random.default_rng(11).normal(size=s0.shape) * 1e-12
    s_pert = s0 + d - np.einsum("ij,ij->i", d, s0)[:, None] * s0
    s_pert /= np.linalg.norm(s_pert, axis=1)[:, None]
    at = {early: {}, steps: {}}
    for name, mode, init in (("fp64", PREC_FP64, s0), ("fp32_field", PREC_FP32_FIELD, s0),
                             ("fp64_perturbed", PREC_FP64, s_pert)):
        eng.set_precision(mode)
        eng.set_spins(init)
        eng.set_iteration(0)
        eng.step(early)
        at[early][name] = eng.get_spins()
        eng.step(steps - early)
        at[steps][name] = eng.get_spins()
    sim.close()
    out = {}
    for k, f in at.items():
        out[f"after_{k}_steps"] = {
            "fp32_field_vs_fp64": precision_errors(f["fp64"], f["fp32_field"], mus),
            "fp64_perturbed_vs_fp64": precision_errors(f["fp64"], f["fp64_perturbed"], mus)}
    return out


def measure_rk4(L: int, steps: int, warmup: int):
    """The C2 system under solver module llg-rk4-gpu (four field evaluations
    per step), FP64 and FP32-field: whole-step HIP-event time."""
    from jams_amd import configs
    from jams_amd.host import Simulation

    out = {"integrator": "llg-rk4-gpu", "steps": steps}
    for prec in ("fp64", "fp32-field"):
        text = configs.c2(L, precision=prec).replace('module = "llg-heun-gpu"',
                                                     'module = "llg-rk4-gpu"')
        sim = Simulation(text, device=0, output_dir="/tmp")
        eng = sim.engine()
        eng.step(warmup)
        ms = eng.time_steps(steps)
        out[prec.replace("-", "_")] = {"ms_per_step": ms / steps,
                                       "value": sim.n * steps / (ms * 1e-3)}
        sim.close()
    return out


def _cpu_model() -> str:
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return platform.processor() or "unknown"


def _nproc() -> int:
    try:
        return int(subprocess.run(["nproc"], capture_output=True, text=True, timeout=10).stdout)
    except (OSError, subprocess.SubprocessError, ValueError):
        return os.cpu_count() or 1


def _time_oracle(b, nthreads: int, budget_s: float, max_steps: int = 10000):
    """Heun steps of the oracle on system b at `nthreads` threads, with the
    noise drawn serially on the host each step, for about budget_s seconds."""
    from oracle import pyoracle as po

    o = b.oracle(nthreads=nthreads)
    sigma = po.thermostat_sigma(b.mus, b.gyro, b.alpha, b.dt)
    s = b.s0.copy()
    t0 = time.perf_counter()
    k = 0
    while True:
        w = po.scale_noise(po.normals(1, k, b.n), sigma, b.temperature)
        o.heun_step(s, b.dt, k * b.dt, w)
        k += 1
        el = time.perf_counter() - t0
        if el > budget_s or k >= max_steps:
            break
    return {"value": b.n * k / el, "steps": k, "seconds": el}


def cpu_baseline(L3: int, L2: int, budget_s: float = 7.0):
    """The oracle (C restatement of llg-heun-cpu, solvers/cpu_llg_heun.cc:15-148)
    timed on this host's cores: the headline system scaled to bcc L3^3 x 2
    (its full 2.19e9-entry CSR does not fit a CPU run) and the C2 system, each
    at 1 thread (the reference's default build, JAMS_BUILD_OMP=OFF) and at
    every thread nproc reports (its OpenMP build), noise drawn serially as in
    cpu_llg_heun.cc:55."""
    from tests import systems as S

    nproc = _nproc()
    model = _cpu_model()
    runs = {}
    for name, b in (("c3_scaled", S.c3(L3, with_csr=True)), ("c2", S.c2(L2))):
        for nt in (1, nproc):
            r = _time_oracle(b, nt, budget_s)
            runs[f"{name}_{nt}_threads"] = {"spins": b.n, **r}
            log(f"  CPU {name} ({b.n} spins) at {nt} threads: {r['value']:.4e} spin-updates/s "
                f"({r['steps']} steps in {r['seconds']:.1f} s)")
    head = runs[f"c3_scaled_{nproc}_threads"]
    return {"value": head["value"], "unit": "spin-updates/s", "cores": nproc, "kind": "port",
            "sample": f"the headline system scaled to bcc {L3}^3 x 2 ({head['spins']} spins, "
                      f"same 58-neighbour DMI tensors, T = 100 K): {head['steps']} Heun steps "
                      f"in {head['seconds']:.1f} s at {nproc} OpenMP threads, noise drawn "
                      f"serially (solvers/cpu_llg_heun.cc:55)",
            "nproc": nproc, "cpu_model": model,
            "one_thread_value": runs["c3_scaled_1_threads"]["value"], "runs": runs}


def _angles(a, b):
    import numpy as np
    cosang = np.einsum("ij,ij->i", a, b) / (np.linalg.norm(a, axis=1) * np.linalg.norm(b, axis=1))
    # atan2 of |a x b| and a.b keeps its precision near zero
    return np.arctan2(np.linalg.norm(np.cross(a, b), axis=1), cosang)


def _phase_times(eng, reps):
    phases = []
    for ph in (0, 1):
        t_ms = eng.time_phase(ph, reps)
        fmt_b, csr_b = eng.phase_bytes(ph)
        phases.append({"phase": ("predictor", "corrector")[ph], "ms": t_ms, "bytes": fmt_b,
                       "gbs": fmt_b / (t_ms * 1e-3) / 1e9,
                       "csr_equiv_gbs": csr_b / (t_ms * 1e-3) / 1e9})
    return phases


SETTLE_S = 0.3


def _boundary_times(eng, s, reps: int = 3):
    """The host-buffer crossings of the C ABI, wall clock with the device idle
    before each (median of `reps`): jh_set_spins (host AoS -> device layout)
    and jh_get_spins (device layout -> a new host AoS array), in ms."""
    set_t, get_t = [], []
    for _ in range(reps):
        eng.synchronize()
        t0 = time.perf_counter()
        eng.set_spins(s)
        eng.synchronize()
        set_t.append((time.perf_counter() - t0) * 1e3)
        t0 = time.perf_counter()
        eng.get_spins()
        get_t.append((time.perf_counter() - t0) * 1e3)
    return sorted(set_t)[reps // 2], sorted(get_t)[reps // 2]


def _settle(eng, seconds: float = SETTLE_S, chunk: int = 20) -> int:
    """Untimed steps for `seconds` of GPU time before a warm-up: the GPU's
    clock falls during host-side work (set-up, spin copies) and takes a few
    hundred milliseconds of load to come back, which a warm-up of a few steps
    does not cover.  Returns the steps taken."""
    done = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        eng.step(chunk)
        eng.synchronize()
        done += chunk
    return done


def _tail_run(eng, mus, steps, tail, every):
    """steps Heun steps; over the last `tail` of them, every `every` steps,
    |m| and the exchange energy per spin (meV)."""
    import numpy as np

    eng.step(steps - tail)
    mabs, energy = [], []
    for _ in range(tail // every):
        eng.step(every)
        m = eng.magnetisation()
        mabs.append(float(np.linalg.norm(m[:3]) / m[3]))
        t = eng.iteration()[1]
        energy.append(eng.term_energy(0, t) / mus.size)
    return np.asarray(mabs), np.asarray(energy)


def _blocked_mean(x, nblocks=10):
    """Mean and its standard error from nblocks contiguous blocks (the samples
    of a trajectory are correlated)."""
    import numpy as np

    k = x.size // nblocks
    blocks = x[:k * nblocks].reshape(nblocks, k).mean(1)
    return float(x.mean()), float(blocks.std(ddof=1) / np.sqrt(nblocks))


def measure_c3(L: int, steps: int, warmup: int, reps: int, accuracy_steps: int,
               tail: int = 1000, every: int = 10):
    """The headline system (bcc L^3 x 2, five shells with DMI, T = 100 K).

    For FP64 and then the FP32-field form:

    Accuracy (untimed, accuracy_steps > 0): from the ordered start (+z) and
    from the random start (FP64 also from the start rotated by ~1e-12 rad per
    spin), all on the same noise stream; afterwards the per-spin angle and
    |dm| of each against FP64 after accuracy_steps, and <|m|> and the
    exchange energy over the last `tail` steps (a statistical comparison,
    which is the meaningful one where the trajectory itself amplifies a
    1e-12 difference).

    Speed: from the workload's own random start, 0.3 s of untimed steps (the
    clock's recovery after the spin copy), `warmup` steps, then `steps` Heun
    steps timed; HIP-event times of each phase kernel; and the median of five
    500-step runs."""
    import numpy as np

    from jams_amd import configs
    from jams_amd.engine import PREC_FP32_FIELD, PREC_FP64
    from jams_amd.host import Simulation

    t0 = time.perf_counter()
    sim = Simulation(configs.c3(L), device=0, output_dir="/tmp")
    eng = sim.engine()
    n = sim.n
    setup_s = time.perf_counter() - t0
    s_random = eng.get_spins()
    mus = sim.sites()[0]
    out = {"n": n, "setup_s": setup_s, "temperature_K": 100.0}
    log(f"  set-up {setup_s:.1f} s")
    # Per precision: its untimed accuracy trajectories, then its timed steps
    # (the timed region then starts on a GPU that has been busy with the same
    # device layout, as in a long simulation, not after the re-layout's idle
    # second).
    rng = np.random.default_rng(11)
    m_of = lambda s: (mus[:, None] * s).sum(0) / mus.sum()
    starts = {"ordered": np.tile([0.0, 0.0, 1.0], (n, 1)), "random": s_random}
    perturbed = {}
    for label, s0 in starts.items():
        d = rng.normal(size=s0.shape) * 1e-12
        sp = s0 + d - np.einsum("ij,ij->i", d, s0)[:, None] * s0
        perturbed[label] = sp / np.linalg.norm(sp, axis=1)[:, None]
    tail = min(tail, accuracy_steps) if accuracy_steps > 0 else 0
    finals = {label: {} for label in starts}
    stats = {label: {} for label in starts}
    runs_of = {"fp64": (("fp64", False), ("fp64_perturbed", True)),
               "fp32_field": (("fp32_field", False),)}
    for name, mode in (("fp64", PREC_FP64), ("fp32_field", PREC_FP32_FIELD)):
        eng.set_precision(mode)
        if accuracy_steps > 0:
            for label, s0 in starts.items():
                for run, pert in runs_of[name]:
                    eng.set_spins(perturbed[label] if pert else s0)
                    eng.set_iteration(0)
                    mabs, energy = _tail_run(eng, mus, accuracy_steps, tail, every)
                    finals[label][run] = eng.get_spins()
                    mm, me = _blocked_mean(mabs)
                    em, ee = _blocked_mean(energy)
                    stats[label][run] = {"mean_abs_m": mm, "mean_abs_m_stderr": me,
                                         "exchange_energy_meV_per_spin": em,
                                         "exchange_energy_stderr": ee}
        eng.set_spins(s_random)
        eng.set_iteration(0)
        settle = _settle(eng)
        eng.step(warmup)
        eng.synchronize()
        w0 = time.perf_counter()
        ms = eng.time_steps(steps)
        wall = time.perf_counter() - w0
        out[name] = {"value": n * steps / (ms * 1e-3), "ms": ms, "ms_per_step": ms / steps,
                     "wall_s": wall, "settle_steps": settle,
                     "phases": _phase_times(eng, reps), "describe": eng.describe()}
        log(f"  C3 {name}: {ms / steps:.4f} ms/step, {out[name]['value']:.4e} spin-updates/s; "
            f"{eng.describe()}")
        # the same run with the spins handed over in host memory at both ends
        # (never the headline value: the timed region starts with them in HBM)
        set_ms, get_ms = _boundary_times(eng, s_random)
        tot = set_ms + ms + get_ms
        out[name]["host_boundary"] = {
            "set_spins_ms": set_ms, "get_spins_ms": get_ms, "bytes_each_way": n * 24,
            "steps": steps, "ms_per_step_with_transfers": tot / steps,
            "value_with_transfers": n * steps / (tot * 1e-3)}
        log(f"  C3 {name} with host buffers at both ends: set_spins {set_ms:.1f} ms, get_spins "
            f"{get_ms:.1f} ms ({n * 24 / 1e6:.1f} MB each way); {steps} steps "
            f"{tot / steps:.4f} ms/step, {n * steps / (tot * 1e-3):.4e} spin-updates/s")
        # steady state: the median of five runs of 500 steps each (reported
        # beside the headline's timing)
        eng.set_spins(s_random)
        eng.set_iteration(0)
        eng.step(50)
        runs = sorted(eng.time_steps(500) / 500 for _ in range(5))
        out[name]["median5_ms_per_step"] = runs[2]
        out[name]["median5_value"] = n / (runs[2] * 1e-3)
        log(f"  C3 {name} steady state: median of 5 x 500 steps {runs[2]:.4f} ms/step")
    out["speedup_fp32_vs_fp64"] = out["fp32_field"]["value"] / out["fp64"]["value"]
    out["median5_speedup_fp32_vs_fp64"] = (out["fp64"]["median5_ms_per_step"]
                                           / out["fp32_field"]["median5_ms_per_step"])
    log(f"  C3 speed-up FP32-field / FP64: {out['speedup_fp32_vs_fp64']:.3f} "
        f"(steady state {out['median5_speedup_fp32_vs_fp64']:.3f})")
    # beside them, FP64 with the gathering row-pattern kernels (the plain
    # path before the whole-line layout), steady state only
    if " line=" in out["fp64"]["describe"]:
        from jams_amd.engine import XFMT_AUTO, XFMT_PATTERN
        eng.set_precision(PREC_FP64)
        eng.set_exchange_format(XFMT_PATTERN)
        eng.set_spins(s_random)
        eng.set_iteration(0)
        _settle(eng)
        runs = sorted(eng.time_steps(500) / 500 for _ in range(5))
        out["fp64_gathering"] = {"median5_ms_per_step": runs[2], "describe": eng.describe()}
        out["median5_speedup_fp32_vs_fp64_gathering"] = (runs[2] /
                                                         out["fp32_field"]["median5_ms_per_step"])
        log(f"  C3 fp64 gathering kernels, steady state: {runs[2]:.4f} ms/step; FP32-field "
            f"(whole lines) / FP64 gathering: {out['median5_speedup_fp32_vs_fp64_gathering']:.3f}")
        eng.set_exchange_format(XFMT_AUTO)
    if accuracy_steps > 0:
        acc = {"steps": accuracy_steps, "tail_steps": tail, "sample_every": every}
        for label in starts:
            fin, stt = finals[label], stats[label]
            m64 = m_of(fin["fp64"])
            st = {"m_fp64": float(np.linalg.norm(m64)), "tail": stt}
            for other in ("fp32_field", "fp64_perturbed"):
                ang = _angles(fin["fp64"], fin[other])
                a, b = stt[other], stt["fp64"]
                st[other + "_vs_fp64"] = {
                    "angle_rad_max": float(ang.max()),
                    "angle_rad_rms": float(np.sqrt((ang ** 2).mean())),
                    "dm_abs": float(np.linalg.norm(m64 - m_of(fin[other]))),
                    "tail_mean_abs_m_diff": a["mean_abs_m"] - b["mean_abs_m"],
                    "tail_mean_abs_m_diff_sigma": float(np.hypot(a["mean_abs_m_stderr"],
                                                                 b["mean_abs_m_stderr"])),
                    "tail_energy_diff": a["exchange_energy_meV_per_spin"]
                    - b["exchange_energy_meV_per_spin"],
                    "tail_energy_diff_sigma": float(np.hypot(a["exchange_energy_stderr"],
                                                             b["exchange_energy_stderr"]))}
            acc[label] = st
            f, c = st["fp32_field_vs_fp64"], st["fp64_perturbed_vs_fp64"]
            log(f"  C3 {label} start, after {accuracy_steps} steps: FP32-field vs FP64 max angle "
                f"{f['angle_rad_max']:.3e} rad (rms {f['angle_rad_rms']:.3e}), |dm| {f['dm_abs']:.3e}, "
                f"tail <|m|> diff {f['tail_mean_abs_m_diff']:.2e} +- {f['tail_mean_abs_m_diff_sigma']:.1e}, "
                f"energy diff {f['tail_energy_diff']:.2e} +- {f['tail_energy_diff_sigma']:.1e} meV; "
                f"FP64 1e-12-perturbed vs FP64 max angle {c['angle_rad_max']:.3e}, |dm| {c['dm_abs']:.3e}")
        out["accuracy"] = acc
    sim.close()
    return out


def measure_dipole(steps: int, reps: int):
    """The dipolar film (sc 256 x 256 x 8 open, 524,288 spins, NN exchange +
    dipole-fft over every pair on the 512 x 512 x 16 padded grid, T = 0):
    step throughput with FP64 and with FP32 transforms from the same random
    state, the dipole evaluation's time and algorithmic bytes, and the
    FP32-vs-FP64 difference after `steps` steps."""
    import numpy as np

    from jams_amd import configs
    from jams_amd.engine import PREC_FP32_FIELD, PREC_FP64
    from jams_amd.host import Simulation

    t0 = time.perf_counter()
    sim = Simulation(configs.c5(), device=0, output_dir="/tmp")
    eng = sim.engine()
    n = sim.n
    setup_s = time.perf_counter() - t0
    s0 = eng.get_spins()
    mus = sim.sites()[0]
    dip = [t for t in range(eng.num_terms()) if eng.term_kind(t) == 3][0]
    p_c = 512 * 512 * (16 // 2 + 1)
    out = {"n": n, "steps": steps, "setup_s": setup_s}
    finals = {}
    for name, mode in (("fp64", PREC_FP64), ("fp32_field", PREC_FP32_FIELD)):
        eng.set_precision(mode)
        eng.set_spins(s0)
        eng.set_iteration(0)
        eng.step(0)
        ms = eng.time_steps(steps)
        finals[name] = eng.get_spins()
        eng.set_spins(s0)
        t_dip = eng.time_term(dip, reps)
        b_dip = eng.term_bytes(dip)
        # the scope table's fixed formula for one evaluation, 48 N + 6 P_c c
        # (tensor) + 12 P_c c (k-space spin and field, written and read),
        # P_c = 512 x 512 x (16/2 + 1), c the bytes of one complex element
        # (16 for the FP64 transforms, 8 for the FP32 ones)
        cb = 16 if mode == PREC_FP64 else 8
        b_formula = 48.0 * n + 6 * p_c * cb + 12 * p_c * cb
        out[name] = {"value": n * steps / (ms * 1e-3), "ms_per_step": ms / steps,
                     "dipole_eval_ms": t_dip, "dipole_eval_bytes": b_dip,
                     "dipole_eval_gbs": b_dip / (t_dip * 1e-3) / 1e9,
                     "dipole_eval_bytes_formula": b_formula,
                     "dipole_eval_gbs_formula": b_formula / (t_dip * 1e-3) / 1e9,
                     "engine": eng.describe()}
        log(f"  C5 {name}: {ms / steps:.4f} ms/step, {out[name]['value']:.4e} spin-updates/s; "
            f"dipole evaluation {t_dip * 1e3:.1f} us, {b_dip / 1e6:.0f} MB -> "
            f"{out[name]['dipole_eval_gbs']:.0f} GB/s")
    ang = _angles(finals["fp64"], finals["fp32_field"])
    m_of = lambda s: (mus[:, None] * s).sum(0) / mus.sum()
    out["fp32_vs_fp64"] = {"angle_rad_max": float(ang.max()),
                           "angle_rad_rms": float(np.sqrt((ang ** 2).mean())),
                           "dm_abs": float(np.linalg.norm(m_of(finals["fp64"]) - m_of(finals["fp32_field"])))}
    out["speedup_fp32_vs_fp64"] = out["fp32_field"]["value"] / out["fp64"]["value"]
    log(f"  C5 FP32 vs FP64 after {steps} steps: max angle {ang.max():.3e} rad, "
        f"|dm| {out['fp32_vs_fp64']['dm_abs']:.3e}; speed-up {out['speedup_fp32_vs_fp64']:.3f}")
    sim.close()
    return out


def measure_batch(steps: int, nrep: int = 256, graph_steps: int = 50):
    """The batching form: nrep independent replicas (sizes drawn from a seed,
    sc L^3 or bcc (L/2)^3 x 2 with L in 8..48, T_r = 1000 r/(nrep-1) K,
    B_r = 2 r/(nrep-1) T along z) in one context, one launch per Heun phase
    for all of them, the step loop replayed from a captured hipGraph, the
    exchange format and workgroup size picked by the autotuner; FP64 and
    FP32-field."""
    from jams_amd import configs
    from jams_amd.batch import ReplicaBatch
    from jams_amd.engine import PREC_FP32_FIELD, PREC_FP64

    reps = configs.c4(nrep=nrep)
    out = {"replicas": nrep, "steps": steps}
    finals = {}
    for name, mode in (("fp64", PREC_FP64), ("fp32_field", PREC_FP32_FIELD)):
        t0 = time.perf_counter()
        b = ReplicaBatch(reps, precision=mode)
        setup_s = time.perf_counter() - t0
        e = b.engine
        tune = e.autotune(5)
        # launches vs graph replay over the same short run, state restored after
        s0 = e.get_spins()
        ms_plain = e.time_steps(200)
        e.set_graph_steps(graph_steps)
        e.set_spins(s0)
        e.set_iteration(0)
        e.step(0)
        ms = e.time_steps(steps)
        m = b.magnetisation()
        finals[name] = (e.get_spins(), m)
        # the same format and block size as one persistent launch per step
        # (predictor, grid barrier, corrector), graph-replayed, from the
        # final state (timing only)
        from jams_amd.engine import STEP_LAUNCH_PERSISTENT, STEP_LAUNCH_TWO
        tuned_launch = "launch=persistent" in tune.splitlines()[-1]
        e.set_step_launch(STEP_LAUNCH_PERSISTENT)
        e.step(graph_steps)
        ms_pers = e.time_steps(1000) / 1000
        e.set_step_launch(STEP_LAUNCH_PERSISTENT if tuned_launch else STEP_LAUNCH_TWO)
        out[name] = {"spins": b.n, "value": b.n * steps / (ms * 1e-3), "ms_per_step": ms / steps,
                     "ms_per_step_plain_launches": ms_plain / 200,
                     "ms_per_step_one_persistent_launch": ms_pers, "setup_s": setup_s,
                     "autotune": tune.splitlines()[-1], "engine": e.describe(),
                     "m_coldest": float(m[0]), "m_hottest": float(m[-1])}
        how = f"as {graph_steps}-step graph replays" if graph_steps > 0 else "in plain launches"
        out[name]["graph_steps"] = graph_steps
        log(f"  C4 {name}: {b.n} spins in {nrep} replicas, {ms / steps:.4f} ms/step {how} "
            f"({ms_plain / 200:.4f} plain, 200 steps; {ms_pers:.4f} as one persistent launch "
            f"per step), {out[name]['value']:.4e} spin-updates/s; "
            f"{tune.splitlines()[-1]}; |m| coldest {m[0]:.4f} hottest {m[-1]:.4f}")
        b.close()
    out["speedup_fp32_vs_fp64"] = out["fp32_field"]["value"] / out["fp64"]["value"]
    # FP32-field against FP64 after the same steps: the per-spin angle over all
    # replicas and the largest per-replica |m| difference (hot replicas amplify
    # any perturbation, as a random start does at C3)
    import numpy as np
    (s64, m64), (s32, m32) = finals["fp64"], finals["fp32_field"]
    ang = _angles(s64, s32)
    out["fp32_vs_fp64"] = {"angle_rad_max": float(ang.max()),
                           "angle_rad_rms": float(np.sqrt((ang ** 2).mean())),
                           "replica_dm_abs_max": float(np.abs(np.asarray(m64) - np.asarray(m32)).max())}
    log(f"  C4 FP32-field vs FP64 after {steps} steps: max angle {ang.max():.3e} rad, "
        f"largest replica |d|m|| {out['fp32_vs_fp64']['replica_dm_abs_max']:.3e}")
    return out


def copy_calibration(nbytes: int = 1 << 30, reps: int = 5) -> int:
    """A plain elementwise copy of nbytes (read nbytes, write nbytes), so that
    rocprofv3's FETCH_SIZE / WRITE_SIZE per dispatch can be scaled to bytes."""
    import torch

    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(reps):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    log(f"copy calibration: {reps} x elementwise copy of {nbytes} bytes")
    return 0


def read_traffic(name: str, phase: str):
    """Counter-measured HBM bytes per launch of the phase kernel, from the
    committed FETCH_SIZE / WRITE_SIZE summary (None if absent)."""
    p = ROOT / "profiles" / name
    if not p.exists():
        return None
    try:
        d = json.loads(p.read_text())
    except ValueError:
        return None
    if phase == "corrector":
        d = d.get("corrector", {})
    return d.get("hbm_bytes_per_launch")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--L3", type=int, default=128, help="headline lattice: bcc L3^3 x 2")
    ap.add_argument("--L", type=int, default=64, help="C2 lattice: bcc L^3 x 2")
    ap.add_argument("--reps", type=int, default=50, help="launches per phase timing")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the headline system's FP64 and FP32-field timings")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c2", action="store_true")
    ap.add_argument("--no-rk4", action="store_true")
    ap.add_argument("--no-c4", action="store_true")
    ap.add_argument("--no-c5", action="store_true")
    ap.add_argument("--c4-steps", type=int, default=5000)
    ap.add_argument("--c5-steps", type=int, default=2000)
    ap.add_argument("--c4-graph-steps", type=int, default=50,
                    help="steps per captured hipGraph in the replica batch (0: plain launches)")
    # FP32-field and perturbed-FP64 runs of this many steps against FP64,
    # untimed, on the headline system and on C2 (0: none)
    ap.add_argument("--accuracy-steps", type=int, default=10000)
    ap.add_argument("--cpu-L3", type=int, default=32)
    ap.add_argument("--copy-calibration", action="store_true")
    args = ap.parse_args()
    # the host layer logs set-up to stdout, as the reference does; keep this
    # process's stdout for the one JSON line
    json_out = os.fdopen(os.dup(1), "w")
    sys.stdout.flush()
    os.dup2(2, 1)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the reference is single-GPU; sharding is a design note only).", file=sys.stderr)
        return 2

    if args.copy_calibration:
        return copy_calibration()

    from jams_amd import configs

    extras = args.headline_only
    log(f"C3 (headline): bcc {args.L3}^3 x 2, {args.steps} timed steps after {args.warmup} "
        f"warm-up, FP64 and FP32-field")
    c3 = measure_c3(args.L3, args.steps, args.warmup, args.reps,
                    0 if extras else args.accuracy_steps)
    c2 = None
    if not (extras or args.no_c2):
        log(f"C2: bcc {args.L}^3 x 2, FP64 and FP32-field")
        c2 = {"fp64": measure(configs.c2(args.L), args.steps, args.warmup, args.reps),
              "fp32_field": measure(configs.c2(args.L, precision="fp32-field"), args.steps,
                                    args.warmup, args.reps)}
        c2["speedup_fp32_vs_fp64"] = c2["fp32_field"]["value"] / c2["fp64"]["value"]
        log(f"  C2 fp64 {c2['fp64']['value']:.4e}, fp32-field {c2['fp32_field']['value']:.4e} "
            f"spin-updates/s")
        if args.accuracy_steps > 100:
            c2["accuracy"] = c2_accuracy(args.L, args.accuracy_steps)
            for when, res in c2["accuracy"].items():
                for k, v in res.items():
                    log(f"  C2 {k} {when}: max angle {v['angle_rad_max']:.3e} rad "
                        f"(rms {v['angle_rad_rms']:.3e}), |dm| {v['dm_abs']:.3e}")
    rk4 = None
    if not (extras or args.no_rk4):
        log("C2 under llg-rk4-gpu (RK4), FP64 and FP32-field")
        rk4 = measure_rk4(args.L, min(args.steps, 1000), min(args.warmup, 50))
        log(f"  fp64 {rk4['fp64']['ms_per_step'] * 1e3:.2f} us/step "
            f"({rk4['fp64']['value']:.4e} spin-updates/s), fp32-field "
            f"{rk4['fp32_field']['ms_per_step'] * 1e3:.2f} us/step")
    c4 = None
    if not (extras or args.no_c4):
        log(f"C4 batching form: 256 replicas, {args.c4_steps} steps FP64 and FP32-field")
        c4 = measure_batch(args.c4_steps, graph_steps=args.c4_graph_steps)
    c5 = None
    if not (extras or args.no_c5):
        log(f"C5 dipolar film: sc 256x256x8 open, {args.c5_steps} steps FP64 and FP32 transforms")
        c5 = measure_dipole(args.c5_steps, args.reps)
    cpu = None
    if not (extras or args.no_cpu):
        log("CPU baseline (oracle): headline system scaled, and C2")
        cpu = cpu_baseline(args.cpu_L3, args.L)

    h = c3["fp64"]
    dom = max(h["phases"], key=lambda p: p["ms"])
    # counter-measured HBM bytes of the dominant phase kernel of this workload
    # (the LINE kernel k_heun_ln<FP64>, or the gathering k_heun<FP64,
    # pattern-full>, predictor or corrector); null otherwise
    traffic = None
    is_line = " line=" in h["describe"]
    kernel = "k_heun_ln<FP64, line>" if is_line else "k_heun<FP64, pattern-full>"
    if "exchange=pattern-full " in h["describe"] and args.L3 == 128:
        traffic = read_traffic("traffic_c3_fp64_line.json" if is_line else "traffic_c3_fp64.json",
                               dom["phase"])
    f32 = c3["fp32_field"]
    line = {
        "metric": METRIC,
        "value": h["value"],
        "unit": "spin-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "settle_s": SETTLE_S,
        "ms_per_step": h["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"C3: bcc {args.L3}^3x2 ({c3['n']} spins), five shells "
                               f"(58 neighbours) J_n=3.0 exp(-r/a) meV + DMI D_n=0.1 J_n, "
                               f"alpha=0.05, dt=0.5 fs, T=100 K (Philox noise), random start, "
                               f"FP64 Heun",
                   "spins": c3["n"], "engine": h["describe"]},
        "roofline": {"bound": "hbm", "kernel": f"{kernel} {dom['phase']}",
                     "achieved": dom["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": dom["gbs"] / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": dom["bytes"],
                     "launch_ms": dom["ms"],
                     "csr_equivalent_gbs": dom["csr_equiv_gbs"]},
        "phases": h["phases"],
        "fp32_field": {"value": f32["value"], "ms_per_step": f32["ms_per_step"],
                       "speedup_vs_fp64": c3["speedup_fp32_vs_fp64"], "phases": f32["phases"],
                       "engine": f32["describe"],
                       "median5_ms_per_step": f32["median5_ms_per_step"],
                       "median5_speedup_vs_fp64": c3["median5_speedup_fp32_vs_fp64"],
                       "median5_speedup_vs_fp64_gathering":
                           c3.get("median5_speedup_fp32_vs_fp64_gathering")},
        "fp64_gathering_median5_ms_per_step": c3.get("fp64_gathering", {}).get("median5_ms_per_step"),
        "fp64_median5_ms_per_step": h["median5_ms_per_step"],
        "host_boundary": h["host_boundary"],
        "cpu_baseline": cpu,
    }
    if "accuracy" in c3:
        line["precision_c3"] = c3["accuracy"]
    if c2 is not None:
        line["c2"] = c2
    if rk4 is not None:
        line["rk4_c2"] = rk4
    if c4 is not None:
        line["batch_c4"] = c4
    if c5 is not None:
        line["dipole_c5"] = c5
    print(json.dumps(line), file=json_out, flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
