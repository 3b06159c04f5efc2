"""Batching form: many independent spin systems in one engine context.

Each replica is a JAMS config set up by the C++ host layer without a device
(lattice, neighbour list, sites, initial spins), plus its own temperature,
noise seed and applied field.  The batch concatenates the replicas into one
engine context in which replica r owns the spins [offsets[r], offsets[r+1])
(jh_set_replicas): one launch per Heun phase covers every replica, the step
loop can replay a captured hipGraph (jh_set_graph_steps), and the exchange
format and workgroup size are picked by timing (jh_autotune).  A replica
evolves bit for bit as it would in a context of its own.
"""
from __future__ import annotations

import numpy as np

from .engine import PREC_FP64, Engine
from .host import Simulation


def concatenate(replicas, output_dir: str = "/tmp") -> dict:
    """Set the replicas up with the host layer (no device) and lay them end to
    end: sites, spins, exchange list (indices shifted by each replica's first
    spin, tensor indices by its first table row), Zeeman fields mus B, and the
    replica table (offsets, temperatures, seeds)."""
    mus, gyro, alpha, spins = [], [], [], []
    pi, pj, vidx, vals = [], [], [], []
    dc = []
    offsets = [0]
    temps, seeds, desc = [], [], []
    dt = None
    nvals = 0
    for rep in replicas:
        text, T, seed, B = rep[:4]
        sim = Simulation(text, device=-1, output_dir=output_dir)
        m, g, a, s = sim.sites()
        d, _, _ = sim.params()
        if dt is None:
            dt = d
        elif d != dt:
            raise ValueError("all replicas must share one time step")
        off = offsets[-1]
        i, j, v, table = sim.exchange_list(0)
        pi.append(i + off)
        pj.append(j + off)
        vidx.append(v + nvals)
        vals.append(table)
        nvals += table.shape[0]
        mus.append(m)
        gyro.append(g)
        alpha.append(a)
        spins.append(s)
        # applied_field.cc:146-148: h = mus_i B
        dc.append(m[:, None] * np.asarray(B, np.float64)[None, :])
        offsets.append(off + sim.n)
        temps.append(float(T))
        seeds.append(int(seed))
        desc.append(rep[4] if len(rep) > 4 else f"{sim.n} spins")
        sim.close()
    return {"offsets": np.asarray(offsets, np.int64), "temperatures": np.asarray(temps),
            "seeds": seeds, "dt": dt, "desc": desc,
            "mus": np.concatenate(mus), "gyro": np.concatenate(gyro),
            "alpha": np.concatenate(alpha), "spins": np.ascontiguousarray(np.concatenate(spins)),
            "pi": np.concatenate(pi), "pj": np.concatenate(pj), "vidx": np.concatenate(vidx),
            "values9": np.concatenate(vals), "zeeman_dc": np.ascontiguousarray(np.concatenate(dc))}


class ReplicaBatch:
    """replicas: iterable of (config text, temperature K, seed, field (3,) T
    [, description])."""

    def __init__(self, replicas, device: int = 0, precision: int = PREC_FP64,
                 output_dir: str = "/tmp"):
        cat = concatenate(replicas, output_dir)
        self.offsets = cat["offsets"]
        self.n = int(self.offsets[-1])
        self.nrep = len(cat["seeds"])
        self.dt = cat["dt"]
        self.desc = cat["desc"]
        e = Engine(device)
        e.set_sites(cat["mus"], cat["gyro"], cat["alpha"])
        e.add_exchange_list(cat["pi"], cat["pj"], cat["vidx"], cat["values9"])
        if np.any(cat["zeeman_dc"] != 0.0):
            e.add_zeeman(cat["zeeman_dc"])
        e.set_time_step(self.dt)
        e.set_precision(precision)
        self.s0 = cat["spins"]
        e.set_spins(self.s0)
        e.set_replicas(self.offsets, cat["temperatures"], cat["seeds"])
        self.temperatures = cat["temperatures"]
        self.seeds = cat["seeds"]
        self.mus, self.gyro, self.alpha = cat["mus"], cat["gyro"], cat["alpha"]
        self.zeeman_dc = cat["zeeman_dc"]
        self.engine = e

    def spins(self, r: int | None = None) -> np.ndarray:
        s = self.engine.get_spins()
        return s if r is None else s[self.offsets[r]:self.offsets[r + 1]]

    def magnetisation(self) -> np.ndarray:
        """|m| per replica: |sum mus s| / sum mus."""
        m = self.engine.replica_magnetisation()
        return np.linalg.norm(m[:, :3], axis=1) / m[:, 3]

    def close(self):
        self.engine.close()
