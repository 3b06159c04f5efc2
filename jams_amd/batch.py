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


class ReplicaBatch:
    """replicas: iterable of (config text, temperature K, seed, field (3,) T
    [, description])."""

    def __init__(self, replicas, device: int = 0, precision: int = PREC_FP64,
                 output_dir: str = "/tmp"):
        mus, gyro, alpha, spins = [], [], [], []
        pi, pj, vidx, vals = [], [], [], []
        dc = []
        offsets = [0]
        temps, seeds, self.desc = [], [], []
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
            self.desc.append(rep[4] if len(rep) > 4 else f"{sim.n} spins")
            sim.close()
        self.offsets = np.asarray(offsets, np.int64)
        self.n = int(self.offsets[-1])
        self.nrep = len(temps)
        self.dt = dt
        e = Engine(device)
        e.set_sites(np.concatenate(mus), np.concatenate(gyro), np.concatenate(alpha))
        e.add_exchange_list(np.concatenate(pi), np.concatenate(pj), np.concatenate(vidx),
                            np.concatenate(vals))
        dcall = np.ascontiguousarray(np.concatenate(dc))
        if np.any(dcall != 0.0):
            e.add_zeeman(dcall)
        e.set_time_step(dt)
        e.set_precision(precision)
        self.s0 = np.ascontiguousarray(np.concatenate(spins))
        e.set_spins(self.s0)
        e.set_replicas(self.offsets, temps, seeds)
        self.temperatures = np.asarray(temps)
        self.seeds = seeds
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
