"""Continuous-batching MD job server on one GPU.

Jobs (structure, potential, ensemble, step count) arrive over time.  The server admits a
queued job as soon as its HBM footprint fits the engine's budget, keeps every admitted
job's state (positions, velocities, neighbour lists, thermostat state) resident on the
device, and advances all resident jobs together: each tick is one batched launch
sequence of `tick_steps` MD steps over every resident job, each job at its own step
count.  A tick is queued on the device before the host admits the jobs that fit
(ndg_engine_step_enqueue): their host-side preparation runs while the tick executes, their
uploads and list builds follow it on the stream (an admission returns once its build is
checked), and they join the next tick.  Finished jobs are fetched and
released after the tick (ndg_engine_step_wait), which frees memory for the queue.
Latency = completion - arrival; throughput = atom-steps of finished work / wall time.
Each job's result does not depend on what else is resident (per-replica kernels,
fixed-order reductions).
"""
from __future__ import annotations

import dataclasses
import time
from collections import deque
from typing import Dict, List, Optional

import numpy as np

from . import engine as G
from . import workloads as W


@dataclasses.dataclass
class Job:
    jid: int
    system: W.System
    nsteps: int
    ensemble: str = "nve"          # "nve" | "nvt"
    temp_k: float = 300.0
    arrival: float = 0.0           # seconds after the server start
    skin: float = 0.5
    # Hold/Disp groups (hold, bc_disp, bc_dir, bc_parm) as Simulation.set_holds; spring
    # pairs and their initial positions come with the system
    holds: Optional[tuple] = None
    # filled by the server
    slot: Optional[int] = None
    admitted: Optional[float] = None
    finished: Optional[float] = None
    result: Optional[dict] = None
    thermo: Optional[list] = None

    @property
    def atom_steps(self) -> int:
        return self.system.np * self.nsteps

    def run_params(self):
        iso = None
        if self.ensemble == "nvt":
            # Berendsen velocity scaling at temp_k, tau = 100 fs (Isokinetic KE/atom tau freq)
            iso = (1.5 * W.KB_EV * self.temp_k, 100 * W.DT_1FS, 1)
        return G.run_params(self.nsteps, W.DT_1FS, out_steps=max(1, self.nsteps), skin=self.skin,
                            trigger=G.TRIGGER_SKIN, isokinetic=iso)


class Server:
    """table: an autotuner table (ndyn_omp_amd.autotune.TuneTable) whose size classes set the
    threads per atom and cell-block shape of admitted jobs, and whose skins replace a job's
    list skin by its size class."""

    def __init__(self, hbm_budget: int = 0, tick_steps: int = 50, path: int = G.PATH_MIXED,
                 graph_steps: int = 0, table=None):
        self.eng = G.BatchEngine(path=path, hbm_budget=hbm_budget, graph_steps=graph_steps)
        self.table = table
        if table is not None:
            table.apply(self.eng)
        self.tick_steps = tick_steps
        self.queue: deque = deque()
        self.active: Dict[int, Job] = {}
        self.done: List[Job] = []
        self.t0 = time.perf_counter()

    def now(self) -> float:
        return time.perf_counter() - self.t0

    def submit(self, job: Job):
        self.queue.append(job)

    def _admit(self):
        """Admit queued jobs in arrival order while their footprint fits."""
        while self.queue:
            job = self.queue[0]
            if self.table is not None:
                job.skin = self.table.skin_for(job.system.np, job.skin)
            slot = self.eng.admit(job.system, job.run_params(), raise_on_full=False,
                                  holds=job.holds)
            if slot is None:
                if not self.active:
                    raise RuntimeError(f"job {job.jid} ({job.system.np} atoms) exceeds the HBM budget")
                return
            self.queue.popleft()
            job.slot = slot
            job.admitted = self.now()
            self.active[slot] = job

    def _retire(self):
        for slot in list(self.active):
            st = self.eng.status(slot)
            if st["err"]:
                raise G.NdgError(st["err"], f"job {self.active[slot].jid} failed")
            if st["finished"]:
                job = self.active.pop(slot)
                job.thermo = self.eng.thermo(slot)
                job.result = self.eng.fetch(slot, job.system.np)
                job.finished = self.now()
                self.eng.release(slot)
                self.done.append(job)

    def run(self, jobs: List[Job], realtime: bool = True, log=None) -> dict:
        """Serve `jobs` (sorted by arrival).  realtime=True honours arrival times on the
        wall clock; False admits each job as soon as it is submitted."""
        pending = deque(sorted(jobs, key=lambda j: j.arrival))
        self.done = []
        self.t0 = time.perf_counter()
        busy = 0.0
        ticks = 0
        # host time by phase: admissions while a tick runs, the wait for the tick, the
        # retirement of finished jobs (status, thermo, fetch, release), idle waiting for work
        t_admit = t_wait = t_retire = t_idle = 0.0
        resident = 0.0   # resident atoms summed over ticks
        next_log = 5.0
        while pending or self.queue or self.active:
            t = self.now()
            while pending and (not realtime or pending[0].arrival <= t):
                self.submit(pending.popleft())
            if not self.active:
                ti = time.perf_counter()
                self._admit()
                if not self.active:
                    if pending:
                        time.sleep(max(0.0, min(0.05, pending[0].arrival - self.now())))
                    t_idle += time.perf_counter() - ti
                    continue
                t_idle += time.perf_counter() - ti
            tb = time.perf_counter()
            resident += sum(j.system.np for j in self.active.values())
            self.eng.step_enqueue(self.tick_steps)
            # jobs admitted now are prepared while the tick runs and join the next one
            self._admit()
            ta = time.perf_counter()
            self.eng.step_wait()
            tw = time.perf_counter()
            busy += tw - tb
            t_admit += ta - tb
            t_wait += tw - ta
            ticks += 1
            self._retire()
            t_retire += time.perf_counter() - tw
            if log and self.now() >= next_log:
                next_log = self.now() + 5.0
                log(f"serve: t={self.now():.1f}s done={len(self.done)} active={len(self.active)} "
                    f"queued={len(self.queue)} pending={len(pending)}")
        wall = self.now()
        lat = np.array([j.finished - j.arrival for j in self.done])
        atom_steps = float(sum(j.atom_steps for j in self.done))
        return dict(jobs=len(self.done), wall_s=wall, busy_s=busy, ticks=ticks,
                    admit_s=t_admit, wait_s=t_wait, retire_s=t_retire, idle_s=t_idle,
                    mean_resident_atoms=resident / ticks if ticks else 0.0,
                    atom_steps=atom_steps, atom_steps_per_s=atom_steps / wall,
                    atom_steps_per_busy_s=atom_steps / busy if busy > 0 else 0.0,
                    p50_latency_s=float(np.percentile(lat, 50)) if len(lat) else None,
                    p99_latency_s=float(np.percentile(lat, 99)) if len(lat) else None)

    def close(self):
        self.eng.close()


def jobs_from_stream(stream) -> List[Job]:
    """Jobs from workloads.serving_jobs tuples (arrival, System, nsteps, ensemble, T)."""
    return [Job(k, s, n, ens, t, arrival=a) for k, (a, s, n, ens, t) in enumerate(stream)]
