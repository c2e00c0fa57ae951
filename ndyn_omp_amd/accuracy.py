"""Accuracy of the mixed-precision path against the FP64 path of the same engine.

The FP64 path reproduces the compiled reference to ~1e-15 (tested against its goldens),
so it serves as the FP64 oracle at sizes the CPU restatement cannot reach.

- `pair_errors`: one force evaluation of both paths on identical positions and lists
  (`thermal_snapshot` gives a thermalised configuration):
  RMS relative force error, largest per-atom energy error (eV), relative errors of the
  potential energy and of the pressure (virial trace / 3V).
- `nve_drift`: total energy of an NVE run, from the thermo records: end-point change
  and least-squares slope, both per atom and scaled to 10^4 steps.
"""
from __future__ import annotations

import numpy as np

from . import engine as G
from . import workloads as W


def thermal_snapshot(s: W.System, steps: int, dt: float, skin: float) -> W.System:
    """The system after `steps` NVE steps of the FP64 path (a perfect lattice has zero
    net forces, so errors are measured on a thermalised configuration)."""
    eng = G.BatchEngine(path=G.PATH_FP64)
    slot = eng.admit(s, G.run_params(steps, dt, out_steps=steps, skin=skin,
                                     trigger=G.TRIGGER_SKIN))
    eng.step(steps)
    st = eng.fetch(slot, s.np)
    eng.close()
    t = s.copy()
    t.pos, t.vel, t.box = st["pos"], st["vel"], st["box"]
    return t


def pair_errors(s: W.System, skin: float) -> dict:
    out = {}
    for path in (G.PATH_FP64, G.PATH_MIXED):
        sim = G.Simulation(s, path=path)
        sim.build_lists(skin)
        pe, vir = sim.compute(True, True)
        st = sim.state()
        out[path] = (pe, np.array(vir), st["force"], st["pene"])
        sim.close()
    pe0, vir0, f0, e0 = out[G.PATH_FP64]
    pe1, vir1, f1, e1 = out[G.PATH_MIXED]
    vol = float(np.prod(s.box))
    p0, p1 = vir0[:3].sum() / (3 * vol), vir1[:3].sum() / (3 * vol)
    return dict(force_rms_rel=float(np.linalg.norm(f1 - f0) / np.linalg.norm(f0)),
                energy_per_atom_max_abs_ev=float(np.abs(e1 - e0).max()),
                pe_rel=float(abs(pe1 - pe0) / abs(pe0)),
                pressure_rel=float(abs(p1 - p0) / max(abs(p0), 1e-300)),
                pressure_ev_a3=float(p0))


def nve_drift(s: W.System, path: int, nsteps: int, dt: float, skin: float,
              out_steps: int = 10) -> dict:
    eng = G.BatchEngine(path=path)
    slot = eng.admit(s, G.run_params(nsteps, dt, out_steps=out_steps, skin=skin,
                                     trigger=G.TRIGGER_SKIN))
    eng.step(nsteps)
    th = eng.thermo(slot)
    st = eng.status(slot)
    eng.close()
    steps = np.array([r["step"] for r in th], dtype=np.float64)
    e = np.array([r["pe"] + r["ke"] for r in th])
    n = s.np
    slope = np.polyfit(steps, e, 1)[0] if len(th) > 2 else 0.0
    return dict(steps=int(steps[-1]), records=len(th),
                drift_per_atom_ev=float((e[-1] - e[0]) / n),
                drift_per_atom_per_1e4_ev=float((e[-1] - e[0]) / n * 1e4 / max(steps[-1] - steps[0], 1)),
                slope_per_atom_per_1e4_ev=float(slope * 1e4 / n),
                max_dev_per_atom_ev=float(np.abs(e - e[0]).max() / n),
                e0_per_atom_ev=float(e[0] / n), err=int(st["err"]))


def nve_drift_batch(systems, path: int, nsteps: int, dt: float, skin: float,
                    out_steps: int = 100) -> dict:
    """nve_drift of several systems advanced together in one engine; the drift is that of
    their summed energy, per atom."""
    eng = G.BatchEngine(path=path)
    slots = [eng.admit(s, G.run_params(nsteps, dt, out_steps=out_steps, skin=skin,
                                       trigger=G.TRIGGER_SKIN)) for s in systems]
    # a system whose list outgrew its capacity pauses while the list is regrown: step until
    # every system has run its nsteps
    for _ in range(8):
        if eng.step(nsteps) == len(slots):
            break
    e = None
    steps = None
    err = 0
    for sl in slots:
        th = eng.thermo(sl, max_rec=nsteps // out_steps + 4)
        err = err or eng.status(sl)["err"]
        es = np.array([r["pe"] + r["ke"] for r in th])
        e = es if e is None else e + es
        steps = np.array([r["step"] for r in th], dtype=np.float64)
    eng.close()
    n = sum(s.np for s in systems)
    slope = np.polyfit(steps, e, 1)[0] if len(steps) > 2 else 0.0
    return dict(steps=int(steps[-1]), records=len(steps),
                drift_per_atom_ev=float((e[-1] - e[0]) / n),
                drift_per_atom_per_1e4_ev=float((e[-1] - e[0]) / n * 1e4 / max(steps[-1] - steps[0], 1)),
                slope_per_atom_per_1e4_ev=float(slope * 1e4 / n),
                max_dev_per_atom_ev=float(np.abs(e - e[0]).max() / n),
                e0_per_atom_ev=float(e[0] / n), err=int(err))
