"""Autotuner for the batched engine: atoms per wavefront (threads per atom in the force
kernel, TPA) and the neighbour-list skin, chosen per size class by timing the batch on
the GPU.

Cell-block shape: cells stored in blocks of 1 x by x bz instead of the linear order change
which atoms share a force tile (ndg_engine_cell_blocks); the results are the same bits.

The settings are tuned on the batch they are for (or an even subsample of it when it is
larger than `max_atoms`), not on each class alone: in a full batch a step's rebuild time
is set by the largest replica that rebuilds in that step, so a larger skin pays for
itself only where it stops the rebuilds that set that time, which an isolated class
cannot show.  The search is coordinate descent: start from TPA 1 and a skin of 0.5 A (or
the first candidate) in every class; then, classes largest first, try each TPA and then each skin for that class
with the others held, keeping a change only if it is more than `gain` faster.  Each trial
advances the whole batch for `warmup + steps` steps (several rebuild intervals) and
counts wall-clock atom-steps/s.

The resulting table is applied with `apply(engine)` (TPA by size class, through
ndg_engine_size_classes) and `skin_for(np)` when a system's run parameters are built.
It round-trips through JSON so a server can reuse it.
"""
from __future__ import annotations

import dataclasses
import json
import time
from typing import Callable, List, Optional, Sequence, Tuple

import numpy as np

from . import engine as G
from . import workloads as W

DEFAULT_CLASSES = (2048, 8192, 1 << 26)
DEFAULT_TPAS = (1, 2, 4)
DEFAULT_SKINS = (0.3, 0.4, 0.5, 0.75, 1.0, 1.25)
START_SKIN = 0.5
# cell-block shapes (by, bz): cells stored in blocks of 1 x by x bz (1 x 1: linear order)
DEFAULT_BLOCKS = ((1, 1), (2, 1), (2, 2))


@dataclasses.dataclass
class TuneTable:
    """Per size class: (np_max, tpa, skin, by, bz), np_max ascending -- threads per atom,
    list skin and cell-block shape (cells stored in blocks of 1 x by x bz); plus the trial
    results."""
    classes: List[Tuple[int, int, float, int, int]]
    trials: List[dict] = dataclasses.field(default_factory=list)

    def __post_init__(self):
        # tables saved before the cell-block shape was tuned hold (np_max, tpa, skin)
        self.classes = [tuple(c) + (1, 1) if len(c) == 3 else tuple(c) for c in self.classes]

    def tpa_classes(self):
        return [(c[0], c[1]) for c in self.classes]

    def block_classes(self):
        return [(c[0], c[3], c[4]) for c in self.classes]

    def skin_for(self, np_atoms: int, default: float = 0.5) -> float:
        for c in self.classes:
            if np_atoms <= c[0]:
                return c[2]
        return default

    def apply(self, eng: G.BatchEngine):
        eng.size_classes(self.tpa_classes())
        eng.cell_blocks(self.block_classes())

    def with_class(self, k: int, tpa: Optional[int] = None, skin: Optional[float] = None,
                   block: Optional[Tuple[int, int]] = None) -> "TuneTable":
        c = list(self.classes)
        npm, t0, s0, by, bz = c[k]
        if block is not None:
            by, bz = block
        c[k] = (npm, t0 if tpa is None else tpa, s0 if skin is None else skin, by, bz)
        return TuneTable(c)

    def to_json(self) -> str:
        return json.dumps({"classes": self.classes, "trials": self.trials})

    @classmethod
    def from_json(cls, text: str) -> "TuneTable":
        d = json.loads(text)
        return cls([tuple(c) for c in d["classes"]], d.get("trials", []))

    def summary(self) -> dict:
        return {f"np<={c[0]}": {"tpa": c[1], "skin": c[2], "cell_block": f"1x{c[3]}x{c[4]}"}
                for c in self.classes}


def subsample(systems: Sequence[W.System], max_atoms: int) -> List[W.System]:
    """An even subsample in size order with about max_atoms atoms (all if fewer)."""
    total = sum(s.np for s in systems)
    if total <= max_atoms:
        return list(systems)
    order = sorted(systems, key=lambda s: s.np)
    k = max(1, int(round(len(order) * max_atoms / total)))
    idx = np.linspace(0, len(order) - 1, k).round().astype(int)
    return [order[i] for i in sorted(set(idx.tolist()))]


def trial(systems: Sequence[W.System], make_rp: Callable[[float, int], G.RunParams],
          table: TuneTable, steps: int, warmup: int, path: int = G.PATH_MIXED) -> float:
    """Wall-clock atom-steps/s of `systems` advanced together under `table`."""
    eng = G.BatchEngine(path=path)
    try:
        table.apply(eng)
        for s in systems:
            eng.admit(s, make_rp(table.skin_for(s.np), warmup + steps))
        eng.step(warmup)
        t0 = time.perf_counter()
        eng.step(steps)
        dt = time.perf_counter() - t0
    finally:
        eng.close()
    return sum(s.np for s in systems) * steps / dt


def tune(systems: Sequence[W.System], make_rp: Callable[[float, int], G.RunParams],
         classes: Sequence[int] = DEFAULT_CLASSES, tpas: Sequence[int] = DEFAULT_TPAS,
         skins: Sequence[float] = DEFAULT_SKINS, blocks: Sequence[Tuple[int, int]] = DEFAULT_BLOCKS,
         steps: int = 120, warmup: int = 20, max_atoms: int = 4_000_000, gain: float = 0.01,
         path: int = G.PATH_MIXED, log: Optional[Callable[[str], None]] = None) -> TuneTable:
    """Coordinate descent over (tpa, skin, cell-block shape) per size class on the batch
    itself."""
    batch = subsample(systems, max_atoms)
    present = [any(lo < s.np <= hi for s in batch)
               for lo, hi in zip((0,) + tuple(classes[:-1]), classes)]
    start = START_SKIN if START_SKIN in skins else skins[0]
    best = TuneTable([(int(c), 1, float(start), 1, 1) for c in classes])
    trials = []

    def measure(t: TuneTable, what: str) -> float:
        rate = trial(batch, make_rp, t, steps, warmup, path)
        trials.append(dict(table=t.classes, atom_steps_per_s=rate, changed=what))
        if log:
            log(f"autotune: {what}: {rate:.4e} atom-steps/s")
        return rate

    best_rate = measure(best, "start")
    for k in reversed(range(len(classes))):
        if not present[k]:
            continue
        for tpa in tpas:
            if tpa == best.classes[k][1]:
                continue
            cand = best.with_class(k, tpa=tpa)
            r = measure(cand, f"np<={classes[k]} tpa={tpa}")
            if r > best_rate * (1.0 + gain):
                best, best_rate = cand, r
        for skin in skins:
            if skin == best.classes[k][2]:
                continue
            cand = best.with_class(k, skin=skin)
            r = measure(cand, f"np<={classes[k]} skin={skin}")
            if r > best_rate * (1.0 + gain):
                best, best_rate = cand, r
        for blk in blocks:
            if tuple(blk) == tuple(best.classes[k][3:5]):
                continue
            cand = best.with_class(k, block=tuple(blk))
            r = measure(cand, f"np<={classes[k]} cell block 1x{blk[0]}x{blk[1]}")
            if r > best_rate * (1.0 + gain):
                best, best_rate = cand, r
    best.trials = trials
    return best
