"""hipGraph replay of the batched step against eager launches.

Two engines advance the same replicas (NVE Cu, NVT alloy, NPT Cu); one launches every
kernel eagerly, the other replays captured graphs of 8 steps (graphs are keyed by the
step sequence, so later calls replay graphs captured by earlier ones).  Between calls
the test does unrelated eager work (state fetches, another engine's steps).  Bar: every
thermo record and the final positions, velocities and box identical bit for bit.
"""
import numpy as np
import pytest

from ndyn_omp_amd import workloads as W

pytestmark = pytest.mark.gpu
G = pytest.importorskip("ndyn_omp_amd.engine")


def replicas():
    kb = W.KB_EV
    return [
        (W.cu_morse(8, temp_k=900.0, seed=51), None, None),
        (W.buck_alloy((6, 6, 6), temp_k=700.0, seed=52), (1.5 * kb * 700.0, 50 * W.DT_1FS, 1), None),
        (W.cu_morse(7, temp_k=500.0, seed=53), None, ((0.0, 0.0, 0.0), 200 * W.DT_1FS, 5)),
    ]


def engine(graph_steps, nsteps):
    eng = G.BatchEngine(path=G.PATH_MIXED, graph_steps=graph_steps)
    slots = []
    for s, iso, isop in replicas():
        rp = G.run_params(nsteps, W.DT_1FS, out_steps=5, skin=0.3, trigger=G.TRIGGER_SKIN,
                          isokinetic=iso, isobaric=isop)
        slots.append((s, eng.admit(s, rp)))
    return eng, slots


def test_graph_replay_equals_eager():
    chunks = [20, 37, 8, 8, 51]
    total = sum(chunks)
    ea, sa = engine(0, total)
    eb, sb = engine(8, total)
    other = G.BatchEngine(path=G.PATH_MIXED)
    o = W.cu_morse(6, temp_k=300.0, seed=54)
    other.admit(o, G.run_params(1000, W.DT_1FS, out_steps=1000))
    rec_a = [[] for _ in sa]
    rec_b = [[] for _ in sb]
    for k, n in enumerate(chunks):
        print(f"case chunk {k} ({n} steps)", flush=True)
        ea.step(n)
        eb.step(n)
        for q in range(len(sa)):
            rec_a[q] += ea.thermo(sa[q][1])
            rec_b[q] += eb.thermo(sb[q][1])
            s = sa[q][0]
            np.testing.assert_array_equal(eb.fetch(sb[q][1], s.np)["pos"],
                                          ea.fetch(sa[q][1], s.np)["pos"])
        other.step(7)      # unrelated eager work between replays
    for q, (s, _) in enumerate(sa):
        assert rec_b[q] == rec_a[q]
        assert len(rec_a[q]) == 1 + total // 5     # step 1, then every 5th step
        fa, fb = ea.fetch(sa[q][1], s.np), eb.fetch(sb[q][1], s.np)
        for key in ("pos", "vel", "box"):
            np.testing.assert_array_equal(fb[key], fa[key])
    for e in (ea, eb, other):
        e.close()
