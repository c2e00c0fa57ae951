"""Rebuild work list over more replicas than one pass of its 1024-slot scan covers.

1100 replicas (1000 of 256 atoms, whose small boxes use the all-pairs list, and 100 of
1372 atoms on a 3 x 3 x 3 cell grid) at 600-1200 K are advanced together with the skin
trigger, so different replicas rebuild their lists at different steps and the rebuild
work list gathers tiles from both 1024-slot chunks of its scan.  A sample of replicas,
including slots on both sides of the chunk boundary, is then advanced alone.  Bar: the
same positions and velocities bit for bit and the same number of list rebuilds (a
replica's result does not depend on which other replicas rebuild with it).  A second
case runs more replicas (4200) than the histogram kernel plans in LDS, so that the work
list is planned by its own kernel.
"""
import numpy as np
import pytest

from ndyn_omp_amd import workloads as W

pytestmark = pytest.mark.gpu
G = pytest.importorskip("ndyn_omp_amd.engine")

NSTEPS = 60


def system(k):
    t = 600.0 + 600.0 * ((k * 37) % 101) / 100.0
    if k % 11 == 5:
        return W.cu_morse(7, temp_k=t, seed=1000 + k)
    return W.cu_morse(4, temp_k=t, seed=1000 + k)


def rp():
    return G.run_params(NSTEPS, W.DT_1FS, out_steps=NSTEPS, skin=0.5, trigger=G.TRIGGER_SKIN)


def test_many_replicas_match_alone():
    n = 1100
    systems = [system(k) for k in range(n)]
    eng = G.BatchEngine(path=G.PATH_MIXED)
    slots = [eng.admit(s, rp()) for s in systems]
    print(f"case batch of {n} replicas, {sum(s.np for s in systems)} atoms", flush=True)
    eng.step(NSTEPS)
    sample = [0, 5, 500, 1022, 1023, 1024, 1025, 1040, 1099]
    batch = {}
    for k in sample:
        st = eng.fetch(slots[k], systems[k].np)
        nup = eng.counters(slots[k])["nupdates"]
        batch[k] = (st["pos"], st["vel"], nup)
    late = [eng.counters(slots[k])["nupdates"] for k in range(1024, n)]
    eng.close()
    assert min(late) >= 1, "replicas past the first 1024 slots rebuilt their lists"
    for k in sample:
        print(f"case alone replica {k} np={systems[k].np}", flush=True)
        e1 = G.BatchEngine(path=G.PATH_MIXED)
        s1 = e1.admit(systems[k], rp())
        e1.step(NSTEPS)
        st = e1.fetch(s1, systems[k].np)
        nup = e1.counters(s1)["nupdates"]
        e1.close()
        np.testing.assert_array_equal(st["pos"], batch[k][0])
        np.testing.assert_array_equal(st["vel"], batch[k][1])
        assert nup == batch[k][2]


def test_replicas_beyond_the_lds_plan_match_alone():
    """More replicas than the histogram kernel plans in LDS (4096): the work list is then
    planned by its own kernel.  4200 replicas of 256 atoms; a sample, including the last
    slots, advanced alone gives the same bits and the same rebuild count."""
    n, nsteps = 4200, 60
    systems = [W.cu_morse(4, temp_k=900.0 + 300.0 * (k % 3), seed=5000 + k) for k in range(n)]

    def rp_short():
        return G.run_params(nsteps, W.DT_1FS, out_steps=nsteps, skin=0.3,
                            trigger=G.TRIGGER_SKIN)

    eng = G.BatchEngine(path=G.PATH_MIXED)
    slots = [eng.admit(s, rp_short()) for s in systems]
    print(f"case batch of {n} replicas, {sum(s.np for s in systems)} atoms", flush=True)
    eng.step(nsteps)
    sample = [0, 4095, 4096, n - 1]
    batch = {}
    for k in sample:
        st = eng.fetch(slots[k], systems[k].np)
        batch[k] = (st["pos"], st["vel"], eng.counters(slots[k])["nupdates"])
    rebuilt = sum(eng.counters(slots[k])["nupdates"] > 0 for k in range(4096, n))
    eng.close()
    assert rebuilt > 0, "replicas past slot 4096 rebuilt their lists"
    for k in sample:
        print(f"case alone replica {k}", flush=True)
        e1 = G.BatchEngine(path=G.PATH_MIXED)
        s1 = e1.admit(systems[k], rp_short())
        e1.step(nsteps)
        st = e1.fetch(s1, systems[k].np)
        nup = e1.counters(s1)["nupdates"]
        e1.close()
        np.testing.assert_array_equal(st["pos"], batch[k][0])
        np.testing.assert_array_equal(st["vel"], batch[k][1])
        assert nup == batch[k][2]
