"""The output path: frames copied on the device behind the queued steps.

ndg_snapshot un-sorts positions, velocities, forces and per-atom energies into caller order
on the device and copies them, with the box and step, into pinned host memory behind the
steps already queued; ndg_run_enqueue / ndg_run_wait split the time loop so that the host
formats a frame while the device runs the next chunk (the driver's ConfOutFreq loop).
Bars: a snapshot equals ndg_get_state bit for bit; a run through enqueue/wait/snapshot gives
the same frames, thermo records, list rebuilds and final state, bit for bit, as the blocking
ndg_run_steps with ndg_get_state after each chunk (thermostat on, so a pending velocity scale
is in flight at every frame; skin trigger, so the atoms are re-sorted between frames).
"""
import numpy as np
import pytest

from ndyn_omp_amd import workloads as W

pytestmark = pytest.mark.gpu
G = pytest.importorskip("ndyn_omp_amd.engine")

PATHS = [("fp64", 0), ("mixed", 1)]


def system():
    return W.buck_alloy((6, 6, 6), temp_k=900.0, seed=141)


def rp(nsteps=120):
    return G.run_params(nsteps, W.DT_1FS, out_steps=10, skin=0.3, trigger=G.TRIGGER_SKIN,
                        isokinetic=(0.05, 0.2, 5))


@pytest.mark.parametrize("path_name,path", PATHS)
def test_snapshot_equals_get_state(path_name, path):
    s = system()
    print(f"case snapshot {path_name} np={s.np}", flush=True)
    sim = G.Simulation(s, path=path)
    sim.run(rp(37))
    a = sim.snapshot()
    b = sim.state()
    sim.close()
    assert a["step"] == 37
    for k in ("pos", "vel", "force", "pene", "box"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


@pytest.mark.parametrize("path_name,path", PATHS)
def test_queued_run_with_frames_equals_blocking_run(path_name, path):
    s = system()
    every = 25
    print(f"case frames {path_name} np={s.np}", flush=True)
    sim = G.Simulation(s, path=path)
    recs_q, frames, nup_q, step_q = sim.run_frames(rp(), every)
    fin_q = sim.state()
    sim.close()

    import ctypes as C
    L = G.load_library()
    sim = G.Simulation(s, path=path)
    r = rp()
    G._check(L.ndg_run_begin(sim.h, C.byref(r), 0))
    cap = r.nsteps // r.out_steps + 2
    buf = (G.Thermo * cap)()
    recs_b, frames_b, left, cur = [], [], r.nsteps, 0
    while left > 0:
        n = min(left, every - cur % every)
        got = C.c_int(0)
        G._check(L.ndg_run_steps(sim.h, n, buf, cap, C.byref(got)))
        recs_b.extend(buf[k].as_dict() for k in range(got.value))
        cur += n
        left -= n
        if cur % every == 0:
            frames_b.append(dict(step=cur, **sim.state()))
    nup, st = C.c_int(0), C.c_int(0)
    G._check(L.ndg_run_end(sim.h, C.byref(nup), C.byref(st)))
    fin_b = sim.state()
    sim.close()

    assert [f["step"] for f in frames] == [f["step"] for f in frames_b] == [25, 50, 75, 100]
    for fa, fb in zip(frames, frames_b):
        for k in ("pos", "vel", "force", "pene", "box"):
            np.testing.assert_array_equal(fa[k], fb[k], err_msg=f"{k} at {fa['step']}")
    assert recs_q == recs_b
    assert (nup_q, step_q) == (nup.value, st.value) and nup_q >= 2
    for k in ("pos", "vel", "acc", "force", "pene", "box"):
        np.testing.assert_array_equal(fin_q[k], fin_b[k], err_msg=k)
