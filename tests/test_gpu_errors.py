"""Error paths of the boundary on the GPU (every kernel still exits normally).

- ndg_set_state with an atom type outside 1..ntypes returns NDG_ERR_ARG before anything is
  uploaded, and the resident state and lists are unchanged (the same forces follow).
- The mixed path's 16-byte packed coordinate holds an atom within skin + 1 A of its
  list-build position.  A Run that keeps its lists (the reference's trigger, which does not
  fire here) while an atom travels 12 A stops with NDG_ERR_RANGE instead of computing pairs
  from clamped coordinates; the FP64 path, which has no such limit, runs the same deck and
  equals the oracle (free flight: x = x0 + v t), and the mixed path with the skin trigger
  rebuilds and runs it too.
"""
import ctypes as C

import numpy as np
import pytest

from ndyn_omp_amd import workloads as W

pytestmark = pytest.mark.gpu
G = pytest.importorskip("ndyn_omp_amd.engine")


def test_set_state_rejects_bad_types():
    s = W.cu_morse(6, temp_k=300.0, seed=5, jitter_a=0.05)
    sim = G.Simulation(s, path=G.PATH_MIXED)
    sim.build_lists(0.5)
    sim.compute(True, True)
    before = sim.state()
    L = G.load_library()
    for bad in (0, 2, -3):
        typ = s.typ.astype(np.int32).copy()
        typ[17] = bad
        print(f"case set_state bad type {bad}", flush=True)
        rc = L.ndg_set_state(sim.h, np.ascontiguousarray(s.pos), None, None, typ,
                             np.ascontiguousarray(s.box))
        assert rc == G.ERR_ARG
        assert b"type" in L.ndg_last_error()
    rc = L.ndg_set_state(sim.h, np.ascontiguousarray(s.pos), None, None, s.typ.astype(np.int32),
                         np.array([s.box[0], -1.0, s.box[2]]))
    assert rc == G.ERR_ARG
    after = sim.state()
    for k in ("pos", "vel", "force", "box"):
        np.testing.assert_array_equal(after[k], before[k])
    sim.compute(True, True)
    np.testing.assert_array_equal(sim.state()["force"], before["force"])
    sim.close()


def dilute():
    """Eight atoms 50 A apart in a 100 A box (no pair within the cutoff); atom 0 travels
    0.3 A per time unit along x."""
    s = W.cu_morse(1, temp_k=0.0)
    g = np.array([[i, j, k] for k in (0, 1) for j in (0, 1) for i in (0, 1)], np.float64)
    s.pos = (g * 0.5 - 0.25) + 0.01
    s.typ = np.ones(8, np.int32)
    s.box = np.array([100.0, 100.0, 100.0])
    s.vel = np.zeros((8, 3))
    s.vel[0, 0] = 0.3
    s.vel[5] = (0.01, -0.02, 0.015)
    return s


def test_packed_range_error_and_fp64_free_flight(oracle):
    s = dilute()
    nsteps, dt = 40, 1.0
    rp = lambda trig: G.run_params(nsteps, dt, out_steps=10, skin=0.5, trigger=trig)
    st_ref, _, err = oracle.run(s, nsteps, dt, skin=0.5, out_steps=10,
                                trigger=oracle.TRIGGER_COMPAT)
    assert err == 0 and st_ref["nupdates"] == 0
    np.testing.assert_allclose(st_ref["pos"], s.pos + s.vel * nsteps * dt / s.box, atol=1e-14)
    print("case free flight fp64 compat", flush=True)
    sim = G.Simulation(s, path=G.PATH_FP64)
    sim.run(rp(G.TRIGGER_COMPAT))
    np.testing.assert_allclose(sim.state()["pos"], st_ref["pos"], rtol=0, atol=1e-14)
    sim.close()
    print("case free flight mixed compat (range error)", flush=True)
    sim = G.Simulation(s, path=G.PATH_MIXED)
    with pytest.raises(G.NdgError) as ei:
        sim.run(rp(G.TRIGGER_COMPAT))
    assert ei.value.code == G.ERR_RANGE
    sim.close()
    print("case free flight mixed skin", flush=True)
    sim = G.Simulation(s, path=G.PATH_MIXED)
    _, nup, _ = sim.run(rp(G.TRIGGER_SKIN))
    assert nup > 0
    np.testing.assert_allclose(sim.state()["pos"], st_ref["pos"], rtol=0, atol=1e-14)
    sim.close()
