"""Asymmetric pair matrices (the reference's compute1, src/nd.f90:364-530) on the device.

An `asym` pair line sets only PairMat(A,B) (src/read_input.f90:127-133).  A pair then uses
the table of the larger id of its two directions, an atom feels a neighbour only when the
neighbour's type acts on its own type, and only pairs with equal ids both ways add to the
energy and virial.  Both force kernels apply these rules when the matrix is asymmetric and
are unchanged otherwise.

Bars: neighbour sets bit-exact; FP64 path forces, energy and virial within 1e-12 relative of
the compiled reference (golden), per-atom energy within 1e-12 of the oracle; mixed path RMS
relative force error <= 1e-5, energy and virial trace <= 1e-6 relative, per-atom energy
within 1e-6 eV; dynamics as in the symmetric parity tests.  Threads per atom (1/2/4/8) only
reorder the FP64 force sums: positions within 1e-9 A of TPA 1.  A batch of asymmetric and
symmetric replicas gives each replica's results bit for bit as when it runs alone.
"""
import math

import numpy as np
import pytest

from conftest import golden, system_from_golden
from ndyn_omp_amd import workloads as W

pytestmark = pytest.mark.gpu

G = pytest.importorskip("ndyn_omp_amd.engine")
A = pytest.importorskip("ndyn_omp_amd.accuracy")

PATHS = [("fp64", 0), ("mixed", 1)]


def rms_rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def full_sets_from_golden(g, n):
    counts = g["counts"]
    ent = g["entries"]
    ii = np.repeat(np.arange(n), counts)
    a = np.concatenate([ii, ent])
    b = np.concatenate([ent, ii])
    order = np.lexsort((b, a))
    a, b = a[order], b[order]
    starts = np.searchsorted(a, np.arange(n + 1))
    return [b[starts[i]:starts[i + 1]] for i in range(n)]


@pytest.mark.parametrize("path_name,path", PATHS)
def test_asym_neighbour_sets_bit_exact(path_name, path):
    """The list of an asymmetric matrix holds every pair with a nonzero id in either
    direction, within the larger id's cutoff plus skin (reference half list, golden)."""
    g = golden("list_asym1.npz")
    s = system_from_golden(g)
    print(f"case list_asym1 {path_name}", flush=True)
    sim = G.Simulation(s, path=path, neimax=int(g["neimax"]))
    sim.build_lists(float(g["skin"]))
    counts, nbr = sim.neighbours()
    ref = full_sets_from_golden(g, s.np)
    for i in range(s.np):
        assert counts[i] == len(ref[i]), i
        np.testing.assert_array_equal(nbr[i, :counts[i]], ref[i])
    sim.close()


def _oracle_pene(oracle, s, skin):
    tables = oracle.pair_tables(s.kinds, s.params, s.ptrmin, s.ptsize)
    _, vnl, nm = oracle.half_list(s, skin)
    return oracle.compute1(s, vnl, nm, tables)


@pytest.mark.parametrize("fname", ["force_asym1.npz", "force_asym2.npz"])
def test_asym_forces_fp64_path(fname, oracle):
    g = golden(fname)
    s = system_from_golden(g)
    assert oracle.asymmetric(s)
    print(f"case {fname} fp64", flush=True)
    sim = G.Simulation(s, path=G.PATH_FP64)
    sim.build_lists(float(g["skin"]))
    pe, vir = sim.compute(True, True)
    st = sim.state()
    sim.close()
    assert rms_rel(st["force"], g["force"]) < 1e-12
    assert np.abs(st["force"] - g["force"]).max() < 1e-11 * np.abs(g["force"]).max()
    assert pe == pytest.approx(float(g["pe"]), rel=1e-12)
    np.testing.assert_allclose(vir, g["virial"], rtol=1e-12, atol=1e-12 * np.abs(g["virial"]).max())
    _, pe_o, _, pene = _oracle_pene(oracle, s, float(g["skin"]))
    np.testing.assert_allclose(st["pene"], pene, rtol=1e-12, atol=1e-14)
    assert st["pene"].sum() == pytest.approx(pe, rel=1e-12)


@pytest.mark.parametrize("fname", ["force_asym1.npz", "force_asym2.npz"])
def test_asym_forces_mixed_path(fname, oracle):
    g = golden(fname)
    s = system_from_golden(g)
    print(f"case {fname} mixed", flush=True)
    sim = G.Simulation(s, path=G.PATH_MIXED)
    sim.build_lists(float(g["skin"]))
    pe, vir = sim.compute(True, True)
    st = sim.state()
    sim.close()
    err = rms_rel(st["force"], g["force"])
    print(f"  rms rel force err {err:.3e}")
    assert err <= 1e-5
    assert pe == pytest.approx(float(g["pe"]), rel=1e-6)
    assert vir[:3].sum() == pytest.approx(float(g["virial"][:3].sum()), rel=1e-6)
    _, _, _, pene = _oracle_pene(oracle, s, float(g["skin"]))
    assert np.abs(st["pene"] - pene).max() < 1e-6


@pytest.mark.parametrize("variant,ptsize", [(1, 2000), (2, 2000), (1, 5000), (2, 14000)])
def test_asym_mixed_vs_fp64_on_thermal_snapshot(variant, ptsize):
    """Thermalised snapshots, with the tables staged in LDS (2001 points: 3-4 pairs of F
    and P fit in 48 KB), the force tables only (5001 points, variant 1) and neither (14001
    points)."""
    s = W.asym_morse(7, variant=variant, temp_k=600.0, seed=70 + variant, ptsize=ptsize)
    print(f"case asym{variant} ptsize={ptsize} np={s.np}", flush=True)
    r = A.pair_errors(A.thermal_snapshot(s, 40, W.DT_1FS, 0.5), 0.5)
    print(r)
    assert r["force_rms_rel"] <= 1e-5
    assert r["energy_per_atom_max_abs_ev"] <= 1e-6
    assert r["pe_rel"] <= 1e-6
    assert r["pressure_rel"] <= 1e-5


def test_asym_dynamics_fp64_path_matches_reference():
    """'dynamics' with compute1 (FP64 path) vs the compiled reference: final positions
    within 1e-10 (scaled), velocities 1e-8 relative, box 1e-12, every thermo line."""
    g = golden("run_asym2_nve.npz")
    s = system_from_golden(g)
    print("case run_asym2_nve fp64 run", flush=True)
    sim = G.Simulation(s, path=G.PATH_FP64)
    rp = G.run_params(int(g["nsteps"]), float(g["dt"]), out_steps=int(g["outsteps"]),
                      skin=float(g["skin"]))
    recs, nup, step = sim.run(rp)
    st = sim.state()
    sim.close()
    assert step == int(g["out_step"]) and nup == int(g["out_nupdates"])
    np.testing.assert_allclose(st["box"], g["out_box"], rtol=1e-12)
    np.testing.assert_allclose(st["pos"], g["out_pos"], rtol=0, atol=1e-10)
    np.testing.assert_allclose(st["vel"], g["out_vel"], rtol=1e-8, atol=1e-10)
    t = g["thermo"]
    assert len(recs) == t.shape[0]
    for row, rec in zip(t, recs):
        assert rec["step"] == int(row[0])
        for k, col in (("pe", 1), ("ke", 2), ("volume", 5)):
            assert rec[k] == pytest.approx(row[col], rel=1e-5, abs=1e-9)


def test_asym_dynamics_mixed_path_tracks_reference():
    """Mixed path 'dynamics' with compute1: positions within 1e-5 A (max) and 1e-6 A (RMS)
    of the reference after the run, thermo PE/KE within 1e-5 relative."""
    g = golden("run_asym2_nve.npz")
    s = system_from_golden(g)
    print("case run_asym2_nve mixed run", flush=True)
    sim = G.Simulation(s, path=G.PATH_MIXED)
    rp = G.run_params(int(g["nsteps"]), float(g["dt"]), out_steps=int(g["outsteps"]),
                      skin=float(g["skin"]))
    recs, nup, step = sim.run(rp)
    st = sim.state()
    sim.close()
    np.testing.assert_allclose(st["box"], g["out_box"], rtol=1e-8)
    dpos = (st["pos"] - g["out_pos"]) * g["out_box"]
    assert np.abs(dpos).max() < 1e-5
    assert math.sqrt((dpos ** 2).mean()) < 1e-6
    t = g["thermo"]
    for row, rec in zip(t, recs):
        assert rec["pe"] == pytest.approx(row[1], rel=1e-5)
        assert rec["ke"] == pytest.approx(row[2], rel=1e-5, abs=1e-9)


def _batch(systems, tpa=0, nsteps=30):
    eng = G.BatchEngine(path=G.PATH_MIXED, tpa=tpa)
    slots = []
    for s in systems:
        rp = G.run_params(nsteps, W.DT_1FS, out_steps=10, skin=0.5, trigger=G.TRIGGER_SKIN)
        slots.append(eng.admit(s, rp))
    eng.step(nsteps)
    out = [dict(st=eng.fetch(k, s.np), th=eng.thermo(k)) for s, k in zip(systems, slots)]
    eng.close()
    return out


def test_asym_threads_per_atom():
    systems = [W.asym_morse(7, variant=1, temp_k=700.0, seed=81),
               W.asym_morse(7, variant=2, temp_k=700.0, seed=82)]
    base = _batch(systems, tpa=1)
    for tpa in (2, 4, 8):
        print(f"case asym tpa={tpa}", flush=True)
        res = _batch(systems, tpa=tpa)
        for s, a, b in zip(systems, base, res):
            d = (a["st"]["pos"] - b["st"]["pos"]) * s.box
            assert np.abs(d).max() < 1e-9, (s.name, tpa)
            assert rms_rel(b["st"]["force"], a["st"]["force"]) < 1e-9
            for ra, rb in zip(a["th"], b["th"]):
                assert rb["pe"] == pytest.approx(ra["pe"], rel=1e-12)


def test_asym_and_symmetric_replicas_in_one_batch():
    systems = [W.asym_morse(6, variant=1, temp_k=900.0, seed=91),
               W.cu_morse(6, temp_k=900.0, seed=92),
               W.asym_morse(8, variant=2, temp_k=500.0, seed=93),
               W.buck_alloy((6, 6, 6), temp_k=800.0, seed=94)]
    print("case batch asym + symmetric", flush=True)
    both = _batch(systems)
    for k, s in enumerate(systems):
        print(f"case alone {s.name}", flush=True)
        alone = _batch([s])[0]
        np.testing.assert_array_equal(both[k]["st"]["pos"], alone["st"]["pos"])
        np.testing.assert_array_equal(both[k]["st"]["vel"], alone["st"]["vel"])
