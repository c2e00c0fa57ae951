"""Spring pairs (the reference's compute2 / compute3, src/nd.f90:531-933) on the device.

A 'spring' pair acts through its strain s = r/r0 - 1 against the initial positions of
initPositionSet (src/nd.f90:1731-1763): force C s Vi^2 along the pair, energy C s^2 Vi^2 r0
/ 2, nothing past the strain limit PairParams(3), and no distance cutoff (every listed pair
interacts).  With an asymmetric PairMat the compute1 rules apply on top (compute3).

Bars: FP64 path forces, energy and virial within 1e-12 relative of the compiled reference
(golden), per-atom energy within 1e-12 of the oracle; mixed path RMS relative force error
<= 1e-5, energy and virial trace <= 1e-6 relative; dynamics as in the symmetric parity
tests; the peridynamic tension deck (examples/testPDspringTension/input.sh) through the
deck driver reproduces the reference log's PE, pressure and volume lines to their printed
6 digits for steps 1 to 3600 (the affine stretch before the first bonds pass the strain
limit; after that the reference's own run is chaotic).
"""
import math
import os

import numpy as np
import pytest

from conftest import golden, system_from_golden
from ndyn_omp_amd import workloads as W

pytestmark = pytest.mark.gpu

G = pytest.importorskip("ndyn_omp_amd.engine")

SPRING_FORCES = ["force_spring_cu7.npz", "force_spring_mix0.npz", "force_spring_mix1.npz",
                 "force_spring_mix2.npz"]

# examples/testPDspringTension/out:38-74 -- step, PE, pressure, volume
PD_LOG = [
    (1, 0.00000, 0.00000, 49836.0), (100, 0.191554, -0.145643E-02, 49970.1),
    (200, 0.774476, -0.292621E-02, 50105.5), (300, 1.74955, -0.439464E-02, 50240.9),
    (400, 3.11752, -0.586172E-02, 50376.4), (500, 4.87914, -0.732749E-02, 50511.8),
    (600, 7.03516, -0.879194E-02, 50647.2), (700, 9.58633, -0.102551E-01, 50782.6),
    (800, 12.5334, -0.117170E-01, 50918.1), (900, 15.8770, -0.131776E-01, 51053.5),
    (1000, 19.6180, -0.146370E-01, 51188.9), (1100, 23.7571, -0.160951E-01, 51324.3),
    (1200, 28.2950, -0.175521E-01, 51459.8), (1300, 33.2324, -0.190078E-01, 51595.2),
    (1400, 38.5700, -0.204623E-01, 51730.6), (1500, 44.3086, -0.219157E-01, 51866.0),
    (1600, 50.4488, -0.233679E-01, 52001.5), (1700, 56.9914, -0.248189E-01, 52136.9),
    (1800, 63.9371, -0.262688E-01, 52272.3), (1900, 71.2865, -0.277176E-01, 52407.7),
    (2000, 79.0404, -0.291653E-01, 52543.2), (2100, 87.1994, -0.306119E-01, 52678.6),
    (2200, 95.7642, -0.320574E-01, 52814.0), (2300, 104.736, -0.335018E-01, 52949.4),
    (2400, 114.114, -0.349452E-01, 53084.9), (2500, 123.901, -0.363875E-01, 53220.3),
    (2600, 134.096, -0.378288E-01, 53355.7), (2700, 144.700, -0.392691E-01, 53491.1),
    (2800, 155.714, -0.407084E-01, 53626.5), (2900, 167.139, -0.421466E-01, 53762.0),
    (3000, 178.974, -0.435839E-01, 53897.4), (3100, 191.222, -0.450202E-01, 54032.8),
    (3200, 203.882, -0.464556E-01, 54168.2), (3300, 216.956, -0.478900E-01, 54303.7),
    (3400, 230.443, -0.493234E-01, 54439.1), (3500, 244.345, -0.507559E-01, 54574.5),
    (3600, 258.661, -0.521875E-01, 54709.9)]


def rms_rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _oracle_force(oracle, s, skin):
    tables = oracle.pair_tables(s.kinds, s.params, s.ptrmin, s.ptsize)
    _, vnl, nm = oracle.half_list(s, skin)
    return oracle.compute(s, vnl, nm, tables, ipos=s.ipos, init_box=s.init_box)


@pytest.mark.parametrize("fname", SPRING_FORCES)
def test_spring_forces_fp64_path(fname, oracle):
    g = golden(fname)
    s = system_from_golden(g)
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
    _, _, _, pene = _oracle_force(oracle, s, float(g["skin"]))
    np.testing.assert_allclose(st["pene"], pene, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("fname", SPRING_FORCES)
def test_spring_forces_mixed_path(fname, oracle):
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
    _, _, _, pene = _oracle_force(oracle, s, float(g["skin"]))
    assert np.abs(st["pene"] - pene).max() < 1e-6


def _run_kw(g):
    kw = {}
    if "box_vel" in g.files:
        kw["box_vel"] = tuple(float(v) for v in g["box_vel"])
    if "renei" in g.files:
        kw["renei"] = bool(int(g["renei"]))
    return kw


@pytest.mark.parametrize("path_name,path", [("fp64", 0), ("mixed", 1)])
@pytest.mark.parametrize("fname", ["run_spring_pd.npz", "run_spring_mix1_nve.npz"])
def test_spring_dynamics_matches_reference(fname, path_name, path):
    """'dynamics' with spring pairs vs the compiled reference (run_spring_pd: BoxVel, Skin 0,
    ReNeighbor F, initial positions set by the Run; run_spring_mix1_nve: compute3 with given
    initial positions).  FP64 path: positions within 1e-10 (scaled), velocities 1e-8
    relative; mixed path: positions within 1e-5 A (max); both: box 1e-12, thermo PE within
    1e-5 relative."""
    g = golden(fname)
    s = system_from_golden(g)
    print(f"case {fname} {path_name} run", flush=True)
    sim = G.Simulation(s, path=path)
    rp = G.run_params(int(g["nsteps"]), float(g["dt"]), out_steps=int(g["outsteps"]),
                      skin=float(g["skin"]), **_run_kw(g))
    recs, nup, step = sim.run(rp)
    st = sim.state()
    sim.close()
    assert step == int(g["out_step"]) and nup == int(g["out_nupdates"])
    np.testing.assert_allclose(st["box"], g["out_box"], rtol=1e-12)
    if path == G.PATH_FP64:
        np.testing.assert_allclose(st["pos"], g["out_pos"], rtol=0, atol=1e-10)
        np.testing.assert_allclose(st["vel"], g["out_vel"], rtol=1e-8, atol=1e-10)
    else:
        assert np.abs((st["pos"] - g["out_pos"]) * g["out_box"]).max() < 1e-5
    t = g["thermo"]
    assert len(recs) == t.shape[0]
    for row, rec in zip(t, recs):
        assert rec["step"] == int(row[0])
        assert rec["pe"] == pytest.approx(row[1], rel=1e-5, abs=1e-9)
        assert rec["volume"] == pytest.approx(row[5], rel=1e-5)


@pytest.mark.parametrize("path_name,path", [("fp64", 0), ("mixed", 1)])
def test_init_positions_mid_deck(path_name, path, oracle):
    """InitPos between two Runs (initPositionSet on the resident state) and initial positions
    set explicitly on a resident system, against the oracle given the same ones."""
    s = W.spring_mix(7, 1, temp_k=400.0, jitter_a=0.0, stretch=(0.0, 0.0, 0.0), seed=171)
    s.ipos = s.init_box = None
    print(f"case init positions {path_name}", flush=True)
    sim = G.Simulation(s, path=path)
    rp = lambda n: G.run_params(n, W.DT_1FS, out_steps=n, skin=0.5)
    sim.run(rp(20))
    mid = sim.state()
    sim.init_positions()
    sim.run(rp(20), step0=20)
    a = sim.state()
    # explicit initial positions: the lattice itself, set while resident
    s2 = W.spring_mix(7, 1, temp_k=0.0, jitter_a=0.0, stretch=(0.0, 0.0, 0.0), seed=171)
    sim.set_initial_positions(s2.ipos, s2.init_box)
    sim.run(rp(20), step0=40)
    b = sim.state()
    sim.close()

    o1, _, _ = oracle.run(s, 20, W.DT_1FS, skin=0.5, out_steps=20)
    t = s.copy()
    t.pos, t.vel, t.box = o1["pos"], o1["vel"], o1["box"]
    o2, _, _ = oracle.run(t, 20, W.DT_1FS, skin=0.5, out_steps=20, step0=20, acc=o1["acc"],
                          ipos=o1["pos"], init_box=o1["box"])
    t.pos, t.vel = o2["pos"], o2["vel"]
    o3, _, _ = oracle.run(t, 20, W.DT_1FS, skin=0.5, out_steps=20, step0=40, acc=o2["acc"],
                          ipos=s2.ipos, init_box=s2.init_box)
    tol = 1e-10 if path == G.PATH_FP64 else 1e-6
    np.testing.assert_allclose(mid["pos"], o1["pos"], rtol=0, atol=tol)
    np.testing.assert_allclose(a["pos"], o2["pos"], rtol=0, atol=tol)
    np.testing.assert_allclose(b["pos"], o3["pos"], rtol=0, atol=tol)
    # the three segments differ: the initial positions matter
    assert np.abs(o3["pos"] - o2["pos"]).max() > 1e-4


DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "ndyn_omp_amd", "bin", "ndyn_gpu")


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built")
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pd_spring_tension_deck(tmp_path, precision):
    """examples/testPDspringTension/input.sh through the deck driver (Run shortened to 3600
    steps): PE, pressure and volume of every thermo line equal the reference log's to the
    printed 6 significant digits."""
    import subprocess
    from test_driver import thermo, write_fcc_dump
    write_fcc_dump(tmp_path / "model.dump")
    deck = f"""START
Model 	model.dump dump
Seed	1234567890
Skin 	0.0
Types 1  #atom and particle types
 Cu 63.546 0.0 # name mass charge
Pairs 1
 Cu Cu sym spring 6.0  2.152295e-3 0.1 12.459008 #rc c s0 Vi
OutFreq 	100
ReNeighbor 	F
ConfOutFreq 	100 dump # configuration output frequency
BoxVel		1.e-2 0.0 0.0 #load the box at this rate Ang/(10 fs)
Precision {precision}
Run		3600 0.1
END
"""
    print(f"case pd tension {precision}", flush=True)
    p = subprocess.run([DRIVER], input=deck, cwd=tmp_path, capture_output=True, text=True,
                       timeout=300)
    assert p.returncode == 0, p.stderr
    rows = thermo(p.stdout)
    assert [int(r[0]) for r in rows] == [r[0] for r in PD_LOG]
    for r, ref in zip(rows, PD_LOG):
        for got, want in ((r[1], ref[1]), (r[4], ref[2]), (r[5], ref[3])):
            assert same_6(got, want), (int(r[0]), got, want)


def same_6(got, want):
    """Equal to a value printed with 6 significant digits, to one unit of its last digit."""
    if want == 0.0:
        return abs(got) < 1e-12
    return abs(got - want) <= 10.0 ** (math.floor(math.log10(abs(want))) - 5)


@pytest.mark.parametrize("path_name,path", [("fp64", 0), ("mixed", 1)])
def test_spring_replicas_in_a_batch(path_name, path):
    """Spring systems admitted into the batched engine (ndg_engine_admit_ex) beside a
    replica without springs, so the spring kernel instance also runs a replica that has
    none.  run_spring_pd takes its initial positions from the admitted state (as
    initPositionSet at the start of 'dynamics'), run_spring_mix1_nve is given them.  Bar:
    every replica bit-identical (positions, velocities, box, thermo) to its run through the
    single-system handle, and the spring replicas within the reference bars above."""
    cases = []
    for fname in ["run_spring_pd.npz", "run_spring_mix1_nve.npz"]:
        g = golden(fname)
        rp = G.run_params(int(g["nsteps"]), float(g["dt"]), out_steps=int(g["outsteps"]),
                          skin=float(g["skin"]), **_run_kw(g))
        cases.append((fname, system_from_golden(g), rp, g))
    plain = W.cu_morse(6, temp_k=700.0, seed=31)
    cases.append(("cu_morse_6", plain, G.run_params(40, W.DT_1FS, out_steps=10, skin=0.5), None))
    alone = {}
    for name, s, rp, _ in cases:
        print(f"case {name} {path_name} alone", flush=True)
        sim = G.Simulation(s.copy(), path=path)
        recs, _, _ = sim.run(rp)
        alone[name] = (sim.state(), recs)
        sim.close()
    print(f"case batch {path_name}", flush=True)
    eng = G.BatchEngine(path=path)
    slots = [eng.admit(s.copy(), rp) for _, s, rp, _ in cases]
    for _ in range(200):
        if eng.step(64) == len(slots):
            break
    for (name, s, rp, g), sl in zip(cases, slots):
        st = eng.status(sl)
        assert st["finished"] and not st["err"], name
        recs = eng.thermo(sl)
        got = eng.fetch(sl, s.np)
        ref, ref_recs = alone[name]
        np.testing.assert_array_equal(got["pos"], ref["pos"], err_msg=name)
        np.testing.assert_array_equal(got["vel"], ref["vel"], err_msg=name)
        np.testing.assert_array_equal(got["box"], ref["box"], err_msg=name)
        assert len(recs) == len(ref_recs), name
        for rec, want in zip(recs, ref_recs):     # 'eerr' is NaN before the first energy
            for k in want:
                np.testing.assert_array_equal(rec[k], want[k], err_msg=f"{name} {k}")
        if g is not None:
            np.testing.assert_allclose(got["box"], g["out_box"], rtol=1e-12)
            assert np.abs((got["pos"] - g["out_pos"]) * g["out_box"]).max() < 1e-5
            for row, rec in zip(g["thermo"], recs):
                assert rec["pe"] == pytest.approx(row[1], rel=1e-5, abs=1e-9)
    eng.close()
