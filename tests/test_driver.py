"""The Fortran deck driver (ndyn_omp_amd/bin/ndyn_gpu) over the C ABI.

CPU: it parses the reference's deck grammar and writes the reference's potential tables
(WritePotentialTable, src/potentials.f90:44-54) equal to the oracle's.
GPU: the reference's example decks reproduce the thermo lines of their committed logs
(examples/MorseTest/out, examples/testT/out) through the driver.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from ndyn_omp_amd import workloads as W

DRIVER = os.path.join(ROOT, "ndyn_omp_amd", "bin", "ndyn_gpu")
needs_driver = pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built")


def run_deck(deck, cwd, timeout=300):
    p = subprocess.run([DRIVER], input=deck, cwd=cwd, capture_output=True, text=True,
                       timeout=timeout)
    return p


def thermo(stdout):
    rows = []
    for line in stdout.splitlines():
        parts = line.split()
        if len(parts) == 6 and not line.lstrip().startswith("#"):
            try:
                rows.append([float(x) for x in parts])
            except ValueError:
                pass
    return rows


def write_fcc_dump(path, n=10, a=3.68):
    """The perfect FCC lattice of examples/MorseTest/model.dump (id order, xs in [0,1))."""
    pos, box = W.fcc(n, n, n, a)
    with open(path, "w") as f:
        f.write("ITEM: TIMESTEP\n0\nITEM: NUMBER OF ATOMS\n%d\nITEM: BOX BOUNDS pp pp pp\n" % len(pos))
        for L in box:
            f.write(f"0 {L:g}\n")
        f.write("ITEM: ATOMS id type xs ys zs\n")
        for k, p in enumerate(pos + 0.5):
            f.write(f"{k + 1} 1 {p[0]:.10g} {p[1]:.10g} {p[2]:.10g}\n")


@needs_driver
def test_driver_writes_reference_tables(tmp_path, oracle):
    deck = """START
Types 1
 Cu 63.546
WritePotentialTable
Pairs 1
 Cu Cu sym morse 6.0 0.3429 1.359 2.866
END
"""
    p = run_deck(deck, tmp_path)
    assert p.returncode == 0, p.stderr
    assert "# pair,rcut2,dr2  1   36.000000    0.035786" in p.stdout   # examples/MorseTest/out:20
    got = np.loadtxt(tmp_path / "forcetab.dat")
    P, F, _, _ = oracle.pair_tables([1], (6.0, 0.3429, 1.359, 2.866), 0.25, 1000)
    # rows i = 1..PTsize of the Fortran table are 0-based rows 0..PTsize-1
    np.testing.assert_allclose(got[:, 1], F[0][0:1000], rtol=1e-6)
    np.testing.assert_allclose(got[:, 0], np.sqrt(np.arange(1, 1001) / 1000.0), rtol=1e-6)
    pot = np.loadtxt(tmp_path / "pottab.dat")
    np.testing.assert_allclose(pot[:, 1], P[0][0:1000], rtol=1e-6)


@needs_driver
def test_driver_rejects_unknown_potential(tmp_path):
    deck = "START\nTypes 1\n Cu 63.546\nPairs 1\n Cu Cu sym tersoff 1.0\nEND\n"
    p = run_deck(deck, tmp_path)
    assert p.returncode != 0
    assert "No potential available for Potential Type" in p.stderr


@pytest.mark.gpu
@needs_driver
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_driver_morsetest_deck(tmp_path, precision):
    """examples/MorseTest/input.sh (Run shortened to 200 steps): thermo lines of
    examples/MorseTest/out:38-43 -- PE -13581.1, pressure 0.358006E-01, volume 49836.0 at
    steps 1, 40, ..., 200; the dump trajectory written at the start and the end."""
    write_fcc_dump(tmp_path / "model.dump")
    deck = f"""START
Model 	model.dump dump
NeiMax  200
Types 1
 Cu 63.546
Pairs 1
 Cu Cu sym morse 6.0 0.3429 1.359 2.866
Precision {precision}
OutFreq 40
ConfOutFreq 0 dump
Run 200 0.0001
END
"""
    print(f"case morsetest {precision}", flush=True)
    p = run_deck(deck, tmp_path)
    assert p.returncode == 0, p.stderr
    rows = thermo(p.stdout)
    assert [int(r[0]) for r in rows] == [1, 40, 80, 120, 160, 200]
    for r in rows:
        assert float(f"{r[1]:.6g}") == -13581.1
        assert float(f"{r[4]:.6g}") == 0.0358006
        assert float(f"{r[5]:.6g}") == 49836.0
    assert "# Number of neighbour list updates:       0" in p.stdout
    dump = (tmp_path / "out.dump").read_text().splitlines()
    assert dump.count("ITEM: TIMESTEP") == 2 and dump[1].strip() == "0" and "200" in dump[dump.index("ITEM: TIMESTEP", 1) + 1]


@pytest.mark.gpu
@needs_driver
def test_driver_testT_npt_deck(tmp_path):
    """examples/testT/input.sh run 1 (Run shortened to 200 steps): examples/testT/out:38-40
    -> step 1 PE -13581.1 V 49836.0; step 100 PE -12798.9 V 56637.3; step 200 PE -12798.8
    V 56637.5; ConfOutFreq 100 writes a dump frame every 100 steps."""
    write_fcc_dump(tmp_path / "model.dump")
    deck = """START
Model 	model.dump dump
Seed	1234567890
Skin 	0.2
Types 1
 Cu 63.546 0.0
Pairs 1
 Cu Cu sym morse 6.0  0.3429 1.359 2.866
OutFreq 	100
ReNeighbor 	F
ConfOutFreq 	100 dump
Isobaric 	0.0 0.0 0.0  0.1 10
Run 		200 0.1
Anisobaric
END
"""
    p = run_deck(deck, tmp_path)
    assert p.returncode == 0, p.stderr
    rows = thermo(p.stdout)
    got = [(int(r[0]), float(f"{r[1]:.6g}"), float(f"{r[5]:.6g}")) for r in rows]
    assert got == [(1, -13581.1, 49836.0), (100, -12798.9, 56637.3), (200, -12798.8, 56637.5)]
    dump = (tmp_path / "out.dump").read_text()
    assert dump.count("ITEM: TIMESTEP") == 3


def write_typed_dump(path, s):
    """A LAMMPS dump of system s whose values are exact in single precision, and the
    system the driver reads back from it (src/read_input.f90:347-371 reads real(4))."""
    f32 = np.float32
    box = np.array([float(f32(L)) for L in s.box])
    xs = (s.pos + 0.5).astype(f32)
    with open(path, "w") as f:
        f.write("ITEM: TIMESTEP\n0\nITEM: NUMBER OF ATOMS\n%d\nITEM: BOX BOUNDS pp pp pp\n" % s.np)
        for L in box:
            f.write(f"0 {float(L)!r}\n")
        f.write("ITEM: ATOMS id type xs ys zs\n")
        for k in range(s.np):
            f.write(f"{k + 1} {int(s.typ[k])} {float(xs[k, 0])!r} {float(xs[k, 1])!r} {float(xs[k, 2])!r}\n")
    t = s.copy()
    t.box = box
    t.pos = (xs - f32(0.5)).astype(np.float64)
    t.vel = np.zeros_like(s.pos)
    return t


@pytest.mark.gpu
@needs_driver
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_driver_asym_pairs_deck(tmp_path, oracle, precision):
    """'asym' pair lines (src/read_input.f90:127-133) select compute1: a two-type deck with
    one directed pair (B atoms feel A atoms, A atoms do not feel B) and its thermo lines
    against the oracle's run of the same configuration: PE and volume within 1e-5 relative
    (the lines print 6 significant digits); the dump frames (ConfOutFreq) at their steps, the
    last one equal to the oracle's final positions to the printed 5 decimals."""
    s = W.asym_morse(6, variant=1, jitter_a=0.05, seed=131)
    t = write_typed_dump(tmp_path / "model.dump", s)
    t.ptsize = 1000   # the deck's default table size (src/read_input.f90:28)
    (A, B, C) = [" ".join(repr(float(v)) for v in s.params[k]) for k in range(3)]
    deck = f"""START
Model 	model.dump dump
Skin 0.5
Types 2
 A 63.5
 B 27.0
Pairs 3
 A A sym morse {A}
 A B asym morse {B}
 B B sym morse {C}
Precision {precision}
OutFreq 20
ConfOutFreq 20 dump
Run 60 {W.DT_1FS!r}
END
"""
    print(f"case asym deck {precision}", flush=True)
    p = run_deck(deck, tmp_path)
    assert p.returncode == 0, p.stderr
    rows = thermo(p.stdout)
    st, ref, err = oracle.run(t, 60, W.DT_1FS, skin=0.5, out_steps=20)
    assert err == 0
    assert [int(r[0]) for r in rows] == [r["step"] for r in ref] == [1, 20, 40, 60]
    for r, q in zip(rows, ref):
        assert r[1] == pytest.approx(q["pe"], rel=1e-5)
        assert r[5] == pytest.approx(q["volume"], rel=1e-5)
    # frames at steps 0, 20, 40, 60 (copied on the device behind each chunk); the last one
    # holds the oracle's final positions to the format's 5 decimals
    lines = (tmp_path / "out.dump").read_text().splitlines()
    heads = [k for k, l in enumerate(lines) if l.startswith("ITEM: TIMESTEP")]
    assert [int(lines[k + 1]) for k in heads] == [0, 20, 40, 60]
    last = np.array([[float(x) for x in l.split()[2:5]] for l in lines[heads[-1] + 9:]])
    assert last.shape == st["pos"].shape
    assert np.abs(last - st["pos"]).max() < 1e-5


@pytest.mark.gpu
@needs_driver
@pytest.mark.parametrize("fmt", ["xyz", "dump"])
def test_driver_engine_frames_equal_write_model(tmp_path, fmt):
    """'FrameWriter engine' (the library's writer thread, formatting while the device steps)
    writes the same bytes as 'FrameWriter fortran' (write_model's edit descriptors) over two
    Runs with frames every 10 steps, the second Run appending to the same file."""
    s = W.asym_morse(6, variant=2, temp_k=600.0, seed=151)
    t = write_typed_dump(tmp_path / "model.dump", s)
    prm = [" ".join(repr(float(v)) for v in s.params[k]) for k in range(4)]
    out = {}
    for writer in ("fortran", "engine"):
        d = tmp_path / writer
        d.mkdir()
        (d / "model.dump").write_bytes((tmp_path / "model.dump").read_bytes())
        deck = f"""START
Model 	model.dump dump
Seed 1234567890
Skin 0.5
Types 2
 Cu 63.5
 Al-x 27.0
Pairs 4
 Cu Cu sym morse {prm[0]}
 Cu Al-x asym morse {prm[1]}
 Al-x Cu asym morse {prm[2]}
 Al-x Al-x sym morse {prm[3]}
Velocity 0.05
OutFreq 10
ConfOutFreq 10 {fmt}
FrameWriter {writer}
Run 30 {W.DT_1FS!r}
Run 25 {W.DT_1FS!r}
END
"""
        print(f"case frames {fmt} {writer}", flush=True)
        p = run_deck(deck, d)
        assert p.returncode == 0, p.stderr
        out[writer] = (d / f"out.{fmt}").read_bytes()
    assert t.np == s.np
    frames = out["fortran"].count(b"ITEM: TIMESTEP") if fmt == "dump" else \
        sum(1 for l in out["fortran"].splitlines() if l.strip() == str(s.np).encode())
    assert frames == 4 + 3      # Run 1: steps 0, 10, 20, 30; Run 2: 30, 40, 50
    assert out["engine"] == out["fortran"]


def dump_frames(path):
    """out.dump -> list of (step, ids, types, xs(n,3))."""
    lines = open(path).read().splitlines()
    out = []
    k = 0
    while k < len(lines):
        if lines[k].startswith("ITEM: TIMESTEP"):
            step = int(lines[k + 1])
            n = int(lines[k + 3])
            rows = [l.split() for l in lines[k + 9:k + 9 + n]]
            out.append((step, np.array([int(r[0]) for r in rows]), np.array([int(r[1]) for r in rows]),
                        np.array([[float(x) for x in r[2:5]] for r in rows])))
            k += 9 + n
        else:
            k += 1
    return out


@pytest.mark.gpu
@needs_driver
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_driver_groups_holds_and_displacements(tmp_path, precision):
    """Group by a sphere domain + ChangeType, Hold by two slab domains (rec) with Disp
    (src/read_input.f90:171-190, src/nd.f90:1625-1730, src/domtype.f90:278-434): the
    sphere's atoms change type, the frozen slab keeps its printed positions, the pulled slab
    moves by 100 * 1e-5 along x, and the atoms selected follow the reference's
    single-precision domain tests."""
    write_fcc_dump(tmp_path / "model.dump")
    deck = f"""START
Model 	model.dump dump
Seed 1234567890
Types 2
 Cu 63.546
 Ag 107.87
Pairs 3
 Cu Cu sym morse 6.0 0.3429 1.359 2.866
 Cu Ag sym morse 6.0 0.3 1.3 2.9
 Ag Ag sym morse 6.0 0.33 1.32 2.95
Group ball dom sph T 0.1 0.1 0.1 0.21
ChangeType ball Ag
Velocity 0.03877798635
Hold fixed dom rec T -0.5 -0.41 0 0 0 0
Hold pull dom rec T 0.39 0.5 0 0 0 0
Hold other typ 2
Disp fixed 0 0 0
Disp pull 1e-5 F F
Precision {precision}
OutFreq 50
ConfOutFreq 100 dump
Run 100 0.05
END
"""
    print(f"case groups/holds {precision}", flush=True)
    p = run_deck(deck, tmp_path)
    assert p.returncode == 0, p.stderr
    frames = dump_frames(tmp_path / "out.dump")
    assert [f[0] for f in frames] == [0, 100]
    (_, ids, t0, x0), (_, _, t1, x1) = frames
    # the model's scaled positions as the driver holds them: real(4) xs - 0.5 (read_model)
    pos, box = W.fcc(10, 10, 10, 3.68)
    xs = ((pos + 0.5).astype(np.float32) - np.float32(0.5)).astype(np.float32)
    c = np.float32(0.1)
    # (boundaries chosen between lattice planes and shells)
    ball = ((xs - c) ** 2).sum(axis=1) <= np.float32(0.21) ** 2
    assert 50 < ball.sum() < len(xs)
    np.testing.assert_array_equal(t0, np.where(ball, 2, 1))
    np.testing.assert_array_equal(t1, t0)
    f32 = np.float32
    fixed = (xs[:, 0] > f32(-0.5)) & (xs[:, 0] <= f32(-0.41))
    pull = (xs[:, 0] > f32(0.39)) & (xs[:, 0] <= f32(0.5))
    assert fixed.sum() > 100 and pull.sum() > 100
    np.testing.assert_array_equal(x1[fixed], x0[fixed])
    np.testing.assert_allclose(x1[pull, 0] - x0[pull, 0], 100 * 1e-5, atol=1.1e-5)
    free = ~fixed & ~pull & ~ball
    assert np.abs(x1[free] - x0[free]).max() > 1e-4


@pytest.mark.gpu
@needs_driver
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_driver_sinetest_soft_deck(tmp_path, precision):
    """examples/SineTest/input.sh (the 'soft' potential on the MorseTest lattice, Skin 1.0;
    Run shortened to 400 steps): examples/SineTest/out:36-46 -> PE 26491.8, pressure
    -55.6990, volume 49836.0 at steps 1, 40, ..., 400.  The pressure follows the
    reference's soft force, which reads parm(1), parm(2) where the energy reads parm(2),
    parm(3) (src/potentials.f90:86, :176)."""
    write_fcc_dump(tmp_path / "model.dump")
    deck = f"""START
Model 	model.dump dump
Skin	1.0
Types 1  #atom and particle types
 Cu 63.546
Pairs 1
 Cu Cu sym soft 6.0 0.5 6.0
Precision {precision}
OutFreq 40
ConfOutFreq 0 dump # configuration output frequency
Run 400 0.00001    # run for [i] steps each at [r] deltat
END
"""
    print(f"case sinetest {precision}", flush=True)
    p = run_deck(deck, tmp_path)
    assert p.returncode == 0, p.stderr
    rows = thermo(p.stdout)
    assert [int(r[0]) for r in rows] == [1] + list(range(40, 401, 40))
    for r in rows:
        assert float(f"{r[1]:.6g}") == 26491.8
        assert float(f"{r[4]:.6g}") == -55.699
        assert float(f"{r[5]:.6g}") == 49836.0
