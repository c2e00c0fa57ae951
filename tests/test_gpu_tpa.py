"""Threads per atom (TPA) and the autotuner through the batched engine's C ABI.

The force kernel splits an atom's list over 1, 2, 4 or 8 lanes; the split only changes
the order of the FP64 force sums.  Bars: after 30 steps of a Cu and an alloy replica,
positions within 1e-9 A and forces within 1e-9 relative (RMS) of TPA 1, and every thermo
record's PE within 1e-12 relative.  A size-class table gives bit-identical results to an
engine-wide TPA of the same class (each replica's kernels depend only on its own TPA).
"""
import numpy as np
import pytest

from ndyn_omp_amd import workloads as W

pytestmark = pytest.mark.gpu
G = pytest.importorskip("ndyn_omp_amd.engine")


def systems():
    return [W.cu_morse(8, temp_k=600.0, seed=41),
            W.buck_alloy((6, 6, 6), temp_k=800.0, seed=42)]


def run(tpa, nsteps=30, classes=None, only=None):
    eng = G.BatchEngine(path=G.PATH_MIXED, tpa=tpa)
    if classes:
        eng.size_classes(classes)
    out = []
    for k, s in enumerate(systems()):
        if only is not None and k != only:
            continue
        rp = G.run_params(nsteps, W.DT_1FS, out_steps=10, skin=0.5, trigger=G.TRIGGER_SKIN)
        out.append((s, eng.admit(s, rp)))
    eng.step(nsteps)
    res = []
    for s, slot in out:
        res.append(dict(st=eng.fetch(slot, s.np), th=eng.thermo(slot)))
    eng.close()
    return res


@pytest.mark.parametrize("tpa", [2, 4, 8])
def test_tpa_variants_match_tpa1(tpa):
    print(f"case tpa={tpa}", flush=True)
    ref, got = run(1), run(tpa)
    for r, g in zip(ref, got):
        dpos = np.abs((g["st"]["pos"] - r["st"]["pos"]) * g["st"]["box"]).max()
        assert dpos < 1e-9
        f0, f1 = r["st"]["force"], g["st"]["force"]
        assert np.linalg.norm(f1 - f0) / np.linalg.norm(f0) < 1e-9
        assert len(g["th"]) == len(r["th"]) > 0
        for a, b in zip(g["th"], r["th"]):
            assert a["step"] == b["step"]
            assert a["pe"] == pytest.approx(b["pe"], rel=1e-12)


def test_size_classes_equal_engine_wide_tpa():
    """The alloy 6^3 (864 atoms) falls in class np<=1000 (TPA 4), Cu 8^3 (2048 atoms) in
    the last class (TPA 2); each matches an engine-wide TPA of its class bit for bit."""
    print("case size classes", flush=True)
    mixed = run(0, classes=[(1000, 4), (1 << 26, 2)])
    alloy_alone = run(4, only=1)[0]
    cu_alone = run(2, only=0)[0]
    for got, want in ((mixed[0], cu_alone), (mixed[1], alloy_alone)):
        np.testing.assert_array_equal(got["st"]["pos"], want["st"]["pos"])
        np.testing.assert_array_equal(got["st"]["force"], want["st"]["force"])
        assert got["th"][-1] == want["th"][-1]


def test_autotune_small():
    from ndyn_omp_amd import autotune as AT
    print("case autotune", flush=True)
    sysl = systems()
    table = AT.tune(sysl, lambda skin, n: G.run_params(n, W.DT_1FS, out_steps=n, skin=skin,
                                                       trigger=G.TRIGGER_SKIN),
                    classes=(1000, 1 << 26), tpas=(1, 2), skins=(0.5, 0.8),
                    blocks=((1, 1), (2, 2)), steps=20, warmup=5)
    assert [c[0] for c in table.classes] == [1000, 1 << 26]
    for _, tpa, skin, by, bz in table.classes:
        assert tpa in (1, 2) and skin in (0.5, 0.8) and (by, bz) in ((1, 1), (2, 2))
    assert len(table.trials) == 7          # start + (one tpa + one skin + one block) per class
    assert all(t["atom_steps_per_s"] > 0 for t in table.trials)
    eng = G.BatchEngine(path=G.PATH_MIXED)
    table.apply(eng)
    for s in sysl:
        eng.admit(s, G.run_params(10, W.DT_1FS, out_steps=10, skin=table.skin_for(s.np),
                                  trigger=G.TRIGGER_SKIN))
    eng.step(10)
    eng.close()
