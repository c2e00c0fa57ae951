"""Cell-block storage order (ndg_engine_cell_blocks): cells stored in blocks of 1 x by x bz
cells instead of the reference's linear order change only which atoms share a tile.  Each
atom's list is the same sequence (stencil cell, then the cell's own order), so forces are
the same bits and NVE trajectories are bit-identical; sums over tiles (energy, virial)
differ only in rounding.  Systems: grids of 3, 4 and 5 x 6 x 7 cells (blocks that do not
divide the grid leave thinner blocks at the edges) and an all-pairs small box."""
import numpy as np
import pytest

from ndyn_omp_amd import workloads as W

pytestmark = pytest.mark.gpu
G = pytest.importorskip("ndyn_omp_amd.engine")

SHAPES = [(1, 1), (2, 2), (3, 2), (4, 4), (1, 3)]


def systems():
    return [W.cu_morse(7, temp_k=900.0, seed=71),
            W.cu_morse(10, temp_k=700.0, seed=72),
            W.buck_alloy((9, 10, 11), temp_k=1000.0, seed=73),
            W.cu_morse(4, temp_k=800.0, seed=74)]


def run(shape, nsteps=60):
    eng = G.BatchEngine(path=G.PATH_MIXED)
    eng.cell_blocks([(1 << 26, shape[0], shape[1])])
    sys_ = systems()
    rp = G.run_params(nsteps, W.DT_1FS, out_steps=20, skin=0.5, trigger=G.TRIGGER_SKIN)
    slots = [eng.admit(s, rp) for s in sys_]
    eng.step(nsteps)
    out = []
    for s, sl in zip(sys_, slots):
        st = eng.fetch(sl, s.np)
        th = eng.thermo(sl)
        out.append((st, th, eng.counters(sl)["nupdates"]))
    eng.close()
    return out


def test_cell_blocks_keep_trajectories():
    ref = run((1, 1))
    assert sum(r[2] for r in ref) > 0, "lists are rebuilt during the run"
    for shape in SHAPES[1:]:
        print(f"case cell block 1x{shape[0]}x{shape[1]}", flush=True)
        got = run(shape)
        for (a, ta, na), (b, tb, nb) in zip(got, ref):
            np.testing.assert_array_equal(a["pos"], b["pos"])
            np.testing.assert_array_equal(a["vel"], b["vel"])
            np.testing.assert_array_equal(a["force"], b["force"])
            assert na == nb
            for x, y in zip(ta, tb):
                assert x["pe"] == pytest.approx(y["pe"], rel=1e-12)
                assert x["ke"] == pytest.approx(y["ke"], rel=1e-12)
                np.testing.assert_allclose(x["virial"], y["virial"], rtol=1e-11, atol=1e-9)


def test_cell_blocks_rejects_bad_shapes():
    eng = G.BatchEngine(path=G.PATH_MIXED)
    with pytest.raises(G.NdgError):
        eng.cell_blocks([(1000, 0, 1)])
    with pytest.raises(G.NdgError):
        eng.cell_blocks([(1000, 2, 9)])
    eng.close()
