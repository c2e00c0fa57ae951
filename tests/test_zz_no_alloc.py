"""ndg_engine_fetch_noalloc: the fetch that allocates nothing (its staging is a fixed
device array of the library) returns exactly what ndg_engine_fetch returns, on two calls in
a row, the second for a smaller system, and refuses a system above its 262144-atom limit
with NDG_ERR_ARG."""
import numpy as np
import pytest

from ndyn_omp_amd import workloads as W

pytestmark = pytest.mark.gpu
G = pytest.importorskip("ndyn_omp_amd.engine")


def _rp(n):
    return G.run_params(n, W.DT_1FS, out_steps=n, skin=0.5, trigger=G.TRIGGER_SKIN,
                        isokinetic=(1.5 * W.KB_EV * 500.0, 20 * W.DT_1FS, 1))


def test_fetch_noalloc_equals_fetch():
    big = W.cu_morse(12, temp_k=700.0, seed=501)                  # 6912 atoms
    small = W.buck_alloy((5, 5, 6), temp_k=400.0, seed=502)       # 600 atoms
    eng = G.BatchEngine(path=G.PATH_MIXED)
    sb = eng.admit(big, _rp(40))
    ss = eng.admit(small, _rp(40))
    eng.step(25)
    for slot, s in ((sb, big), (ss, small)):
        print(f"case fetch_noalloc np={s.np}", flush=True)
        a = eng.fetch_noalloc(slot, s.np)
        b = eng.fetch(slot, s.np)
        for k in ("pos", "vel", "force", "box"):
            np.testing.assert_array_equal(a[k], b[k])
        assert np.abs(a["force"]).max() > 0 and np.abs(a["vel"]).max() > 0
    # the staging keeps the larger system's data beyond the smaller one's: a second fetch of
    # the larger one after more steps is again equal
    eng.step(10)
    a = eng.fetch_noalloc(sb, big.np)
    b = eng.fetch(sb, big.np)
    for k in ("pos", "vel", "force", "box"):
        np.testing.assert_array_equal(a[k], b[k])
    eng.close()


def test_fetch_noalloc_refuses_above_its_limit():
    s = W.cu_morse(41, temp_k=0.0)                               # 275684 atoms
    assert s.np > 1 << 18
    eng = G.BatchEngine(path=G.PATH_MIXED)
    slot = eng.admit(s, _rp(2))
    with pytest.raises(G.NdgError) as ei:
        eng.fetch_noalloc(slot, s.np)
    assert ei.value.code == G.ERR_ARG
    assert "262144" in str(ei.value)
    eng.close()
