"""Type pairs with different potentials and list cutoffs (Buckingham A-A rc 8.0, Morse
A-B rc 6.5, LJ B-B rc 7.2): the list builder's per-pair cutoff table and the force
kernels' per-pair tables against the oracle.  Bars: neighbour sets bit-exact (both
paths); FP64 forces within 1e-12 relative, mixed within 1e-5 RMS relative; PE within
1e-12 (FP64) and 1e-6 (mixed) relative.
"""
import numpy as np
import pytest

from ndyn_omp_amd import workloads as W

pytestmark = pytest.mark.gpu
G = pytest.importorskip("ndyn_omp_amd.engine")


def system():
    s = W.buck_alloy((6, 6, 7), temp_k=500.0, seed=81, jitter_a=0.08)
    s.kinds = np.array([W.BUCK, W.MORSE, W.LJ], dtype=np.int32)
    s.params = np.array([[8.0, 1000.0, 0.3, 10.0],
                         [6.5, 0.2, 1.4, 2.9],
                         [7.2, 0.01, 3.4, 0.0]], dtype=np.float64)
    return s


@pytest.mark.parametrize("path_name,path", [("fp64", 0), ("mixed", 1)])
def test_mixed_cutoffs_lists_and_forces(path_name, path, oracle):
    s, skin = system(), 0.5
    print(f"case mixed cutoffs {path_name} np={s.np}", flush=True)
    _, vnl, nm = oracle.half_list(s, skin)
    ref = oracle.neighbour_sets_from_half(vnl)
    tables = oracle.pair_tables(s.kinds, s.params, s.ptrmin, s.ptsize)
    f_ref, pe_ref, _, _ = oracle.compute0(s, vnl, nm, tables)
    sim = G.Simulation(s, path=path)
    sim.build_lists(skin)
    counts, nbr = sim.neighbours()
    for i in range(s.np):
        np.testing.assert_array_equal(nbr[i, :counts[i]], ref[i])
    pe, _ = sim.compute(True, True)
    f = sim.state()["force"]
    err = float(np.linalg.norm(f - f_ref) / np.linalg.norm(f_ref))
    print(f"  force err {err:.3e}")
    if path == 0:
        assert err <= 1e-12 and pe == pytest.approx(pe_ref, rel=1e-12)
    else:
        assert err <= 1e-5 and pe == pytest.approx(pe_ref, rel=1e-6)
    sim.close()
