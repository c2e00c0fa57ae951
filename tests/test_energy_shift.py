"""The energy-shifted tables (kind | NDG_ESHIFT; not in the reference).

The reference truncates every pair potential at its cutoff (src/potentials.f90:21-32: the
table stops at rc), so the potential energy jumps by V(rc) whenever a pair crosses the
cutoff while the force does not.  Measured with the oracle below: that jump is the whole NVE
"drift" of the LJ and Buckingham configurations -- the dynamics themselves conserve
PE_shifted + KE, where PE_shifted = sum of V(r) - V(rc), to ~1e-6 eV/atom per 10^4 steps.

- tables: the energy table is the reference's minus V(rc) (to rounding), the force table is
  bit-identical (so forces and trajectories are the reference's), from the oracle and the
  engine's host builder alike; NDG_ESHIFT is refused for spring pairs;
- NVE (oracle, FP64): LJ argon at 2 fs and the Buckingham alloy at 1 fs -- the reference
  tables drift by more than 1e-4 eV/atom per 10^4 steps, the shifted tables by less than
  1e-5, along the identical trajectory;
- the deck driver reads '<name>_shift' pair names (WritePotentialTable output shifted).
"""
import os

import numpy as np
import pytest

from conftest import ROOT
from ndyn_omp_amd import workloads as W


def _v(kind, prm, r):
    if kind == W.LJ:
        eps, sig = prm[1], prm[2]
        return 4 * eps * ((sig / r) ** 12 - (sig / r) ** 6)
    A, rho, C = prm[1], prm[2], prm[3]
    return A * np.exp(-r / rho) - C / r ** 6


@pytest.mark.parametrize("kind,prm", [(W.LJ, W.AR_LJ), (W.BUCK, W.ALLOY_BUCK),
                                      (W.MORSE, W.CU_MORSE)])
def test_shifted_tables(oracle, kind, prm):
    P0, F0, rc2, dr2 = oracle.pair_tables([kind], prm, 0.25, 5000)
    P1, F1, rc2b, dr2b = oracle.pair_tables([kind | W.ESHIFT], prm, 0.25, 5000)
    np.testing.assert_array_equal(F1, F0)
    assert (rc2b, dr2b) == (rc2, dr2)
    shift = P0 - P1
    assert np.ptp(shift) <= 1e-15 * np.abs(P0).max()
    # the table point at rc (index ptsize - 1, s = rc^2): the shift is V(rc) there, and the
    # shifted energy vanishes
    if kind != W.MORSE:
        assert shift[0, 4999] == pytest.approx(_v(kind, prm, prm[0]), rel=1e-12)
    assert abs(P1[0, 4999]) <= 1e-15
    lib = os.path.join(ROOT, "ndyn_omp_amd", "libndyn_gpu.so")
    if os.path.exists(lib):
        from ndyn_omp_amd import engine as G
        P2, F2, _, _ = G.pair_tables([kind | W.ESHIFT], np.array([prm]), 0.25, 5000)
        np.testing.assert_array_equal(F2, F0)
        np.testing.assert_allclose(P2, P1, rtol=0, atol=1e-13 * np.abs(P0).max())
        with pytest.raises(G.NdgError):
            G.pair_tables([W.SPRING | W.ESHIFT], np.array([W.PD_SPRING]), 0.25, 1000)


def _drift(oracle, s, tables, dt, nsteps, skin):
    st, th, err = oracle.run(s, nsteps, dt, skin=skin, out_steps=100,
                             trigger=oracle.TRIGGER_SKIN, tables=tables)
    assert err == 0
    e = np.array([r["pe"] + r["ke"] for r in th])
    steps = np.array([r["step"] for r in th])
    return (e[-1] - e[0]) / s.np * 1e4 / (steps[-1] - steps[0]), st


@pytest.mark.parametrize("case", ["lj", "buck"])
def test_shift_removes_the_nve_drift(oracle, case):
    if case == "lj":
        s, dt, skin, n = W.ar_lj(7, temp_k=90.0, seed=73), W.DT_2FS, 0.4, 2000
    else:
        s, dt, skin, n = W.buck_alloy((7, 7, 7), temp_k=900.0, seed=74), W.DT_1FS, 0.5, 2000
    t0 = oracle.pair_tables(s.kinds, s.params, s.ptrmin, s.ptsize)
    t1 = oracle.pair_tables(np.asarray(s.kinds) | W.ESHIFT, s.params, s.ptrmin, s.ptsize)
    d0, st0 = _drift(oracle, s, t0, dt, n, skin)
    d1, st1 = _drift(oracle, s, t1, dt, n, skin)
    print(f"{case}: drift per atom per 1e4 steps: reference tables {d0:.2e}, shifted {d1:.2e}")
    np.testing.assert_array_equal(st1["pos"], st0["pos"])      # the same trajectory
    assert abs(d0) > 1e-4
    assert abs(d1) < 1e-5


DRIVER = os.path.join(ROOT, "ndyn_omp_amd", "bin", "ndyn_gpu")


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built")
def test_driver_reads_shift_names(tmp_path, oracle):
    from test_driver import run_deck
    deck = """START
Types 1
 Ar 39.948
WritePotentialTable
Pairs 1 0.25 1000
 Ar Ar sym LJ_shift 8.5125 0.0104 3.405
END
"""
    p = run_deck(deck, tmp_path)
    assert p.returncode == 0, p.stderr
    pot = np.loadtxt(tmp_path / "pottab.dat")
    P1, F1, _, _ = oracle.pair_tables([W.LJ | W.ESHIFT], W.AR_LJ, 0.25, 1000)
    np.testing.assert_allclose(pot[:, 1], P1[0][0:1000], rtol=1e-6, atol=1e-9)
