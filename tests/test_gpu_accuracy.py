"""Mixed path against the FP64 path (which reproduces the compiled reference to ~1e-15)
on thermalised configurations, and NVE energy drift.

Bars: RMS relative force error <= 1e-5, per-atom energy within 1e-6 eV, pressure within
1e-5 relative (Morse Cu, LJ argon, Buckingham alloy).  NVE with the energy-consistent
Morse table, 10^4 steps of 1 fs at 300 K: |E(end) - E(0)| <= 1e-4 eV/atom on both paths,
and the mixed path's drift within 10 % of the FP64 path's (precision adds no drift).
"""
import pytest

from ndyn_omp_amd import workloads as W

pytestmark = pytest.mark.gpu
A = pytest.importorskip("ndyn_omp_amd.accuracy")
G = pytest.importorskip("ndyn_omp_amd.engine")


@pytest.mark.parametrize("name", ["cu", "ar", "alloy"])
def test_pair_errors_on_thermal_snapshot(name):
    s, skin, dt = {
        "cu": (W.cu_morse(9, temp_k=600.0, seed=61), 0.5, W.DT_1FS),
        "ar": (W.ar_lj(7, temp_k=90.0, seed=62), 0.4, W.DT_2FS),
        "alloy": (W.buck_alloy((7, 7, 7), temp_k=900.0, seed=63), 0.5, W.DT_1FS),
    }[name]
    print(f"case {name} np={s.np}", flush=True)
    r = A.pair_errors(A.thermal_snapshot(s, 100, dt, skin), skin)
    print(r)
    assert r["force_rms_rel"] <= 1e-5
    assert r["energy_per_atom_max_abs_ev"] <= 1e-6
    assert r["pressure_rel"] <= 1e-5


def test_nve_drift_energy_consistent_morse():
    s = W.cu_morse(8, temp_k=300.0, seed=64, kind=W.MORSE_PHYS)
    print(f"case nve np={s.np}", flush=True)
    d = {p: A.nve_drift(s, p, 10000, W.DT_1FS, 0.5) for p in (G.PATH_MIXED, G.PATH_FP64)}
    print(d)
    for p in d:
        assert d[p]["err"] == 0 and d[p]["steps"] == 10000
        assert abs(d[p]["drift_per_atom_ev"]) <= 1e-4
    m, f = d[G.PATH_MIXED]["drift_per_atom_ev"], d[G.PATH_FP64]["drift_per_atom_ev"]
    assert abs(m - f) <= 0.1 * abs(f) + 1e-7


@pytest.mark.parametrize("ptsize", [5000, 8000, 14000])
def test_energy_virial_with_staged_and_global_tables(ptsize):
    """The energy/virial force variant stages the FP32 force and potential tables in LDS
    when both fit in 48 KB (5001 points: 40 KB), reads the potential table from global
    memory when only the force table fits (8001 points), and reads both from global
    memory beyond that (14001 points).  Same bars in all three cases."""
    s = W.cu_morse(9, temp_k=600.0, seed=65, ptsize=ptsize)
    print(f"case ptsize={ptsize} np={s.np}", flush=True)
    r = A.pair_errors(A.thermal_snapshot(s, 50, W.DT_1FS, 0.5), 0.5)
    print(r)
    assert r["force_rms_rel"] <= 1e-5
    assert r["energy_per_atom_max_abs_ev"] <= 1e-6
    assert r["pe_rel"] <= 1e-6
    assert r["pressure_rel"] <= 1e-5
