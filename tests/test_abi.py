"""CPU checks of the boundary: the library loads and exports every entry point that
include/ndyn_gpu.h declares; the pair-table helper (host-only) matches the oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "ndyn_gpu.h")
LIB = os.path.join(ROOT, "ndyn_omp_amd", "libndyn_gpu.so")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(ndg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for must in ("ndg_create", "ndg_compute", "ndg_integrate1", "ndg_integrate2",
                 "ndg_build_lists", "ndg_run_steps", "ndg_engine_admit", "ndg_engine_step"):
        assert must in syms


@pytest.mark.skipif(not os.path.exists(LIB), reason="HIP library not built")
def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if l.strip())
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="HIP library not built")
def test_library_loads_and_tables_match_oracle(oracle):
    from ndyn_omp_amd import engine as G
    from ndyn_omp_amd import workloads as W
    G.load_library()
    for kind, prm in ((1, W.CU_MORSE), (2, W.ALLOY_BUCK), (3, W.AR_LJ), (4, W.CU_MORSE)):
        a = G.pair_tables([kind], prm, 0.25, 5000)
        b = oracle.pair_tables([kind], prm, 0.25, 5000)
        for x, y in zip(a, b):
            np.testing.assert_allclose(x, y, rtol=1e-14, atol=1e-15)


@pytest.mark.skipif(not os.path.exists(LIB), reason="HIP library not built")
def test_library_builds_every_reference_potential():
    """The engine's host table builder against the compiled reference's tables for sine,
    cosine, soft, John, MorseC, BuckC and LJC (golden), as the oracle is."""
    from conftest import golden
    from ndyn_omp_amd import engine as G
    g = golden("tables_more.npz")
    for name in ("sine", "cosine", "soft", "john", "morsec", "buckc", "ljc"):
        P, F, rc2, dr2 = G.pair_tables([int(g[f"{name}_kind"])], g[f"{name}_params"],
                                       float(g["ptrmin"]), int(g["ptsize"]),
                                       qprd=[float(g[f"{name}_qprd"])])
        assert rc2[0] == g[f"{name}_rc2"] and dr2[0] == g[f"{name}_dr2"], name
        for got, want in ((P[0], g[f"{name}_P"]), (F[0], g[f"{name}_F"])):
            np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-13 * np.abs(want).max(),
                                       err_msg=name)


def test_fortran_module_binds_every_entry_point():
    """ndyn_omp_amd/fortran/ndg_iface.f90 declares a bind(C) interface for every
    function of include/ndyn_gpu.h."""
    import re
    hdr = open(os.path.join(ROOT, "include", "ndyn_gpu.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    decl = set(re.findall(r"^\s*(?:int|size_t|const char \*)\s*(ndg_\w+)\s*\(", hdr, flags=re.M))
    f90 = open(os.path.join(ROOT, "ndyn_omp_amd", "fortran", "ndg_iface.f90")).read()
    bound = set(re.findall(r"name='(ndg_\w+)'", f90))
    assert decl and decl - bound == set()


def test_hold_arrays_are_checked_before_the_call():
    """The Hold/Disp arrays the library reads (hold[np], bc_disp[nhold], bc_dir and
    bc_parm (nhold,3)) are shape-checked on the host before any call."""
    from ndyn_omp_amd import engine as G
    ok = G._hold_arrays(4, [0, 1, 1, 0], [1], [[1, 0, 0]], [[1e-4, 0, 0]])
    assert [a.shape for a in ok] == [(4,), (1,), (3,), (3,)]
    with pytest.raises(ValueError):
        G._hold_arrays(5, [0, 1, 1, 0], [1], [[1, 0, 0]], [[1e-4, 0, 0]])
    with pytest.raises(ValueError):
        G._hold_arrays(4, [0, 1, 1, 0], [1, 0], [[1, 0, 0]], [[1e-4, 0, 0]])
    with pytest.raises(ValueError):
        G._hold_arrays(4, [0, 1, 1, 0], [1], [[1, 0]], [[1e-4, 0, 0]])
