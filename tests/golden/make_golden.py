"""Writes the golden fixtures in tests/golden/ from the compiled Ndyn reference.

Run in the build container only (it needs oracle/_ref/ndyn_ref_harness, built by
`make -C oracle ref` from the reference sources):

    python tests/golden/make_golden.py

Each fixture is an .npz holding the inputs and the reference's outputs at full
precision.  Systems come from ndyn_omp_amd.workloads (seeded), so the inputs can also
be regenerated; they are stored anyway so the fixtures stand alone.  The harness runs
the reference with one OpenMP thread (its force sum order then equals a serial sweep).
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from ndyn_omp_amd import workloads as W  # noqa: E402

from oracle import refrun  # noqa: E402

HARNESS = refrun.HARNESS
OUT = os.path.dirname(os.path.abspath(__file__))
run_harness = refrun.run_harness


def sys_arrays(s):
    d = dict(pos=s.pos, vel=s.vel, typ=s.typ, box=s.box, masses=s.masses,
             pairmat=s.pairmat, kinds=s.kinds, params=s.params,
             ptrmin=np.float64(s.ptrmin), ptsize=np.int64(s.ptsize))
    if s.ipos is not None:
        d.update(ipos=s.ipos, init_box=s.init_box)
    return d


def ipos_kw(s):
    return {} if s.ipos is None else dict(ipos=s.ipos, init_box=s.init_box)


def thermo_lines(stdout):
    rows = []
    for line in stdout.splitlines():
        parts = line.split()
        if len(parts) == 6 and not line.startswith("#"):
            try:
                rows.append([float(x) for x in parts])
            except ValueError:
                pass
    return np.array(rows)


def golden_tables():
    out = {}
    for name, kind, prm in [("morse", W.MORSE, W.CU_MORSE), ("lj", W.LJ, W.AR_LJ),
                            ("buck", W.BUCK, W.ALLOY_BUCK)]:
        s = W.cu_morse(1, temp_k=0.0)
        s.kinds = np.array([kind], np.int32)
        s.params = np.array([prm], np.float64)
        raw, _ = run_harness(1, s)
        n = s.ptsize + 1
        a = np.frombuffer(raw, "<f8")
        out[f"{name}_params"] = np.array(prm)
        out[f"{name}_rc2"], out[f"{name}_dr2"] = a[0], a[1]
        out[f"{name}_P"] = a[2:2 + n]
        out[f"{name}_F"] = a[2 + n:2 + 2 * n]
    out["ptrmin"], out["ptsize"] = np.float64(0.25), np.int64(5000)
    np.savez_compressed(os.path.join(OUT, "tables.npz"), **out)


# the other potentials of buildPairTables (src/potentials.f90:82-147): name, kind, params
# {rc, p2, p3, p4}, charge product
MORE_TABLES = [("sine", W.SINE, (6.0, 0.0, 0.0, 0.0), 0.0),
               ("cosine", W.COSINE, (6.0, 0.0, 0.0, 0.0), 0.0),
               ("soft", W.SOFT, (6.0, 0.5, 6.0, 0.0), 0.0),          # examples/SineTest
               ("john", W.JOHNSON, (6.0, 0.5, 2.5, 5.5), 0.0),
               ("morsec", W.MORSE_C, W.CU_MORSE, 0.8),
               ("buckc", W.BUCK_C, W.ALLOY_BUCK, -1.2),
               ("ljc", W.LJ_C, W.AR_LJ, 0.35)]


def golden_tables_more():
    out = {}
    for name, kind, prm, q in MORE_TABLES:
        s = W.cu_morse(1, temp_k=0.0)
        s.kinds = np.array([kind], np.int32)
        s.params = np.array([prm], np.float64)
        raw, _ = run_harness(1, s, qprd=[q])
        n = s.ptsize + 1
        a = np.frombuffer(raw, "<f8")
        out[f"{name}_kind"], out[f"{name}_params"], out[f"{name}_qprd"] = kind, np.array(prm), q
        out[f"{name}_rc2"], out[f"{name}_dr2"] = a[0], a[1]
        out[f"{name}_P"] = a[2:2 + n]
        out[f"{name}_F"] = a[2 + n:2 + 2 * n]
    out["ptrmin"], out["ptsize"] = np.float64(0.25), np.int64(5000)
    np.savez_compressed(os.path.join(OUT, "tables_more.npz"), **out)


def golden_list(fname, s, skin):
    raw, _ = run_harness(2, s, skin=skin)
    hdr = np.frombuffer(raw[:16], "<i4")
    mcell, neimax = hdr[:3], int(hdr[3])
    vnl = np.frombuffer(raw[16:], "<i4").reshape(s.np, neimax + 1)
    counts = vnl[:, 0] - 1
    ent = np.concatenate([vnl[i, 1:1 + counts[i]] - 1 for i in range(s.np)]).astype(np.int32)
    np.savez_compressed(os.path.join(OUT, fname), **sys_arrays(s), skin=np.float64(skin),
                        mcell=mcell, neimax=np.int64(neimax), counts=counts.astype(np.int32),
                        entries=ent)


def golden_force(fname, s, skin):
    raw, _ = run_harness(3, s, skin=skin, **ipos_kw(s))
    a = np.frombuffer(raw, "<f8")
    n = s.np
    np.savez_compressed(os.path.join(OUT, fname), **sys_arrays(s), skin=np.float64(skin),
                        force=a[:3 * n].reshape(n, 3), pe=a[3 * n], virial=a[3 * n + 1:3 * n + 7])


def golden_run(fname, s, nsteps, dt, skin, outsteps, **kw):
    raw, stdout = run_harness(4, s, nsteps=nsteps, dt=dt, skin=skin, outsteps=outsteps,
                              **ipos_kw(s), **kw)
    a = np.frombuffer(raw, "<f8")
    n = s.np
    o = 0
    res = {}
    for k in ("pos", "vel", "acc", "force"):
        res["out_" + k] = a[o:o + 3 * n].reshape(n, 3)
        o += 3 * n
    res["out_pe"], res["out_ke"] = a[o], a[o + 1]
    res["out_virial"] = a[o + 2:o + 8]
    res["out_box"] = a[o + 8:o + 11]
    res["out_e0"], res["out_nupdates"], res["out_step"] = a[o + 11], a[o + 12], a[o + 13]
    meta = dict(nsteps=np.int64(nsteps), dt=np.float64(dt), skin=np.float64(skin),
                outsteps=np.int64(outsteps))
    holds = kw.pop("holds", None)
    if holds is not None:
        meta.update(hold=np.asarray(holds[0], np.int32), bc_disp=np.asarray(holds[1], np.int32),
                    bc_dir=np.asarray(holds[2], np.int32), bc_parm=np.asarray(holds[3], np.float64))
    for k, v in kw.items():
        if v is None or isinstance(v, bool):
            meta[k] = np.array(-1 if v is None else int(v))
        else:
            meta[k] = np.array(v, dtype=object if isinstance(v, tuple) and
                               any(isinstance(x, tuple) for x in v) else None)
    if "iso_p" in kw and kw["iso_p"]:
        meta["iso_p"] = np.array(list(kw["iso_p"][0]) + [kw["iso_p"][1], kw["iso_p"][2]])
    res["thermo"] = thermo_lines(stdout)
    np.savez_compressed(os.path.join(OUT, fname), **sys_arrays(s), **meta, **res)


def asym_fixtures():
    """Asymmetric pair matrices (compute1): one-way A->B, and two directed tables."""
    golden_list("list_asym1.npz", W.asym_morse(7, 1, jitter_a=0.05, seed=121), 0.5)
    golden_force("force_asym1.npz", W.asym_morse(7, 1, jitter_a=0.05, seed=122), 0.5)
    golden_force("force_asym2.npz", W.asym_morse(7, 2, jitter_a=0.05, seed=123), 0.5)
    golden_run("run_asym2_nve.npz", W.asym_morse(7, 2, temp_k=300.0, seed=12351), 60,
               W.DT_1FS, 0.5, 20)


def spring_fixtures():
    """Spring pairs (compute2 / compute3) against initial positions, and the peridynamic
    tension deck examples/testPDspringTension/input.sh (4000 atoms, BoxVel, ReNeighbor F,
    initial positions set by dynamics) for its first 300 steps."""
    golden_force("force_spring_cu7.npz", W.spring_cu(7, stretch=(0.03, 0.0, 0.0), jitter_a=0.15,
                                                     seed=131), 0.0)
    golden_force("force_spring_mix0.npz", W.spring_mix(7, 0, seed=132), 0.5)
    golden_force("force_spring_mix1.npz", W.spring_mix(7, 1, seed=133), 0.5)
    golden_force("force_spring_mix2.npz", W.spring_mix(7, 2, seed=134), 0.5)
    pd = W.spring_cu(10, a=3.68)
    pd.ipos = pd.init_box = None
    golden_run("run_spring_pd.npz", pd, 300, 0.1, 0.0, 100, renei=False,
               box_vel=(1.0e-2, 0.0, 0.0))
    golden_run("run_spring_mix1_nve.npz", W.spring_mix(7, 1, temp_k=300.0, jitter_a=0.0,
                                                       stretch=(0.0, 0.0, 0.0), seed=12362),
               60, W.DT_1FS, 0.5, 20)


def hold_fixtures():
    """Hold/Disp boundary conditions (src/nd.f90:1048-1075) on a thermalised Cu block: a
    slab frozen (Disp 0 0 0), a slab pulled along x (Disp 2e-5 F F), a hold group without
    Disp (free), the rest free."""
    s = W.cu_morse(7, temp_k=600.0, seed=12370)
    x = s.pos[:, 0]
    hold = np.where(x < -0.35, 1, np.where(x > 0.35, 2, np.where(np.abs(x) < 0.05, 3, 0)))
    holds = (hold.astype(np.int32), [1, 1, 0], [[1, 1, 1], [1, 0, 0], [1, 1, 1]],
             [[0.0, 0.0, 0.0], [2.0e-5, 0.0, 0.0], [0.0, 0.0, 0.0]])
    golden_run("run_hold_cu7.npz", s, 60, W.DT_1FS, 0.5, 20, holds=holds)


def main():
    if not os.path.exists(HARNESS):
        raise SystemExit("build the reference harness first: make -C oracle ref")
    if "--asym-only" in sys.argv:
        asym_fixtures()
        return
    if "--spring-only" in sys.argv:
        spring_fixtures()
        return
    if "--hold-only" in sys.argv:
        hold_fixtures()
        return
    if "--tables-more" in sys.argv:
        golden_tables_more()
        return
    golden_tables()
    golden_tables_more()
    # half lists: Cu 7^3 (MCELL 3), c0-size Cu 10^3 (MCELL 4), Buckingham alloy (3 pairs)
    golden_list("list_cu7.npz", W.cu_morse(7, jitter_a=0.05, seed=101), 0.5)
    golden_list("list_cu10.npz", W.cu_morse(10, jitter_a=0.05, seed=102), 0.5)
    golden_list("list_alloy5.npz", W.buck_alloy((5, 5, 5), jitter_a=0.1, seed=103), 0.5)
    golden_list("list_ar6.npz", W.ar_lj(6, seed=104), 0.4)
    # forces
    golden_force("force_cu10.npz", W.cu_morse(10, jitter_a=0.05, seed=105), 0.5)
    golden_force("force_ar6.npz", W.ar_lj(6, jitter_a=0.05, seed=106), 0.4)
    golden_force("force_alloy5.npz", W.buck_alloy((5, 5, 5), jitter_a=0.1, seed=107), 0.5)
    # dynamics: NVE Cu (config-0 system, 100 steps), NVT Ar, NPT Cu, isoenergetic alloy
    golden_run("run_cu10_nve.npz", W.cu_morse(10, temp_k=300.0, seed=12345), 100, W.DT_1FS,
               0.5, 25)
    golden_run("run_ar6_nvt.npz", W.ar_lj(6, seed=12346), 100, W.DT_2FS, 0.4, 20,
               iso_ke=(1.5 * W.KB_EV * 90.0, W.DT_2FS, 1))
    golden_run("run_cu7_npt.npz", W.cu_morse(7, temp_k=600.0, seed=12347), 100, W.DT_1FS,
               0.5, 20, iso_ke=(1.5 * W.KB_EV * 600.0, 100 * W.DT_1FS, 1),
               iso_p=((0.0, 0.0, 0.0), 1000 * W.DT_1FS, 10))
    golden_run("run_alloy5_nve.npz", W.buck_alloy((5, 5, 5), temp_k=900.0, seed=12348), 60,
               W.DT_1FS, 0.5, 20, iso_e=(50 * W.DT_1FS, 5))
    asym_fixtures()
    spring_fixtures()
    hold_fixtures()
    for f in sorted(os.listdir(OUT)):
        if f.endswith(".npz"):
            print(f"{f:28s} {os.path.getsize(os.path.join(OUT, f)) / 1024:8.1f} KiB")


if __name__ == "__main__":
    main()
