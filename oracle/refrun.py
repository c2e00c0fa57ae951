"""Runs the compiled Ndyn reference (oracle/_ref/ndyn_ref_harness) on a System.

TEST INFRASTRUCTURE ONLY: used by tests/golden/make_golden.py to write golden vectors
and by bench.py's cpu_baseline leg to time the reference's own step loop on the host.
"""
from __future__ import annotations

import os
import re
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HARNESS = os.path.join(HERE, "_ref", "ndyn_ref_harness")
# the same harness at the reference Makefile's optimisation level (timing only, no goldens)
HARNESS_FAST = os.path.join(HERE, "_ref_fast", "ndyn_ref_harness")

MODE_TABLES, MODE_LIST, MODE_FORCE, MODE_RUN = 1, 2, 3, 4


def available(fast: bool = False) -> bool:
    return os.path.exists(HARNESS_FAST if fast else HARNESS)


def write_input(path, mode, s, nsteps=0, outsteps=1, step0=0, neimax=0, skin=-1.0, dt=0.0,
                renei=True, renei_freq=1, iso_ke=None, iso_p=None, iso_e=None, acc=None,
                nthreads=1, box_vel=None, ipos=None, init_box=None, holds=None, qprd=None):
    """Binary input of ref_harness.f90 (its header documents the layout)."""
    iv = np.array([mode, s.np, s.ntypes, s.npairs, s.ptsize, neimax, nsteps, outsteps, step0,
                   int(renei), renei_freq,
                   1 if iso_ke else 0, iso_ke[2] if iso_ke else 1,
                   1 if iso_p else 0, iso_p[2] if iso_p else 1,
                   1 if iso_e else 0, iso_e[1] if iso_e else 1, nthreads], dtype="<i4")
    dv = np.zeros(13, dtype="<f8")
    dv[0:3] = s.box
    dv[3] = dt
    dv[4] = skin
    dv[5] = s.ptrmin
    if iso_ke:
        dv[6], dv[7] = iso_ke[0], iso_ke[1]
    if iso_p:
        dv[8:11] = iso_p[0]
        dv[11] = iso_p[1]
    if iso_e:
        dv[12] = iso_e[0]
    with open(path, "wb") as f:
        f.write(iv.tobytes())
        f.write(dv.tobytes())
        f.write(np.asarray(s.masses, "<f8").tobytes())
        f.write(np.asarray(s.pairmat.T, "<i4").tobytes())          # column-major
        f.write(np.asarray(s.kinds, "<i4").tobytes())
        f.write(np.asarray(s.params, "<f8").tobytes())             # 4 x npairs column-major
        f.write(np.asarray(s.typ, "<i4").tobytes())
        f.write(np.asarray(s.pos, "<f8").tobytes())                # 3 x np column-major
        f.write(np.asarray(s.vel, "<f8").tobytes())
        a = np.zeros_like(s.pos) if acc is None else acc
        f.write(np.asarray(a, "<f8").tobytes())
        f.write(np.array([0 if box_vel is None else 1, 0 if ipos is None else 1], "<i4").tobytes())
        f.write(np.asarray([0.0, 0.0, 0.0] if box_vel is None else box_vel, "<f8").tobytes())
        if ipos is not None:
            f.write(np.asarray(ipos, "<f8").tobytes())
            f.write(np.asarray(init_box, "<f8").tobytes())
        f.write(np.array([0 if holds is None else len(holds[1])], "<i4").tobytes())
        if holds is not None:
            f.write(np.asarray(holds[0], "<i4").tobytes())
            f.write(np.asarray(holds[1], "<i4").tobytes())
            f.write(np.asarray(holds[2], "<i4").tobytes())       # (nhold,3) = bcDir(3,nhold)
            f.write(np.asarray(holds[3], "<f8").tobytes())
        f.write(np.asarray(np.zeros(s.npairs) if qprd is None else qprd, "<f8").tobytes())


def run_harness(mode, s, timeout=3600, fast=False, **kw):
    """Returns (raw output bytes, stdout text).  fast=True runs the build at the reference
    Makefile's optimisation level (for timing; goldens come from the strict build)."""
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        write_input(fin, mode, s, **kw)
        env = dict(os.environ, OMP_NUM_THREADS=str(kw.get("nthreads", 1)),
                   OMP_PROC_BIND=os.environ.get("OMP_PROC_BIND", "close"))
        p = subprocess.run([HARNESS_FAST if fast else HARNESS, fin, fout], cwd=d, env=env,
                           capture_output=True,
                           text=True, timeout=timeout)
        # a Fortran STOP (e.g. 'too many neighbours in list') exits with status 0
        if p.returncode != 0 or "STOP" in p.stderr:
            raise RuntimeError(p.stdout[-2000:] + p.stderr)
        return open(fout, "rb").read(), p.stdout


def step_loop_seconds(stdout: str) -> float:
    """The reference's own timer of its step loop ('Elapsed time for main dynamics',
    src/nd.f90:129,186-190)."""
    m = re.search(r"Elapsed time for main dynamics:\s*\n#\s*([0-9.Ee+-]+)", stdout)
    if not m:
        raise RuntimeError("no step-loop timer in reference output")
    return float(m.group(1))
