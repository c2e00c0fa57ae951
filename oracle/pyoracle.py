"""ctypes wrapper of oracle/libndyn_oracle.so (TEST INFRASTRUCTURE ONLY).

The oracle is the CPU FP64 restatement of the reference MD step used to check the
HIP engine.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
import this module; the product package never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libndyn_oracle.so")
_lib = None

dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")

MORSE, BUCK, LJ, MORSE_PHYS, SPRING = 1, 2, 3, 4, 5
SINE, COSINE, SOFT, JOHNSON, MORSE_C, BUCK_C, LJ_C = 6, 7, 8, 9, 10, 11, 12
ESHIFT = 0x100
TRIGGER_COMPAT, TRIGGER_SKIN = 0, 1


class RunParams(C.Structure):
    _fields_ = [("nsteps", C.c_int), ("dt", C.c_double), ("out_steps", C.c_int),
                ("step0", C.c_int), ("skin", C.c_double), ("neimax", C.c_int),
                ("renei", C.c_int), ("renei_freq", C.c_int), ("trigger", C.c_int),
                ("brute", C.c_int),
                ("isokinetic", C.c_int), ("ke_target", C.c_double), ("ke_tau", C.c_double),
                ("ke_freq", C.c_int),
                ("isobaric", C.c_int), ("p_target", C.c_double * 3), ("p_tau", C.c_double),
                ("p_freq", C.c_int),
                ("isoenergetic", C.c_int), ("e_tau", C.c_double), ("e_freq", C.c_int),
                ("box_load", C.c_int), ("box_vel", C.c_double * 3),
                ("springs", C.c_int), ("ipos", C.c_void_p), ("init_box", C.c_double * 3),
                ("ipos_set", C.c_int), ("nhold", C.c_int), ("hold", C.c_void_p),
                ("bc_disp", C.c_void_p), ("bc_dir", C.c_void_p), ("bc_parm", C.c_void_p)]


class Thermo(C.Structure):
    _fields_ = [("step", C.c_int), ("pe", C.c_double), ("ke", C.c_double),
                ("eerr", C.c_double), ("pressure", C.c_double), ("volume", C.c_double),
                ("virial", C.c_double * 6), ("box", C.c_double * 3)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.ndo_pair_table_q.argtypes = [C.c_int, dp, C.c_double, C.c_double, C.c_int, dp, dp,
                                       C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.ndo_pair_table.argtypes = [C.c_int, dp, C.c_double, C.c_int, dp, dp,
                                     C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.ndo_neimax_estimate.argtypes = [C.c_double, C.c_double, C.c_int, dp]
        L.ndo_cell_grid.argtypes = [dp, C.c_double, C.c_double, ip]
        L.ndo_cell_grid.restype = None
        L.ndo_build_half_list.argtypes = [C.c_int, dp, ip, dp, ip, C.c_int, ip, dp, C.c_double,
                                          C.c_double, C.c_int, C.c_int, ip, C.c_void_p]
        L.ndo_compute0.argtypes = [C.c_int, dp, ip, dp, ip, C.c_int, C.c_int, ip, dp, dp, dp, dp,
                                   C.c_double, C.c_int, C.c_int, C.c_int, dp,
                                   C.POINTER(C.c_double), dp, C.c_void_p]
        L.ndo_compute0.restype = None
        L.ndo_compute1.argtypes = L.ndo_compute0.argtypes
        L.ndo_compute1.restype = None
        L.ndo_compute23.argtypes = [C.c_int, C.c_int, dp, ip, dp, ip, C.c_int, C.c_int, ip, dp,
                                    dp, dp, dp, C.c_double, C.c_int, ip, dp, dp, dp, C.c_int,
                                    C.c_int, dp, C.POINTER(C.c_double), dp, C.c_void_p]
        L.ndo_compute23.restype = None
        L.ndo_asymmetric.argtypes = [C.c_int, ip]
        L.ndo_asymmetric.restype = C.c_int
        L.ndo_integrate1.argtypes = [C.c_int, dp, dp, dp, dp, dp, C.c_double]
        L.ndo_integrate1.restype = None
        L.ndo_integrate2.argtypes = [C.c_int, dp, dp, dp, ip, dp, C.c_double, C.c_int,
                                     C.POINTER(C.c_double)]
        L.ndo_integrate2.restype = None
        L.ndo_mk_iso_ke.argtypes = [C.c_int, dp, C.c_double, C.c_double, C.c_double, C.c_double]
        L.ndo_mk_iso_ke.restype = C.c_double
        L.ndo_mk_iso_e.argtypes = [C.c_int, dp, C.c_double, C.c_double, C.c_double, C.c_double,
                                   C.c_double]
        L.ndo_mk_iso_e.restype = C.c_double
        L.ndo_mk_iso_p.argtypes = [dp, dp, C.c_double, C.c_double, dp]
        L.ndo_mk_iso_p.restype = None
        L.ndo_check_list_update.argtypes = [C.c_int, dp, C.c_double, dp, dp, C.c_double, C.c_int]
        L.ndo_run.argtypes = [C.POINTER(RunParams), C.c_int, C.c_int, C.c_int, C.c_int,
                              C.c_double, dp, ip, dp, dp, dp, dp, dp, ip, dp, dp, dp, dp, dp,
                              C.c_void_p, C.POINTER(Thermo), C.c_int, C.POINTER(C.c_int),
                              C.POINTER(C.c_int)]
        _lib = L
    return _lib


def pair_tables(kinds, params, ptrmin, ptsize, qprd=None):
    """Tables for all pairs: P, F (npairs, ptsize+1); rc2, dr2 (npairs,)."""
    kinds = np.asarray(kinds, dtype=np.int32)
    params = np.ascontiguousarray(params, dtype=np.float64).reshape(-1, 4)
    npairs = kinds.shape[0]
    P = np.zeros((npairs, ptsize + 1))
    F = np.zeros((npairs, ptsize + 1))
    rc2 = np.zeros(npairs)
    dr2 = np.zeros(npairs)
    for p in range(npairs):
        a, b = C.c_double(), C.c_double()
        Pp = np.zeros(ptsize + 1)
        Fp = np.zeros(ptsize + 1)
        q = 0.0 if qprd is None else float(qprd[p])
        rc = lib().ndo_pair_table_q(int(kinds[p]), np.ascontiguousarray(params[p]), q, ptrmin,
                                    ptsize, Pp, Fp, C.byref(a), C.byref(b))
        if rc != 0:
            raise ValueError(f"ndo_pair_table failed ({rc})")
        P[p], F[p] = Pp, Fp
        rc2[p], dr2[p] = a.value, b.value
    return P, F, rc2, dr2


def cell_grid(box, rcut, skin):
    m = np.zeros(3, dtype=np.int32)
    lib().ndo_cell_grid(np.ascontiguousarray(box, dtype=np.float64), rcut, skin, m)
    return m


def neimax_estimate(rcut, skin, npart, box):
    return lib().ndo_neimax_estimate(rcut, skin, npart, np.ascontiguousarray(box, dtype=np.float64))


def half_list(sys, skin, neimax=None, brute=False, mcell=None):
    """Reference half list as (counts[np], vnl[np, neimax+1]) with 0-based j."""
    rc = np.ascontiguousarray(sys.params[:, 0])
    rcut = float(rc.max())
    if mcell is None:
        mcell = cell_grid(sys.box, rcut, skin)
    if neimax is None:
        neimax = neimax_estimate(rcut, skin, sys.np, sys.box)
    vnl = np.zeros((sys.np, neimax + 1), dtype=np.int32)
    err = lib().ndo_build_half_list(sys.np, np.ascontiguousarray(sys.pos), sys.typ.astype(np.int32),
                                    np.ascontiguousarray(sys.box), np.asarray(mcell, np.int32),
                                    sys.ntypes, np.ascontiguousarray(sys.pairmat.T).ravel().astype(np.int32),
                                    rc, skin, sys.ptrmin, neimax, int(brute), vnl, None)
    return err, vnl, neimax


def pair_set_from_half(vnl) -> set:
    s = set()
    for i in range(vnl.shape[0]):
        for j in vnl[i, 1:1 + vnl[i, 0]]:
            a, b = (i, int(j)) if i < j else (int(j), i)
            s.add((a, b))
    return s


def neighbour_sets_from_half(vnl):
    """Full neighbour sets per atom (sorted int arrays) from a half list."""
    npart = vnl.shape[0]
    counts = vnl[:, 0]
    ii = np.repeat(np.arange(npart), counts)
    jj = np.concatenate([vnl[i, 1:1 + counts[i]] for i in range(npart)]) if npart else np.zeros(0, int)
    a = np.concatenate([ii, jj])
    b = np.concatenate([jj, ii])
    order = np.lexsort((b, a))
    a, b = a[order], b[order]
    starts = np.searchsorted(a, np.arange(npart + 1))
    return [b[starts[i]:starts[i + 1]] for i in range(npart)]


def asymmetric(sys) -> bool:
    """The reference's 'asym' flag: some PairMat(a,b) != PairMat(b,a)."""
    pm = np.ascontiguousarray(sys.pairmat.T).ravel().astype(np.int32)
    return bool(lib().ndo_asymmetric(sys.ntypes, pm))


def springs(sys) -> int:
    """Bit mask of the 'spring' pair ids (bit p: pair p+1), the reference's springpots."""
    return int(sum(1 << p for p, k in enumerate(sys.kinds) if int(k) == SPRING))


def compute(sys, vnl, neimax, tables, cv=True, cu=True, ipos=None, init_box=None):
    """The force routine 'dynamics' would call (src/nd.f90:155-163): compute2/compute3
    when spring pairs exist, else compute1 for asymmetric pair matrices, else compute0."""
    if springs(sys):
        return compute23(sys, vnl, neimax, tables, ipos, init_box, cv, cu)
    fn = compute1 if asymmetric(sys) else compute0
    return fn(sys, vnl, neimax, tables, cv, cu)


def compute23(sys, vnl, neimax, tables, ipos, init_box, cv=True, cu=True, asym=None):
    """compute2 (symmetric PairMat) or compute3 (asymmetric) with the initial positions
    ipos (np,3 scaled) and box init_box of initPositionSet."""
    P, F, rc2, dr2 = tables
    asym = asymmetric(sys) if asym is None else asym
    spring = np.array([1 if int(k) == SPRING else 0 for k in sys.kinds], np.int32)
    force = np.zeros((sys.np, 3))
    pe = C.c_double(0.0)
    vir = np.zeros(6)
    pene = np.zeros(sys.np)
    lib().ndo_compute23(int(asym), sys.np, np.ascontiguousarray(sys.pos), sys.typ.astype(np.int32),
                        np.ascontiguousarray(sys.box), np.ascontiguousarray(vnl), neimax,
                        sys.ntypes, np.ascontiguousarray(sys.pairmat.T).ravel().astype(np.int32),
                        np.ascontiguousarray(P), np.ascontiguousarray(F), rc2, dr2, sys.ptrmin,
                        sys.ptsize, spring, np.ascontiguousarray(sys.params, dtype=np.float64).ravel(),
                        np.ascontiguousarray(ipos, dtype=np.float64),
                        np.ascontiguousarray(init_box, dtype=np.float64), int(cv), int(cu),
                        force, C.byref(pe), vir, pene.ctypes.data_as(C.c_void_p))
    return force, pe.value, vir, pene


def compute1(sys, vnl, neimax, tables, cv=True, cu=True):
    return _compute(lib().ndo_compute1, sys, vnl, neimax, tables, cv, cu)


def compute0(sys, vnl, neimax, tables, cv=True, cu=True):
    return _compute(lib().ndo_compute0, sys, vnl, neimax, tables, cv, cu)


def _compute(fn, sys, vnl, neimax, tables, cv, cu):
    P, F, rc2, dr2 = tables
    force = np.zeros((sys.np, 3))
    pe = C.c_double(0.0)
    vir = np.zeros(6)
    pene = np.zeros(sys.np)
    fn(sys.np, np.ascontiguousarray(sys.pos), sys.typ.astype(np.int32),
                       np.ascontiguousarray(sys.box), np.ascontiguousarray(vnl), neimax,
                       sys.ntypes, np.ascontiguousarray(sys.pairmat.T).ravel().astype(np.int32),
                       np.ascontiguousarray(P), np.ascontiguousarray(F), rc2, dr2,
                       sys.ptrmin, sys.ptsize, int(cv), int(cu), force, C.byref(pe), vir,
                       pene.ctypes.data_as(C.c_void_p))
    return force, pe.value, vir, pene


def run(sys, nsteps, dt, skin=-1.0, out_steps=None, step0=0, neimax=0, renei=True,
        renei_freq=1, trigger=TRIGGER_COMPAT, brute=False, acc=None,
        isokinetic=None, isobaric=None, isoenergetic=None, box_vel=None, tables=None,
        ipos=None, init_box=None, holds=None):
    """ndo_run on a copy of sys. isokinetic=(ke_target, tau, freq); isobaric=((px,py,pz),
    tau, freq); isoenergetic=(tau, freq).  With spring pairs, ipos/init_box are the initial
    positions and box (default: pos and box at the start of the Run, as dynamics sets them).
    holds=(hold[np], bc_disp[nhold], bc_dir[nhold,3], bc_parm[nhold,3]) sets Hold/Disp.
    Returns (state dict, thermo list, err); the state holds ipos/init_box."""
    if tables is None:
        tables = pair_tables(sys.kinds, sys.params, sys.ptrmin, sys.ptsize, sys.qprd)
    P, F, rc2, dr2 = tables
    rp = RunParams()
    rp.nsteps, rp.dt = nsteps, dt
    rp.out_steps = out_steps if out_steps else max(1, nsteps)
    rp.step0, rp.skin, rp.neimax = step0, skin, neimax
    rp.renei, rp.renei_freq, rp.trigger, rp.brute = int(renei), renei_freq, trigger, int(brute)
    if isokinetic:
        rp.isokinetic, rp.ke_target, rp.ke_tau, rp.ke_freq = 1, isokinetic[0], isokinetic[1], isokinetic[2]
    else:
        rp.ke_freq = 1
    if isobaric:
        rp.isobaric = 1
        for a in range(3):
            rp.p_target[a] = isobaric[0][a]
        rp.p_tau, rp.p_freq = isobaric[1], isobaric[2]
    else:
        rp.p_freq = 1
    if isoenergetic:
        rp.isoenergetic, rp.e_tau, rp.e_freq = 1, isoenergetic[0], isoenergetic[1]
    else:
        rp.e_freq = 1
    if box_vel is not None:
        rp.box_load = 1
        for a in range(3):
            rp.box_vel[a] = box_vel[a]
    rp.springs = springs(sys)
    ipos_buf = np.zeros((sys.np, 3))
    if ipos is not None:
        ipos_buf[:] = ipos
        for a in range(3):
            rp.init_box[a] = init_box[a]
        rp.ipos_set = 1
    rp.ipos = ipos_buf.ctypes.data
    if holds is not None:
        hk = [np.ascontiguousarray(holds[0], dtype=np.int32),
              np.ascontiguousarray(holds[1], dtype=np.int32),
              np.ascontiguousarray(holds[2], dtype=np.int32).ravel(),
              np.ascontiguousarray(holds[3], dtype=np.float64).ravel()]
        rp.nhold = len(hk[1])
        rp.hold, rp.bc_disp, rp.bc_dir, rp.bc_parm = (a.ctypes.data for a in hk)
    pos = np.ascontiguousarray(sys.pos, dtype=np.float64).copy()
    vel = np.ascontiguousarray(sys.vel, dtype=np.float64).copy()
    accv = np.zeros_like(pos) if acc is None else np.ascontiguousarray(acc, dtype=np.float64).copy()
    box = np.ascontiguousarray(sys.box, dtype=np.float64).copy()
    force = np.zeros_like(pos)
    pene = np.zeros(sys.np)
    nrec_max = nsteps // rp.out_steps + 2
    recs = (Thermo * nrec_max)()
    nrec = C.c_int(0)
    nup = C.c_int(0)
    err = lib().ndo_run(C.byref(rp), sys.np, sys.ntypes, sys.npairs, sys.ptsize, sys.ptrmin,
                        np.ascontiguousarray(sys.masses, dtype=np.float64),
                        np.ascontiguousarray(sys.pairmat.T).ravel().astype(np.int32),
                        np.ascontiguousarray(sys.params, dtype=np.float64).ravel(),
                        np.ascontiguousarray(P), np.ascontiguousarray(F), rc2, dr2,
                        sys.typ.astype(np.int32), pos, vel, accv, box, force,
                        pene.ctypes.data_as(C.c_void_p), recs, nrec_max, C.byref(nrec),
                        C.byref(nup))
    thermo = [dict(step=r.step, pe=r.pe, ke=r.ke, eerr=r.eerr, pressure=r.pressure,
                   volume=r.volume, virial=list(r.virial), box=list(r.box))
              for r in recs[:nrec.value]]
    state = dict(pos=pos, vel=vel, acc=accv, box=box, force=force, pene=pene,
                 nupdates=nup.value, ipos=ipos_buf, init_box=np.array(rp.init_box[:]))
    return state, thermo, err
