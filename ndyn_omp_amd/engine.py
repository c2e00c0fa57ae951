"""Python host binding of libndyn_gpu.so (the C ABI in include/ndyn_gpu.h).

Mirrors the reference's engine layer (Ndyn src/nd.f90: dynamics, compute0,
integrate1/2, buildCellList/buildVerletList, CheckListUpdate, mkIsoKE/P/E) as methods of
`Simulation` (one system) and `BatchEngine` (many replicas per launch).  The HIP library
must be built (python -c "import __graft_entry__ as g; g.build()"); nothing here falls
back to CPU code.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import List, Optional, Sequence

import numpy as np

from . import workloads as W

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libndyn_gpu.so")

OK, ERR_NEIMAX, ERR_OUTSIDE, ERR_HIP, ERR_ARG, ERR_NOMEM, ERR_RANGE = 0, 1, 2, 3, 4, 5, 6
PATH_FP64, PATH_MIXED = 0, 1
TRIGGER_COMPAT, TRIGGER_SKIN = 0, 1

dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class NdgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ndg error {code}: {msg}")
        self.code = code


class RunParams(C.Structure):
    _fields_ = [("nsteps", C.c_int), ("dt", C.c_double), ("out_steps", C.c_int),
                ("skin", C.c_double), ("renei", C.c_int), ("renei_freq", C.c_int),
                ("trigger", C.c_int),
                ("isokinetic", C.c_int), ("ke_target", C.c_double), ("ke_tau", C.c_double),
                ("ke_freq", C.c_int),
                ("isobaric", C.c_int), ("p_target", C.c_double * 3), ("p_tau", C.c_double),
                ("p_freq", C.c_int),
                ("isoenergetic", C.c_int), ("e_tau", C.c_double), ("e_freq", C.c_int),
                ("box_load", C.c_int), ("box_vel", C.c_double * 3)]


class Thermo(C.Structure):
    _fields_ = [("step", C.c_int), ("pad_", C.c_int), ("pe", C.c_double), ("ke", C.c_double),
                ("eerr", C.c_double), ("pressure", C.c_double), ("volume", C.c_double),
                ("virial", C.c_double * 6), ("box", C.c_double * 3)]

    def as_dict(self):
        return dict(step=self.step, pe=self.pe, ke=self.ke, eerr=self.eerr,
                    pressure=self.pressure, volume=self.volume, virial=list(self.virial),
                    box=list(self.box))


_lib = None


def load_library():
    """Load the HIP engine; raises if it was not built (no CPU fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build the HIP engine first "
                           "(python -c 'import __graft_entry__ as g; g.build()')")
    L = C.CDLL(LIB_PATH)
    V = C.c_void_p
    P = C.POINTER
    sig = {
        "ndg_pair_table": [C.c_int, dp, C.c_double, C.c_int, dp, dp, P(C.c_double), P(C.c_double)],
        "ndg_pair_table_q": [C.c_int, dp, C.c_double, C.c_double, C.c_int, dp, dp, P(C.c_double),
                             P(C.c_double)],
        "ndg_create": [P(V), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int],
        "ndg_set_tables": [V, dp, dp, dp, dp, C.c_double, ip, dp, dp],
        "ndg_set_state": [V, dp, V, V, ip, dp],
        "ndg_set_pair_kinds": [V, ip],
        "ndg_init_positions": [V],
        "ndg_set_initial_positions": [V, dp, dp],
        "ndg_set_holds": [V, V, C.c_int, V, V, V],
        "ndg_get_state": [V, V, V, V, V, V, V],
        "ndg_set_box": [V, dp],
        "ndg_get_box": [V, dp],
        "ndg_build_lists": [V, C.c_double],
        "ndg_get_neighbours": [V, ip, V, C.c_int],
        "ndg_check_rebuild": [V, C.c_double, C.c_int, P(C.c_int)],
        "ndg_integrate1": [V, C.c_double],
        "ndg_compute": [V, C.c_int, C.c_int, P(C.c_double), dp],
        "ndg_integrate2": [V, C.c_double, C.c_int, P(C.c_double)],
        "ndg_scale_velocities": [V, C.c_double],
        "ndg_run_begin": [V, P(RunParams), C.c_int],
        "ndg_run_steps": [V, C.c_int, P(Thermo), C.c_int, P(C.c_int)],
        "ndg_run_end": [V, P(C.c_int), P(C.c_int)],
        "ndg_run_enqueue": [V, C.c_int],
        "ndg_run_wait": [V, P(Thermo), C.c_int, P(C.c_int)],
        "ndg_snapshot": [V],
        "ndg_snapshot_get": [V, V, V, V, V, V, P(C.c_int)],
        "ndg_destroy": [V],
        "ndg_engine_create": [P(V), C.c_int, C.c_size_t],
        "ndg_engine_destroy": [V],
        "ndg_engine_admit": [V, P(C.c_int), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, dp, dp,
                             dp, dp, C.c_double, ip, dp, dp, dp, V, ip, dp, P(RunParams)],
        "ndg_engine_admit_ex": [V, P(C.c_int), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                dp, dp, dp, dp, C.c_double, ip, dp, dp, dp, V, ip, dp,
                                P(RunParams), V, V, V, V, C.c_int, V, V, V],
        "ndg_engine_step": [V, C.c_int, P(C.c_int)],
        "ndg_engine_step_enqueue": [V, C.c_int],
        "ndg_engine_step_wait": [V, P(C.c_int)],
        "ndg_engine_status": [V, C.c_int, P(C.c_int), P(C.c_int), P(Thermo), P(C.c_int)],
        "ndg_engine_thermo": [V, C.c_int, P(Thermo), C.c_int, P(C.c_int)],
        "ndg_engine_fetch": [V, C.c_int, V, V, V, V],
        "ndg_engine_fetch_noalloc": [V, C.c_int, V, V, V, V],
        "ndg_engine_release": [V, C.c_int],
        "ndg_engine_config": [V, C.c_int, C.c_int],
        "ndg_engine_size_classes": [V, C.c_int, ip, ip],
        "ndg_engine_cell_blocks": [V, C.c_int, ip, ip, ip],
        "ndg_engine_timing": [V, C.c_int],
        "ndg_engine_kernel_time": [V, C.c_int, P(C.c_double), P(C.c_longlong), P(C.c_double),
                                   P(C.c_double)],
        "ndg_engine_list_stats": [V, P(C.c_double), P(C.c_longlong)],
        "ndg_engine_counters": [V, C.c_int, P(C.c_int), P(C.c_int), P(C.c_double)],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    L.ndg_copy_bandwidth.argtypes = [C.c_size_t, C.c_int, P(C.c_double)]
    L.ndg_copy_bandwidth.restype = C.c_int
    L.ndg_footprint.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    L.ndg_footprint.restype = C.c_size_t
    L.ndg_last_error.argtypes = []
    L.ndg_last_error.restype = C.c_char_p
    _lib = L
    return L


def _check(rc):
    if rc != OK:
        raise NdgError(rc, load_library().ndg_last_error().decode())


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _hold_arrays(np_atoms, hold, bc_disp, bc_dir, bc_parm):
    """The Hold/Disp arrays in the C layout, checked against the sizes the library reads
    (hold[np], bc_disp[nhold], bc_dir and bc_parm (nhold, 3))."""
    hk = [np.ascontiguousarray(hold, dtype=np.int32).ravel(),
          np.ascontiguousarray(bc_disp, dtype=np.int32).ravel(),
          np.ascontiguousarray(bc_dir, dtype=np.int32).reshape(-1),
          np.ascontiguousarray(bc_parm, dtype=np.float64).reshape(-1)]
    nhold = hk[1].shape[0]
    if hk[0].shape[0] != np_atoms:
        raise ValueError(f"hold has {hk[0].shape[0]} entries for {np_atoms} atoms")
    if hk[2].shape[0] != 3 * nhold or hk[3].shape[0] != 3 * nhold:
        raise ValueError(f"bc_dir and bc_parm must be ({nhold}, 3)")
    return hk


def pair_tables(kinds, params, ptrmin, ptsize, qprd=None):
    """Tables of every pair with the engine's builder: P, F (npairs, ptsize+1), rc2, dr2."""
    L = load_library()
    kinds = np.asarray(kinds, dtype=np.int32)
    params = np.ascontiguousarray(params, dtype=np.float64).reshape(-1, 4)
    npairs = kinds.shape[0]
    P = np.zeros((npairs, ptsize + 1))
    F = np.zeros((npairs, ptsize + 1))
    rc2 = np.zeros(npairs)
    dr2 = np.zeros(npairs)
    for p in range(npairs):
        a, b = C.c_double(), C.c_double()
        Pp, Fp = np.zeros(ptsize + 1), np.zeros(ptsize + 1)
        q = 0.0 if qprd is None else float(qprd[p])
        _check(L.ndg_pair_table_q(int(kinds[p]), np.ascontiguousarray(params[p]), q,
                                  float(ptrmin), int(ptsize), Pp, Fp, C.byref(a), C.byref(b)))
        P[p], F[p], rc2[p], dr2[p] = Pp, Fp, a.value, b.value
    return P, F, rc2, dr2


def run_params(nsteps, dt, out_steps=None, skin=-1.0, renei=True, renei_freq=1,
               trigger=TRIGGER_COMPAT, isokinetic=None, isobaric=None, isoenergetic=None,
               box_vel=None) -> RunParams:
    """isokinetic=(KE_target per atom, tau, freq); isobaric=((px,py,pz), tau, freq);
    isoenergetic=(tau, freq) -- the deck's Isokinetic / Isobaric / Isoenergetic."""
    rp = RunParams()
    rp.nsteps, rp.dt = int(nsteps), float(dt)
    rp.out_steps = int(out_steps) if out_steps else max(1, int(nsteps))
    rp.skin, rp.renei, rp.renei_freq, rp.trigger = float(skin), int(renei), int(renei_freq), int(trigger)
    rp.ke_freq = rp.p_freq = rp.e_freq = 1
    if isokinetic:
        rp.isokinetic, rp.ke_target, rp.ke_tau, rp.ke_freq = 1, isokinetic[0], isokinetic[1], int(isokinetic[2])
    if isobaric:
        rp.isobaric = 1
        for a in range(3):
            rp.p_target[a] = isobaric[0][a]
        rp.p_tau, rp.p_freq = isobaric[1], int(isobaric[2])
    if isoenergetic:
        rp.isoenergetic, rp.e_tau, rp.e_freq = 1, isoenergetic[0], int(isoenergetic[1])
    if box_vel is not None:
        rp.box_load = 1
        for a in range(3):
            rp.box_vel[a] = box_vel[a]
    return rp


class Simulation:
    """One system on the device (the ndg_t handle)."""

    def __init__(self, system: W.System, path=PATH_MIXED, neimax=0, acc=None):
        L = load_library()
        self.sys = system
        self.path = path
        h = C.c_void_p()
        _check(L.ndg_create(C.byref(h), system.np, system.ntypes, system.npairs, system.ptsize,
                            int(neimax), int(path)))
        self.h = h
        self.tables = pair_tables(system.kinds, system.params, system.ptrmin, system.ptsize,
                                  system.qprd)
        P, F, rc2, dr2 = self.tables
        _check(L.ndg_set_tables(h, np.ascontiguousarray(P), np.ascontiguousarray(F), rc2, dr2,
                                float(system.ptrmin),
                                np.ascontiguousarray(system.pairmat.T).ravel().astype(np.int32),
                                np.ascontiguousarray(system.params, dtype=np.float64).ravel(),
                                np.ascontiguousarray(system.masses, dtype=np.float64)))
        if system.springs:
            _check(L.ndg_set_pair_kinds(h, np.ascontiguousarray(system.kinds, dtype=np.int32)))
        pos = np.ascontiguousarray(system.pos, dtype=np.float64)
        vel = np.ascontiguousarray(system.vel, dtype=np.float64)
        a = None if acc is None else np.ascontiguousarray(acc, dtype=np.float64)
        self._keep = (pos, vel, a)
        _check(L.ndg_set_state(h, pos, _ptr(vel), _ptr(a), system.typ.astype(np.int32),
                               np.ascontiguousarray(system.box, dtype=np.float64)))
        if system.springs and system.ipos is not None:
            self.set_initial_positions(system.ipos, system.init_box)

    def set_initial_positions(self, ipos, init_box):
        """ipos (np,3 scaled) and initBox of the spring pairs (ndg_set_initial_positions)."""
        _check(load_library().ndg_set_initial_positions(
            self.h, np.ascontiguousarray(ipos, dtype=np.float64),
            np.ascontiguousarray(init_box, dtype=np.float64)))

    def set_holds(self, hold, bc_disp, bc_dir, bc_parm):
        """Hold/Disp boundary conditions (ndg_set_holds): hold group per atom (0: none) and,
        per group, bcTyp 'd' flag, constrained axes and scaled displacement per step."""
        hk = _hold_arrays(self.sys.np, hold, bc_disp, bc_dir, bc_parm)
        self._holds = hk
        _check(load_library().ndg_set_holds(self.h, _ptr(hk[0]), len(hk[1]), _ptr(hk[1]),
                                            _ptr(hk[2]), _ptr(hk[3])))

    def init_positions(self):
        """initPositionSet: the current positions and box become the initial ones."""
        _check(load_library().ndg_init_positions(self.h))

    def close(self):
        if getattr(self, "h", None):
            load_library().ndg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- reference routines ---------------------------------------------------------
    def build_lists(self, skin):
        _check(load_library().ndg_build_lists(self.h, float(skin)))

    def neighbours(self, cap=None):
        n = self.sys.np
        counts = np.zeros(n, np.int32)
        _check(load_library().ndg_get_neighbours(self.h, counts, None, 0))
        cap = int(counts.max()) if cap is None else cap
        nbr = np.full((n, max(cap, 1)), -1, np.int32)
        _check(load_library().ndg_get_neighbours(self.h, counts, _ptr(nbr), nbr.shape[1]))
        return counts, nbr

    def integrate1(self, dt):
        _check(load_library().ndg_integrate1(self.h, float(dt)))

    def compute(self, cv=True, cu=True):
        pe = C.c_double(0.0)
        vir = np.zeros(6)
        _check(load_library().ndg_compute(self.h, int(cv), int(cu), C.byref(pe), vir))
        return pe.value, vir

    def integrate2(self, dt, ck=True):
        ke = C.c_double(0.0)
        _check(load_library().ndg_integrate2(self.h, float(dt), int(ck), C.byref(ke)))
        return ke.value

    def scale_velocities(self, lam):
        _check(load_library().ndg_scale_velocities(self.h, float(lam)))

    def check_rebuild(self, skin, trigger=TRIGGER_COMPAT):
        r = C.c_int(0)
        _check(load_library().ndg_check_rebuild(self.h, float(skin), int(trigger), C.byref(r)))
        return bool(r.value)

    def state(self):
        n = self.sys.np
        pos, vel, acc, force = (np.zeros((n, 3)) for _ in range(4))
        pene, box = np.zeros(n), np.zeros(3)
        _check(load_library().ndg_get_state(self.h, _ptr(pos), _ptr(vel), _ptr(acc), _ptr(force),
                                             _ptr(pene), _ptr(box)))
        return dict(pos=pos, vel=vel, acc=acc, force=force, pene=pene, box=box)

    def set_box(self, box):
        _check(load_library().ndg_set_box(self.h, np.ascontiguousarray(box, dtype=np.float64)))

    def snapshot(self):
        """ndg_snapshot + ndg_snapshot_get: the output state in caller order."""
        L = load_library()
        n = self.sys.np
        pos, vel, force = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        pene, box, st = np.zeros(n), np.zeros(3), C.c_int(0)
        _check(L.ndg_snapshot(self.h))
        _check(L.ndg_snapshot_get(self.h, _ptr(pos), _ptr(vel), _ptr(force), _ptr(pene),
                                  _ptr(box), C.byref(st)))
        return dict(step=st.value, pos=pos, vel=vel, force=force, pene=pene, box=box)

    # --- the dynamics loop ----------------------------------------------------------------
    def run(self, rp: RunParams, step0=0, chunk=None):
        L = load_library()
        _check(L.ndg_run_begin(self.h, C.byref(rp), int(step0)))
        recs = []
        left = rp.nsteps
        cap = rp.nsteps // max(1, rp.out_steps) + 2
        buf = (Thermo * cap)()
        chunk = chunk or rp.nsteps
        while left > 0:
            n = min(chunk, left)
            got = C.c_int(0)
            _check(L.ndg_run_steps(self.h, n, buf, cap, C.byref(got)))
            recs.extend(buf[k].as_dict() for k in range(got.value))
            left -= n
        nup, step = C.c_int(0), C.c_int(0)
        _check(L.ndg_run_end(self.h, C.byref(nup), C.byref(step)))
        return recs, nup.value, step.value

    def run_frames(self, rp: RunParams, frame_every, step0=0):
        """The Fortran driver's output loop: chunks that end at multiples of frame_every,
        each frame copied on the device behind its chunk (ndg_snapshot) and read after the
        next chunk is queued (ndg_run_enqueue / ndg_run_wait / ndg_snapshot_get)."""
        L = load_library()
        _check(L.ndg_run_begin(self.h, C.byref(rp), int(step0)))
        recs, frames = [], []
        cap = rp.nsteps // max(1, rp.out_steps) + 2
        buf = (Thermo * cap)()
        left, cur = rp.nsteps, step0
        chunk = min(left, frame_every - cur % frame_every)
        _check(L.ndg_run_enqueue(self.h, chunk))
        while left > 0:
            frame = (cur + chunk) % frame_every == 0
            if frame:
                _check(L.ndg_snapshot(self.h))
            got = C.c_int(0)
            _check(L.ndg_run_wait(self.h, buf, cap, C.byref(got)))
            recs.extend(buf[k].as_dict() for k in range(got.value))
            cur += chunk
            left -= chunk
            if left > 0:
                chunk = min(left, frame_every - cur % frame_every)
                _check(L.ndg_run_enqueue(self.h, chunk))
            if frame:
                n = self.sys.np
                pos, vel, force = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
                pene, box, st = np.zeros(n), np.zeros(3), C.c_int(0)
                _check(L.ndg_snapshot_get(self.h, _ptr(pos), _ptr(vel), _ptr(force), _ptr(pene),
                                          _ptr(box), C.byref(st)))
                frames.append(dict(step=st.value, pos=pos, vel=vel, force=force, pene=pene, box=box))
        nup, step = C.c_int(0), C.c_int(0)
        _check(L.ndg_run_end(self.h, C.byref(nup), C.byref(step)))
        return recs, frames, nup.value, step.value


class BatchEngine:
    """Many independent systems advanced per launch (ndg_engine_t)."""

    def __init__(self, path=PATH_MIXED, hbm_budget=0, tpa=0, graph_steps=0):
        L = load_library()
        e = C.c_void_p()
        _check(L.ndg_engine_create(C.byref(e), int(path), C.c_size_t(hbm_budget)))
        self.e = e
        self.path = path
        self._tables = {}
        _check(L.ndg_engine_config(e, int(tpa), int(graph_steps)))

    def close(self):
        if getattr(self, "e", None):
            load_library().ndg_engine_destroy(self.e)
            self.e = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def config(self, tpa=0, graph_steps=0):
        _check(load_library().ndg_engine_config(self.e, int(tpa), int(graph_steps)))

    def size_classes(self, classes):
        """Threads per atom by size class for systems admitted from now on:
        [(np_max, tpa), ...] with np_max ascending."""
        npm = np.array([c[0] for c in classes], dtype=np.int32)
        tpa = np.array([c[1] for c in classes], dtype=np.int32)
        _check(load_library().ndg_engine_size_classes(self.e, len(classes), npm, tpa))

    def cell_blocks(self, classes):
        """Cell-block shape by size class for systems admitted from now on:
        [(np_max, by, bz), ...] with np_max ascending."""
        npm = np.array([c[0] for c in classes], dtype=np.int32)
        by = np.array([c[1] for c in classes], dtype=np.int32)
        bz = np.array([c[2] for c in classes], dtype=np.int32)
        _check(load_library().ndg_engine_cell_blocks(self.e, len(classes), npm, by, bz))

    def admit(self, s: W.System, rp: RunParams, neimax=0, raise_on_full=True, holds=None):
        """Admit one system; holds = (hold, bc_disp, bc_dir, bc_parm) as Simulation.set_holds.
        Spring systems carry their pair kinds and initial positions (None: the admitted
        state, as initPositionSet)."""
        q = None if s.qprd is None else np.asarray(s.qprd, dtype=np.float64)
        key = (s.kinds.tobytes(), s.params.tobytes(), s.ptrmin, s.ptsize,
               None if q is None else q.tobytes())
        if key not in self._tables:
            self._tables[key] = pair_tables(s.kinds, s.params, s.ptrmin, s.ptsize, q)
        P, F, rc2, dr2 = self._tables[key]
        slot = C.c_int(-1)
        vel = np.ascontiguousarray(s.vel, dtype=np.float64)
        args = (self.e, C.byref(slot), s.np, s.ntypes, s.npairs, s.ptsize, int(neimax),
                np.ascontiguousarray(P), np.ascontiguousarray(F), rc2, dr2, float(s.ptrmin),
                np.ascontiguousarray(s.pairmat.T).ravel().astype(np.int32),
                np.ascontiguousarray(s.params, dtype=np.float64).ravel(),
                np.ascontiguousarray(s.masses, dtype=np.float64),
                np.ascontiguousarray(s.pos, dtype=np.float64), _ptr(vel), s.typ.astype(np.int32),
                np.ascontiguousarray(s.box, dtype=np.float64), C.byref(rp))
        if s.springs or holds is not None:
            kinds = ipos = ibox = None
            if s.springs:
                kinds = np.ascontiguousarray(s.kinds, dtype=np.int32)
                if s.ipos is not None:
                    ipos = np.ascontiguousarray(s.ipos, dtype=np.float64)
                    ibox = np.ascontiguousarray(s.init_box, dtype=np.float64)
            hk = [None] * 4
            if holds is not None:
                hk = _hold_arrays(s.np, *holds)
            nhold = 0 if holds is None else len(hk[1])
            rc = load_library().ndg_engine_admit_ex(
                *args, _ptr(kinds), _ptr(ipos), _ptr(ibox), _ptr(hk[0]), nhold, _ptr(hk[1]),
                _ptr(hk[2]), _ptr(hk[3]))
        else:
            rc = load_library().ndg_engine_admit(*args)
        if rc == ERR_NOMEM and not raise_on_full:
            return None
        _check(rc)
        return slot.value

    def step(self, nsteps):
        nd = C.c_int(0)
        _check(load_library().ndg_engine_step(self.e, int(nsteps), C.byref(nd)))
        return nd.value

    def step_enqueue(self, nsteps):
        """Queue nsteps (<= 512) batched steps and return; admissions may follow while
        they run (they join the next steps)."""
        _check(load_library().ndg_engine_step_enqueue(self.e, int(nsteps)))

    def step_wait(self):
        nd = C.c_int(0)
        _check(load_library().ndg_engine_step_wait(self.e, C.byref(nd)))
        return nd.value

    def status(self, slot):
        sd, fin, err = C.c_int(0), C.c_int(0), C.c_int(0)
        last = Thermo()
        _check(load_library().ndg_engine_status(self.e, int(slot), C.byref(sd), C.byref(fin),
                                                C.byref(last), C.byref(err)))
        return dict(steps=sd.value, finished=bool(fin.value), err=err.value, last=last.as_dict())

    def thermo(self, slot, max_rec=4096):
        buf = (Thermo * max_rec)()
        n = C.c_int(0)
        _check(load_library().ndg_engine_thermo(self.e, int(slot), buf, max_rec, C.byref(n)))
        return [buf[k].as_dict() for k in range(n.value)]

    def fetch(self, slot, n):
        pos, vel, force = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        box = np.zeros(3)
        _check(load_library().ndg_engine_fetch(self.e, int(slot), _ptr(pos), _ptr(vel),
                                               _ptr(force), _ptr(box)))
        return dict(pos=pos, vel=vel, force=force, box=box)

    def fetch_noalloc(self, slot, n):
        """fetch through ndg_engine_fetch_noalloc (no allocation; at most 262144 atoms)."""
        pos, vel, force = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        box = np.zeros(3)
        _check(load_library().ndg_engine_fetch_noalloc(self.e, int(slot), _ptr(pos), _ptr(vel),
                                                       _ptr(force), _ptr(box)))
        return dict(pos=pos, vel=vel, force=force, box=box)

    def release(self, slot):
        _check(load_library().ndg_engine_release(self.e, int(slot)))

    def timing(self, on=True):
        _check(load_library().ndg_engine_timing(self.e, int(on)))

    def kernel_time(self, reset=False):
        ms, cnt = C.c_double(0), C.c_longlong(0)
        pairs, atoms = C.c_double(0), C.c_double(0)
        _check(load_library().ndg_engine_kernel_time(self.e, int(reset), C.byref(ms), C.byref(cnt),
                                                     C.byref(pairs), C.byref(atoms)))
        return dict(avg_ms=ms.value, count=cnt.value, list_entries=pairs.value, atoms=atoms.value)

    def counters(self, slot):
        nup, step, mean = C.c_int(0), C.c_int(0), C.c_double(0)
        _check(load_library().ndg_engine_counters(self.e, int(slot), C.byref(nup), C.byref(step),
                                                  C.byref(mean)))
        return dict(nupdates=nup.value, step=step.value, mean_nbr=mean.value)

    def list_stats(self):
        m, a = C.c_double(0), C.c_longlong(0)
        _check(load_library().ndg_engine_list_stats(self.e, C.byref(m), C.byref(a)))
        return m.value, a.value


def footprint(s: W.System, neimax: int) -> int:
    return int(load_library().ndg_footprint(s.np, s.ntypes, s.npairs, s.ptsize, int(neimax)))


def copy_bandwidth(nbytes=1 << 30, iters=20) -> float:
    """Measured device copy bandwidth in GB/s (read + write bytes)."""
    g = C.c_double(0)
    _check(load_library().ndg_copy_bandwidth(C.c_size_t(nbytes), int(iters), C.byref(g)))
    return g.value
