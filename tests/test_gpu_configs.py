"""The benchmark configurations themselves against the oracle (GPU).

- c3, the 512-replica mix (atoms log-uniform in [500, 20000], Morse Cu or Buckingham alloy,
  100-1200 K), advanced by the batched mixed engine with the saved autotuner table (skin per
  size class), the skin trigger and NVE, for 100 steps.  Eight replicas spread over the
  size range are compared with oracle runs of the same settings: the FP64 path (batched
  engine, one system per launch) rebuilds its lists at the same steps and ends within
  1e-9 A of the oracle; the mixed path rebuilds within one rebuild of the oracle, ends within
  1e-6 A (max) of it (measured: the same rebuild steps, 2e-8 A), and its final forces are within 1e-5 RMS relative of the oracle's
  compute0 on a fresh list at the same positions (so every rebuilt list was complete).
- c2's operating mode (src/nd.f90:1183-1213, :1447-1499 with the skin trigger): Berendsen
  thermostat and the barostat every step on a hot Cu block, skin trigger; the box grows until
  the cell grid changes (3 -> 4 cells per axis), so rebuilds re-grid.  FP64 path: the
  oracle's rebuild count, box within 1e-10, positions within 1e-9 A, thermo PE within 1e-10;
  mixed path: rebuilds within one, box within 1e-6, positions within 1e-6 A, PE within 1e-5.
  After the run the device list holds every pair inside the cutoff, and the last forces
  equal the oracle's compute0 on a fresh list (FP64 1e-12, mixed 1e-5 RMS relative).
- Full size: c1 (256,000 LJ argon atoms, the configuration's own 0.005 A jitter) and c2
  (2,048,000 Cu atoms, jittered 0.1 A): neighbour sets bit-exact against the oracle's half
  list (per atom: count and two 64-bit multiset hashes of the neighbour indices, plus exact
  rows for the first 50,000 atoms), forces against the oracle's compute0 (FP64 1e-12, mixed
  1e-5 RMS relative), PE within 1e-12 / 1e-6.
"""
import os

import numpy as np
import pytest

from conftest import ROOT
from ndyn_omp_amd import workloads as W

pytestmark = pytest.mark.gpu
G = pytest.importorskip("ndyn_omp_amd.engine")


def rms_rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def max_disp_a(pos_a, pos_b, box):
    """Largest distance (A) between two scaled position sets, minimum image."""
    d = pos_a - pos_b
    d -= np.round(d)
    return float(np.abs(d * box[None, :]).max())


def brute_for(s, skin):
    rc = float(s.params[:, 0].max())
    return min(int(L / (rc + skin)) for L in s.box) < 3


def fresh_forces(oracle, s, pos, box, skin=0.5):
    s2 = s.copy()
    s2.pos, s2.box = np.ascontiguousarray(pos), np.ascontiguousarray(box, dtype=np.float64)
    tables = oracle.pair_tables(s.kinds, s.params, s.ptrmin, s.ptsize)
    _, vnl, nm = oracle.half_list(s2, skin, brute=brute_for(s2, skin))
    return oracle.compute0(s2, vnl, nm, tables)


# ---------------------------------------------------------------- c3: the replica mix

C3_STEPS = 100


def c3_table():
    from ndyn_omp_amd import autotune as AT
    with open(os.path.join(ROOT, "ndyn_omp_amd", "tuning", "c3.json")) as f:
        return AT.TuneTable.from_json(f.read())


def c3_rp(skin):
    return G.run_params(C3_STEPS, W.DT_1FS, out_steps=C3_STEPS, skin=skin,
                        trigger=G.TRIGGER_SKIN)


def c3_sample(systems):
    """Eight replicas spread over the size range (both kinds among them)."""
    order = np.argsort([s.np for s in systems], kind="stable")
    pick = [int(order[k]) for k in np.linspace(0, len(order) - 1, 8).round().astype(int)]
    kinds = {int(systems[r].npairs) for r in pick}
    assert kinds == {1, 3}, "the sample holds Morse and Buckingham replicas"
    return pick


@pytest.fixture(scope="module")
def c3_oracle(oracle):
    systems = W.replica_mix(512)
    table = c3_table()
    pick = c3_sample(systems)
    ref = {}
    for r in pick:
        s = systems[r]
        skin = table.skin_for(s.np, 0.5)
        st, _, err = oracle.run(s, C3_STEPS, W.DT_1FS, skin=skin, out_steps=C3_STEPS,
                                trigger=oracle.TRIGGER_SKIN, brute=brute_for(s, skin))
        assert err == 0
        ref[r] = st
    return systems, table, pick, ref


def test_c3_replica_mix_mixed_path(c3_oracle, oracle):
    systems, table, pick, ref = c3_oracle
    eng = G.BatchEngine(path=G.PATH_MIXED)
    table.apply(eng)
    slots = [eng.admit(s, c3_rp(table.skin_for(s.np, 0.5))) for s in systems]
    print(f"case c3 mix mixed path: {len(systems)} replicas, "
          f"{sum(s.np for s in systems)} atoms, {C3_STEPS} steps", flush=True)
    eng.step(C3_STEPS)
    total_nup = 0
    for r in pick:
        s = systems[r]
        st = eng.fetch(slots[r], s.np)
        nup = eng.counters(slots[r])["nupdates"]
        total_nup += ref[r]["nupdates"]
        dx = max_disp_a(st["pos"], ref[r]["pos"], s.box)
        f_ref, _, _, _ = fresh_forces(oracle, s, st["pos"], st["box"])
        err = rms_rel(st["force"], f_ref)
        print(f"  replica {r} np={s.np} pairs={s.npairs} rebuilds {nup}/{ref[r]['nupdates']} "
              f"max|dx| {dx:.2e} A force err {err:.2e}", flush=True)
        assert abs(nup - ref[r]["nupdates"]) <= 1
        assert dx < 1e-6
        assert err <= 1e-5
    eng.close()
    assert total_nup > 0, "the sample rebuilds its lists during the run"


def test_c3_replica_mix_fp64_path(c3_oracle, oracle):
    systems, table, pick, ref = c3_oracle
    eng = G.BatchEngine(path=G.PATH_FP64)
    slots = {r: eng.admit(systems[r], c3_rp(table.skin_for(systems[r].np, 0.5))) for r in pick}
    print(f"case c3 sample fp64 path: {len(pick)} replicas", flush=True)
    eng.step(C3_STEPS)
    for r in pick:
        s = systems[r]
        st = eng.fetch(slots[r], s.np)
        nup = eng.counters(slots[r])["nupdates"]
        dx = max_disp_a(st["pos"], ref[r]["pos"], s.box)
        print(f"  replica {r} np={s.np} rebuilds {nup}/{ref[r]['nupdates']} max|dx| {dx:.2e} A",
              flush=True)
        assert nup == ref[r]["nupdates"]
        assert dx < 1e-9
        np.testing.assert_allclose(st["vel"], ref[r]["vel"], rtol=0,
                                   atol=1e-9 * np.abs(ref[r]["vel"]).max())
    eng.close()


# ---------------------------------------------------------------- c2's operating mode

C2M_STEPS = 300
C2M_SKIN = 0.5


def c2_mode_system():
    return W.cu_morse(8, temp_k=600.0, seed=31)


def c2_mode_kwargs():
    return dict(isokinetic=(1.5 * W.KB_EV * 600.0, 100 * W.DT_1FS, 1),
                isobaric=((0.0, 0.0, 0.0), 100 * W.DT_1FS, 1))


@pytest.fixture(scope="module")
def c2_mode_oracle(oracle):
    s = c2_mode_system()
    st, thermo, err = oracle.run(s, C2M_STEPS, W.DT_1FS, skin=C2M_SKIN, out_steps=50,
                                 trigger=oracle.TRIGGER_SKIN, **c2_mode_kwargs())
    assert err == 0
    rc = 7.0 + C2M_SKIN
    m0 = int(s.box[0] / rc)
    m1 = int(st["box"][0] / rc)
    assert (m0, m1) == (3, 4), "the box grows through a change of the cell grid"
    assert st["nupdates"] >= 5
    return s, st, thermo


@pytest.mark.parametrize("path_name,path", [("fp64", 0), ("mixed", 1)])
def test_c2_mode_npt_skin_trigger(c2_mode_oracle, oracle, path_name, path):
    s, ref, thermo = c2_mode_oracle
    print(f"case c2 mode NPT skin trigger {path_name}", flush=True)
    sim = G.Simulation(s, path=path)
    recs, nup, step = sim.run(G.run_params(C2M_STEPS, W.DT_1FS, out_steps=50, skin=C2M_SKIN,
                                           trigger=G.TRIGGER_SKIN, **c2_mode_kwargs()))
    st = sim.state()
    dx = max_disp_a(st["pos"], ref["pos"], ref["box"])
    print(f"  rebuilds {nup}/{ref['nupdates']} box {st['box']} / {ref['box']} max|dx| {dx:.2e}",
          flush=True)
    assert step == C2M_STEPS and len(recs) == len(thermo)
    fp64 = path == G.PATH_FP64
    if fp64:
        assert nup == ref["nupdates"]
    else:
        assert abs(nup - ref["nupdates"]) <= 1
    np.testing.assert_allclose(st["box"], ref["box"], rtol=1e-10 if fp64 else 1e-6)
    assert dx < (1e-9 if fp64 else 1e-6)
    for rec, row in zip(recs, thermo):
        assert rec["step"] == row["step"]
        assert rec["pe"] == pytest.approx(row["pe"], rel=1e-10 if fp64 else 1e-5)
        np.testing.assert_allclose(rec["box"], row["box"], rtol=1e-10 if fp64 else 1e-6)
    # the list after the last rebuild holds every pair inside the cutoff at the end
    s_end = s.copy()
    s_end.pos, s_end.box = st["pos"], st["box"]
    _, vnl0, _ = oracle.half_list(s_end, 0.0)
    counts, nbr = sim.neighbours()
    for i in range(s.np):
        need = vnl0[i, 1:1 + vnl0[i, 0]]
        have = nbr[i, :counts[i]]
        assert np.isin(need, have).all(), i
    # the last forces were evaluated in the box of the last thermo record (the barostat
    # changes the box after the force call, src/nd.f90:176)
    f_ref, _, _, _ = fresh_forces(oracle, s, st["pos"], np.array(recs[-1]["box"]))
    err = rms_rel(st["force"], f_ref)
    print(f"  final force err {err:.2e}", flush=True)
    assert err <= (1e-12 if fp64 else 1e-5)
    sim.close()


# ---------------------------------------------------------------- full-size c1 and c2

def _mix(x, salt):
    return W._splitmix64(x.astype(np.uint64) ^ np.uint64(salt))


def full_set_digest_from_half(vnl):
    """Per atom: size and two 64-bit sums of hashes of its full neighbour set, from a
    reference half list (row 0 = count)."""
    n = vnl.shape[0]
    cnt = vnl[:, 0].astype(np.int64)
    width = vnl.shape[1] - 1
    mask = np.arange(width)[None, :] < cnt[:, None]
    jj = vnl[:, 1:][mask].astype(np.int64)
    ii = np.repeat(np.arange(n, dtype=np.int64), cnt)
    full = cnt + np.bincount(jj, minlength=n)
    out = [full]
    with np.errstate(over="ignore"):
        for salt in (0x1234567, 0xABCDEF01):
            h = np.zeros(n, np.uint64)
            hj = np.zeros(vnl.shape[:1] + (width,), np.uint64)
            hj[mask] = _mix(jj, salt)
            h += hj.sum(axis=1, dtype=np.uint64)
            del hj
            np.add.at(h, jj, _mix(ii, salt))
            out.append(h)
    return out


def full_set_digest_from_rows(counts, nbr):
    n = counts.shape[0]
    mask = np.arange(nbr.shape[1])[None, :] < counts[:, None]
    out = [counts.astype(np.int64)]
    with np.errstate(over="ignore"):
        for salt in (0x1234567, 0xABCDEF01):
            h = np.zeros(nbr.shape, np.uint64)
            h[mask] = _mix(nbr[mask].astype(np.int64), salt)
            out.append(h.sum(axis=1, dtype=np.uint64))
    assert all(o.shape == (n,) for o in out)
    return out


def check_full_size(oracle, s, skin, path, nexact=50000):
    tables = oracle.pair_tables(s.kinds, s.params, s.ptrmin, s.ptsize)
    _, vnl, nm = oracle.half_list(s, skin)
    f_ref, pe_ref, vir_ref, _ = oracle.compute0(s, vnl, nm, tables)
    print(f"  oracle: {s.np} atoms, {int(vnl[:, 0].sum())} half pairs", flush=True)
    sim = G.Simulation(s, path=path)
    sim.build_lists(skin)
    pe, vir = sim.compute(True, True)
    f = sim.state()["force"]
    err = rms_rel(f, f_ref)
    print(f"  path {path}: force err {err:.2e}, pe {pe:.10f} / {pe_ref:.10f}", flush=True)
    if path == G.PATH_FP64:
        assert err <= 1e-12
        assert pe == pytest.approx(pe_ref, rel=1e-12)
    else:
        assert err <= 1e-5
        assert pe == pytest.approx(pe_ref, rel=1e-6)
    counts, nbr = sim.neighbours()
    sim.close()
    got = full_set_digest_from_rows(counts, nbr)
    want = full_set_digest_from_half(vnl)
    for k, (a, b) in enumerate(zip(got, want)):
        bad = np.nonzero(a != b)[0]
        assert bad.size == 0, f"digest {k} differs at atoms {bad[:10]}"
    # exact rows of the first atoms: their half-list rows and the rows that name them
    cnt = vnl[:, 0]
    ii = np.repeat(np.arange(s.np, dtype=np.int32), cnt)
    jj = vnl[:, 1:][np.arange(vnl.shape[1] - 1)[None, :] < cnt[:, None]].astype(np.int32)
    k1, k2 = ii < nexact, jj < nexact
    a = np.concatenate([ii[k1], jj[k2]])
    b = np.concatenate([jj[k1], ii[k2]])
    order = np.lexsort((b, a))
    a, b = a[order], b[order]
    starts = np.searchsorted(a, np.arange(nexact + 1))
    for i in range(nexact):
        np.testing.assert_array_equal(nbr[i, :counts[i]], b[starts[i]:starts[i + 1]])


@pytest.mark.parametrize("path_name,path", [("fp64", 0), ("mixed", 1)])
def test_full_size_c1_lists_and_forces(oracle, path_name, path):
    s = W.ar_lj(40)
    print(f"case c1 full size {path_name}: {s.np} atoms", flush=True)
    check_full_size(oracle, s, 0.4, path)


@pytest.mark.parametrize("path_name,path", [("fp64", 0), ("mixed", 1)])
def test_full_size_c2_lists_and_forces(oracle, path_name, path):
    s = W.cu_morse(80, temp_k=600.0, seed=12347, jitter_a=0.1)
    print(f"case c2 full size {path_name}: {s.np} atoms", flush=True)
    check_full_size(oracle, s, 0.5, path)
