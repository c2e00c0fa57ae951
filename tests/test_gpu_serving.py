"""Serving mode: jobs served alone and mixed with others give the same result.

Four jobs of different size, potential, ensemble and length are served (a) each alone
and (b) together through one server whose HBM budget cannot hold all of them at once, with
an odd tick length so jobs finish mid-tick and queued jobs are admitted while others are
on the GPU.  Bar: final positions, velocities and the last thermo record identical bit
for bit (per-replica kernels and fixed-order reductions make a job independent of its
neighbours in the batch).
"""
import numpy as np
import pytest

from ndyn_omp_amd import workloads as W

pytestmark = pytest.mark.gpu
G = pytest.importorskip("ndyn_omp_amd.engine")
S = pytest.importorskip("ndyn_omp_amd.serve")


def make_jobs():
    specs = [
        (W.cu_morse(6, temp_k=700.0, seed=31), 37, "nvt", 700.0),       # small box: brute list
        (W.buck_alloy((5, 5, 5), temp_k=900.0, seed=32), 53, "nve", 900.0),
        (W.cu_morse(8, temp_k=300.0, seed=33), 80, "nve", 300.0),
        (W.cu_morse(7, temp_k=1100.0, seed=34), 45, "nvt", 1100.0),
    ]
    return [S.Job(k, s, n, ens, t, arrival=0.0) for k, (s, n, ens, t) in enumerate(specs)]


def test_served_alone_equals_served_in_mix():
    alone = {}
    for job in make_jobs():
        print(f"case alone job {job.jid} np={job.system.np}", flush=True)
        srv = S.Server(tick_steps=16)
        srv.run([job], realtime=False)
        alone[job.jid] = job
        srv.close()
    jobs = make_jobs()
    fp = sorted(G.footprint(j.system, 200) for j in jobs)
    budget = fp[-1] + fp[-2] + (8 << 20)        # at most about two jobs resident at a time
    print(f"case mixed budget={budget}", flush=True)
    srv = S.Server(hbm_budget=budget, tick_steps=7)
    stats = srv.run(jobs, realtime=False)
    srv.close()
    assert stats["jobs"] == 4
    admitted = sorted(j.admitted for j in jobs)
    assert admitted[-1] > admitted[0]            # some job waited for memory
    for j in jobs:
        a = alone[j.jid]
        np.testing.assert_array_equal(j.result["pos"], a.result["pos"])
        np.testing.assert_array_equal(j.result["vel"], a.result["vel"])
        np.testing.assert_array_equal(j.result["box"], a.result["box"])
        assert j.thermo[-1] == a.thermo[-1]


def test_serving_stream_latency_accounting():
    """A short Poisson stream: every job completes, latency >= its own service time."""
    stream = W.serving_jobs(12, seed=7, size_scale=0.02, step_scale=0.004, rate_hz=50.0)
    jobs = S.jobs_from_stream(stream)
    srv = S.Server(tick_steps=25)
    stats = srv.run(jobs, realtime=True)
    srv.close()
    assert stats["jobs"] == 12
    for j in jobs:
        assert j.finished >= j.admitted >= j.arrival - 1e-3
        assert j.thermo and j.thermo[-1]["step"] == j.nsteps
    assert stats["p99_latency_s"] >= stats["p50_latency_s"] > 0


def test_config_stream_alone_equals_mix():
    """The first 16 jobs of the configuration's own stream (W.serving_jobs, seed 12349:
    sizes log-uniform in [1e3, 2e5] atoms, 1e3-5e4 steps, Morse Cu or Buckingham alloy,
    NVE/NVT 100-1200 K; 1.34e10 atom-steps in all) served by continuous batching under an
    HBM budget that holds about a third of them at once, so most jobs are admitted while
    others are on the GPU.  Every job completes its Run; five of them (1176 to 27360 atoms)
    served alone give bit-identical positions, velocities, box and last thermo record."""
    stream = W.serving_jobs(16, seed=12349)
    jobs = S.jobs_from_stream(stream)
    fp = [G.footprint(j.system, 200) for j in jobs]
    budget = int(sum(fp) / 3)
    assert budget > max(fp)
    print(f"case config stream: 16 jobs, {sum(j.atom_steps for j in jobs):.3e} atom-steps, "
          f"budget {budget}", flush=True)
    srv = S.Server(hbm_budget=budget, tick_steps=50)
    stats = srv.run(jobs, realtime=False)
    srv.close()
    assert stats["jobs"] == 16
    for j in jobs:
        assert j.thermo and j.thermo[-1]["step"] == j.nsteps
    late = [j for j in jobs if j.admitted > min(k.admitted for k in jobs) + 1e-3]
    assert len(late) >= 8, "most jobs waited for memory and joined a running batch"
    sample = [2, 4, 6, 9, 12]
    for k in sample:
        job = S.jobs_from_stream([stream[k]])[0]
        print(f"case alone job {k} np={job.system.np} steps={job.nsteps}", flush=True)
        srv = S.Server(tick_steps=50)
        srv.run([job], realtime=False)
        srv.close()
        mixed = jobs[k]
        np.testing.assert_array_equal(job.result["pos"], mixed.result["pos"])
        np.testing.assert_array_equal(job.result["vel"], mixed.result["vel"])
        np.testing.assert_array_equal(job.result["box"], mixed.result["box"])
        assert job.thermo[-1] == mixed.thermo[-1]


def test_admission_while_steps_are_queued():
    """ndg_engine_step_enqueue / ndg_engine_step_wait: a system admitted while 20 steps of
    another are queued joins the steps that follow; both end bit-identical to runs alone
    (50 and 30 steps), and enqueue + wait equals ndg_engine_step."""
    x = W.cu_morse(7, temp_k=800.0, seed=41)
    y = W.buck_alloy((6, 6, 6), temp_k=500.0, seed=42)
    rp = lambda n: G.run_params(n, W.DT_1FS, out_steps=10, skin=0.5, trigger=G.TRIGGER_SKIN)
    eng = G.BatchEngine(path=G.PATH_MIXED)
    sx = eng.admit(x, rp(50))
    eng.step_enqueue(20)
    sy = eng.admit(y, rp(30))            # while the 20 steps are queued
    assert eng.step_wait() == 0
    assert eng.counters(sx)["step"] == 20 and eng.counters(sy)["step"] == 0
    eng.step_enqueue(30)
    assert eng.step_wait() == 2
    got = {k: eng.fetch(s, m.np) for k, s, m in (("x", sx, x), ("y", sy, y))}
    eng.close()
    for k, m, n in (("x", x, 50), ("y", y, 30)):
        e1 = G.BatchEngine(path=G.PATH_MIXED)
        s1 = e1.admit(m, rp(n))
        e1.step(n)
        ref = e1.fetch(s1, m.np)
        e1.close()
        for q in ("pos", "vel", "box"):
            np.testing.assert_array_equal(got[k][q], ref[q])


def test_served_with_size_class_table_alone_equals_mix():
    """A server with an autotuner table: jobs of different size classes run with different
    threads per atom (1 and 2) and list skins (0.5 and 1.0 A) in the same ticks.  Each job
    served alone under the same table gives bit-identical results, and the skin of its size
    class is the one it ran with (a job's kernels depend only on its own class)."""
    from ndyn_omp_amd import autotune as AT
    table = AT.TuneTable([(2000, 1, 0.5, 1, 1), (1 << 26, 2, 1.0, 1, 2)])
    alone = {}
    for job in make_jobs():
        print(f"case table alone job {job.jid} np={job.system.np}", flush=True)
        srv = S.Server(tick_steps=16, table=table)
        srv.run([job], realtime=False)
        alone[job.jid] = job
        srv.close()
    jobs = make_jobs()
    print("case table mixed", flush=True)
    srv = S.Server(tick_steps=9, table=table)
    stats = srv.run(jobs, realtime=False)
    srv.close()
    assert stats["jobs"] == 4
    for j in jobs:
        assert j.skin == (0.5 if j.system.np <= 2000 else 1.0)
        a = alone[j.jid]
        np.testing.assert_array_equal(j.result["pos"], a.result["pos"])
        np.testing.assert_array_equal(j.result["vel"], a.result["vel"])
        assert j.thermo[-1] == a.thermo[-1]
    assert stats["ticks"] > 0 and stats["mean_resident_atoms"] > 0
    assert stats["wait_s"] >= 0 and stats["admit_s"] >= 0 and stats["retire_s"] >= 0
