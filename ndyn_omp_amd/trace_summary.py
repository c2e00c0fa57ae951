"""Per-kernel summary of a rocprofv3 kernel trace (CSV) over the batched stepping window.

    python -m ndyn_omp_amd.trace_summary <prefix>_kernel_trace.csv [--pmc fetch.csv write.csv]

The window is the last contiguous run of `k_drift_mixed` launches with the largest grid
seen (the batched launches over every replica; runs are split where launches are more
than 50 ms apart), which leaves out autotuner trials, admission-time list builds and the
single-replica launches of the force-error check.  For each kernel it
prints launches, total, average and per-step time (steps = batched drift launches).

With --pmc, it also prints FETCH_SIZE / WRITE_SIZE per dispatch (KiB) of each kernel from
two `rocprofv3 --pmc` counter collections, and HBM bytes per dispatch scaled by what the
same counters read for `k_copy`, which reads and writes --copy-bytes per dispatch
(ndg_copy_bandwidth's 1 GiB copy): factor = known bytes / counter bytes.

With --sq COUNTERS_CSV AGENT_CSV (a collection of SQ_INSTS_VALU, SQ_WAVES, SQ_WAVE_CYCLES,
SQ_BUSY_CYCLES, SQ_INSTS_VMEM_RD, SQ_INSTS_LDS, SQ_INSTS_SALU and GRBM_GUI_ACTIVE, and the
run's agent info), it prints per kernel (largest-grid dispatches): wave instructions, the
kernel's cycles (GRBM_GUI_ACTIVE over the XCDs), VALU busy = 4 cycles per wave instruction
over SIMDs x cycles, and waves resident on average (SQ_WAVE_CYCLES counts 4-cycle units).
--pairs N,K gives kernel K's work units per dispatch (e.g. list entries of the force
kernel) to print VALU instructions per unit (64 units per wave instruction).
"""
from __future__ import annotations

import argparse
import csv
import statistics
from collections import defaultdict


def kernel_name(raw):
    """'void ndg::k_force_mixed<1, false>(ndg::EngineDev, int, ...)' -> 'k_force_mixed<1, false>'."""
    n = raw.strip()
    if n.startswith("void "):
        n = n[5:]
    n = n.split("(")[0].replace("ndg::", "")
    return n


def load_trace(path):
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            rows.append(dict(name=kernel_name(r["Kernel_Name"]), t0=int(r["Start_Timestamp"]),
                             t1=int(r["End_Timestamp"]), grid=int(r["Grid_Size_X"]),
                             wg=int(r["Workgroup_Size_X"])))
    rows.sort(key=lambda r: r["t0"])
    return rows


def window(rows, anchor="k_drift_mixed", gap_ms=50.0):
    """The last contiguous run of batched steps: largest-grid anchor launches, split where
    consecutive ones are more than gap_ms apart (admission of a new batch between runs)."""
    # a templated kernel matches by its name without the template arguments
    drifts = [r for r in rows if r["name"].split("<")[0] == anchor]
    if not drifts:
        return rows, 0, 0
    gmax = max(r["grid"] for r in drifts)
    big = [r for r in drifts if r["grid"] == gmax]
    start = 0
    for k in range(1, len(big)):
        if (big[k]["t0"] - big[k - 1]["t1"]) * 1e-6 > gap_ms:
            start = k
    big = big[start:]
    t_lo, t_hi = big[0]["t0"], big[-1]["t1"]
    # the last step ends with its reduction
    last_reduce = [r for r in rows if r["name"] == "k_reduce" and r["t0"] >= t_hi]
    if last_reduce:
        t_hi = last_reduce[0]["t1"]
    return [r for r in rows if r["t0"] >= t_lo and r["t1"] <= t_hi], len(big), gmax


def summarize(rows, steps):
    by = defaultdict(list)
    for r in rows:
        by[r["name"]].append(r)
    out = []
    for name, rs in by.items():
        d = [(r["t1"] - r["t0"]) * 1e-6 for r in rs]
        full = [x for x, r in zip(d, rs) if r["grid"] == max(q["grid"] for q in rs)]
        out.append(dict(name=name, calls=len(d), total_ms=sum(d), avg_ms=statistics.mean(d),
                        full_grid=max(r["grid"] for r in rs), full_calls=len(full),
                        full_avg_ms=statistics.mean(full),
                        per_step_ms=sum(d) / steps if steps else 0.0))
    out.sort(key=lambda x: -x["total_ms"])
    return out


def load_pmc(path, counter):
    """Counter value (summed over its instances) per dispatch: {kernel: [(grid, value)]}."""
    per = defaultdict(float)
    meta = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            if r.get("Counter_Name") == counter:
                key = (kernel_name(r["Kernel_Name"]), r["Dispatch_Id"])
                per[key] += float(r["Counter_Value"])
                meta[key] = int(r["Grid_Size"])
    vals = defaultdict(list)
    for (name, _), v in per.items():
        vals[name].append((meta[(name, _)], v))
    return vals


def full_grid_mean(pairs):
    g = max(p[0] for p in pairs)
    v = [p[1] for p in pairs if p[0] == g]
    return g, statistics.mean(v), len(v)


def load_counters(path):
    """{(kernel, dispatch): {counter: value summed over instances}} and grids."""
    per = defaultdict(lambda: defaultdict(float))
    grid = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            key = (kernel_name(r["Kernel_Name"]), r["Dispatch_Id"])
            per[key][r["Counter_Name"]] += float(r["Counter_Value"])
            grid[key] = int(r["Grid_Size"])
    return per, grid


def gpu_agent(path):
    with open(path) as f:
        for r in csv.DictReader(f):
            if r.get("Agent_Type") == "GPU":
                return int(r["Simd_Count"]), int(r["Num_Xcc"])
    raise SystemExit("no GPU agent in " + path)


def print_sq(counters_csv, agent_csv, pairs):
    simds, xcc = gpu_agent(agent_csv)
    per, grid = load_counters(counters_csv)
    by = defaultdict(list)
    for key, c in per.items():
        by[key[0]].append((grid[key], c))
    print(f"\nSQ counters per dispatch (largest grid; {simds} SIMDs, {xcc} XCDs): kernel, grid, "
          f"dispatches, VALU wave-insts, cycles, VALU busy, waves resident, VMEM_RD, LDS, SALU"
          + (", VALU insts per unit" if pairs else ""))
    for name in sorted(by):
        g = max(x[0] for x in by[name])
        cs = [x[1] for x in by[name] if x[0] == g]
        m = {k: statistics.mean(c[k] for c in cs) for k in cs[0]}
        cyc = m.get("GRBM_GUI_ACTIVE", 0.0) / xcc
        valu = m.get("SQ_INSTS_VALU", 0.0)
        busy = 4.0 * valu / (simds * cyc) if cyc else 0.0
        res = 4.0 * m.get("SQ_WAVE_CYCLES", 0.0) / cyc if cyc else 0.0
        line = (f"{name[:28]:28s} {g:9d} {len(cs):4d} {valu:12.4e} {cyc:11.4e} {busy:7.3f} "
                f"{res:8.0f} {m.get('SQ_INSTS_VMEM_RD', 0.0):11.4e} {m.get('SQ_INSTS_LDS', 0.0):11.4e} "
                f"{m.get('SQ_INSTS_SALU', 0.0):11.4e}")
        if pairs and name.startswith(pairs[1]):
            line += f" {valu * 64.0 / pairs[0]:8.2f}"
        print(line)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("trace")
    ap.add_argument("--anchor", default="k_drift_mixed")
    ap.add_argument("--pmc", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"))
    ap.add_argument("--copy-bytes", type=float, default=float(1 << 30))
    ap.add_argument("--sq", nargs=2, metavar=("COUNTERS_CSV", "AGENT_CSV"))
    ap.add_argument("--pairs", default=None, help="N,KERNEL: work units per dispatch of KERNEL")
    a = ap.parse_args()
    rows = load_trace(a.trace)
    win, steps, gmax = window(rows, a.anchor)
    span = (win[-1]["t1"] - win[0]["t0"]) * 1e-6 if win else 0.0
    print(f"window: {steps} batched steps ({a.anchor} grid {gmax}), {span:.2f} ms of GPU time "
          f"span, {span / max(steps, 1):.4f} ms/step")
    print(f"{'kernel':28s} {'calls':>7s} {'total ms':>10s} {'avg ms':>9s} {'full-grid avg':>14s} "
          f"{'ms/step':>9s}")
    busy = 0.0
    for k in summarize(win, steps):
        busy += k["total_ms"]
        print(f"{k['name'][:28]:28s} {k['calls']:7d} {k['total_ms']:10.3f} {k['avg_ms']:9.4f} "
              f"{k['full_avg_ms']:9.4f} ({k['full_calls']:3d}) {k['per_step_ms']:9.4f}")
    print(f"{'(sum of kernel time)':28s} {'':7s} {busy:10.3f} {'':9s} {'':14s} "
          f"{busy / max(steps, 1):9.4f}")
    if a.pmc:
        fetch = load_pmc(a.pmc[0], "FETCH_SIZE")
        write = load_pmc(a.pmc[1], "WRITE_SIZE")
        kf = kw = None
        if "k_copy" in fetch and "k_copy" in write:
            kf = a.copy_bytes / (full_grid_mean(fetch["k_copy"])[1] * 1024.0)
            kw = a.copy_bytes / (full_grid_mean(write["k_copy"])[1] * 1024.0)
            print(f"\ncopy calibration ({a.copy_bytes:.0f} B read + written per k_copy dispatch): "
                  f"FETCH_SIZE x {kf:.3f}, WRITE_SIZE x {kw:.3f}")
        print("\nlargest-grid dispatches: kernel, grid, dispatches, FETCH_SIZE KiB, WRITE_SIZE KiB"
              + (", HBM bytes (calibrated)" if kf else ""))
        for name in sorted(set(fetch) | set(write)):
            if name not in fetch or name not in write:
                continue
            g, fv, n = full_grid_mean(fetch[name])
            _, wv, _ = full_grid_mean(write[name])
            line = f"{name[:28]:28s} {g:9d} {n:4d} {fv:14.1f} {wv:14.1f}"
            if kf:
                line += f" {fv * 1024 * kf + wv * 1024 * kw:16.4e}"
            print(line)
    if a.sq:
        pairs = None
        if a.pairs:
            n, k = a.pairs.split(",", 1)
            pairs = (float(n), k)
        print_sq(a.sq[0], a.sq[1], pairs)


if __name__ == "__main__":
    main()
