"""Summaries of rocprofv3 output for the solver's kernels.

Summaries produced from rocprofv3's CSV output:

* ``trace``: per-kernel dispatch count and average / min / max duration from
  ``--kernel-trace`` (``*_kernel_trace.csv``), restricted to the dispatches at
  each kernel's most frequent launch shape, i.e. the run's steps over the
  whole particle set (not the relaxation steps of a scene build nor the
  launch autotuner's trial steps).
* ``timeline``: per step, the span, the GPU busy time (union of dispatch
  intervals) and the idle time between dispatches.
* ``pmc``: per-kernel HBM traffic from two ``--pmc`` runs (FETCH_SIZE, then
  WRITE_SIZE, in KiB per dispatch), each scaled by what the same counter reads
  for a device copy of known size.
* ``units``: per-kernel busy fractions of the vector-memory path (TA, TD, L1
  tag lookups) and VALU issue rates from three more ``--pmc`` runs.

Usage::

    python -m fluids_amd.profiling copy --bytes 268435456   # calibration run, under --pmc
    python -m fluids_amd.profiling trace <trace_dir> [--traffic traffic.md] > summary.md
    python -m fluids_amd.profiling pmc <fetch_dir> <write_dir> <copy_fetch_dir> <copy_write_dir> \
        --copy-bytes 268435456 [--tcc-dir <hit_miss_dir>] > traffic.md
    python -m fluids_amd.profiling stats <trace_dir or *kernel_stats.csv> > stats.md
    python -m fluids_amd.profiling timeline <trace_dir> [--skip 10] > timeline.md
    python -m fluids_amd.profiling units <ta_td_dir> <sq_dir> <tcp_dir> > units.md

``bench.py`` reads the machine-readable ``traffic`` rows of a ``pmc`` summary
(``read_traffic``) to fill its roofline's ``traffic`` field.
"""
import argparse
import collections
import csv
import glob
import os
import re
import sys

# kernel symbol -> the name the solver's own profiler reports
PROFILER_NAMES = {
    "k_gather_fluid": "gather_ranges",
}


def short_name(kernel_name):
    """'void sph::k_lists_density<false>(int, ...)' -> 'lists_density'."""
    m = re.search(r"sph::(k_\w+)", kernel_name)
    if not m:
        return kernel_name.split("(")[0].strip()
    sym = m.group(1)
    return PROFILER_NAMES.get(sym, sym[2:])


def _csv(path_or_dir, suffix):
    if os.path.isdir(path_or_dir):
        hits = sorted(glob.glob(os.path.join(path_or_dir, "**", f"*{suffix}"), recursive=True))
        if not hits:
            raise FileNotFoundError(f"no *{suffix} under {path_or_dir}")
        return hits[0]
    return path_or_dir


def _grid(r):
    # counter collection CSVs have Grid_Size, kernel traces Grid_Size_X/Y/Z
    if "Grid_Size" in r:
        return int(r["Grid_Size"])
    return int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"])


def _wg(r):
    for f in ("Workgroup_Size_X", "Workgroup_Size"):
        if f in r and r[f]:
            return int(r[f])
    return 0


def _full_grid(rows, key, min_count=16):
    """Labels each dispatch by launch shape and keeps the run's own: per
    kernel, shapes (grid, workgroup size) with at least `min_count`
    dispatches (the scene build's relaxation steps and the launch
    autotuner's trial steps have fewer); the shape with the most total time
    keeps the kernel's name, others are named `kernel@grid/workgroup` (the
    correction launches past the third, which return at once, are such a
    shape). Returns (label, row) pairs."""
    count = collections.Counter((key(r), _grid(r), _wg(r)) for r in rows)
    total = collections.defaultdict(float)
    for r in rows:
        total[(key(r), _grid(r), _wg(r))] += _dur(r)
    main = {}
    for (k, g, w), c in count.items():
        if c < min_count:
            continue
        if k not in main or total[(k, g, w)] > total[(k,) + main[k]]:
            main[k] = (g, w)
    out = []
    for r in rows:
        k, g, w = key(r), _grid(r), _wg(r)
        if count[(k, g, w)] < min_count:
            continue
        out.append((k if main.get(k) == (g, w) else f"{k}@{g}/{w}", r))
    return out


def _dur(r):
    if "End_Timestamp" in r:
        return float(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    return 1.0


def trace_summary(trace_dir, traffic_file=None):
    rows = list(csv.DictReader(open(_csv(trace_dir, "kernel_trace.csv"))))
    rows = [r for r in rows if "sph::" in r["Kernel_Name"]]
    dur = collections.defaultdict(list)
    grid = {}
    for k, r in _full_grid(rows, lambda r: short_name(r["Kernel_Name"])):
        dur[k].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
        grid[k] = _grid(r)
    out = ["| kernel | grid (threads) | dispatches | avg us | min us | max us | total ms | HBM bytes | HBM GB/s |",
           "|---|---|---|---|---|---|---|---|---|"]
    for k, v in sorted(dur.items(), key=lambda kv: -sum(kv[1])):
        avg = sum(v) / len(v)
        t = read_traffic(traffic_file, k) if traffic_file else None
        tb = "-" if t is None else f"{t:.0f}"
        bw = "-" if t is None else f"{t / (avg * 1e-6) / 1e9:.0f}"
        out.append(f"| {k} | {grid[k]} | {len(v)} | {avg:.1f} | {min(v):.1f} | {max(v):.1f} | "
                   f"{sum(v) / 1e3:.3f} | {tb} | {bw} |")
    return "\n".join(out)


def timeline_summary(trace_dir, marker="gather_ranges", skip=10, exclude=("stats",)):
    """Step-by-step timeline of the solver's dispatches from a kernel trace:
    a step starts at each full-grid dispatch of `marker` (one per step
    whichever sort ran); the first `skip` steps (scene build, warm-up) and
    the steps that hold a kernel named in `exclude` (the diagnostic counters
    read between steps) are dropped. Per step: span from the first
    dispatch's start to the next step's first start, GPU busy time (the
    union of dispatch intervals), the idle time between dispatches, and the
    dispatch count; means and medians over the steps."""
    rows = list(csv.DictReader(open(_csv(trace_dir, "kernel_trace.csv"))))
    rows = [r for r in rows if "sph::" in r["Kernel_Name"]]
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    big = max(_grid(r) for r in rows if short_name(r["Kernel_Name"]) == marker)
    starts = [i for i, r in enumerate(rows) if short_name(r["Kernel_Name"]) == marker and _grid(r) == big]
    spans, busy, counts = [], [], []
    for a, b in zip(starts[skip:], starts[skip + 1:]):
        seg = rows[a:b]
        if any(short_name(r["Kernel_Name"]) in exclude for r in seg):
            continue
        t0, t1 = int(seg[0]["Start_Timestamp"]), int(rows[b]["Start_Timestamp"])
        iv = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"])) for r in seg)
        u, cs, ce = 0, iv[0][0], iv[0][1]
        for s0, e0 in iv[1:]:
            if s0 > ce:
                u += ce - cs
                cs, ce = s0, e0
            else:
                ce = max(ce, e0)
        u += ce - cs
        spans.append((t1 - t0) / 1e3)
        busy.append(u / 1e3)
        counts.append(len(seg))
    if not spans:
        return "no complete steps in the trace"
    n = len(spans)
    span, bsy = sum(spans) / n, sum(busy) / n
    med = lambda v: sorted(v)[len(v) // 2]
    return "\n".join([
        "| steps | dispatches / step | span us / step | GPU busy us / step | idle between dispatches us / step | idle per dispatch us |",
        "|---|---|---|---|---|---|",
        f"| {n} (mean) | {sum(counts) / n:.1f} | {span:.1f} | {bsy:.1f} | {span - bsy:.1f} | {(span - bsy) / (sum(counts) / n):.2f} |",
        f"| {n} (median) | {med(counts)} | {med(spans):.1f} | {med(busy):.1f} | "
        f"{med([a - b for a, b in zip(spans, busy)]):.1f} | - |",
    ])


def _counter_means(d, counter):
    rows = list(csv.DictReader(open(_csv(d, "counter_collection.csv"))))
    rows = [r for r in rows if r["Counter_Name"] == counter]
    return rows


def pmc_summary(fetch_dir, write_dir, copy_fetch_dir, copy_write_dir, copy_bytes, tcc_dir=None):
    def per_kernel(d, counter):
        rows = [r for r in _counter_means(d, counter) if "sph::" in r["Kernel_Name"]]
        acc = collections.defaultdict(list)
        for k, r in _full_grid(rows, lambda r: short_name(r["Kernel_Name"])):
            acc[k].append(float(r["Counter_Value"]))
        return {k: sum(v) / len(v) for k, v in acc.items()}, {k: len(v) for k, v in acc.items()}

    def copy_kib(d, counter):
        rows = [r for r in _counter_means(d, counter) if "copyBuffer" in r["Kernel_Name"]]
        vals = sorted(float(r["Counter_Value"]) for r in rows)
        return vals[len(vals) // 2]

    f, nf = per_kernel(fetch_dir, "FETCH_SIZE")
    w, nw = per_kernel(write_dir, "WRITE_SIZE")
    hit = miss = None
    if tcc_dir:
        hit, _ = per_kernel(tcc_dir, "TCC_HIT_sum")
        miss, _ = per_kernel(tcc_dir, "TCC_MISS_sum")
    cf, cw = copy_kib(copy_fetch_dir, "FETCH_SIZE"), copy_kib(copy_write_dir, "WRITE_SIZE")
    kf = copy_bytes / 1024.0 / cf
    kw = copy_bytes / 1024.0 / cw
    out = [f"Copy calibration: a {copy_bytes} B device copy reads FETCH_SIZE = {cf:.1f} KiB and "
           f"WRITE_SIZE = {cw:.1f} KiB, so fetch KiB x {kf:.3f} and write KiB x {kw:.3f} give bytes.", "",
           "| kernel | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB | traffic bytes | L2 hit rate |",
           "|---|---|---|---|---|---|"]
    for k in sorted(set(f) & set(w), key=lambda k: -(f[k] * kf + w[k] * kw)):
        t = 1024.0 * (f[k] * kf + w[k] * kw)
        rate = "-"
        if hit is not None and k in hit and hit[k] + miss.get(k, 0) > 0:
            rate = f"{hit[k] / (hit[k] + miss[k]):.3f}"
        out.append(f"| {k} | {nf[k]} | {f[k]:.1f} | {w[k]:.1f} | {t:.0f} | {rate} |")
    out.append("")
    out.append("Machine-readable rows (kernel, traffic bytes per dispatch):")
    out.append("")
    for k in sorted(set(f) & set(w)):
        out.append(f"traffic {k} {1024.0 * (f[k] * kf + w[k] * kw):.0f}")
    return "\n".join(out)


# Counter passes of `units`: each list fits one rocprofv3 --pmc run (at most
# 8 SQ, 4 TCP, 2 TA, 2 TD, 1 GRBM hardware counters per run).
UNIT_PASSES = {
    "ta_td": ["GRBM_GUI_ACTIVE", "TA_BUSY_avr", "TA_BUSY_max", "TA_ADDR_STALLED_BY_TC_CYCLES_sum",
              "TD_TD_BUSY_sum", "TD_TC_STALL_sum"],
    "sq": ["GRBM_GUI_ACTIVE", "SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_VMEM_RD", "SQ_INSTS_VMEM_WR",
           "SQ_ACTIVE_INST_VALU", "SQ_INSTS_SALU", "SQ_INSTS_LDS", "SQ_BUSY_CU_CYCLES"],
    "tcp": ["GRBM_GUI_ACTIVE", "TCP_TOTAL_CACHE_ACCESSES_sum", "TCP_TCC_READ_REQ_sum",
            "TCP_PENDING_STALL_CYCLES_sum", "TCP_TCP_TA_DATA_STALL_CYCLES_sum"],
}


def _pass_means(d):
    """Per kernel label: mean of every counter of one --pmc run, plus the
    mean dispatch duration in ns (from the rows' timestamps)."""
    rows = [r for r in csv.DictReader(open(_csv(d, "counter_collection.csv"))) if "sph::" in r["Kernel_Name"]]
    by_counter = collections.defaultdict(list)
    for r in rows:
        by_counter[r["Counter_Name"]].append(r)
    out = collections.defaultdict(dict)
    for cname, rs in by_counter.items():
        acc = collections.defaultdict(list)
        durs = collections.defaultdict(list)
        for k, r in _full_grid(rs, lambda r: short_name(r["Kernel_Name"])):
            acc[k].append(float(r["Counter_Value"]))
            durs[k].append(_dur(r))
        for k, v in acc.items():
            out[k][cname] = sum(v) / len(v)
            out[k]["_n"] = len(v)
            out[k]["_ns"] = sum(durs[k]) / len(durs[k])
        for k, r in _full_grid(rs, lambda r: short_name(r["Kernel_Name"])):
            out[k]["_shape"] = f"{_grid(r)}/{_wg(r)}"
    return out


def units_summary(ta_td_dir, sq_dir, tcp_dir, cus=256, xcds=8):
    """Per-kernel use of the vector-memory path (TA address unit, TD data
    return, the L1 tag lookups) and of the VALU, from the three counter
    passes of UNIT_PASSES.

    GRBM_GUI_ACTIVE is read per XCD and summed over the `xcds` XCDs, so one
    XCD's busy cycles are GRBM_GUI_ACTIVE / xcds (the derived metrics that
    rocprofiler ships take its maximum over XCDs for the same reason); the
    clock this implies (cycles / dispatch ns) is printed as a check. There
    is one TA, one TD and one L1 (TCP) per CU. A SIMD issues one wave64 VALU
    instruction per 4 cycles, so the CU's 4 SIMDs issue at most one per
    cycle: VALU issue = SQ_INSTS_VALU / (cus x cycles)."""
    passes = {"ta_td": _pass_means(ta_td_dir), "sq": _pass_means(sq_dir), "tcp": _pass_means(tcp_dir)}
    kernels = set(passes["ta_td"]) & set(passes["sq"]) & set(passes["tcp"])

    def cyc(p, k):
        return passes[p][k]["GRBM_GUI_ACTIVE"] / xcds

    out = ["| kernel | grid/block | dispatches | us | clock GHz | TA busy | TA busy (max CU) | TA addr stalled by L1/L2 | "
           "TD busy | TD stalled by L1/L2 | VALU issue | VALUBusy | VALU / wave | VMEM rd / wave | "
           "VMEM wr / wave | L1 tag lookups / VMEM rd | L1 → L2 reads / dispatch | L1 pending-L2 stall |",
           "|---|---|---|---|---|---|---|---|---|---|---|---|---|---|---|---|---|---|"]

    def pct(x):
        return f"{100.0 * x:.0f} %"

    rows = []
    for k in kernels:
        a, s, t = passes["ta_td"][k], passes["sq"][k], passes["tcp"][k]
        ca, cs, ct = cyc("ta_td", k), cyc("sq", k), cyc("tcp", k)
        waves = max(s["SQ_WAVES"], 1.0)
        rows.append((a["_ns"] * a["_n"], (
            f"| {k} | {a['_shape']} | {a['_n']} | {a['_ns'] / 1e3:.1f} | {ca / a['_ns']:.2f} | "
            f"{pct(a['TA_BUSY_avr'] / ca)} | {pct(a['TA_BUSY_max'] / ca)} | "
            f"{pct(a['TA_ADDR_STALLED_BY_TC_CYCLES_sum'] / cus / ca)} | "
            f"{pct(a['TD_TD_BUSY_sum'] / cus / ca)} | {pct(a['TD_TC_STALL_sum'] / cus / ca)} | "
            f"{pct(s['SQ_INSTS_VALU'] / (cus * cs))} | {pct(s['SQ_ACTIVE_INST_VALU'] / cus / cs)} | "
            f"{s['SQ_INSTS_VALU'] / waves:.0f} | {s['SQ_INSTS_VMEM_RD'] / waves:.1f} | "
            f"{s['SQ_INSTS_VMEM_WR'] / waves:.1f} | "
            f"{t['TCP_TOTAL_CACHE_ACCESSES_sum'] / max(s['SQ_INSTS_VMEM_RD'], 1.0):.1f} | "
            f"{t['TCP_TCC_READ_REQ_sum']:.0f} | {pct(t['TCP_PENDING_STALL_CYCLES_sum'] / cus / ct)} |")))
    out += [r for _, r in sorted(rows, key=lambda x: -x[0])]
    return "\n".join(out)


def read_traffic(path, kernel):
    """Traffic bytes per dispatch of `kernel` from a pmc summary, or None."""
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) == 3 and parts[0] == "traffic" and parts[1] == kernel:
            return float(parts[2])
    return None


def copy_calibration(nbytes):
    """Three device-to-device copies of `nbytes`, the known-size kernel whose
    FETCH_SIZE / WRITE_SIZE readings scale the solver's."""
    import torch
    a = torch.ones(nbytes // 4, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {nbytes} B per launch", flush=True)


def stats_summary(stats_csv, top=25):
    """The rows of rocprofv3's --stats kernel table (all dispatches), by total time."""
    rows = list(csv.DictReader(open(_csv(stats_csv, "kernel_stats.csv"))))
    agg = collections.OrderedDict()
    for r in rows:
        k = short_name(r["Name"])
        a = agg.setdefault(k, [0, 0.0, 0.0])
        a[0] += int(r["Calls"])
        a[1] += float(r["TotalDurationNs"])
        a[2] += float(r["Percentage"])
    out = ["| kernel | calls | total ms | avg us | % |", "|---|---|---|---|---|"]
    for k, (calls, tot, pct) in sorted(agg.items(), key=lambda kv: -kv[1][1])[:top]:
        out.append(f"| {k} | {calls} | {tot / 1e6:.3f} | {tot / calls / 1e3:.1f} | {pct:.1f} |")
    return "\n".join(out)


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("copy")
    c.add_argument("--bytes", type=int, default=1 << 28)
    t = sub.add_parser("trace")
    t.add_argument("trace_dir")
    t.add_argument("--traffic", default=None, help="pmc summary whose traffic rows give HBM bytes per dispatch")
    p = sub.add_parser("pmc")
    p.add_argument("fetch_dir")
    p.add_argument("write_dir")
    p.add_argument("copy_fetch_dir")
    p.add_argument("copy_write_dir")
    p.add_argument("--copy-bytes", type=int, required=True)
    p.add_argument("--tcc-dir", default=None, help="run with TCC_HIT_sum and TCC_MISS_sum, for L2 hit rates")
    tl = sub.add_parser("timeline")
    tl.add_argument("trace_dir")
    tl.add_argument("--skip", type=int, default=10, help="steps dropped at the start (build, warm-up)")
    st = sub.add_parser("stats")
    st.add_argument("stats_csv", help="a *kernel_stats.csv or a directory holding one")
    un = sub.add_parser("units")
    un.add_argument("ta_td_dir", help="run with the counters of UNIT_PASSES['ta_td']")
    un.add_argument("sq_dir", help="run with the counters of UNIT_PASSES['sq']")
    un.add_argument("tcp_dir", help="run with the counters of UNIT_PASSES['tcp']")
    un.add_argument("--cus", type=int, default=256)
    un.add_argument("--xcds", type=int, default=8)
    pc = sub.add_parser("passes", help="print the --pmc counter list of one units pass")
    pc.add_argument("name", choices=sorted(UNIT_PASSES))
    a = ap.parse_args(argv)
    if a.cmd == "copy":
        copy_calibration(a.bytes)
    elif a.cmd == "units":
        print(units_summary(a.ta_td_dir, a.sq_dir, a.tcp_dir, a.cus, a.xcds))
    elif a.cmd == "passes":
        print(" ".join(UNIT_PASSES[a.name]))
    elif a.cmd == "trace":
        print(trace_summary(a.trace_dir, a.traffic))
    elif a.cmd == "timeline":
        print(timeline_summary(a.trace_dir, skip=a.skip))
    elif a.cmd == "stats":
        print(stats_summary(a.stats_csv))
    else:
        print(pmc_summary(a.fetch_dir, a.write_dir, a.copy_fetch_dir, a.copy_write_dir, a.copy_bytes, a.tcc_dir))


if __name__ == "__main__":
    sys.exit(main())
