"""Summaries of rocprofv3 CSV output for the tile path's kernels.

    python -m frontierlabproject_amd.metrics.rocprof pmc DIR OUT.md --command "..."
    python -m frontierlabproject_amd.metrics.rocprof trace DIR OUT.md --command "..."
    python -m frontierlabproject_amd.metrics.rocprof traffic FETCH_DIR WRITE_DIR OUT.json \\
        --state-bytes B --circuits C --calib-bytes G --commands "..." "..."

``pmc``: per-dispatch hardware counters (``--pmc`` runs, ``*_counter_collection.csv``)
averaged per kernel class, with the derived figures: instructions per wave,
LDS bank-conflict and VALU-busy percentages of GPU time, occupancy as a
percentage of the maximum, and the per-dispatch table of the chunk passes in
launch order.  ``traffic``: HBM bytes per launch from a FETCH_SIZE run and a
WRITE_SIZE run (KiB per dispatch), each scaled by expected/measured for the
scaling pass over G bytes of complex64 state the same run makes (``bench.py
--calib-gib``: it reads and writes every amplitude once).
The specialised chunk-pass kernels (``qsim_jit_<hash>``, one per pass) form
one class."""
import argparse
import csv
import glob
import json
import os
from collections import OrderedDict, defaultdict

CU_NUM = 256
XCD_NUM = 8
MAX_WAVES_PER_CU = 32  # the OccupancyPercent normalisation of rocprofv3 on gfx950


def kernel_class(name):
    if name.startswith("qsim_jit_"):
        return "qsim_jit_* (specialised chunk passes)"
    if "k_permute" in name:
        return "k_permute"
    if "copyBuffer" in name:
        return "__amd_rocclr_copyBuffer"
    if "k_scale" in name:
        return "k_scale"
    return name


def _counter_csv(d):
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise FileNotFoundError(f"no *counter_collection.csv under {d}")
    return files[0]


def read_dispatches(path):
    """{dispatch id: {"name", "start", "end", counter: value}} in launch order."""
    out = OrderedDict()
    with open(path) as f:
        for r in csv.DictReader(f):
            d = out.setdefault(int(r["Dispatch_Id"]), {"name": r["Kernel_Name"],
                                                       "start": int(r["Start_Timestamp"]),
                                                       "end": int(r["End_Timestamp"]),
                                                       "vgpr": int(r["VGPR_Count"]),
                                                       "lds": int(r["LDS_Block_Size"]),
                                                       "scratch": int(r["Scratch_Size"])})
            d[r["Counter_Name"]] = d.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    return OrderedDict(sorted(out.items(), key=lambda kv: kv[1]["start"]))


def derived(d):
    """Per-dispatch figures from the raw counters that are present."""
    # the CSV sums GRBM_GUI_ACTIVE over the XCDs; rocprofv3's derived metrics
    # take its maximum over them, which for one kernel on the whole chip is
    # the per-XCD count
    g = d.get("GRBM_GUI_ACTIVE", 0.0) / XCD_NUM
    w = d.get("SQ_WAVES")
    out = {}
    for k, name in (("SQ_INSTS_VALU", "valu_per_wave"), ("SQ_INSTS_LDS", "lds_per_wave"),
                    ("SQ_INSTS_SALU", "salu_per_wave")):
        if w and k in d:
            out[name] = d[k] / w
    for k in ("LDSBankConflict", "OccupancyPercent", "VALUBusy"):
        if k in d:
            out[k] = d[k]
    if g:
        if "SQ_LDS_BANK_CONFLICT" in d and "LDSBankConflict" not in out:
            out["LDSBankConflict"] = 100.0 * d["SQ_LDS_BANK_CONFLICT"] / g / CU_NUM
        if "SQ_WAVE_CYCLES" in d and "OccupancyPercent" not in out:
            out["OccupancyPercent"] = 400.0 * d["SQ_WAVE_CYCLES"] / g / CU_NUM / MAX_WAVES_PER_CU
        if "SQ_ACTIVE_INST_VALU" in d and "VALUBusy" not in out:
            out["VALUBusy"] = 100.0 * d["SQ_ACTIVE_INST_VALU"] / CU_NUM / g
    return out


def pmc_markdown(dispatches, command, title):
    counters = sorted({k for d in dispatches.values() for k in d
                       if k not in ("name", "start", "end", "vgpr", "lds", "scratch")})
    groups = defaultdict(list)
    for d in dispatches.values():
        groups[kernel_class(d["name"])].append(d)
    lines = [f"# {title}", "", "Command (one MI355X, from the repository root):", "", "```", command, "```", "",
             "Per-dispatch averages per kernel class (raw counters summed over the chip).", ""]
    lines.append("| kernel | dispatches | " + " | ".join(counters) + " |")
    lines.append("|---|---|" + "---|" * len(counters))
    for cls, ds in sorted(groups.items(), key=lambda kv: -len(kv[1])):
        avg = [sum(d.get(c, 0.0) for d in ds) / len(ds) for c in counters]
        lines.append(f"| `{cls}` | {len(ds)} | " + " | ".join(f"{v:.4g}" for v in avg) + " |")
    der_keys = ["valu_per_wave", "lds_per_wave", "salu_per_wave", "LDSBankConflict", "OccupancyPercent", "VALUBusy"]
    lines += ["", "Derived (per dispatch, then averaged): instructions per wave; LDS bank-conflict stall and VALU "
              "busy as % of GPU time; occupancy as % of the maximum waves.", "",
              "| kernel | " + " | ".join(der_keys) + " |", "|---|" + "---|" * len(der_keys)]
    for cls, ds in sorted(groups.items(), key=lambda kv: -len(kv[1])):
        ders = [derived(d) for d in ds]
        vals = []
        for k in der_keys:
            xs = [x[k] for x in ders if k in x]
            vals.append(f"{sum(xs) / len(xs):.4g}" if xs else "")
        lines.append(f"| `{cls}` | " + " | ".join(vals) + " |")
    jit = [d for d in dispatches.values() if d["name"].startswith("qsim_jit_")]
    if jit:
        lines += ["", "Chunk passes in launch order (durations are under the profiler, which serialises "
                  "dispatches):", "", "| # | ms | " + " | ".join(der_keys) + " |",
                  "|---|---|" + "---|" * len(der_keys)]
        for i, d in enumerate(jit):
            x = derived(d)
            lines.append(f"| {i} | {(d['end'] - d['start']) / 1e6:.2f} | " +
                         " | ".join(f"{x[k]:.4g}" if k in x else "" for k in der_keys) + " |")
    return "\n".join(lines) + "\n"


def trace_markdown(d, command, title):
    """Kernel-trace summary (``--kernel-trace --stats`` output): time per
    kernel class and each class's resources."""
    stats = glob.glob(os.path.join(d, "**", "*kernel_stats.csv"), recursive=True)
    trace = glob.glob(os.path.join(d, "**", "*kernel_trace.csv"), recursive=True)
    if not stats or not trace:
        raise FileNotFoundError(f"no kernel stats/trace under {d}")
    agg = OrderedDict()
    with open(stats[0]) as f:
        for r in csv.DictReader(f):
            cls = kernel_class(r["Name"])
            a = agg.setdefault(cls, {"calls": 0, "total": 0.0, "min": float("inf"), "max": 0.0, "pct": 0.0,
                                     "distinct": 0})
            a["calls"] += int(r["Calls"])
            a["total"] += float(r["TotalDurationNs"])
            a["min"] = min(a["min"], float(r["MinNs"]))
            a["max"] = max(a["max"], float(r["MaxNs"]))
            a["pct"] += float(r["Percentage"])
            a["distinct"] += 1
    res = defaultdict(lambda: {"vgpr": set(), "agpr": set(), "sgpr": set(), "lds": set(), "scratch": set(),
                               "grid": set(), "wg": set()})
    with open(trace[0]) as f:
        for r in csv.DictReader(f):
            x = res[kernel_class(r["Kernel_Name"])]
            x["vgpr"].add(int(r["VGPR_Count"]))
            x["agpr"].add(int(r["Accum_VGPR_Count"]))
            x["sgpr"].add(int(r["SGPR_Count"]))
            x["lds"].add(int(r["LDS_Block_Size"]))
            x["scratch"].add(int(r["Scratch_Size"]))
            x["grid"].add(int(r["Grid_Size_X"]))
            x["wg"].add(int(r["Workgroup_Size_X"]))

    def span(v):
        return str(min(v)) if min(v) == max(v) else f"{min(v)}-{max(v)}"

    lines = [f"# {title}", "", "Command (one MI355X, from the repository root):", "", "```", command, "```", "",
             "| kernel | calls | distinct kernels | total ms | avg ms | min ms | max ms | % |",
             "|---|---|---|---|---|---|---|---|"]
    for cls, a in sorted(agg.items(), key=lambda kv: -kv[1]["total"]):
        lines.append(f"| `{cls}` | {a['calls']} | {a['distinct']} | {a['total'] / 1e6:.3f} | "
                     f"{a['total'] / a['calls'] / 1e6:.4f} | {a['min'] / 1e6:.4f} | {a['max'] / 1e6:.4f} | "
                     f"{a['pct']:.2f} |")
    lines += ["", "| kernel | VGPR | AGPR | SGPR | LDS B | scratch B | grid_x | wg_x |",
              "|---|---|---|---|---|---|---|---|"]
    for cls in sorted(res):
        x = res[cls]
        lines.append(f"| `{cls}` | " + " | ".join(span(x[k]) for k in ("vgpr", "agpr", "sgpr", "lds", "scratch",
                                                                      "grid", "wg")) + " |")
    # every chunk pass's duration, in launch order (one row per 16)
    passes = []
    with open(trace[0]) as f:
        for r in csv.DictReader(f):
            if r["Kernel_Name"].startswith("qsim_jit_"):
                passes.append((int(r["Start_Timestamp"]), (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6))
    if passes:
        passes.sort()
        lines += ["", f"Chunk-pass durations in launch order, ms ({len(passes)} dispatches):", "", "```"]
        for i in range(0, len(passes), 16):
            lines.append(" ".join(f"{ms:6.2f}" for _, ms in passes[i:i + 16]))
        lines.append("```")
    return "\n".join(lines) + "\n"


def traffic_json(fetch_dir, write_dir, state_bytes, circuits, calib_bytes, commands):
    fetch = read_dispatches(_counter_csv(fetch_dir))
    write = read_dispatches(_counter_csv(write_dir))

    def by_class(ds, counter):
        g = defaultdict(list)
        for d in ds.values():
            g[kernel_class(d["name"])].append(d[counter])
        return g

    fk, wk = by_class(fetch, "FETCH_SIZE"), by_class(write, "WRITE_SIZE")
    calib = "k_scale"
    cf = max(fk.get(calib, [0.0]))
    cw = max(wk.get(calib, [0.0]))
    if not cf or not cw:
        raise RuntimeError("no calibration pass in the runs")
    fscale, wscale = calib_bytes / 1024 / cf, calib_bytes / 1024 / cw
    out = {"note": "HBM bytes per launch from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (KiB per dispatch), each "
                   f"scaled by the ratio expected/measured for a scaling pass (k_scale) over a "
                   f"{calib_bytes / 2**30:g} GiB complex64 state, which reads and writes every amplitude once, run in "
                   "the same process. A circuit run from reset() has one chunk pass that generates |0...0> and only "
                   "writes: its algorithmic bytes are half a pass.",
           "commands": list(commands),
           "calibration": {"copy_bytes_read": calib_bytes, "copy_bytes_written": calib_bytes,
                           "fetch_kib_measured": cf, "write_kib_measured": cw,
                           "fetch_scale": round(fscale, 4), "write_scale": round(wscale, 4)},
           "kernels": {}}
    for key, cls in (("chunk_pass", "qsim_jit_* (specialised chunk passes)"), ("permute", "k_permute")):
        f, w = fk.get(cls, []), wk.get(cls, [])
        if not f or not w:
            out["kernels"][key] = {"kernel": cls, "dispatches": 0}
            continue
        gen = circuits if key == "chunk_pass" else 0
        favg, wavg = sum(f) / len(f), sum(w) / len(w)
        fb, wb = favg * 1024 * fscale, wavg * 1024 * wscale
        alg = state_bytes * 2 * (len(f) - 0.5 * gen) / len(f)
        out["kernels"][key] = {"kernel": cls, "dispatches": len(f), "generating_dispatches": gen,
                               "fetch_kib_avg": favg, "write_kib_avg": wavg,
                               "fetch_bytes_per_launch": fb, "write_bytes_per_launch": wb,
                               "traffic_bytes_per_launch": fb + wb, "algorithmic_bytes_per_launch": alg,
                               "traffic_over_algorithmic": round((fb + wb) / alg, 4)}
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    p = sub.add_parser("pmc")
    p.add_argument("dir")
    p.add_argument("out")
    p.add_argument("--command", required=True)
    p.add_argument("--title", default="Hardware counters: 33-qubit depth-30 random circuit, complex64, tile path")
    k = sub.add_parser("trace")
    k.add_argument("dir")
    k.add_argument("out")
    k.add_argument("--command", required=True)
    k.add_argument("--title", default="Kernel trace: 33-qubit depth-30 random circuit, complex64, tile path")
    t = sub.add_parser("traffic")
    t.add_argument("fetch_dir")
    t.add_argument("write_dir")
    t.add_argument("out")
    t.add_argument("--state-bytes", type=int, required=True)
    t.add_argument("--circuits", type=int, required=True, help="circuits run from reset() in the profiled run")
    t.add_argument("--calib-bytes", type=int, required=True)
    t.add_argument("--commands", nargs="+", required=True)
    a = ap.parse_args(argv)
    if a.cmd in ("pmc", "trace"):
        text = (pmc_markdown(read_dispatches(_counter_csv(a.dir)), a.command, a.title) if a.cmd == "pmc"
                else trace_markdown(a.dir, a.command, a.title))
        with open(a.out, "w") as f:
            f.write(text)
    else:
        js = traffic_json(a.fetch_dir, a.write_dir, a.state_bytes, a.circuits, a.calib_bytes, a.commands)
        with open(a.out, "w") as f:
            json.dump(js, f, indent=1)


if __name__ == "__main__":
    main()
