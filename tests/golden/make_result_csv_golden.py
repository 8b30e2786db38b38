"""Writes result_csv_custom.json: the known answers of the reference's
published tile-path runs, taken from the data file it ships
(qgpusim results/result.csv; columns as written by run_benchmarks.py:551-566
and runner.py:136-174).

Only the rows of the custom (LQR tile) backend are kept, one per distinct set
of answers: circuit shape after transpile (nqubits_t, depth_t, total_gates,
twoq_count, cx_count) and the metrics run_custom_backend returned
(backend.py:505-520), parsed; the shape of the same circuits as built
("aer_rows": the Aer baseline rows, transpiled at optimisation level 0, so
before any gate merging); plus each kept row verbatim ("csv_row", every
column as written but the four naming the machine and time) and the file's
header, which pin the schema and the number formats of a result row.

Run once where the reference is present:
    python tests/golden/make_result_csv_golden.py /root/reference/results/result.csv
"""
import csv
import json
import os
import sys

SHAPE_FIELDS = ("nqubits", "depth", "nqubits_t", "depth_t", "total_gates", "twoq_count", "cx_count", "cz_count",
                "swap_count", "rz_count", "rx_count")
INT_FIELDS = SHAPE_FIELDS + ("nL", "num_tiles", "num_gates_custom", "num_fused_ops", "num_reorders",
                             "num_kernel_launches", "num_graph_replays", "avoided_launches")
FLOAT_FIELDS = ("fusion_ratio", "t_launch_overhead_saved_us")
STR_FIELDS = ("circuit", "task", "tile_mode", "backend")
ENV_FIELDS = ("timestamp", "host", "os", "python")  # where and when the run was made: not kept


def main(src, dst):
    rows, aer_rows = [], []
    seen, seen_aer = set(), set()
    with open(src, newline="") as f:
        reader = csv.DictReader(f)
        header = list(reader.fieldnames)
        for line, r in enumerate(reader, start=2):
            if r.get("backend") == "aer" and r.get("transpiler") == "baseline":
                # the circuit as built, transpiled at optimisation level 0
                rec = {k: int(r[k]) for k in SHAPE_FIELDS}
                rec["circuit"] = r["circuit"]
                key = json.dumps(rec, sort_keys=True)
                if key not in seen_aer:
                    seen_aer.add(key)
                    rec["csv_line"] = line
                    aer_rows.append(rec)
                continue
            if r.get("backend") != "custom" or not r.get("num_tiles"):
                continue
            rec = {k: int(r[k]) for k in INT_FIELDS}
            rec.update({k: float(r[k]) for k in FLOAT_FIELDS})
            rec.update({k: r[k] for k in STR_FIELDS})
            key = json.dumps(rec, sort_keys=True)
            if key in seen:
                continue
            seen.add(key)
            rec["csv_line"] = line
            rec["csv_row"] = {k: v for k, v in r.items() if k not in ENV_FIELDS}
            rows.append(rec)
    with open(dst, "w") as f:
        json.dump({"source": "results/result.csv", "header": header, "rows": rows, "aer_rows": aer_rows}, f,
                  indent=1, sort_keys=True)
        f.write("\n")
    print(f"{len(rows)} rows -> {dst}")


if __name__ == "__main__":
    here = os.path.dirname(os.path.abspath(__file__))
    main(sys.argv[1], os.path.join(here, "result_csv_custom.json"))
