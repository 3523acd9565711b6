"""Summarise rocprofv3 results as Markdown tables.

Kernel trace: reads either the SQLite database rocprofv3 writes by default
(``*_results.db``, view ``top_kernels``) or a ``*kernel_stats.csv`` file.

    python -m vlxmi.profsum RESULT [--command "the profiled command"] [--top N]

Counters: reads the ``mem``, ``fetch`` and ``write`` passes (``--output-format
csv``) of one command under PMC_DIR and the per-class JSON bench.py writes.

    python -m vlxmi.profsum --counters PMC_DIR --per-class PERCLASS.json [--top N]
"""

from __future__ import annotations

import argparse
import csv
import sqlite3


def _short(name: str) -> str:
    return name.split("(")[0]


def rows_from(path: str):
    if path.endswith(".db"):
        con = sqlite3.connect(path)
        cur = con.cursor()
        # the top_kernels view reports microseconds; rows are kept in nanoseconds
        out = [(_short(n), int(c), 1e3 * float(t), 1e3 * float(a), float(p))
               for n, c, t, a, p in cur.execute("select name,total_calls,total_duration,average,percentage from top_kernels")]
        con.close()
        return out
    out = []
    with open(path) as f:
        for r in csv.DictReader(f):
            out.append((_short(r["Name"]), int(r["Calls"]), float(r["TotalDurationNs"]), float(r["AverageNs"]),
                        float(r["Percentage"])))
    return out


def markdown(path: str, command: str = "", top: int = 30) -> str:
    rows = sorted(rows_from(path), key=lambda r: -r[2])
    total = sum(r[2] for r in rows)
    o = []
    if command:
        o.append("Command: `%s`" % command)
        o.append("")
    o.append("| kernel | calls | total ms | avg ms | % |")
    o.append("|---|---:|---:|---:|---:|")
    for n, c, t, a, p in rows[:top]:
        o.append("| %s | %d | %.3f | %.3f | %.2f |" % (n, c, t * 1e-6, a * 1e-6, p))
    o.append("")
    o.append("All kernels: %d, total %.1f ms." % (len(rows), total * 1e-6))
    return "\n".join(o) + "\n"


# --------------------------------------------------------------------------
# counter collections (rocprofv3 --pmc ... --output-format csv)
# --------------------------------------------------------------------------

def counters_from(path: str):
    """{(dispatch id, kernel): {counter: value, "_t": seconds}} from run_counter_collection.csv."""
    out = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            key = (int(r["Dispatch_Id"]), _short(r["Kernel_Name"]))
            d = out.setdefault(key, {})
            d[r["Counter_Name"]] = d.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
            d["_t"] = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-9
    return out


def flops_f64(c):
    """Executed FP64 FLOPs of a dispatch (wave64 instructions x 64 lanes; FMA = 2)."""
    return 64.0 * (c.get("SQ_INSTS_VALU_ADD_F64", 0) + c.get("SQ_INSTS_VALU_MUL_F64", 0)
                   + 2 * c.get("SQ_INSTS_VALU_FMA_F64", 0) + c.get("SQ_INSTS_VALU_TRANS_F64", 0))


def flops_f32(c):
    return 64.0 * (c.get("SQ_INSTS_VALU_ADD_F32", 0) + c.get("SQ_INSTS_VALU_MUL_F32", 0)
                   + 2 * c.get("SQ_INSTS_VALU_FMA_F32", 0) + c.get("SQ_INSTS_VALU_TRANS_F32", 0))


def copy_calibration(fetch_path, write_path, copy_bytes):
    """Bytes per FETCH_SIZE / WRITE_SIZE unit measured on the device copy kernel
    (the largest elementwise copy dispatch) of ``copy_bytes`` bytes."""
    fe = counters_from(fetch_path)
    wr = counters_from(write_path)

    def pick(d, cname):
        cands = [(v.get(cname, 0.0), k) for k, v in d.items() if "copy" in k[1].lower()]
        return max(cands)[0] if cands else None

    f, w = pick(fe, "FETCH_SIZE"), pick(wr, "WRITE_SIZE")
    return (copy_bytes / f if f else None), (copy_bytes / w if w else None)


def counters_markdown(pmc_dir, per_class_path, top=25, copy_bytes=512 << 20):
    """Per-kernel memory-pipeline and HBM summary of the counter passes.

    TA/TD busy: TA_TA_BUSY / TD_TD_BUSY summed over the per-CU instances,
    divided by (GRBM_GUI_ACTIVE / 8 XCDs) x 256 CUs -- the fraction of the
    kernel's cycles the address / data units of an average CU were busy.
    HBM bytes: FETCH_SIZE and WRITE_SIZE scaled by what they read for the copy
    kernel of ``copy_bytes`` bytes in the same runs."""
    import json
    import os
    mem = counters_from(os.path.join(pmc_dir, "mem", "run_counter_collection.csv"))
    fe_p = os.path.join(pmc_dir, "fetch", "run_counter_collection.csv")
    wr_p = os.path.join(pmc_dir, "write", "run_counter_collection.csv")
    kf, kw = copy_calibration(fe_p, wr_p, copy_bytes)
    fe, wr = counters_from(fe_p), counters_from(wr_p)
    with open(per_class_path) as f:
        pc = json.load(f)

    def by_name(d):
        out = {}
        for (_, name), v in d.items():
            out.setdefault(name, v)
        return out

    mem, fe, wr = by_name(mem), by_name(fe), by_name(wr)
    names = sorted((n for n in mem if n in pc and not n.endswith("_mixed")), key=lambda n: -pc[n]["avg_s"])[:top]
    o = ["Copy-kernel calibration: %.1f bytes per FETCH_SIZE unit, %.1f bytes per WRITE_SIZE unit." % (kf, kw), "",
         "| kernel | ms | TFLOP/s (alg.) | TA busy | TD busy | VALU instr / quartet | VMEM rd / quartet "
         "| HBM read GB | HBM write GB | HBM GB/s |",
         "|---|---:|---:|---:|---:|---:|---:|---:|---:|---:|"]
    for n in names:
        c = mem[n]
        g = c.get("GRBM_GUI_ACTIVE", 0.0) / 8.0 * 256.0
        q = pc[n]["quartets"] / 64.0 or 1.0
        rb = fe.get(n, {}).get("FETCH_SIZE", 0.0) * kf
        wb = wr.get(n, {}).get("WRITE_SIZE", 0.0) * kw
        t = pc[n]["avg_s"]
        o.append("| %s | %.1f | %.2f | %.2f | %.2f | %.1f | %.2f | %.2f | %.2f | %.0f |" % (
            n, 1e3 * t, pc[n]["tflops"], c.get("TA_TA_BUSY", 0.0) / g if g else 0.0,
            c.get("TD_TD_BUSY", 0.0) / g if g else 0.0, c.get("SQ_INSTS_VALU", 0.0) / q,
            c.get("SQ_INSTS_VMEM_RD", 0.0) / q, rb / 1e9, wb / 1e9, (rb + wb) / t / 1e9))
    return "\n".join(o) + "\n"


def trace_by_grid(path, kernels, top=None):
    """Per (kernel, grid size) durations from a kernel trace
    (run_kernel_trace.csv): a run that autotunes a kernel's grid launches it
    at several sizes; the tuned size's average is the one timed runs see."""
    rows = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            name = _short(r["Kernel_Name"])
            if name not in kernels:
                continue
            g = r.get("Grid_Size_X") or r.get("Grid_Size") or "?"
            wg = r.get("Workgroup_Size_X") or r.get("Workgroup_Size") or "?"
            d = rows.setdefault((name, g, wg), [])
            d.append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-6)
    o = ["| kernel | grid (threads) | workgroup | calls | avg ms | min ms | max ms |", "|---|---:|---:|---:|---:|---:|---:|"]
    for (name, g, wg), v in sorted(rows.items(), key=lambda kv: (kv[0][0], -len(kv[1]))):
        o.append("| %s | %s | %s | %d | %.3f | %.3f | %.3f |" % (name, g, wg, len(v), sum(v) / len(v), min(v), max(v)))
    return "\n".join(o) + "\n"


def traffic_json(fetch_path, write_path, copy_bytes, command=""):
    """Per-kernel HBM bytes per dispatch from a FETCH_SIZE pass and a
    WRITE_SIZE pass of one command, each scaled by what the counter reads for
    the device copy of ``copy_bytes`` bytes in the same run (bytes per counter
    unit).  Returns {"calibration": ..., "kernels": {name: {"read", "write",
    "dispatches"}}}."""
    kf, kw = copy_calibration(fetch_path, write_path, copy_bytes)
    out = {}
    for path, cname, scale, key in ((fetch_path, "FETCH_SIZE", kf, "read"), (write_path, "WRITE_SIZE", kw, "write")):
        acc = {}
        for (_, name), v in counters_from(path).items():
            a = acc.setdefault(name, [0.0, 0])
            a[0] += v.get(cname, 0.0)
            a[1] += 1
        for name, (tot, n) in acc.items():
            d = out.setdefault(name, {"read": 0.0, "write": 0.0, "dispatches": n})
            d[key] = tot / n * scale if scale else None
    return {"command": command, "calibration": {"copy_bytes": copy_bytes, "bytes_per_fetch_unit": kf,
                                                "bytes_per_write_unit": kw},
            "kernels": out}


def busy_markdown(path, top=30):
    """Per-kernel issue and memory-pipeline figures of one counter pass
    (SQ_WAVES, SQ_WAVE_CYCLES, SQ_INSTS_VALU, SQ_ACTIVE_INST_VALU,
    SQ_INSTS_VMEM_RD, SQ_INSTS_LDS, SQ_WAIT_INST_ANY, TA_TA_BUSY, TD_TD_BUSY,
    GRBM_GUI_ACTIVE), averaged over each kernel's dispatches:

    * VALU busy: SQ_ACTIVE_INST_VALU / SQ_WAVE_CYCLES, the share of a wave's
      life spent issuing vector ALU instructions;
    * waiting: SQ_WAIT_INST_ANY / SQ_WAVE_CYCLES, the share spent waiting for
      any instruction's inputs (memory, mostly);
    * TA / TD busy: summed over the per-CU units, divided by
      (GRBM_GUI_ACTIVE / 8 XCDs) x 256 CUs."""
    d = counters_from(path)
    agg = {}
    for (_, name), v in d.items():
        a = agg.setdefault(name, {"_n": 0})
        a["_n"] += 1
        for k, x in v.items():
            a[k] = a.get(k, 0.0) + x
    rows = sorted(agg.items(), key=lambda kv: -kv[1].get("_t", 0.0))[:top]
    o = ["| kernel | dispatches | ms / dispatch | waves | VALU instr / wave | VMEM rd / wave | LDS instr / wave "
         "| VALU busy | waiting | TA busy | TD busy |",
         "|---|---:|---:|---:|---:|---:|---:|---:|---:|---:|---:|"]
    for name, a in rows:
        n = a["_n"]
        w = a.get("SQ_WAVES", 0.0) or 1.0
        cyc = a.get("SQ_WAVE_CYCLES", 0.0) or 1.0
        g = a.get("GRBM_GUI_ACTIVE", 0.0) / 8.0 * 256.0
        o.append("| %s | %d | %.3f | %.0f | %.1f | %.2f | %.2f | %.2f | %.2f | %.2f | %.2f |" % (
            name, n, 1e3 * a.get("_t", 0.0) / n, w / n, a.get("SQ_INSTS_VALU", 0.0) / w,
            a.get("SQ_INSTS_VMEM_RD", 0.0) / w, a.get("SQ_INSTS_LDS", 0.0) / w,
            a.get("SQ_ACTIVE_INST_VALU", 0.0) / cyc, a.get("SQ_WAIT_INST_ANY", 0.0) / cyc,
            a.get("TA_TA_BUSY", 0.0) / g if g else 0.0, a.get("TD_TD_BUSY", 0.0) / g if g else 0.0))
    return "\n".join(o) + "\n"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("result", nargs="?")
    ap.add_argument("--command", default="")
    ap.add_argument("--top", type=int, default=30)
    ap.add_argument("--counters", default=None, help="directory holding the mem/fetch/write counter passes")
    ap.add_argument("--per-class", default=None, help="per-class JSON written by bench.py --per-class-out")
    ap.add_argument("--busy", default=None, help="run_counter_collection.csv of one issue/memory counter pass")
    ap.add_argument("--traffic", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"), default=None,
                    help="FETCH_SIZE and WRITE_SIZE passes of one command: per-kernel HBM bytes (JSON)")
    ap.add_argument("--copy-bytes", type=int, default=512 << 20, help="size of the calibration copy in those runs")
    ap.add_argument("--trace", default=None, help="run_kernel_trace.csv: per (kernel, grid size) durations")
    ap.add_argument("--kernels", nargs="*", default=[], help="with --trace: the kernels to list")
    a = ap.parse_args()
    if a.trace:
        print(trace_by_grid(a.trace, set(a.kernels)), end="")
    elif a.traffic:
        import json
        print(json.dumps(traffic_json(a.traffic[0], a.traffic[1], a.copy_bytes, a.command), indent=1, sort_keys=True))
    elif a.busy:
        print(busy_markdown(a.busy, a.top), end="")
    elif a.counters:
        print(counters_markdown(a.counters, a.per_class, a.top), end="")
    else:
        print(markdown(a.result, a.command, a.top), end="")


if __name__ == "__main__":
    main()
