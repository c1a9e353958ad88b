r"""Summaries of rocprofv3 CSV output for the hip backend's kernels.

    python -m mhd_gpu_solver_amd.profiling trace  <kernel_trace.csv> [last N]
    python -m mhd_gpu_solver_amd.profiling traffic <fetch counter_collection.csv> \
        <write counter_collection.csv> <copy bytes> <interior cells per launch> [kernel filter]
    python -m mhd_gpu_solver_amd.profiling valu <SQ counter_collection.csv> <kernel_trace.csv> \
        <SIMDs> <clock GHz> <last N> [kernel filter]

``trace`` prints per-kernel launch statistics. Launches shorter than ``NOOP_US`` are the
early-exit steps at the end of a fixed-size batch and are counted apart from the active ones.
With ``last N`` it adds the mean of each kernel's last N active launches: the timed steps of a
bench.py leg come last, after its autotuning and warm-up launches.
``traffic`` converts FETCH_SIZE / WRITE_SIZE (KiB per dispatch) into bytes per launch, scaled
by what the same counter reads for a plain device copy of known size in the same run.
``valu`` takes SQ_INSTS_VALU / SQ_INSTS_SALU (wave instructions per dispatch, the mean of each
kernel's last N active launches) from a counter run and the mean duration of the same kernels'
last N active launches from a trace run without counters, and gives the VALU issue
utilisation: instructions x 4 cycles (one wave64 instruction on a 16-lane SIMD) over SIMDs x
clock x duration. Quarter-rate instructions (transcendentals, 32-bit integer multiplies) hold
a SIMD for 16 cycles, so the figure is a lower bound on the SIMDs' busy time.
"""
from __future__ import annotations

import collections
import csv
import json
import sys

NOOP_US = 50.0


def _short(name: str) -> str:
    """``void mhdhip::(anonymous namespace)::k_x<float>(float const*, ...)`` -> ``k_x<float>``."""
    name = name.replace("(anonymous namespace)::", "")
    return name.split("(")[0].replace("void ", "").replace("mhdhip::", "")


def kernel_table(trace_csv: str, last: int = 0) -> str:
    rows = sorted(csv.DictReader(open(trace_csv)), key=lambda r: int(r["Start_Timestamp"]))
    durs = collections.defaultdict(list)
    for r in rows:
        d = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
        durs[_short(r["Kernel_Name"])].append(d)
    tail = f" mean of last {last} (us) |" if last else ""
    out = ["| kernel | launches | active | mean active (us) | min (us) | max (us) | no-op mean (us) |"
           + tail, "|---|---|---|---|---|---|---|" + ("---|" if last else "")]
    for k, v in sorted(durs.items(), key=lambda kv: -sum(kv[1])):
        act = [x for x in v if x > NOOP_US]
        noop = [x for x in v if x <= NOOP_US]
        if act:
            line = (f"| {k} | {len(v)} | {len(act)} | {sum(act) / len(act):.1f} | "
                    f"{min(act):.1f} | {max(act):.1f} | "
                    f"{(sum(noop) / len(noop)) if noop else 0:.2f} |")
            if last:
                sel = act[-last:]
                line += f" {sum(sel) / len(sel):.1f} |"
        else:
            line = f"| {k} | {len(v)} | 0 | - | - | - | {sum(noop) / len(noop):.2f} |"
            if last:
                line += " - |"
        out.append(line)
    return "\n".join(out)


def _counter(path: str, counter: str):
    per = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        d = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
        per[_short(r["Kernel_Name"])].append((float(r["Counter_Value"]), d))
    return per


def traffic(fetch_csv: str, write_csv: str, copy_bytes: int, cells: int,
            name_filter: str = "stage") -> dict:
    fetch = _counter(fetch_csv, "FETCH_SIZE")
    write = _counter(write_csv, "WRITE_SIZE")
    kib = copy_bytes / 1024.0
    copy_f = [v for v, d in fetch["__amd_rocclr_copyBuffer"] if v > 0.1 * kib]
    copy_w = [v for v, d in write["__amd_rocclr_copyBuffer"] if v > 0.1 * kib]
    f_scale = kib / (sum(copy_f) / len(copy_f))
    w_scale = kib / (sum(copy_w) / len(copy_w))
    res = {"copy_bytes": copy_bytes, "fetch_scale": round(f_scale, 4),
           "write_scale": round(w_scale, 4), "kernels": {}}
    for k in sorted(set(fetch) | set(write)):
        if name_filter not in k:
            continue
        f = [v for v, d in fetch.get(k, []) if d > NOOP_US]
        w = [v for v, d in write.get(k, []) if d > NOOP_US]
        if not f or not w:
            continue
        fb = sum(f) / len(f) * f_scale * 1024
        wb = sum(w) / len(w) * w_scale * 1024
        elem = 8 if "double" in k else 4
        stage = 1 if "stage<1" in k else 2
        alg_r = (9 if stage == 1 else 18) * elem * cells
        alg_w = 9 * elem * cells
        res["kernels"][k] = {
            "fetch_bytes": round(fb), "write_bytes": round(wb), "traffic_bytes": round(fb + wb),
            "algorithmic_read": alg_r, "algorithmic_write": alg_w,
            "fetch_over_algorithmic": round(fb / alg_r, 3),
            "write_over_algorithmic": round(wb / alg_w, 3),
        }
    return res


def _last_active(values, last):
    act = [v for v in values if v[1] > NOOP_US]
    return act[-last:] if last else act


def valu(counter_csv: str, trace_csv: str, simds: int, clock_ghz: float, last: int,
         name_filter: str = "stage") -> str:
    per = collections.defaultdict(lambda: collections.defaultdict(list))
    rows = sorted(csv.DictReader(open(counter_csv)), key=lambda r: int(r["Start_Timestamp"]))
    for r in rows:
        k = _short(r["Kernel_Name"])
        if name_filter not in k:
            continue
        d = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
        per[k][r["Counter_Name"]].append((float(r["Counter_Value"]), d))
    durs = collections.defaultdict(list)
    for r in sorted(csv.DictReader(open(trace_csv)), key=lambda r: int(r["Start_Timestamp"])):
        d = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
        durs[_short(r["Kernel_Name"])].append((0.0, d))
    out = ["| kernel | VALU instructions per launch | SALU per launch | mean duration (us) | "
           "VALU issue utilisation |", "|---|---|---|---|---|"]
    for k in sorted(per):
        v = _last_active(per[k].get("SQ_INSTS_VALU", []), last)
        sa = _last_active(per[k].get("SQ_INSTS_SALU", []), last)
        dd = _last_active(durs.get(k, []), last)
        if not v or not dd:
            continue
        nv = sum(x for x, _ in v) / len(v)
        ns = sum(x for x, _ in sa) / len(sa) if sa else float("nan")
        dur = sum(d for _, d in dd) / len(dd)
        util = nv * 4.0 / (simds * clock_ghz * 1e3 * dur)
        out.append(f"| {k} | {nv / 1e6:.1f} M | {ns / 1e6:.1f} M | {dur:.1f} | {util * 100:.1f} % |")
    return "\n".join(out)


def main(argv):
    if len(argv) >= 2 and argv[0] == "trace":
        last = int(argv[3]) if len(argv) >= 4 and argv[2] == "last" else 0
        print(kernel_table(argv[1], last))
    elif len(argv) >= 5 and argv[0] == "traffic":
        filt = argv[5] if len(argv) >= 6 else "stage"
        print(json.dumps(traffic(argv[1], argv[2], int(argv[3]), int(argv[4]), filt), indent=1))
    elif len(argv) >= 6 and argv[0] == "valu":
        filt = argv[6] if len(argv) >= 7 else "stage"
        print(valu(argv[1], argv[2], int(argv[3]), float(argv[4]), int(argv[5]), filt))
    else:
        print(__doc__)
        return 2
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
