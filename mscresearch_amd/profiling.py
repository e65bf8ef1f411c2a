"""HBM traffic per kernel-family launch from rocprofv3 counter runs, and
kernel durations from kernel traces, over the bench's marked phases.

Inputs: the counter_collection.csv files of four ``rocprofv3 --pmc`` runs
(FETCH_SIZE and WRITE_SIZE of the bench, FETCH_SIZE and WRITE_SIZE of
``bench.py --calibrate-copy BYTES``) and the bench's JSON line from the
counter runs.  Each counter is scaled by known bytes / counted bytes of the
calibration copy, summed over the kernels of a family and divided by that
family's launches in the same run (the bench run with --profile-all, so the family
launch counts cover every pass the counters saw, warm-up included; a c2-only
run, --only-grid, sums its region counts' families).

Output: a JSON file per workload that bench.py reads for roofline.traffic.
"""
from __future__ import annotations

import csv
import json
import os
import re
from collections import defaultdict

# kernel name (demangled, template arguments dropped) -> engine family.
# k_reset (counter and per-tile resets, a few KB) belongs to none.
FAMILY_OF = {
    "k_init": "init", "k_preflow": "init", "k_init_marked": "init", "k_marks_clear": "init",
    "k_gr_seed": "bfs", "k_bfs_small": "bfs", "k_bfs_multi": "bfs", "k_bfs_td": "bfs", "k_bfs_td_heavy": "bfs", "k_bfs_td_giant": "bfs",
    "k_bfs_advance": "bfs", "k_bfs_seed1": "bfs", "k_local_mark": "bfs", "k_local_bound": "bfs",
    "k_local_bucket": "bfs", "k_local_bfs": "bfs", "k_local_finish": "bfs", "k_mask_combine": "bfs",
    "k_bfs_bu": "bfs_bu", "k_bfs_bu_med": "bfs_bu", "k_bfs_bu_heavy": "bfs_bu", "k_bu_bits": "bfs_bu",
    "k_active": "active", "k_active_sparse": "active", "k_list_copy": "active",
    "k_discharge_list": "discharge", "k_discharge_tile": "discharge", "k_giant_prep": "discharge",
    "k_giant_push": "discharge", "k_giant_relabel": "discharge", "k_path_push": "discharge",
    "k_hot_select": "discharge",
    "k_rm": "rm",
    "k_cert": "cert", "k_cert_fp": "cert", "k_cert_chunks": "cert", "k_cert_fin": "cert",
    "k_cert_sp_pairs": "cert", "k_cert_sp_rows": "cert",
}


def kernel_key(name: str) -> str:
    """'void mf::k_bfs_bu<0>(mf::SlotView const*, int const*)' -> 'k_bfs_bu'."""
    base = name.split("(")[0]
    base = re.sub(r"<.*", "", base)
    return base.split("::")[-1].replace("void ", "").strip()


MARKER = "k_trace_marker"


def marked_rows(path: str) -> list:
    """The CSV's dispatch rows, in dispatch order, that lie between a pair of
    trace markers (the bench emits one before and one after each measured
    phase); every row when the file holds no marker."""
    with open(path) as f:
        rows = list(csv.DictReader(f))
    key = "Dispatch_Id" if rows and "Dispatch_Id" in rows[0] else "Correlation_Id"
    rows.sort(key=lambda r: int(r[key]))
    if not any(kernel_key(r["Kernel_Name"]) == MARKER for r in rows):
        return rows
    out, inside = [], False
    for r in rows:
        if kernel_key(r["Kernel_Name"]) == MARKER:
            inside = not inside
            continue
        if inside:
            out.append(r)
    return out


def per_kernel(path: str) -> dict:
    """kernel key -> (dispatches, summed counter value in KiB), marked phases only."""
    out = defaultdict(lambda: [0, 0.0])
    for r in marked_rows(path):
        k = kernel_key(r["Kernel_Name"])
        out[k][0] += 1
        out[k][1] += float(r["Counter_Value"])
    return out


def kernel_stats(trace_csv: str) -> list:
    """Per kernel over the marked phases of a --kernel-trace run: (name,
    calls, total ms, average us, min us, max us), by total time."""
    acc = defaultdict(list)
    for r in marked_rows(trace_csv):
        acc[r["Kernel_Name"].split("(")[0]].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    out = [(k, len(v), sum(v) / 1e3, sum(v) / len(v), min(v), max(v)) for k, v in acc.items()]
    return sorted(out, key=lambda x: -x[2])


def calibration(path: str, nbytes: int) -> float:
    """known bytes / counted bytes for the copy kernels of a --calibrate-copy run."""
    vals = []
    with open(path) as f:
        for r in csv.DictReader(f):
            if "copyBuffer" in r["Kernel_Name"] or "copy" in r["Kernel_Name"].lower():
                vals.append(float(r["Counter_Value"]) * 1024.0)
    if not vals:
        raise ValueError(f"no copy kernel in {path}")
    return nbytes / (sum(vals) / len(vals))


def families_of(bench_json: str) -> dict:
    """Kernel families (ms, launches, alg_bytes) of a bench run's JSON line:
    the headline's, or a c2-only run's summed over its region counts."""
    with open(bench_json) as f:
        line = json.loads([l for l in f if l.startswith("{")][-1])
    if "kernel_families" in line:
        return line["kernel_families"]
    fam = {}
    for rec in line["c2_grid"]["per_R"]:
        for k, st in rec["kernel_families"].items():
            a = fam.setdefault(k, {"ms": 0.0, "launches": 0, "alg_bytes": 0.0})
            for f in a:
                a[f] += st[f]
    return fam


def summarize(fetch_csv, write_csv, cfetch_csv, cwrite_csv, copy_bytes, fetch_json, write_json, command, out_json):
    """Per family launch: FETCH_SIZE bytes over the FETCH run's launches plus
    WRITE_SIZE bytes over the WRITE run's (each run's own family table), and
    algorithmic bytes per launch averaged over the two runs."""
    fs = calibration(cfetch_csv, copy_bytes)
    ws = calibration(cwrite_csv, copy_bytes)
    F, W = per_kernel(fetch_csv), per_kernel(write_csv)
    famF, famW = families_of(fetch_json), families_of(write_json)
    kernels = {}
    fetch_b, write_b = defaultdict(float), defaultdict(float)
    for k in sorted(set(F) | set(W)):
        fb = F[k][1] * 1024.0 * fs
        wb = W[k][1] * 1024.0 * ws
        kernels[k] = {"family": FAMILY_OF.get(k), "dispatches_fetch_run": F[k][0], "dispatches_write_run": W[k][0],
                      "fetch_bytes_per_dispatch": fb / max(1, F[k][0]),
                      "write_bytes_per_dispatch": wb / max(1, W[k][0])}
        if k in FAMILY_OF:
            fetch_b[FAMILY_OF[k]] += fb
            write_b[FAMILY_OF[k]] += wb
    families = {}
    for name in famF:
        lf, lw = famF[name]["launches"], famW.get(name, {}).get("launches", 0)
        if lf > 0 and lw > 0:
            alg = 0.5 * (famF[name]["alg_bytes"] / lf + famW[name]["alg_bytes"] / lw)
            families[name] = {"launches_fetch_run": lf, "launches_write_run": lw,
                              "fetch_bytes_per_launch": fetch_b.get(name, 0.0) / lf,
                              "write_bytes_per_launch": write_b.get(name, 0.0) / lw,
                              "hbm_bytes_per_launch": fetch_b.get(name, 0.0) / lf + write_b.get(name, 0.0) / lw,
                              "alg_bytes_per_launch": alg}
    out = {"command": command, "fetch_scale": fs, "write_scale": ws, "copy_bytes": copy_bytes,
           "families": families, "kernels": kernels}
    with open(out_json, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    return out


def traffic_for(workload: str, family: str, root: str | None = None, full: bool = False):
    """Measured HBM bytes per launch of `family` for `workload` (with
    full=True the family's whole record of the counter runs: HBM and
    algorithmic bytes per launch, launches), or None."""
    root = root or os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "profiles")
    path = os.path.join(root, f"{workload}_hbm_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    fam = d.get("families", {}).get(family)
    if fam is None:
        return None
    return fam if full else fam["hbm_bytes_per_launch"]
