"""HBM traffic per kernel-family launch from rocprofv3 counter runs.

Inputs: the counter_collection.csv files of four ``rocprofv3 --pmc`` runs
(FETCH_SIZE and WRITE_SIZE of the bench, FETCH_SIZE and WRITE_SIZE of
``bench.py --calibrate-copy BYTES``) and the bench's JSON line from the
counter runs.  Each counter is scaled by known bytes / counted bytes of the
calibration copy, summed over the kernels of a family and divided by that
family's launches in the same run (the bench run with --profile-all, so the family
launch counts cover every pass the counters saw, warm-up included).

Output: a JSON file per workload that bench.py reads for roofline.traffic.
"""
from __future__ import annotations

import csv
import json
import os
import re
from collections import defaultdict

# kernel name (demangled, template arguments dropped) -> engine family
FAMILY_OF = {
    "k_init": "init", "k_preflow": "init",
    "k_gr_seed": "bfs", "k_bfs_small": "bfs", "k_bfs_td": "bfs", "k_bfs_td_heavy": "bfs", "k_bfs_td_giant": "bfs",
    "k_bfs_advance": "bfs", "k_bfs_seed1": "bfs", "k_local_mark": "bfs", "k_local_bound": "bfs",
    "k_local_bucket": "bfs", "k_local_bfs": "bfs", "k_local_finish": "bfs", "k_mask_combine": "bfs",
    "k_bfs_bu": "bfs_bu", "k_bfs_bu_med": "bfs_bu", "k_bfs_bu_heavy": "bfs_bu",
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


def per_kernel(path: str) -> dict:
    """kernel key -> (dispatches, summed counter value in KiB)."""
    out = defaultdict(lambda: [0, 0.0])
    with open(path) as f:
        for r in csv.DictReader(f):
            k = kernel_key(r["Kernel_Name"])
            out[k][0] += 1
            out[k][1] += float(r["Counter_Value"])
    return out


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


def summarize(fetch_csv, write_csv, cfetch_csv, cwrite_csv, copy_bytes, bench_json, command, out_json):
    fs = calibration(cfetch_csv, copy_bytes)
    ws = calibration(cwrite_csv, copy_bytes)
    F, W = per_kernel(fetch_csv), per_kernel(write_csv)
    with open(bench_json) as f:
        line = [l for l in f if l.startswith("{")][-1]
    fam = json.loads(line)["kernel_families"]
    kernels = {}
    fam_bytes = defaultdict(float)
    for k in sorted(set(F) | set(W)):
        nd = max(F[k][0], W[k][0])
        fb = F[k][1] * 1024.0 * fs
        wb = W[k][1] * 1024.0 * ws
        kernels[k] = {"family": FAMILY_OF.get(k), "dispatches": nd,
                      "fetch_bytes_per_dispatch": fb / max(1, F[k][0]),
                      "write_bytes_per_dispatch": wb / max(1, W[k][0])}
        if k in FAMILY_OF:
            fam_bytes[FAMILY_OF[k]] += fb + wb
    families = {}
    for name, st in fam.items():
        if st["launches"] > 0:
            families[name] = {"launches": st["launches"],
                              "hbm_bytes_per_launch": fam_bytes.get(name, 0.0) / st["launches"],
                              "alg_bytes_per_launch": st["alg_bytes"] / st["launches"],
                              "ms_per_launch": st["ms"] / st["launches"]}
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
