"""rocprofv3 summaries: HBM traffic per launch from two counter runs, and a
markdown table of a ``--kernel-trace --stats`` run.

FETCH_SIZE and WRITE_SIZE (KiB per dispatch) are collected in separate
``rocprofv3 --pmc`` passes over ``bench.py --copy-probe``; the probe's plain
copy kernel (``k_copy``, known bytes read and written) calibrates each counter
(factor = known bytes / counter bytes), and the step kernel's average is scaled
by that factor.

    python -m muphfasa_amd.pmc FETCH_CSV WRITE_CSV --n 256 --out profiles/...json
    python -m muphfasa_amd.pmc --stats KERNEL_STATS_CSV
"""
from __future__ import annotations

import argparse
import csv
import json
from collections import defaultdict


def per_kernel(path, counter):
    """{kernel name: [values per dispatch]} for one counter."""
    out = defaultdict(list)
    with open(path, newline="") as fh:
        for row in csv.DictReader(fh):
            if row["Counter_Name"] == counter:
                out[row["Kernel_Name"]].append(float(row["Counter_Value"]))
    return out


def _pick(d, needle):
    names = [k for k in d if needle in k]
    if not names:
        raise SystemExit(f"no kernel matching {needle!r}; have {sorted(d)[:10]}")
    return max(names, key=lambda k: len(d[k]))


def summarize(fetch_csv, write_csv, n, Q=19, elem=8, step="k_step", copy="k_copy",
              copy_bytes=None, commands=(), steps_per_launch=1, workload=None, commit=None,
              cells=None):
    """cells: lattice cells one launch updates (default n^3).  The algorithmic
    bytes count one read and one write of every population per step; a
    launch of two steps that keeps the intermediate state on chip needs at
    least one read and one write per launch (minimum_bytes_per_launch)."""
    fe = per_kernel(fetch_csv, "FETCH_SIZE")
    wr = per_kernel(write_csv, "WRITE_SIZE")
    ks, kc = _pick(fe, step), _pick(fe, copy)
    cells = n ** 3 if cells is None else cells
    algo = steps_per_launch * 2 * Q * elem * cells
    cb = copy_bytes if copy_bytes is not None else Q * elem * cells
    mean = lambda v: sum(v) / len(v)
    # factor = known bytes / counter bytes for the copy kernel
    ff = cb / (mean(fe[kc]) * 1024)
    wf = cb / (mean(wr[_pick(wr, copy)]) * 1024)
    rd = mean(fe[ks]) * 1024 * ff
    wt = mean(wr[_pick(wr, step)]) * 1024 * wf
    return {
        "kernel": ks, "workload": workload or f"D3Q19 {n}^3 FP64 forced periodic (bench.py)",
        "steps_per_launch": steps_per_launch,
        "fetch_size_kib_per_dispatch": mean(fe[ks]),
        "write_size_kib_per_dispatch": mean(wr[_pick(wr, step)]),
        "dispatches": len(fe[ks]),
        "calibration": {"kernel": kc, "known_bytes_read": cb, "known_bytes_written": cb,
                        "fetch_size_kib": mean(fe[kc]), "write_size_kib": mean(wr[_pick(wr, copy)]),
                        "fetch_factor": ff, "write_factor": wf},
        "hbm_read_bytes_per_launch": rd, "hbm_write_bytes_per_launch": wt,
        "hbm_bytes_per_launch": rd + wt, "algorithmic_bytes_per_launch": algo,
        "minimum_bytes_per_launch": 2 * Q * elem * cells,
        "traffic_over_algorithmic": (rd + wt) / algo, "commands": list(commands),
        "commit": commit,
    }


def stats_table(path, top=10, width=110):
    """Markdown table of a rocprofv3 ``*_kernel_stats.csv``."""
    rows = []
    with open(path, newline="") as fh:
        for r in csv.DictReader(fh):
            name = r["Name"].split("(")[0] if r["Name"].startswith("void ") else r["Name"]
            rows.append((name[:width], int(r["Calls"]), float(r["AverageNs"]) / 1e3,
                         float(r["MinNs"]) / 1e3, float(r["MaxNs"]) / 1e3, float(r["Percentage"])))
    rows.sort(key=lambda t: -t[5])
    out = ["| Kernel | Calls | Average (us) | Min (us) | Max (us) | % of GPU time |",
           "|---|---|---|---|---|---|"]
    out += [f"| `{n}` | {c} | {a:.1f} | {lo:.1f} | {hi:.1f} | {pc:.2f} |"
            for n, c, a, lo, hi, pc in rows[:top]]
    return "\n".join(out)


def main():
    p = argparse.ArgumentParser()
    p.add_argument("fetch_csv", nargs="?")
    p.add_argument("write_csv", nargs="?")
    p.add_argument("--stats", default="")
    p.add_argument("--top", type=int, default=10)
    p.add_argument("--n", type=int, default=256)
    p.add_argument("--out", default="")
    p.add_argument("--command", action="append", default=[])
    p.add_argument("--step", default="k_step", help="name fragment of the measured kernel")
    p.add_argument("--steps-per-launch", type=int, default=1)
    p.add_argument("--workload", default=None)
    p.add_argument("--commit", default=None, help="source commit the counters were measured at")
    p.add_argument("--cells", type=int, default=None, help="cells per launch (default n^3)")
    p.add_argument("--q", type=int, default=19)
    p.add_argument("--elem", type=int, default=8, help="bytes per stored population")
    p.add_argument("--copy", default="k_copy", help="name fragment of the calibration copy kernel")
    p.add_argument("--copy-bytes", type=int, default=None,
                   help="bytes the calibration copy reads (and writes) per launch")
    a = p.parse_args()
    if a.stats:
        print(stats_table(a.stats, a.top))
        return
    r = summarize(a.fetch_csv, a.write_csv, a.n, Q=a.q, elem=a.elem, step=a.step, copy=a.copy,
                  copy_bytes=a.copy_bytes, commands=a.command, steps_per_launch=a.steps_per_launch,
                  workload=a.workload, commit=a.commit, cells=a.cells)
    s = json.dumps(r, indent=1)
    if a.out:
        with open(a.out, "w") as fh:
            fh.write(s + "\n")
    print(s)


if __name__ == "__main__":
    main()
