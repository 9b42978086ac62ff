"""Summaries of rocprofv3 CSV output for the profiles/ folder.

    python -m gary_amd.profiling kernels <kernel_stats.csv> --cmd "<command>" > profiles/X.md
    python -m gary_amd.profiling counters <counter_collection.csv>... --kernel K --cmd "<commands>"
    python -m gary_amd.profiling traffic <fetch_counter_collection.csv> <write_counter_collection.csv>
        --kernel tc_pivot_kernel<4 --copy-bytes N --cmd "<command>" > profiles/Y.md

The traffic summary scales the kernel's FETCH_SIZE / WRITE_SIZE (KiB per dispatch) by the
ratio bytes / counter that the same counters read for a device-to-device copy of known size
(the `__amd_rocclr_copyBuffer` dispatch), and ends with a machine-readable line
`traffic_bytes_per_launch: N` that bench.py reads.
"""
from __future__ import annotations

import argparse
import collections
import csv
import re
import sys
from pathlib import Path


def kernel_table(stats_csv: str) -> str:
    rows = list(csv.DictReader(open(stats_csv)))
    out = ["| kernel | calls | avg (us) | min (us) | max (us) | share |", "|---|---|---|---|---|---|"]
    for r in rows:
        name = r["Name"]
        short = re.sub(r"\([^()]*\)$", "", name.replace("(anonymous namespace)", "{anon}"))
        short = short.replace("void ", "")
        out.append(f"| `{short}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                   f"{float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | "
                   f"{float(r['Percentage']):.2f}% |")
    return "\n".join(out)


def _per_dispatch(csv_path: str, counter: str) -> dict[str, list[float]]:
    agg = collections.defaultdict(list)
    for r in csv.DictReader(open(csv_path)):
        if r["Counter_Name"] == counter:
            agg[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return agg


def traffic(fetch_csv: str, write_csv: str, kernel: str, copy_bytes: int) -> tuple[str, int]:
    f = _per_dispatch(fetch_csv, "FETCH_SIZE")
    w = _per_dispatch(write_csv, "WRITE_SIZE")

    def pick(agg, pat):
        ks = [k for k in agg if pat in k]
        if not ks:
            raise SystemExit(f"no dispatch of {pat!r}")
        return ks[0], agg[ks[0]]

    _, fc = pick(f, "copyBuffer")
    _, wc = pick(w, "copyBuffer")
    fcopy, wcopy = max(fc), max(wc)  # the calibration copy is the largest copy dispatch
    f_scale = copy_bytes / (fcopy * 1024.0)
    w_scale = copy_bytes / (wcopy * 1024.0)
    kname, fk = pick(f, kernel)
    _, wk = pick(w, kernel)
    fk_avg = sum(fk) / len(fk)
    wk_avg = sum(wk) / len(wk)
    t = int(round(fk_avg * 1024 * f_scale + wk_avg * 1024 * w_scale))
    lines = [
        f"kernel: `{re.sub(r'[(][^()]*[)]$', '', kname)}` ({len(fk)} dispatches)",
        "",
        "| counter | kernel KiB/dispatch | copy KiB ({:.0f} MiB copy) | scale (bytes per counted byte) |"
        .format(copy_bytes / 2**20),
        "|---|---|---|---|",
        f"| FETCH_SIZE | {fk_avg:.0f} | {fcopy:.0f} | {f_scale:.3f} |",
        f"| WRITE_SIZE | {wk_avg:.0f} | {wcopy:.0f} | {w_scale:.3f} |",
        "",
        f"HBM-side bytes per launch = FETCH*scale + WRITE*scale = {t / 1e9:.3f} GB",
        "",
        f"traffic_bytes_per_launch: {t}",
    ]
    return "\n".join(lines), t


def counters_table(csvs: list[str], kernel: str) -> str:
    """Average per dispatch of every counter found in the given counter_collection CSVs, for the
    kernels whose name contains `kernel`."""
    rows = []
    for path in csvs:
        agg = collections.defaultdict(list)
        for r in csv.DictReader(open(path)):
            if kernel in r["Kernel_Name"]:
                agg[r["Counter_Name"]].append(float(r["Counter_Value"]))
        for name, vals in sorted(agg.items()):
            rows.append(f"| {name} | {len(vals)} | {sum(vals) / len(vals):.0f} |")
    return "\n".join(["| counter | dispatches | average per dispatch |", "|---|---|---|", *rows])


def read_traffic(md_path: str | Path) -> int | None:
    p = Path(md_path)
    if not p.exists():
        return None
    m = re.search(r"^traffic_bytes_per_launch:\s*(\d+)", p.read_text(), re.M)
    return int(m.group(1)) if m else None


def read_commit(md_path: str | Path) -> str | None:
    """The `source commit:` line of a summary (the tree the profiled command ran)."""
    p = Path(md_path)
    if not p.exists():
        return None
    m = re.search(r"^source commit:\s*(\S+)", p.read_text(), re.M)
    return m.group(1) if m else None


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="what", required=True)
    k = sub.add_parser("kernels")
    k.add_argument("stats_csv")
    k.add_argument("--cmd", required=True)
    k.add_argument("--commit", default=None, help="commit of the profiled tree")
    k.add_argument("--title", default="rocprofv3 kernel trace")
    c = sub.add_parser("counters")
    c.add_argument("csvs", nargs="+")
    c.add_argument("--kernel", required=True)
    c.add_argument("--cmd", required=True)
    c.add_argument("--commit", default=None, help="commit of the profiled tree")
    c.add_argument("--title", default="PMC counters per dispatch")
    t = sub.add_parser("traffic")
    t.add_argument("fetch_csv")
    t.add_argument("write_csv")
    t.add_argument("--kernel", required=True)
    t.add_argument("--copy-bytes", type=int, required=True)
    t.add_argument("--cmd", required=True)
    t.add_argument("--commit", default=None, help="commit of the profiled tree")
    t.add_argument("--title", default="HBM traffic per launch")
    a = ap.parse_args(argv)
    if a.commit:
        a.cmd = f"{a.cmd}\n\nsource commit: {a.commit}"
    if a.what == "kernels":
        print(f"# {a.title}\n\ncommand:\n\n    {a.cmd}\n\n{kernel_table(a.stats_csv)}")
    elif a.what == "counters":
        print(f"# {a.title}\n\ncommands (one PMC pass each):\n\n    {a.cmd}\n\n"
              f"kernel: `{a.kernel}`\n\n{counters_table(a.csvs, a.kernel)}")
    else:
        body, _ = traffic(a.fetch_csv, a.write_csv, a.kernel, a.copy_bytes)
        print(f"# {a.title}\n\ncommands:\n\n    {a.cmd}\n\n{body}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
