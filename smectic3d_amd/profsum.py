"""Summaries of rocprofv3 CSV output (kernel statistics and hardware counters).

    python -m smectic3d_amd.profsum stats <..._kernel_stats.csv> --cmd "<command>" -o out.txt
    python -m smectic3d_amd.profsum pmc <..._counter_collection.csv> [...] --cmd "<command>" \
        -o out.txt [--json out.json] [--copy-bytes N]

`stats` turns the per-kernel statistics of a --kernel-trace --stats run into a short
table (calls, total and average duration in microseconds).  `pmc` averages each
counter per kernel over its dispatches; FETCH_SIZE and WRITE_SIZE (KiB per dispatch)
are converted to bytes, and when --copy-bytes gives the bytes one launch of the plain
copy kernel k_copy_f4 reads (and writes), that kernel's two counters calibrate one
factor each (counted bytes / moved bytes) that the other kernels' values are divided by.
"""
import argparse
import csv
import json
import sys
from collections import defaultdict


def _col(row, *names):
    for n in names:
        if n in row:
            return row[n]
    low = {k.lower(): k for k in row}
    for n in names:
        if n.lower() in low:
            return row[low[n.lower()]]
    raise KeyError(f"none of {names} in {list(row)}")


def short_name(name):
    """Kernel name without the argument list (template arguments kept)."""
    depth = 0
    for i, ch in enumerate(name):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            return name[:i].replace("void ", "", 1).strip()
    return name.replace("void ", "", 1).strip()


def stats(path, cmd, out):
    rows = list(csv.DictReader(open(path)))
    recs = []
    for r in rows:
        recs.append((short_name(_col(r, "Name", "KernelName", "Kernel_Name")), int(_col(r, "Calls")),
                     float(_col(r, "TotalDurationNs", "Total_Duration_Ns")),
                     float(_col(r, "AverageNs", "Average_Ns")), float(_col(r, "Percentage"))))
    recs.sort(key=lambda x: -x[2])
    with open(out, "w") as f:
        f.write("# rocprofv3 --kernel-trace --stats summary (MI355X, one GPU)\n")
        f.write(f"# command: {cmd}\n")
        f.write("# columns: calls, total_us, avg_us, percent, kernel\n")
        for name, calls, tot, avg, pct in recs:
            f.write(f"{calls:8d} {tot / 1e3:12.1f} {avg / 1e3:9.2f} {pct:6.2f}  {name}\n")


def pmc(paths, cmd, out, json_out=None, copy_bytes=None):
    acc = defaultdict(lambda: [0.0, 0])
    for p in paths:
        for r in csv.DictReader(open(p)):
            k = (short_name(_col(r, "Kernel_Name", "KernelName", "Name")), _col(r, "Counter_Name", "CounterName"))
            a = acc[k]
            a[0] += float(_col(r, "Counter_Value", "CounterValue"))
            a[1] += 1
    per = defaultdict(dict)
    for (kname, cname), (s, n) in acc.items():
        per[kname][cname] = (s / n, n)
    factor = {}
    if copy_bytes:
        for kname, cs in per.items():
            if "k_copy_f4" in kname:
                for cname in ("FETCH_SIZE", "WRITE_SIZE"):
                    if cname in cs:
                        factor[cname] = cs[cname][0] * 1024.0 / float(copy_bytes)
    res = {}
    with open(out, "w") as f:
        f.write("# rocprofv3 --pmc summary (MI355X, one GPU): counter averages per dispatch\n")
        f.write(f"# command: {cmd}\n")
        for cname, fac in sorted(factor.items()):
            f.write(f"# calibration: k_copy_f4 moves {copy_bytes} B read and {copy_bytes} B written per launch; "
                    f"{cname} counts {fac:.4f} x the bytes moved (divided out below)\n")
        f.write("# columns: dispatches, counter, average (KiB), calibrated MB, kernel\n")
        for kname in sorted(per):
            for cname, (avg, n) in sorted(per[kname].items()):
                b = avg * 1024.0 / factor.get(cname, 1.0)
                mb = b / 1e6 if cname in ("FETCH_SIZE", "WRITE_SIZE") else float("nan")
                f.write(f"{n:8d} {cname:12s} {avg:14.1f} {mb:10.2f}  {kname}\n")
                res.setdefault(kname, {})[cname + "_bytes"] = b
    if json_out:
        json.dump({"command": cmd, "copy_factors": factor, "kernels": res}, open(json_out, "w"), indent=1)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["stats", "pmc"])
    ap.add_argument("csv", nargs="+")
    ap.add_argument("--cmd", required=True)
    ap.add_argument("-o", "--out", required=True)
    ap.add_argument("--json")
    ap.add_argument("--copy-bytes", type=int)
    a = ap.parse_args(argv)
    if a.mode == "stats":
        stats(a.csv[0], a.cmd, a.out)
    else:
        pmc(a.csv, a.cmd, a.out, a.json, a.copy_bytes)


if __name__ == "__main__":
    sys.exit(main())
