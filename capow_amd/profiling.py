"""Summaries of rocprofv3 output for the P_l(k) path (measurement support).

Two inputs, both written by rocprofv3 with ``--output-format csv``:

* ``*_kernel_stats.csv`` (``--kernel-trace --stats``): per-kernel call
  counts and durations -> a markdown table;
* ``*_counter_collection.csv`` (``--pmc ...``): one row per (dispatch,
  counter) -> per-stage values of the last profiled step, with FETCH_SIZE /
  WRITE_SIZE converted to bytes through the calibration copy.

Calibration: FETCH_SIZE and WRITE_SIZE count in units whose relation to HBM
bytes depends on the GPU (on gfx950 FETCH_SIZE reads about half the bytes a
kernel streams).  The profiling pass of bench.py copies a buffer of known
size with hipMemcpy (kernel ``__amd_rocclr_copyBuffer``); the largest such
dispatch gives bytes-per-count factors for reads and writes.

Usage:
    python -m capow_amd.profiling stats <kernel_stats.csv> [--title T] [--command C]
    python -m capow_amd.profiling traffic <fetch counter csv> <write counter csv>
        --copy-bytes B [--command C]
"""
import argparse
import csv
import json
import re
import sys

# stage names of one step, in launch order (capow_power)
STAGE_PATTERNS = [
    # k_bucket<T, ORDER, SCATTER, EDGE[, IDX]>: the third argument is the pass
    ("bucket_count", r"k_bucket<[^,]+, \d+, false"),
    ("sort_clear", r"k_zero_u32"),
    ("stats_final", r"k_stats_final"),
    ("scan", r"k_scan_"),
    ("bucket_scatter", r"k_bucket<[^,]+, \d+, true"),
    ("deposit", r"k_deposit"),
    # the x pass fused with the binning (FP32), its quanta and final sums
    ("fft_x_bin", r"k_fft_x_bin"),
    ("xbin_scale", r"k_xbin_scale"),
    ("bin_final", r"k_xbin_final"),
    ("bin_counts_clear", r"k_fill_zero"),
    # transposed pipeline: the r2c row pass (5th template argument true) is
    # z, the two c2c passes y then x
    ("fft_yx", r"k_fft_rows_t<[^,]+, \d+, \d+, \d+, false"),
    ("fft_z", r"k_fft_rows"),
    ("fft_yx", r"k_fft_cols"),
    ("bin_counts", r"k_bin_counts"),
    ("bin", r"k_bin(_rows|_mirror)?<"),
    ("bin_final", r"k_bin_final"),
]


def _col(row, *names):
    low = {k.lower(): v for k, v in row.items()}
    for n in names:
        if n.lower() in low:
            return low[n.lower()]
    raise KeyError(names)


def _short(name, width=70):
    name = re.sub(r"^void ", "", name)
    name = re.sub(r"\(.*$", "", name)
    return name if len(name) <= width else name[:width]


def stats_table(path):
    rows = []
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            rows.append(dict(
                name=_short(_col(r, "Name", "KernelName", "Kernel_Name")),
                calls=int(_col(r, "Calls")),
                total=float(_col(r, "TotalDurationNs")) / 1e6,
                avg=float(_col(r, "AverageNs")) / 1e6,
                mn=float(_col(r, "MinNs")) / 1e6,
                mx=float(_col(r, "MaxNs")) / 1e6,
                pct=float(_col(r, "Percentage")),
            ))
    rows.sort(key=lambda r: -r["total"])
    out = ["| kernel | calls | total ms | avg ms | min ms | max ms | % |",
           "|---|---|---|---|---|---|---|"]
    for r in rows:
        out.append("| %s | %d | %.3f | %.3f | %.3f | %.3f | %.1f |" % (
            r["name"], r["calls"], r["total"], r["avg"], r["mn"], r["mx"], r["pct"]))
    return rows, "\n".join(out)


def counter_dispatches(path, counter):
    """[(dispatch_id, kernel_name, value)] in dispatch order (values summed
    over the counter's instances)."""
    acc = {}
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            if _col(r, "Counter_Name", "CounterName") != counter:
                continue
            d = int(_col(r, "Dispatch_Id", "DispatchId"))
            name = _col(r, "Kernel_Name", "KernelName")
            v = float(_col(r, "Counter_Value", "CounterValue"))
            if d in acc:
                acc[d][1] += v
            else:
                acc[d] = [name, v]
    return [(d, acc[d][0], acc[d][1]) for d in sorted(acc)]


def last_step(disp):
    """The capow dispatches of the last step, labelled by stage."""
    caps = [(d, n, v) for d, n, v in disp if "capow::" in n]
    starts = [i for i, (_, n, _) in enumerate(caps) if re.search(STAGE_PATTERNS[0][1], n)]
    if not starts:
        return []
    step = caps[starts[-1]:]
    out = []
    nfft = 0
    for d, n, v in step:
        label = None
        for lab, pat in STAGE_PATTERNS:
            if re.search(pat, n):
                label = lab
                break
        if label == "fft_yx":
            label = "fft_y" if nfft == 0 else "fft_x"
            nfft += 1
        out.append((label or _short(n), n, v))
    return out


def copy_factor(disp, copy_bytes):
    """Bytes per counted unit (KiB) from the largest copyBuffer dispatch."""
    vals = [v for _, n, v in disp if "copyBuffer" in n]
    if not vals or max(vals) <= 0:
        return None
    return copy_bytes / max(vals)


def traffic(fetch_csv, write_csv, copy_bytes):
    fd = counter_dispatches(fetch_csv, "FETCH_SIZE")
    wd = counter_dispatches(write_csv, "WRITE_SIZE")
    ff, wf = copy_factor(fd, copy_bytes), copy_factor(wd, copy_bytes)
    fs, ws = last_step(fd), last_step(wd)
    rows = []
    merged = {}
    for (lab, n, fv), (_, _, wv) in zip(fs, ws):
        rb, wb = fv * ff, wv * wf
        key = lab
        q = 2
        while key in merged:
            key = "%s#%d" % (lab, q)
            q += 1
        merged[key] = rb + wb
        rows.append((key, _short(n, 48), fv, wv, rb, wb))
    return ff, wf, rows, merged


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("--title", default="Kernel trace")
    s.add_argument("--command", default="")
    t = sub.add_parser("traffic")
    t.add_argument("fetch_csv")
    t.add_argument("write_csv")
    t.add_argument("--copy-bytes", type=float, required=True)
    t.add_argument("--command", default="")
    a = ap.parse_args(argv)
    if a.cmd == "stats":
        _, table = stats_table(a.csv)
        print("# %s\n" % a.title)
        if a.command:
            print("Command:\n\n```\n%s\n```\n" % a.command)
        print(table)
        return
    ff, wf, rows, merged = traffic(a.fetch_csv, a.write_csv, a.copy_bytes)
    print("# HBM traffic per launch (FETCH_SIZE / WRITE_SIZE, calibrated)\n")
    if a.command:
        print("Commands:\n\n```\n%s\n```\n" % a.command)
    print("Calibration: a %.0f-byte hipMemcpy (copyBuffer) reads %.4g counted KiB of "
          "FETCH_SIZE and writes %.4g of WRITE_SIZE: %.1f bytes per FETCH_SIZE unit, "
          "%.1f bytes per WRITE_SIZE unit.\n" % (a.copy_bytes, a.copy_bytes / ff,
                                                a.copy_bytes / wf, ff, wf))
    print("| stage | kernel | FETCH_SIZE | WRITE_SIZE | read GB | written GB |")
    print("|---|---|---|---|---|---|")
    for key, n, fv, wv, rb, wb in rows:
        print("| %s | %s | %.4g | %.4g | %.3f | %.3f |" % (key, n, fv, wv, rb / 1e9, wb / 1e9))
    print("\n```json")
    print(json.dumps({"traffic_bytes": {k: round(v) for k, v in merged.items()}}))
    print("```")


if __name__ == "__main__":
    sys.exit(main())
