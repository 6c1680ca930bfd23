#!/usr/bin/env python3
"""Summarise the rocprofv3 runs of profile_kernels.sh into markdown.

Writes OUT/summary/kernels_cornell1024_c1_job.md (per-kernel durations) and
OUT/summary/k_trace_cornell1024_c1_traffic.md (HBM bytes per launch of each
engine kernel from FETCH_SIZE and WRITE_SIZE, each scaled by the factor that
makes a plain 1 GiB device copy read and write exactly 1 GiB).
"""
import csv
import glob
import os
import sys
from collections import defaultdict

CMD = "rpd-statmc-denoiser_amd/profile_kernels.sh (rocprofv3 on: python3 bench.py --steps 1 --warmup 0 --no-cpu-baseline --no-quality --no-served)"
COPY_BYTES = 1 << 30


def rows(d, name):
    f = glob.glob(os.path.join(d, "**", name), recursive=True)
    if not f:
        raise SystemExit(f"missing {name} under {d}")
    with open(f[0]) as fh:
        return list(csv.DictReader(fh))


def short(k):
    return k.split("(")[0].replace("void ", "")


def counter_per_dispatch(d):
    per = defaultdict(list)
    for r in rows(d, "run_counter_collection.csv"):
        per[short(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return per


def main():
    out = sys.argv[1]
    summ = os.path.join(out, "summary")
    os.makedirs(summ, exist_ok=True)
    stats = rows(os.path.join(out, "trace"), "run_kernel_stats.csv")
    with open(os.path.join(summ, "kernels_cornell1024_c1_job.md"), "w") as f:
        f.write("# Kernel time of one cornell_box 1024x1024 job (16 + 3x4 spp, StatMC r10, var 3, aperture 0.14)\n\n")
        f.write(f"Command: `{CMD}`, step 1 (`--kernel-trace --stats`).\n\n")
        shapes = [short(r["Name"]) for r in stats if short(r["Name"]).startswith("rpd::k_statmc<")]
        if len(shapes) > 1:
            f.write(f"The run is the process's first job, so the StatMC window kernel runs twice in each of its "
                    f"{len(shapes)} tile shapes (the context times them, in order and then in reverse, and keeps "
                    f"the fastest for this radius and image size); a later job launches only the fastest.\n\n")
        f.write("| kernel | calls | total ms | avg us | % |\n|---|---|---|---|---|\n")
        for r in stats[:20]:
            f.write(f"| {short(r['Name'])} | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.3f} | "
                    f"{float(r['AverageNs']) / 1e3:.1f} | {float(r['Percentage']):.2f} |\n")
    fetch = counter_per_dispatch(os.path.join(out, "FETCH_SIZE"))
    write = counter_per_dispatch(os.path.join(out, "WRITE_SIZE"))
    cal_f = counter_per_dispatch(os.path.join(out, "cal_FETCH_SIZE"))
    cal_w = counter_per_dispatch(os.path.join(out, "cal_WRITE_SIZE"))
    copy_k = [k for k in cal_f if "copy" in k.lower() or "elementwise" in k.lower()]
    big = max(copy_k, key=lambda k: max(cal_f[k])) if copy_k else max(cal_f, key=lambda k: max(cal_f[k]))
    f_read = (COPY_BYTES / 1024.0) / (sum(cal_f[big]) / len(cal_f[big]))
    f_write = (COPY_BYTES / 1024.0) / (sum(cal_w[big]) / len(cal_w[big]))
    avg_ns = {short(r["Name"]): float(r["AverageNs"]) for r in stats}
    with open(os.path.join(summ, "k_trace_cornell1024_c1_traffic.md"), "w") as f:
        f.write("# HBM traffic per launch, cornell_box 1024x1024 job\n\n")
        f.write(f"Command: `{CMD}`, steps 2-3 (`--pmc FETCH_SIZE`, `--pmc WRITE_SIZE`, one run each) and the same "
                "counters on `python3 bench.py --calibrate-copy`.\n\n")
        f.write(f"Calibration kernel `{big}` (1 GiB read + 1 GiB written): "
                f"read factor {f_read:.4f}, write factor {f_write:.4f} (bytes = KiB counter x 1024 x factor).\n\n")
        f.write("| kernel | launches | read MB/launch | written MB/launch | avg us | GB/s |\n|---|---|---|---|---|---|\n")
        for k in sorted(fetch, key=lambda k: -sum(fetch[k])):
            if not k.startswith("rpd::"):
                continue
            rd = sum(fetch[k]) / len(fetch[k]) * 1024 * f_read
            wr = sum(write.get(k, [0])) / max(1, len(write.get(k, []))) * 1024 * f_write
            t = avg_ns.get(k)
            gbs = (rd + wr) / t if t else float("nan")
            f.write(f"| {k} | {len(fetch[k])} | {rd / 1e6:.1f} | {wr / 1e6:.1f} | "
                    f"{(t or 0) / 1e3:.1f} | {gbs:.0f} |\n")
        tr = [k for k in fetch if k.startswith("rpd::k_trace")][0]
        rd = sum(fetch[tr]) / len(fetch[tr]) * 1024 * f_read
        wr = sum(write[tr]) / len(write[tr]) * 1024 * f_write
        f.write(f"\ntraffic_bytes_per_launch: {rd + wr:.0f}\n")
    print(open(os.path.join(summ, "k_trace_cornell1024_c1_traffic.md")).read())


if __name__ == "__main__":
    main()
