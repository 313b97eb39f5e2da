"""Helpers for rocprofv3 measurements of the sigma kernel.

    python -m dice_amd.profiling copy [--mib N]
        runs an elementwise kernel that reads N MiB (default 1024) and writes
        N MiB, five times: the calibration program for the FETCH_SIZE /
        WRITE_SIZE counters (known bytes read and written per dispatch).

    python -m dice_amd.profiling stats <kernel_stats.csv> [--top N]
        per-kernel table (calls, total ms, average us, share) of a
        `rocprofv3 --kernel-trace --stats` run.

    python -m dice_amd.profiling traffic --kernel NAME --fetch DIR --write DIR
            --copy-fetch DIR --copy-write DIR [--copy-mib N] [--out FILE]
        HBM bytes per dispatch of NAME from a FETCH_SIZE run and a WRITE_SIZE
        run, each scaled by bytes / counter of the copy kernel measured the
        same way; writes a JSON summary.

    python -m dice_amd.profiling requests --kernel NAME --read DIR --write DIR
        the same bytes counted from the L2's memory-side requests by size:
        read = 128 TCC_EA0_RDREQ_128B + 32 TCC_EA0_RDREQ_32B + 64 (the rest)
        from a run of those counters (and TCC_EA0_RDREQ); write = 64
        TCC_EA0_WRREQ_64B + 32 (the rest).  An independent check of the
        FETCH_SIZE scale: on gfx950 FETCH_SIZE counts 128-byte reads through
        TCC_BUBBLE, which stays 0, so each is counted as 64 bytes.
"""
import argparse
import csv
import glob
import json
import os
import sys


def _counter_rows(directory):
    files = glob.glob(os.path.join(directory, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise FileNotFoundError("no *counter_collection.csv under %s" % directory)
    rows = []
    for f in files:
        with open(f) as fh:
            rows += list(csv.DictReader(fh))
    return rows


def counter_per_dispatch(directory, kernel_substr, counter):
    """Mean counter value per dispatch over kernels whose name contains kernel_substr."""
    per = {}
    for r in _counter_rows(directory):
        if r.get("Counter_Name") != counter or kernel_substr not in r.get("Kernel_Name", ""):
            continue
        key = r.get("Dispatch_Id") or r.get("Correlation_Id") or str(len(per))
        per[key] = per.get(key, 0.0) + float(r["Counter_Value"])
    if not per:
        raise ValueError("no %s samples for kernels matching %r in %s" % (counter, kernel_substr, directory))
    vals = list(per.values())
    return sum(vals) / len(vals), len(vals)


def copy_main(mib):
    import torch
    n = mib * (1 << 20) // 8
    # filled on the host and copied: the only kernels are the five products
    a = torch.ones(n, dtype=torch.float64).to("cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        torch.mul(a, 2.0, out=b)
    torch.cuda.synchronize()
    print("copied %d MiB x5" % mib)


def stats_main(path, top):
    with open(path) as fh:
        rows = list(csv.DictReader(fh))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    print("%-70s %7s %12s %12s %7s" % ("kernel", "calls", "total ms", "avg us", "share"))
    for r in rows[:top]:
        print("%-70s %7s %12.2f %12.1f %6.1f%%" % (r["Name"][:70], r["Calls"], float(r["TotalDurationNs"]) / 1e6,
                                                   float(r["AverageNs"]) / 1e3, float(r["Percentage"])))


def traffic_main(a):
    copy_bytes = a.copy_mib * (1 << 20)
    # FETCH_SIZE and WRITE_SIZE are reported in KiB
    cf, _ = counter_per_dispatch(a.copy_fetch, a.copy_kernel, "FETCH_SIZE")
    cw, _ = counter_per_dispatch(a.copy_write, a.copy_kernel, "WRITE_SIZE")
    kf, nf = counter_per_dispatch(a.fetch, a.kernel, "FETCH_SIZE")
    kw, nw = counter_per_dispatch(a.write, a.kernel, "WRITE_SIZE")
    f_fetch = copy_bytes / (cf * 1024.0)
    f_write = copy_bytes / (cw * 1024.0)
    out = {
        "kernel": a.kernel,
        "dispatches": {"fetch": nf, "write": nw},
        "fetch_kib_per_dispatch": kf,
        "write_kib_per_dispatch": kw,
        "copy_kernel": a.copy_kernel,
        "copy_bytes": copy_bytes,
        "copy_fetch_kib": cf,
        "copy_write_kib": cw,
        "scale_fetch": f_fetch,
        "scale_write": f_write,
        "traffic_bytes_per_dispatch": kf * 1024.0 * f_fetch + kw * 1024.0 * f_write,
    }
    text = json.dumps(out, indent=1)
    if a.out:
        with open(a.out, "w") as fh:
            fh.write(text + "\n")
    print(text)


def requests_main(a):
    r128, n = counter_per_dispatch(a.read, a.kernel, "TCC_EA0_RDREQ_128B_sum")
    r32, _ = counter_per_dispatch(a.read, a.kernel, "TCC_EA0_RDREQ_32B_sum")
    rall, _ = counter_per_dispatch(a.read, a.kernel, "TCC_EA0_RDREQ_sum")
    w64, nw = counter_per_dispatch(a.write, a.kernel, "TCC_EA0_WRREQ_64B_sum")
    wall, _ = counter_per_dispatch(a.write, a.kernel, "TCC_EA0_WRREQ_sum")
    read = 128.0 * r128 + 32.0 * r32 + 64.0 * (rall - r128 - r32)
    write = 64.0 * w64 + 32.0 * (wall - w64)
    out = {"kernel": a.kernel, "dispatches": {"read": n, "write": nw}, "read_bytes_per_dispatch": read,
           "write_bytes_per_dispatch": write, "traffic_bytes_per_dispatch": read + write,
           "rdreq": rall, "rdreq_128b": r128, "rdreq_32b": r32, "wrreq": wall, "wrreq_64b": w64}
    print(json.dumps(out, indent=1))


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m dice_amd.profiling")
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("copy")
    c.add_argument("--mib", type=int, default=1024)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("--top", type=int, default=25)
    t = sub.add_parser("traffic")
    t.add_argument("--kernel", required=True)
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--copy-fetch", required=True)
    t.add_argument("--copy-write", required=True)
    t.add_argument("--copy-kernel", default="MulFunctor")
    t.add_argument("--copy-mib", type=int, default=1024)
    t.add_argument("--out")
    q = sub.add_parser("requests")
    q.add_argument("--kernel", required=True)
    q.add_argument("--read", required=True)
    q.add_argument("--write", required=True)
    a = ap.parse_args(argv)
    if a.cmd == "copy":
        copy_main(a.mib)
    elif a.cmd == "stats":
        stats_main(a.csv, a.top)
    elif a.cmd == "requests":
        requests_main(a)
    else:
        traffic_main(a)


if __name__ == "__main__":
    sys.exit(main())
