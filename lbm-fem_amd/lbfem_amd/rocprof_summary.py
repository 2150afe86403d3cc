"""Short text summaries of rocprofv3 CSV output (kernel statistics, counters,
HBM traffic per launch), the form in which profiles are kept with the sources.

    python -m lbfem_amd.rocprof_summary stats DIR
    python -m lbfem_amd.rocprof_summary counters DIR [--kernel REGEX]
    python -m lbfem_amd.rocprof_summary traffic FETCH_DIR WRITE_DIR --kernel REGEX --copy-bytes B

traffic: FETCH_SIZE and WRITE_SIZE (KiB per dispatch, summed over the L2
channels) of the kernels matching REGEX, each scaled by the factor that makes
the copy kernel k_stream_copy of B bytes read and write exactly B bytes.
"""
import argparse
import collections
import csv
import glob
import os
import re
import sys


def short(name):
    name = re.sub(r"^void ", "", name)
    name = name.replace("lbfem_amd::tile::", "").replace("lbfem_amd::", "")
    depth, out = 0, []
    for ch in name:  # drop the argument list
        if ch == "(" and depth == 0:
            break
        depth += ch == "<"
        depth -= ch == ">"
        out.append(ch)
    return "".join(out).strip()


def _csv(d, suffix):
    files = sorted(glob.glob(os.path.join(d, "**", f"*{suffix}"), recursive=True))
    if not files:
        sys.exit(f"no *{suffix} under {d}")
    return files[0]


def stats(d):
    rows = list(csv.DictReader(open(_csv(d, "kernel_stats.csv"))))
    print(f"{'kernel':72s} {'calls':>7s} {'avg_us':>10s} {'total_ms':>10s} {'pct':>6s}")
    for r in rows:
        print(f"{short(r['Name'])[:72]:72s} {int(r['Calls']):7d} {float(r['AverageNs']) / 1e3:10.3f} "
              f"{float(r['TotalDurationNs']) / 1e6:10.3f} {float(r['Percentage']):6.2f}")


def per_kernel(d):
    vals = collections.defaultdict(lambda: collections.defaultdict(list))
    for r in csv.DictReader(open(_csv(d, "counter_collection.csv"))):
        vals[short(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    return vals


def counters(d, kernel):
    for k, cs in sorted(per_kernel(d).items()):
        if kernel and not re.search(kernel, k):
            continue
        n = max(len(v) for v in cs.values())
        print(f"{k}  ({n} dispatches; mean per dispatch)")
        for c, v in sorted(cs.items()):
            print(f"    {c:28s} {sum(v) / len(v):16.1f}")


def traffic(fetch_dir, write_dir, kernel, copy_bytes):
    fk, wk = per_kernel(fetch_dir), per_kernel(write_dir)
    mean = lambda v: sum(v) / len(v)
    cf = mean(fk["k_stream_copy"]["FETCH_SIZE"])
    cw = mean(wk["k_stream_copy"]["WRITE_SIZE"])
    sf, sw = copy_bytes / 1024 / cf, copy_bytes / 1024 / cw
    print(f"copy kernel k_stream_copy of {copy_bytes} B: FETCH_SIZE {cf:.1f} KiB, WRITE_SIZE {cw:.1f} KiB "
          f"-> scale {sf:.4f} (fetch), {sw:.4f} (write)")
    for k in sorted(fk):
        if not re.search(kernel, k) or k not in wk:
            continue
        f, w = mean(fk[k]["FETCH_SIZE"]), mean(wk[k]["WRITE_SIZE"])
        t = (f * sf + w * sw) * 1024
        print(f"{k}: FETCH_SIZE {f:.1f} KiB, WRITE_SIZE {w:.1f} KiB per dispatch "
              f"({len(fk[k]['FETCH_SIZE'])} / {len(wk[k]['WRITE_SIZE'])} dispatches)")
        print(f"traffic_bytes_per_launch[{k}]: {t:.0f}")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("stats")
    a.add_argument("dir")
    a = sub.add_parser("counters")
    a.add_argument("dir")
    a.add_argument("--kernel", default="")
    a = sub.add_parser("traffic")
    a.add_argument("fetch_dir")
    a.add_argument("write_dir")
    a.add_argument("--kernel", required=True)
    a.add_argument("--copy-bytes", type=int, required=True)
    args = ap.parse_args(argv)
    if args.cmd == "stats":
        stats(args.dir)
    elif args.cmd == "counters":
        counters(args.dir, args.kernel)
    else:
        traffic(args.fetch_dir, args.write_dir, args.kernel, args.copy_bytes)


if __name__ == "__main__":
    main()
