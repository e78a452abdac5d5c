"""Summaries of rocprofv3 output for the profiles/ directory.

  python cella-nova_amd/prof_summary.py stats <run_kernel_stats.csv>
  python cella-nova_amd/prof_summary.py traffic <fetch run_counter_collection.csv> <write ...csv> [calib MiB] [--json DIR TAG]
  python cella-nova_amd/prof_summary.py requests <N> <C> <K> <counter csv> [<counter csv> ...]

`requests` averages L1/L2 request counters per kernel launch over any number of counter passes
and divides each kernel's algorithmic read and write bytes at an N x N grid with C channels and
K kernels (compact kernel spectra) by its L1-to-L2 read and write requests.

`traffic` averages FETCH_SIZE / WRITE_SIZE (KiB per dispatch) per kernel and scales them by what
the same counters read for the calibration copy (a device copy of `calib` MiB, dispatched as
__amd_rocclr_copyBuffer): factor = expected KiB / measured KiB.  With --json it also writes
DIR/traffic_<kernel>_<TAG>.json (bytes per launch) for bench.py.
"""
import collections
import csv
import json
import os
import sys

SHORT = {"k_row_fwd": "row_fwd_r2c", "k_col_fused": "col_fused", "k_col_segments": "col_fused",
         "k_row_inv_growth": "row_inv_growth",
         "k_row_inv_growth_pair": "row_inv_growth",
         "k_diffusion": "diffusion", "k_diffusion_adj": "diffusion_adj", "k_growth_adj_row": "growth_adj_row",
         "k_col_adj": "col_adj", "k_row_inv_add": "row_inv_add", "k_param_reduce": "param_reduce",
         "k_mse_grad": "mse_grad", "k_zero_gparam": "zero_gparam", "k_adam": "adam"}


def short(name):
    """Short name of one of this library's kernels (exact base name, template arguments dropped)."""
    i = name.find("cn::")
    if i < 0:
        return None
    base = name[i + 4:].split("<")[0].split("(")[0]
    return SHORT.get(base)


def variant(name):
    """Kernel name with its template arguments, e.g. k_col_fused<2048, 4, 16>."""
    base = name.split("(")[0].replace("void ", "").replace("cn::", "")
    return base


def stats(path):
    rows = list(csv.DictReader(open(path)))
    print(f"{'kernel':<72} {'calls':>6} {'avg_us':>9} {'min_us':>9} {'max_us':>9} {'pct':>6}")
    for r in rows:
        print(f"{r['Name'][:72]:<72} {int(r['Calls']):>6} {float(r['AverageNs']) / 1e3:>9.2f} "
              f"{float(r['MinNs']) / 1e3:>9.2f} {float(r['MaxNs']) / 1e3:>9.2f} {float(r['Percentage']):>6.1f}")


def _per_kernel(path, counter):
    agg = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] == counter:
            agg[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return agg


def traffic(fetch_csv, write_csv, calib_mib=256, json_dir=None, tag=None):
    f, w = _per_kernel(fetch_csv, "FETCH_SIZE"), _per_kernel(write_csv, "WRITE_SIZE")
    exp = calib_mib * 1024.0
    cf = [v for v in f.get("__amd_rocclr_copyBuffer", []) if v > 0.25 * exp]
    cw = [v for v in w.get("__amd_rocclr_copyBuffer", []) if v > 0.5 * exp]
    ff = exp / (sum(cf) / len(cf)) if cf else 1.0
    fw = exp / (sum(cw) / len(cw)) if cw else 1.0
    print(f"calibration copy {calib_mib} MiB: FETCH factor {ff:.3f}, WRITE factor {fw:.3f}")
    print(f"{'kernel':<44} {'launches':>8} {'fetch_MB':>10} {'write_MB':>10} {'total_MB':>10}")
    most = {}
    for name in f:
        s = short(name)
        if s and len(f[name]) > most.get(s, (0, None))[0]:
            most[s] = (len(f[name]), name)
    for name in f:
        s = short(name)
        if not s:
            continue
        fb = sum(f[name]) / len(f[name]) * 1024 * ff
        wb = sum(w[name]) / len(w[name]) * 1024 * fw if name in w else 0.0
        print(f"{variant(name):<44} {len(f[name]):>8} {fb / 1e6:>10.1f} {wb / 1e6:>10.1f} {(fb + wb) / 1e6:>10.1f}")
        if json_dir and most[s][1] == name:
            with open(os.path.join(json_dir, f"traffic_{s}_{tag}.json"), "w") as fh:
                json.dump({"kernel": s, "variant": variant(name), "workload": tag, "fetch_bytes": round(fb),
                           "write_bytes": round(wb),
                           "bytes_per_launch": round(fb + wb), "fetch_factor": round(ff, 4),
                           "write_factor": round(fw, 4)}, fh)
                fh.write("\n")


def alg_bytes(n, c, k):
    """{kernel: (read bytes, write bytes)} per launch: the minimum each kernel must move."""
    cells = n * n
    return {"row_fwd_r2c": (4 * c * cells, 4 * c * cells), "col_fused": ((4 * c + k) * cells, 4 * k * cells),
            "row_inv_growth": (4 * k * cells, 4 * c * cells), "diffusion": (8 * c * cells, 4 * c * cells)}


def requests(n, c, k, paths):
    vals = collections.defaultdict(lambda: collections.defaultdict(list))
    names = {}
    for path in paths:
        for r in csv.DictReader(open(path)):
            s = short(r["Kernel_Name"])
            if s:
                names.setdefault(s, variant(r["Kernel_Name"]))
                vals[s][r["Counter_Name"]].append(float(r["Counter_Value"]))
    alg = alg_bytes(n, c, k)

    def mean(s, ctr):
        v = vals[s].get(ctr) or vals[s].get(ctr + "_sum")
        return sum(v) / len(v) if v else float("nan")

    print(f"{'kernel':<36} {'rd_req M':>9} {'B/req':>6} {'rd_lat':>7} {'wr_req M':>9} {'B/req':>6} {'wr_lat':>7} "
          f"{'pend_stall M':>12} {'TLB miss M':>10} {'L2->mem wr M (64B)':>19}")
    for s in ("row_fwd_r2c", "col_fused", "row_inv_growth", "diffusion"):
        if s not in vals:
            continue
        rd, wr = mean(s, "TCP_TCC_READ_REQ"), mean(s, "TCP_TCC_WRITE_REQ")
        rl = mean(s, "TCP_TCC_READ_REQ_LATENCY") / rd
        wl = mean(s, "TCP_TCC_WRITE_REQ_LATENCY") / wr
        ea, ea64 = mean(s, "TCC_EA0_WRREQ"), mean(s, "TCC_EA0_WRREQ_64B")
        print(f"{names[s]:<36} {rd / 1e6:>9.3f} {alg[s][0] / rd:>6.0f} {rl:>7.0f} {wr / 1e6:>9.3f} "
              f"{alg[s][1] / wr:>6.0f} {wl:>7.0f} {mean(s, 'TCP_PENDING_STALL_CYCLES') / 1e6:>12.2f} "
              f"{mean(s, 'TCP_UTCL1_TRANSLATION_MISS') / 1e6:>10.3f} {ea / 1e6:>10.3f} ({ea64 / 1e6:.3f})")


if __name__ == "__main__":
    if sys.argv[1] == "stats":
        stats(sys.argv[2])
    elif sys.argv[1] == "requests":
        requests(int(sys.argv[2]), int(sys.argv[3]), int(sys.argv[4]), sys.argv[5:])
    else:
        args = sys.argv[2:]
        jd = tg = None
        if "--json" in args:
            i = args.index("--json")
            jd, tg = args[i + 1], args[i + 2]
            args = args[:i]
        traffic(args[0], args[1], int(args[2]) if len(args) > 2 else 256, jd, tg)
