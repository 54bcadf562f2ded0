"""Summaries of rocprofv3 CSV output for the per-step solves.

    python -m pecs_amd.rocprof_summary trace <dir>/run_kernel_trace.csv
    python -m pecs_amd.rocprof_summary pmc <counter_collection.csv> [--copy-bytes B]

A triangular solve is a sequence of launches (per supernode level: a gather and
a sweep, forward then backward, then one copy).  Its dispatches are found in
the trace as the window between the marker kernels that bracket it in every
step: k_schur_rhs ... k_schur_recover for the carrier solves of both regions
(which run concurrently on two streams) and k_poisson_face ...
k_poisson_recover for the Poisson multipliers.

`trace` prints per-kernel statistics and, per window, the sum of kernel
durations and the span from the first dispatch's start to the last one's end
(what a hipEvent pair around the concurrent solves measures).
`pmc` sums FETCH_SIZE / WRITE_SIZE (KiB per dispatch) over each window; with
--copy-bytes it also reports the counter reading of the largest plain copy
dispatch (__amd_rocclr_copyBuffer) against its
known size, the scale factor that converts the counters into HBM bytes
(torch's device-to-device copy_ of 1 GiB, run by bench.py --copy-calibration-gib 1).
"""
import argparse
import csv
import sys
from collections import OrderedDict, defaultdict

OPEN = {"k_schur_rhs": "carrier", "k_poisson_face": "poisson"}
CLOSE = {"k_schur_recover": "carrier", "k_poisson_recover": "poisson"}


def _short(name):
    for pre in ("void ", "pecs::(anonymous namespace)::"):
        name = name.replace(pre, "")
    return name.split("(")[0].strip()


def _marker(name, table):
    s = _short(name)
    for k, v in table.items():
        if s.startswith(k + "<") or s == k:
            return v
    return None


def windows(dispatches):
    """dispatches: list of (dispatch_id, name, payload) in dispatch order.
    Returns {label: [list of payload lists]}: 'carrier' (both regions' Cholesky
    solves of one step, between the last k_schur_rhs and the first
    k_schur_recover) and 'poisson'."""
    out = defaultdict(list)
    cur, kind = None, None
    for _, name, payload in dispatches:
        close = _marker(name, CLOSE)
        if cur is not None and close == kind:
            out[kind].append(cur)
            cur, kind = None, None
            continue
        opened = _marker(name, OPEN)
        if opened:
            cur, kind = [], opened
            continue
        if cur is not None:
            cur.append(payload)
    return out


def _read(path):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def trace(path, top=25, start="k_rhs_carrier"):
    rows = _read(path)
    rows.sort(key=lambda r: int(r["Dispatch_Id"]))
    first = next((i for i, r in enumerate(rows) if _short(r["Kernel_Name"]).startswith(start)), 0)
    print(f"per-kernel statistics from the first {start} dispatch on (setup dispatches before it, "
          f"including the tile-shape timing, are left out): {len(rows) - first} of {len(rows)} dispatches")
    stats = OrderedDict()
    for r in rows[first:]:
        d = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
        s = stats.setdefault(_short(r["Kernel_Name"]), [0, 0])
        s[0] += 1
        s[1] += d
    total = sum(v[1] for v in stats.values())
    print(f"total kernel time {total / 1e6:.3f} ms")
    print(f"{'kernel':60s} {'calls':>6s} {'total ms':>10s} {'avg us':>9s} {'%':>6s}")
    for k, (n, t) in sorted(stats.items(), key=lambda kv: -kv[1][1])[:top]:
        print(f"{k[:60]:60s} {n:6d} {t / 1e6:10.3f} {t / n / 1e3:9.1f} {100 * t / total:6.2f}")
    disp = [(int(r["Dispatch_Id"]), r["Kernel_Name"], (int(r["End_Timestamp"]) - int(r["Start_Timestamp"]),
             int(r["Start_Timestamp"]), int(r["End_Timestamp"]))) for r in rows]
    for label, ws in windows(disp).items():
        sums = [sum(p[0] for p in w) / 1e6 for w in ws]
        spans = [(w[-1][2] - w[0][1]) / 1e6 for w in ws if w]
        print(f"solve window {label}: {len(ws)} solves, {len(ws[0]) if ws else 0} dispatches each; "
              f"kernel-time sum avg {sum(sums) / len(sums):.4f} ms (min {min(sums):.4f}, max {max(sums):.4f}); "
              f"first-start..last-end span avg {sum(spans) / max(len(spans), 1):.4f} ms")


def pmc(path, copy_bytes=None):
    rows = _read(path)
    per = OrderedDict()
    for r in rows:
        d = int(r["Dispatch_Id"])
        e = per.setdefault(d, {"name": r["Kernel_Name"], "c": defaultdict(float)})
        e["c"][r["Counter_Name"]] += float(r["Counter_Value"])
    disp = [(d, e["name"], e["c"]) for d, e in sorted(per.items())]
    counters = sorted({c for _, _, cs in disp for c in cs})
    print(f"dispatches with counters: {len(disp)}; counters: {', '.join(counters)}")
    for label, ws in windows(disp).items():
        for c in counters:
            vals = [sum(p.get(c, 0.0) for p in w) for w in ws]
            print(f"solve window {label}: {c} per solve avg {sum(vals) / len(vals):.6g} KiB "
                  f"= {sum(vals) / len(vals) * 1024:.6g} B over {len(ws)} solves")
    if copy_bytes:
        copies = [(d, n, cs) for d, n, cs in disp if "copyBuffer" in n or "copy_kernel" in n]
        if copies:
            d, n, cs = max(copies, key=lambda x: max(x[2].values()))
            for c in counters:
                kib = cs.get(c, 0.0)
                if kib > 0:
                    print(f"copy calibration: dispatch {d} {_short(n)[:40]}: {c} {kib:.6g} KiB for "
                          f"{copy_bytes:.6g} B copied -> factor {copy_bytes / (kib * 1024):.4f} B per counted B")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["trace", "pmc"])
    ap.add_argument("csv")
    ap.add_argument("--copy-bytes", type=float, default=None)
    a = ap.parse_args(argv)
    if a.mode == "trace":
        trace(a.csv)
    else:
        pmc(a.csv, a.copy_bytes)


if __name__ == "__main__":
    sys.exit(main())
