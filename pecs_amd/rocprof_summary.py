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
`pmc` sums each counter over each window (FETCH_SIZE / WRITE_SIZE in KiB per
dispatch); with the FP64 MFMA counters it also gives the window's MFMA flops
(SQ_INSTS_VALU_MFMA_MOPS_F64 x 512, rocprofv3's MfmaFlopsF64) and MFMA
utilisation (SQ_VALU_MFMA_BUSY_CYCLES / (GRBM_GUI_ACTIVE x SIMDs), its
MfmaUtil, GRBM_GUI_ACTIVE summed over the serialised dispatches).  With
--copy-bytes it also reports the counter reading of the largest plain copy
dispatch (__amd_rocclr_copyBuffer) against its
known size, the scale factor that converts the counters into HBM bytes
(torch's device-to-device copy_ of 1 GiB, run by bench.py --copy-calibration-gib 1).
`pmc ... --kernel NAME` prints every counter's average per dispatch of the
kernels whose name starts with NAME (all their dispatches).
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


def _derived(label, sums, simds):
    if "SQ_INSTS_VALU_MFMA_MOPS_F64" in sums:
        print(f"{label}: FP64 MFMA flops {sums['SQ_INSTS_VALU_MFMA_MOPS_F64'] * 512:.6g}")
    if "SQ_VALU_MFMA_BUSY_CYCLES" in sums and sums.get("GRBM_GUI_ACTIVE", 0) > 0:
        u = sums["SQ_VALU_MFMA_BUSY_CYCLES"] / (sums["GRBM_GUI_ACTIVE"] * simds)
        print(f"{label}: MFMA busy {100 * u:.2f}% of {simds} SIMDs x {sums['GRBM_GUI_ACTIVE']:.6g} active cycles")


def pmc(path, copy_bytes=None, kernel=None, simds=1024):
    rows = _read(path)
    per = OrderedDict()
    for r in rows:
        d = int(r["Dispatch_Id"])
        e = per.setdefault(d, {"name": r["Kernel_Name"], "c": defaultdict(float)})
        e["c"][r["Counter_Name"]] += float(r["Counter_Value"])
    disp = [(d, e["name"], e["c"]) for d, e in sorted(per.items())]
    counters = sorted({c for _, _, cs in disp for c in cs})
    print(f"dispatches with counters: {len(disp)}; counters: {', '.join(counters)}")
    if kernel:
        sel = [cs for _, n, cs in disp if _short(n).startswith(kernel)]
        print(f"kernel {kernel}: {len(sel)} dispatches")
        avg = {c: sum(cs.get(c, 0.0) for cs in sel) / max(len(sel), 1) for c in counters}
        for c in counters:
            unit = f" KiB = {avg[c] * 1024:.6g} B" if c in ("FETCH_SIZE", "WRITE_SIZE") else ""
            print(f"kernel {kernel}: {c} per dispatch avg {avg[c]:.6g}{unit}")
        _derived(f"kernel {kernel} per dispatch", avg, simds)
        return
    for label, ws in windows(disp).items():
        avg = {}
        for c in counters:
            vals = [sum(p.get(c, 0.0) for p in w) for w in ws]
            avg[c] = sum(vals) / len(vals)
            unit = f" KiB = {avg[c] * 1024:.6g} B" if c in ("FETCH_SIZE", "WRITE_SIZE") else ""
            print(f"solve window {label}: {c} per solve avg {avg[c]:.6g}{unit} over {len(ws)} solves")
        _derived(f"solve window {label} per solve", avg, simds)
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
    ap.add_argument("--kernel", default=None)
    ap.add_argument("--simds", type=int, default=1024, help="SIMDs of the GPU (MI355X: 256 CUs x 4)")
    a = ap.parse_args(argv)
    if a.mode == "trace":
        trace(a.csv)
    else:
        pmc(a.csv, a.copy_bytes, a.kernel, a.simds)


if __name__ == "__main__":
    sys.exit(main())
