"""Summaries of rocprofv3 runs of the time step, as committed under profiles/.

Two kinds:

``stats``: per-kernel time of a ``--kernel-trace`` run of ``bench.py --profile-steps K``:
dispatches, mean duration and time per step of every kernel, and the busy union of the
step (the time during which at least one kernel runs).

``hbm``: HBM bytes per step from two ``--pmc`` runs (FETCH_SIZE, then WRITE_SIZE) of
``bench.py --profile-steps K --calibrate-copy``.  The counters are scaled by what they
read for the plain 256 MiB device copy that run ends with (three copies of 2^25
doubles, ``__amd_rocclr_copyBuffer``): bytes read = factor x FETCH_SIZE, bytes written =
factor x WRITE_SIZE, each factor the copy's known bytes over its counter sum.

    python -m gatt_dns_amd.profile_summary stats --trace T.csv --steps 10 --out profiles/NAME
    python -m gatt_dns_amd.profile_summary hbm --fetch F.csv --write W.csv --steps 4 --points 16777216 \\
        --out profiles/NAME

Each writes NAME.md (and NAME.json for ``hbm``).
"""
import argparse
import collections
import csv
import json
import sys

COPY_KERNEL = "__amd_rocclr_copyBuffer"
COPY_BYTES = 3 * (1 << 25) * 8     # three copies of 2^25 doubles, read and written once each


def _rows(path):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def _lib_kernel(name):
    return name.startswith("gatt::") or name.startswith("void gatt::")


def counter_sums(path, counter):
    """{kernel name: (dispatches, summed counter value)} of one --pmc run."""
    out = collections.defaultdict(lambda: [0, 0.0])
    seen = set()
    for r in _rows(path):
        if r.get("Counter_Name", counter) != counter:
            continue
        key = (r.get("Dispatch_Id"), r.get("Counter_Name"))
        name = r["Kernel_Name"]
        if key not in seen:
            out[name][0] += 1
            seen.add(key)
        out[name][1] += float(r["Counter_Value"])
    return out


def hbm(fetch_csv, write_csv, steps, points, cmd):
    fetch = counter_sums(fetch_csv, "FETCH_SIZE")
    write = counter_sums(write_csv, "WRITE_SIZE")
    fcopy = sum(v[1] for k, v in fetch.items() if k.startswith(COPY_KERNEL))
    wcopy = sum(v[1] for k, v in write.items() if k.startswith(COPY_KERNEL))
    if fcopy <= 0 or wcopy <= 0:
        raise SystemExit("no calibration copy in the counter files")
    rf, wf = COPY_BYTES / fcopy, COPY_BYTES / wcopy
    kernels = {}
    rd_tot = wr_tot = 0.0
    for name in set(fetch) | set(write):
        if not _lib_kernel(name):
            continue
        nd = max(fetch.get(name, (0, 0))[0], write.get(name, (0, 0))[0])
        rb = fetch.get(name, (0, 0.0))[1] * rf / steps
        wb = write.get(name, (0, 0.0))[1] * wf / steps
        kernels[name] = {"dispatches_per_step": nd / steps, "read_bytes_per_step": rb, "write_bytes_per_step": wb}
        rd_tot += rb
        wr_tot += wb
    arr = 8.0 * points
    res = {"workload": "turbulence box 256^3, default shapes, %d steps (bench.py --profile-steps %d)" % (steps, steps),
           "command": cmd,
           "calibration": {"read_bytes_per_fetch_kib": rf, "write_bytes_per_write_kib": wf},
           "bytes_per_step": rd_tot + wr_tot, "read_bytes_per_step": rd_tot, "write_bytes_per_step": wr_tot,
           "array_touches_per_point_step": (rd_tot + wr_tot) / arr, "kernels": kernels}
    lines = ["# HBM bytes per step, turbulence box 256^3 (default shapes)", "",
             "Commands (each counter in its own run, from the repository root on one MI355X):", "", "```", cmd, "```",
             "", "Counter sums over every dispatch of the library's kernels, divided by the %d steps, scaled by the "
             "counters of the calibration copy (3 x 256 MiB read and written): read = %.3f B per FETCH_SIZE KiB, "
             "write = %.3f B per WRITE_SIZE KiB.  One array is %d MiB." % (steps, rf, wf, arr / 2 ** 20), "",
             "| kernel | dispatches per step | arrays read / written per dispatch | arrays read + written per step |",
             "|---|---:|---:|---:|"]
    for name, v in sorted(kernels.items(), key=lambda kv: -(kv[1]["read_bytes_per_step"] + kv[1]["write_bytes_per_step"])):
        nd = v["dispatches_per_step"] or 1
        lines.append("| `%s` | %g | %.2f / %.2f | %.1f |" % (
            name, v["dispatches_per_step"], v["read_bytes_per_step"] / nd / arr, v["write_bytes_per_step"] / nd / arr,
            (v["read_bytes_per_step"] + v["write_bytes_per_step"]) / arr))
    lines += ["", "per step: read %.3f GB, write %.3f GB, total %.3f GB = %.1f array touches per point-step"
              % (rd_tot / 1e9, wr_tot / 1e9, (rd_tot + wr_tot) / 1e9, (rd_tot + wr_tot) / arr)]
    return res, "\n".join(lines) + "\n"


def _grid(r):
    return int(r.get("Grid_Size_X", 1) or 1) * int(r.get("Grid_Size_Y", 1) or 1) * int(r.get("Grid_Size_Z", 1) or 1)


def stats(trace_csv, steps, cmd, title=None, start_grid=0, note=""):
    """Per-kernel table of a kernel trace.  With start_grid, only the dispatches from the first
    one whose grid holds at least start_grid work-items onwards are counted (a run that sets up a
    smaller case first)."""
    rows = sorted(_rows(trace_csv), key=lambda r: int(r["Start_Timestamp"]))
    if start_grid:
        first = next((i for i, r in enumerate(rows) if _grid(r) >= start_grid), len(rows))
        rows = rows[first:]
    rows = [r for r in rows if _lib_kernel(r["Kernel_Name"])]
    per = collections.defaultdict(lambda: [0, 0.0])
    iv = []
    for r in rows:
        t0, t1 = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
        per[r["Kernel_Name"]][0] += 1
        per[r["Kernel_Name"]][1] += (t1 - t0) / 1e3
        iv.append((t0, t1))
    iv.sort()
    busy, cur0, cur1 = 0, None, None
    for a, b in iv:
        if cur1 is None or a > cur1:
            if cur1 is not None:
                busy += cur1 - cur0
            cur0, cur1 = a, b
        else:
            cur1 = max(cur1, b)
    if cur1 is not None:
        busy += cur1 - cur0
    lines = ["# " + (title or "Kernel time of the step, turbulence box 256^3 (FP64 RK4 step, default shapes)"), "",
             "Command (from the repository root, on one MI355X):", "", "```", cmd, "```", "",
             "Per step: kernel time sum %.4f ms; busy union %.4f ms (the time during which at least one kernel of "
             "the step runs; the direction solves of a pass run concurrently on forked streams, so a kernel's mean "
             "here can exceed its isolated launch time).%s" % (sum(v[1] for v in per.values()) / steps / 1e3,
                                                                busy / 1e6 / steps, (" " + note) if note else ""),
             "",
             "| kernel | dispatches per step | mean us | us per step |", "|---|---:|---:|---:|"]
    for name, (n, us) in sorted(per.items(), key=lambda kv: -kv[1][1]):
        lines.append("| `%s` | %g | %.1f | %.1f |" % (name, n / steps, us / n, us / steps))
    return "\n".join(lines) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="kind", required=True)
    a = sub.add_parser("stats")
    a.add_argument("--trace", required=True)
    a.add_argument("--steps", type=int, required=True)
    a.add_argument("--cmd", default="")
    a.add_argument("--title", default=None)
    a.add_argument("--note", default="")
    a.add_argument("--start-grid", type=int, default=0)
    a.add_argument("--out", required=True)
    b = sub.add_parser("hbm")
    b.add_argument("--fetch", required=True)
    b.add_argument("--write", required=True)
    b.add_argument("--steps", type=int, required=True)
    b.add_argument("--points", type=int, default=256 ** 3)
    b.add_argument("--cmd", default="")
    b.add_argument("--out", required=True)
    args = ap.parse_args(argv)
    if args.kind == "stats":
        open(args.out + ".md", "w").write(stats(args.trace, args.steps, args.cmd, args.title, args.start_grid,
                                                args.note))
    else:
        res, md = hbm(args.fetch, args.write, args.steps, args.points, args.cmd)
        json.dump(res, open(args.out + ".json", "w"), indent=1)
        open(args.out + ".md", "w").write(md)
    return 0


if __name__ == "__main__":
    sys.exit(main())
