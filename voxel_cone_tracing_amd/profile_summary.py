"""Short text summaries of rocprofv3 output for the profiles/ folder.

Kernel times (from `rocprofv3 --kernel-trace --stats --output-format csv`):

    python -m voxel_cone_tracing_amd.profile_summary times --trace DIR/run_kernel_trace.csv \\
        --cmd "<the profiled command>" --out profiles/kernel_times_soup256.md

Counter ratios (VALU active, TA busy, stalls, instructions per wave) from
`rocprofv3 --pmc` runs of SQ, TA and GRBM counters:

    python -m voxel_cone_tracing_amd.profile_summary counters --dirs DIR1 DIR2 ... \\
        --cmd "<the profiled commands>" --out profiles/kernel_counters_soup256.md

HBM bytes per launch (from two `rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE`
runs of a program that also copied a buffer of known size with the runtime's
copy kernel, bench.py --copy-calibration BYTES), optionally with a third run
of cache counters:

    python -m voxel_cone_tracing_amd.profile_summary traffic --fetch DIR1 --write DIR2 [--cache DIR3] \\
        --copy-bytes 268435456 --cmd "<command of each run>" --out profiles/hbm_traffic_soup256.md

Kernels are keyed by name and grid size, so that launches of one kernel over
different work (the full frame, a quarter-resolution frame, tile lists) are
kept apart.  The traffic summary ends with one machine-readable line per
(kernel, grid):  `traffic_bytes_per_launch <kernel> <grid> <bytes>`, which
bench.py reads for its roofline, and a `kernel_source_sha256 <hash>` line: the
hash of the kernel sources the counters were taken with.  bench.py uses the
counters only while the sources still hash the same.
"""
from __future__ import annotations

import argparse
import collections
import csv
import hashlib
import re
from pathlib import Path

CSRC = Path(__file__).resolve().parent / "csrc"


def source_hash() -> str:
    """SHA-256 over the kernel sources (names and contents of csrc/*.hip, *.h, *.cpp)."""
    h = hashlib.sha256()
    for f in sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.h")) + list(CSRC.glob("*.cpp"))):
        h.update(f.name.encode() + b"\0" + f.read_bytes() + b"\0")
    return h.hexdigest()


def short_name(name: str) -> str:
    """`void vct::trace_frame_kernel<1, vct::Shape<128, 11, 2> >(vct::TraceArgs)` ->
    `trace_frame_kernel<1,Shape<128,11,2>>` (no namespaces, no spaces)."""
    n = name.strip().replace("(anonymous namespace)::", "").replace("vct::", "")
    n = re.sub(r"\(.*$", "", n)
    n = re.sub(r"^void ", "", n)
    return n.replace(" ", "")


def _times(trace_csv: Path):
    rows = list(csv.DictReader(open(trace_csv)))
    agg = collections.defaultdict(list)
    for r in rows:
        grid = int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"])
        agg[(short_name(r["Kernel_Name"]), grid)].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    return agg


def cmd_times(a):
    agg = _times(Path(a.trace))
    total = sum(sum(v) for v in agg.values())
    lines = [f"# Kernel times: {a.title}", "", "Command (from the repository root):", "", "```", a.cmd, "```", "",
             "Summarised by `python -m voxel_cone_tracing_amd.profile_summary times --trace "
             "<dir>/run_kernel_trace.csv`.  Launches keyed by kernel and grid size (threads); times from the "
             "kernel trace.", "",
             "| kernel | grid (threads) | calls | avg us | total ms | share |", "|---|---:|---:|---:|---:|---:|"]
    for (k, g), v in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
        if sum(v) < 0.001 * total:
            continue
        lines.append(f"| `{k}` | {g} | {len(v)} | {sum(v) / len(v) / 1e3:.1f} | {sum(v) / 1e6:.3f} | "
                     f"{sum(v) / total:.1%} |")
    Path(a.out).write_text("\n".join(lines) + "\n")
    print(f"wrote {a.out}")


def _counters(d: Path):
    files = list(Path(d).glob("**/*counter_collection.csv"))
    if not files:
        raise SystemExit(f"no counter_collection.csv under {d}")
    agg = collections.defaultdict(lambda: collections.defaultdict(list))
    for f in files:
        for r in csv.DictReader(open(f)):
            agg[(short_name(r["Kernel_Name"]), int(r["Grid_Size"]))][r["Counter_Name"]].append(
                float(r["Counter_Value"]))
    return agg


def cmd_traffic(a):
    fetch, write = _counters(Path(a.fetch)), _counters(Path(a.write))
    cache = _counters(Path(a.cache)) if a.cache else {}
    # calibration: the largest __amd_rocclr_copyBuffer launch copies copy_bytes
    cal = [(g, c) for (k, g), c in fetch.items() if k.startswith("__amd_rocclr_copyBuffer")]
    g_cal = max(g for g, _ in cal)
    f_kib = sum(fetch[("__amd_rocclr_copyBuffer", g_cal)]["FETCH_SIZE"]) / len(
        fetch[("__amd_rocclr_copyBuffer", g_cal)]["FETCH_SIZE"])
    w_kib = sum(write[("__amd_rocclr_copyBuffer", g_cal)]["WRITE_SIZE"]) / len(
        write[("__amd_rocclr_copyBuffer", g_cal)]["WRITE_SIZE"])
    ff = a.copy_bytes / (f_kib * 1024.0)
    fw = a.copy_bytes / (w_kib * 1024.0)
    lines = [f"# HBM traffic per launch: {a.title}", "", "Commands (from the repository root), one per counter run:",
             "", "```", a.cmd, "```", "",
             "Summarised by `python -m voxel_cone_tracing_amd.profile_summary traffic --fetch <dir1> --write <dir2> "
             f"--cache <dir3> --copy-bytes {a.copy_bytes}`.", "",
             f"Calibration: a device copy of {a.copy_bytes} bytes reads FETCH_SIZE {f_kib:.0f} KiB and WRITE_SIZE "
             f"{w_kib:.0f} KiB per dispatch, so FETCH_SIZE is scaled by {ff:.3f} and WRITE_SIZE by {fw:.3f}.", "",
             "| kernel | grid (threads) | launches | FETCH KiB | WRITE KiB | HBM bytes (scaled) | L2 hit |",
             "|---|---:|---:|---:|---:|---:|---:|"]
    machine = []
    for key in sorted(fetch, key=lambda k: -sum(fetch[k]["FETCH_SIZE"]) / len(fetch[k]["FETCH_SIZE"])):
        k, g = key
        if k.startswith("__amd_rocclr") or key not in write:
            continue
        fv = fetch[key]["FETCH_SIZE"]
        wv = write[key]["WRITE_SIZE"]
        f, w = sum(fv) / len(fv), sum(wv) / len(wv)
        if f + w < 64:
            continue
        traffic = f * 1024.0 * ff + w * 1024.0 * fw
        hit = ""
        machine.append(f"traffic_bytes_per_launch {k} {g} {traffic:.0f}")
        if key in cache and "TCC_HIT_sum" in cache[key]:
            h, m = sum(cache[key]["TCC_HIT_sum"]), sum(cache[key].get("TCC_MISS_sum", [0.0]))
            hit = f"{h / max(h + m, 1.0):.1%}"
            machine.append(f"l2_hit_rate {k} {g} {h / max(h + m, 1.0):.4f}")
        lines.append(f"| `{k}` | {g} | {len(fv)} | {f:.0f} | {w:.0f} | {traffic:.4g} | {hit} |")
    machine.append(f"kernel_source_sha256 {source_hash()}")
    lines += ["", "Machine-readable (kernel, grid, scaled HBM bytes per launch; the hash of the kernel sources "
              "profiled):", "", "```"] + machine + ["```"]
    Path(a.out).write_text("\n".join(lines) + "\n")
    print(f"wrote {a.out}")


def cmd_counters(a):
    """Utilisation of the listed kernels from SQ / TA / GRBM counter runs.

    SQ counters are sums over the shader engines, GRBM_GUI_ACTIVE a sum over
    the XCDs; with N CUs and X XCDs:
      VALU active = SQ_ACTIVE_INST_VALU / N / (GRBM_GUI_ACTIVE / X)
      TA busy   = TA_TA_BUSY_sum / N / (GRBM_GUI_ACTIVE / X)
      wait      = SQ_WAIT_ANY / SQ_WAVE_CYCLES (share of wave time stalled)
    and instructions per wave are the SQ_INSTS_* counts over SQ_WAVES.
    VALU active is a ratio of counter sums, not a utilisation bounded by 1:
    it passes 100 % on kernels whose waves keep VALU work in flight on every
    SIMD, and such a kernel can still be bound by latency."""
    agg = collections.defaultdict(lambda: collections.defaultdict(list))
    for d in a.dirs:
        for k, c in _counters(Path(d)).items():
            for n, v in c.items():
                agg[k][n].extend(v)
    lines = [f"# Kernel counters: {a.title}", "", "Commands (from the repository root), one per counter set:", "",
             "```", a.cmd, "```", "",
             "Summarised by `python -m voxel_cone_tracing_amd.profile_summary counters --dirs <dir>...`, means "
             f"over launches, {a.cus} CUs in {a.xcds} XCDs.  VALU active = SQ_ACTIVE_INST_VALU / CUs / "
             "(GRBM_GUI_ACTIVE / XCDs), a ratio of counter sums that can pass 100 %: it shows where VALU work "
             "fills the issue slots, not that the ALUs bound the kernel; TA busy = TA_TA_BUSY_sum / CUs / "
             "(GRBM_GUI_ACTIVE / XCDs); stalled = SQ_WAIT_ANY / SQ_WAVE_CYCLES; per-wave counts = "
             "SQ_INSTS_* / SQ_WAVES.", "",
             "| kernel | grid (threads) | VALU active | TA busy | stalled | VALU / wave | SALU / wave | LDS / wave "
             "| VMEM rd / wave | LDS bank conflicts / LDS instr |",
             "|---|---:|---:|---:|---:|---:|---:|---:|---:|---:|"]

    def mean(c, n):
        v = c.get(n)
        return sum(v) / len(v) if v else None

    for key in sorted(agg, key=lambda k: -(mean(agg[k], "GRBM_GUI_ACTIVE") or 0.0)):
        k, g = key
        if a.kernels and not any(k.startswith(x) for x in a.kernels):
            continue
        c = agg[key]
        gui = mean(c, "GRBM_GUI_ACTIVE")
        waves = mean(c, "SQ_WAVES")

        def frac(n):
            v = mean(c, n)
            return f"{v / a.cus / (gui / a.xcds):.0%}" if v is not None and gui else ""

        def per_wave(n):
            v = mean(c, n)
            return f"{v / waves:.0f}" if v is not None and waves else ""

        wc, wa = mean(c, "SQ_WAVE_CYCLES"), mean(c, "SQ_WAIT_ANY")
        li, bc = mean(c, "SQ_INSTS_LDS"), mean(c, "SQ_LDS_BANK_CONFLICT")
        lines.append(f"| `{k}` | {g} | {frac('SQ_ACTIVE_INST_VALU')} | {frac('TA_TA_BUSY_sum')} | "
                     f"{f'{wa / wc:.0%}' if wc and wa is not None else ''} | {per_wave('SQ_INSTS_VALU')} | "
                     f"{per_wave('SQ_INSTS_SALU')} | {per_wave('SQ_INSTS_LDS')} | {per_wave('SQ_INSTS_VMEM_RD')} | "
                     f"{f'{bc / li:.2f}' if li and bc is not None else ''} |")
    lines += ["", f"kernel_source_sha256 {source_hash()}"]
    Path(a.out).write_text("\n".join(lines) + "\n")
    print(f"wrote {a.out}")


def summary_is_current(path: Path) -> bool:
    """True if the summary was taken with the kernel sources as they are now."""
    if not Path(path).exists():
        return False
    for line in Path(path).read_text().splitlines():
        p = line.split()
        if len(p) == 2 and p[0] == "kernel_source_sha256":
            return p[1] == source_hash()
    return False


def read_summary(path: Path, what: str, kernel: str, grid: int | None = None):
    """A machine-readable value of a traffic summary: what = "traffic_bytes_per_launch"
    (scaled HBM bytes per launch) or "l2_hit_rate", for `kernel` (and grid), or
    None; None also when the summary is stale (taken with other kernel sources)."""
    if not summary_is_current(path):
        return None
    for line in Path(path).read_text().splitlines():
        p = line.split()
        if len(p) == 4 and p[0] == what and p[1] == kernel and (grid is None or int(p[2]) == grid):
            return float(p[3])
    return None


def read_traffic(path: Path, kernel: str, grid: int | None = None):
    """Scaled HBM bytes per launch of `kernel` (and grid) from a traffic summary, or None."""
    v = read_summary(path, "traffic_bytes_per_launch", kernel, grid)
    return None if v is None else int(v)


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="what", required=True)
    t = sub.add_parser("times")
    t.add_argument("--trace", required=True)
    t.add_argument("--cmd", required=True)
    t.add_argument("--title", default="soup256 (1M triangles, 256^3, 2560x1440)")
    t.add_argument("--out", required=True)
    r = sub.add_parser("traffic")
    r.add_argument("--fetch", required=True)
    r.add_argument("--write", required=True)
    r.add_argument("--cache", default=None)
    r.add_argument("--copy-bytes", type=int, required=True)
    r.add_argument("--cmd", required=True)
    r.add_argument("--title", default="soup256 (1M triangles, 256^3, 2560x1440)")
    r.add_argument("--out", required=True)
    c = sub.add_parser("counters")
    c.add_argument("--dirs", nargs="+", required=True)
    c.add_argument("--kernels", nargs="*", default=None, help="kernel name prefixes to list (default: all)")
    c.add_argument("--cus", type=int, default=256)
    c.add_argument("--xcds", type=int, default=8)
    c.add_argument("--cmd", required=True)
    c.add_argument("--title", default="soup256 (1M triangles, 256^3, 2560x1440)")
    c.add_argument("--out", required=True)
    a = ap.parse_args()
    {"times": cmd_times, "traffic": cmd_traffic, "counters": cmd_counters}[a.what](a)


if __name__ == "__main__":
    main()
