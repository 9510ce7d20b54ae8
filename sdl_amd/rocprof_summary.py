"""Summarises rocprofv3 CSV output of bench.py runs into the tables kept under
profiles/ (per-stage and per-kernel time per chain, HBM bytes per stage from
FETCH_SIZE / WRITE_SIZE scaled by a copy-kernel calibration, VALU / LDS / MFMA
counters).

    python -m sdl_amd.rocprof_summary --trace DIR/kt_kernel_trace.csv \
        [--fetch F.csv --write W.csv --calib-fetch CF.csv --calib-write CW.csv --calib-bytes N] \
        [--counters C.csv ...] [--kernel-counters K.csv ...]

A chain is the run of kernels that ends with k_counters; its stages start at
a marker kernel of each stage (the order of sdl_amd/csrc/sdl_amd.hip's
runStage; a scan that opens a stage, such as the triplet counts' scan, is
counted with the stage before it).  Chains that launch k_tl_stage1_scored are the learned form.
"""
import argparse
import collections
import csv

STAGE_START = (("k_md_cap", "md"), ("k_sg_cap", "sg"), ("k_tp_write", "tp"), ("k_slot_md_sg_list", "tl"),
               ("k_tpe_count", "tcp"), ("k_tcl_count", "tcl"), ("k_counters", "counters"))
STAGES = ("prep", "md", "sg", "tp", "tl", "tcp", "tcl", "counters")


def short(name):
    n = name.replace("(anonymous namespace)::", "").replace("void ", "")
    return n.split("(")[0]


def chains_of(rows, key):
    """Splits dispatch rows (one or several per dispatch: one per counter) into
    chains in dispatch order; returns [(learned, [(stage, row), ...])]."""
    by_dispatch = collections.OrderedDict()
    for r in sorted(rows, key=key):
        by_dispatch.setdefault(r["Dispatch_Id"], []).append(r)
    out, cur, stage = [], [], "prep"
    for rs in by_dispatch.values():
        name = rs[0]["Kernel_Name"]
        for marker, s in STAGE_START:
            if marker in name:
                stage = s
        cur += [(stage, r) for r in rs]
        if "k_counters" in name:
            out.append((any("k_tl_stage1_scored" in x["Kernel_Name"] for _, x in cur), cur))
            cur, stage = [], "prep"
    return out


def dur_us(r):
    return (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3


def trace_tables(path, last=10):
    rows = list(csv.DictReader(open(path)))
    chains = chains_of(rows, lambda r: int(r["Start_Timestamp"]))
    lines = []
    for learned in (False, True):
        sel = [c for l, c in chains if l == learned][-last:]
        if not sel:
            continue
        per_stage = collections.defaultdict(float)
        per_kernel = collections.defaultdict(lambda: [0.0, 0])
        walls = []
        for c in sel:
            for s, r in c:
                per_stage[s] += dur_us(r) / len(sel)
                k = per_kernel[(s, short(r["Kernel_Name"]))]
                k[0] += dur_us(r) / len(sel)
                k[1] += 1
            walls.append((int(c[-1][1]["End_Timestamp"]) - int(c[0][1]["Start_Timestamp"])) / 1e3)
        form = "learned proposal" if learned else "plain"
        lines.append(f"### {form} chain: {len(sel)} chains, per chain (us)\n")
        lines.append(f"first-to-last kernel span per chain: median {sorted(walls)[len(walls) // 2]:.1f} us; "
                     f"sum of kernel durations {sum(per_stage.values()):.1f} us\n")
        lines.append("| stage | kernel time per chain (us) |")
        lines.append("|---|---|")
        for s in STAGES:
            if s in per_stage:
                lines.append(f"| {s} | {per_stage[s]:.1f} |")
        lines.append("")
        lines.append("| stage | kernel | launches per chain | us per chain | us per launch |")
        lines.append("|---|---|---|---|---|")
        for (s, k), (t, n) in sorted(per_kernel.items(), key=lambda kv: -kv[1][0]):
            if t < 2.0:
                continue
            lines.append(f"| {s} | `{k}` | {n / len(sel):.0f} | {t:.1f} | {t / (n / len(sel)):.1f} |")
        lines.append("")
    return lines


def counter_rows(path):
    rows = list(csv.DictReader(open(path)))
    return rows


def per_stage_counter(path, counter, last=2):
    """{stage: mean counter value per plain chain} over the last `last` plain chains."""
    rows = [r for r in counter_rows(path) if r["Counter_Name"] == counter]
    chains = chains_of(rows, lambda r: int(r["Dispatch_Id"]))
    sel = [c for l, c in chains if not l][-last:]
    out = collections.defaultdict(float)
    for c in sel:
        for s, r in c:
            out[s] += float(r["Counter_Value"]) / len(sel)
    return out


def calib(path, counter, nbytes):
    rows = [r for r in counter_rows(path) if r["Counter_Name"] == counter and "copy" in r["Kernel_Name"]]
    v = [float(r["Counter_Value"]) for r in rows]
    kib = sorted(v)[len(v) // 2]
    return nbytes / (kib * 1024.0), kib


# MI355X: 256 CUs x 4 SIMDs; a wave64 VALU instruction occupies its SIMD (16
# lanes) for 4 cycles at up to 2.4 GHz.
N_SIMD = 1024
CLOCK_HZ = 2.4e9
VALU_CYCLES = 4


def kernel_table(trace_path, counter_paths, top=12):
    """Per-kernel counters of the plain chains (mean per launch) with the
    kernel's duration from the trace: VALU instructions per wave and the VALU
    issue utilisation, SQ_INSTS_VALU x 4 cycles / (duration x 2.4 GHz x 1024
    SIMDs), the fraction of the chip's VALU issue slots the kernel filled."""
    dur = collections.defaultdict(list)
    for learned, c in chains_of(list(csv.DictReader(open(trace_path))), lambda r: int(r["Start_Timestamp"])):
        if learned:
            continue
        for s, r in c:
            dur[short(r["Kernel_Name"])].append(dur_us(r))
    acc = collections.defaultdict(lambda: collections.defaultdict(list))
    for path in counter_paths:
        rows = counter_rows(path)
        for learned, c in chains_of(rows, lambda r: int(r["Dispatch_Id"])):
            if learned:
                continue
            per = collections.defaultdict(lambda: collections.defaultdict(float))
            for s, r in c:
                per[(short(r["Kernel_Name"]), r["Dispatch_Id"])][r["Counter_Name"]] += float(r["Counter_Value"])
            for (k, _), vals in per.items():
                for n, v in vals.items():
                    acc[k][n].append(v)
    mean = lambda v: sum(v) / len(v) if v else 0.0
    # per-wave columns: instruction and wait counters (not the chip-wide cycle counters)
    names = sorted({n for k in acc for n in acc[k]} - {"SQ_WAVES", "SQ_INSTS_VALU", "GRBM_GUI_ACTIVE", "SQ_BUSY_CYCLES"})
    kern = sorted((k for k in acc if k in dur), key=lambda k: -mean(dur[k]))[:top]
    lines = ["### per kernel, plain chain (mean per launch; duration from the kernel trace)\n",
             "| kernel | us | waves | VALU / wave | VALU issue utilisation | " + " | ".join(
                 n.replace("SQ_INSTS_VALU_", "").replace("SQ_INSTS_", "") + " / wave" for n in names) + " |",
             "|---|---|---|---|---|" + "---|" * len(names)]
    for k in kern:
        d = sorted(dur[k])[len(dur[k]) // 2]
        w = mean(acc[k]["SQ_WAVES"])
        valu = mean(acc[k]["SQ_INSTS_VALU"])
        util = valu * VALU_CYCLES / (d * 1e-6 * CLOCK_HZ * N_SIMD) if d > 0 else 0.0
        extra = " | ".join(f"{mean(acc[k][n]) / w:.1f}" if w else "-" for n in names)
        lines.append(f"| `{k}` | {d:.1f} | {w:.0f} | {valu / w if w else 0:.1f} | {util:.3f} | {extra} |")
    lines.append("")
    return lines


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--trace")
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--calib-fetch")
    ap.add_argument("--calib-write")
    ap.add_argument("--calib-bytes", type=float, default=268435456.0)
    ap.add_argument("--counters", nargs="*", default=[])
    ap.add_argument("--kernel-counters", nargs="*", default=[],
                    help="counter CSVs (each with SQ_WAVES and SQ_INSTS_VALU) for the per-kernel table (needs --trace)")
    a = ap.parse_args()
    lines = []
    if a.trace:
        lines += trace_tables(a.trace)
    if a.fetch and a.write and a.calib_fetch and a.calib_write:
        ff, fk = calib(a.calib_fetch, "FETCH_SIZE", a.calib_bytes)
        wf, wk = calib(a.calib_write, "WRITE_SIZE", a.calib_bytes)
        fetch = per_stage_counter(a.fetch, "FETCH_SIZE")
        write = per_stage_counter(a.write, "WRITE_SIZE")
        lines.append("### HBM traffic per plain chain (FETCH_SIZE, WRITE_SIZE)\n")
        lines.append(f"Calibration: a copy kernel moving {a.calib_bytes:.0f} B reads FETCH_SIZE {fk:.0f} KiB "
                     f"(factor {ff:.3f}) and WRITE_SIZE {wk:.0f} KiB (factor {wf:.3f}); bytes = KiB x 1024 x factor.\n")
        lines.append("| stage | fetched MB | written MB |")
        lines.append("|---|---|---|")
        for s in STAGES:
            if s in fetch or s in write:
                lines.append(f"| {s} | {fetch.get(s, 0) * 1024 * ff / 1e6:.2f} | {write.get(s, 0) * 1024 * wf / 1e6:.2f} |")
        lines.append("")
    for path in a.counters:
        rows = counter_rows(path)
        names = sorted({r["Counter_Name"] for r in rows})
        chains = chains_of(rows, lambda r: int(r["Dispatch_Id"]))
        lines.append(f"### counters {', '.join(names)} (per chain, summed over the stage's kernels)\n")
        lines.append("| form | stage | " + " | ".join(names) + " |")
        lines.append("|---|---|" + "---|" * len(names))
        for learned in (False, True):
            sel = [c for l, c in chains if l == learned][-2:]
            if not sel:
                continue
            acc = collections.defaultdict(lambda: collections.defaultdict(float))
            for c in sel:
                for s, r in c:
                    acc[s][r["Counter_Name"]] += float(r["Counter_Value"]) / len(sel)
            for s in STAGES:
                if s in acc:
                    vals = " | ".join(f"{acc[s][n]:.4g}" for n in names)
                    lines.append(f"| {'learned' if learned else 'plain'} | {s} | {vals} |")
        lines.append("")
    if a.kernel_counters and a.trace:
        lines += kernel_table(a.trace, a.kernel_counters)
    print("\n".join(lines))


if __name__ == "__main__":
    main()
