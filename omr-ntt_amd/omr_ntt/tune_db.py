"""Regenerates the engine's shipped tuning table with the exhaustive timing
autotuner (every schedule class of every shape timed on the GPU).

    python -m omr_ntt.tune_db --out omr-ntt_amd/tuning/mi355x.tune

Full transforms: every field, log2 N = 4 .. min(26, two-adicity), batch 1.
Batched full transforms (the table's slot for batches of 32 MB or more):
log2 N = 10 .. min(26, two-adicity), as many vectors as make 512 MB (at least
two, at most 65535), so that the timed batch streams from HBM rather than the
256 MB Infinity Cache.
Partial prefixes: log2 N in {16, 18, 20, 22, 24}, every log2 out_len < log2 N.

    python -m omr_ntt.tune_db --out omr-ntt_amd/tuning/mi355x.tune --partial-only BabyBear,KoalaBear

re-times only the partial-prefix shapes of the named fields, keeps every
other entry of the table at --out (and of its .json report), and rewrites
both; --batched-only does the same for the batched full-transform slots.
"""
import argparse
import json
import os
import sys
import time

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import omr_ntt as E  # noqa: E402


PARTIAL_LOG_N = (16, 18, 20, 22, 24)


def tune_partials(F, reps, report):
    for ln in PARTIAL_LOG_N:
        if ln > F.TWO_ADICITY:
            continue
        for lo in range(1, ln):
            n = 1 << ln
            c, timed = E.autotune_shape(F, n, 1 << lo, 1, use_model=False, reps=reps)
            report.append({"field": F.name, "log_n": ln, "out": lo, "class": c, "timed": timed,
                           "schedule": E.schedule_class(F, n, 1 << lo, c)})


def batch_of(F, n):
    """Vectors in a tuning batch: as many as make 512 MB, at least two, at
    most 65535."""
    return min(65535, max(2, (512 << 20) // (n * F.bytes)))


def tune_batched(F, max_log, reps, report):
    for ln in range(10, min(max_log, F.TWO_ADICITY) + 1):
        n = 1 << ln
        batch = batch_of(F, n)
        c, timed = E.autotune_shape(F, n, n, batch, use_model=False, reps=reps)
        report.append({"field": F.name, "log_n": ln, "out": ln, "batch": batch, "class": c, "timed": timed,
                       "schedule": E.schedule_class(F, n, n, c)})
        print(f"{F.name} 2^{ln} x {batch}: class {c} {report[-1]['schedule']}", flush=True)


def retune(out, names, what, reps=5, max_log=26):
    """Re-times the partial-prefix shapes (what = "partial") or the batched
    slots (what = "batched") of the named fields on top of the table at
    `out`; every other entry stays as it is."""
    fields = [getattr(E, nm) for nm in names]
    E.context()
    E.tune_clear()
    E.tune_load(out)
    with open(out + ".json") as fh:
        old = json.load(fh)
    if what == "partial":
        mine = lambda r: r["field"] in names and r["out"] < r["log_n"] and "batch" not in r  # noqa: E731
    else:
        mine = lambda r: r["field"] in names and "batch" in r  # noqa: E731
    keep = [r for r in old if not mine(r)]
    report = []
    t0 = time.time()
    for F in fields:
        if what == "partial":
            tune_partials(F, reps, report)
        else:
            tune_batched(F, max_log, reps, report)
        print(f"{F.name}: {what} shapes re-tuned ({time.time() - t0:.0f} s)", flush=True)
    before = {(r["field"], r["log_n"], r["out"]): r["class"] for r in old if r not in keep}
    changed = sum(1 for r in report if before.get((r["field"], r["log_n"], r["out"])) != r["class"])
    E.tune_save(out)
    with open(out + ".json", "w") as fh:
        json.dump(keep + report, fh, indent=0)
    print(f"wrote {out}: {len(report)} {what} shapes re-timed, {changed} changed class")


def build(out, max_log=26, reps=5, partial=True):
    os.environ["OMR_TUNE_FILE"] = "none"  # start from the untuned defaults
    E.context()
    E.tune_clear()
    report = []
    t0 = time.time()
    for F in E.FIELDS:
        for ln in range(4, min(max_log, F.TWO_ADICITY) + 1):
            n = 1 << ln
            c, timed = E.autotune_shape(F, n, n, 1, use_model=False, reps=reps)
            report.append({"field": F.name, "log_n": ln, "out": ln, "class": c, "timed": timed,
                           "schedule": E.schedule_class(F, n, n, c)})
            print(f"{F.name} 2^{ln}: class {c} {report[-1]['schedule']}", flush=True)
        tune_batched(F, max_log, reps, report)
        if partial:
            tune_partials(F, reps, report)
            print(f"{F.name}: partial shapes tuned ({time.time() - t0:.0f} s)", flush=True)
    E.tune_save(out)
    with open(out + ".json", "w") as fh:
        json.dump(report, fh, indent=0)
    print(f"wrote {out} ({len(report)} shapes, {time.time() - t0:.0f} s)")


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--max-log", type=int, default=26)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--no-partial", action="store_true")
    ap.add_argument("--partial-only", default=None, metavar="FIELDS",
                    help="comma-separated field names: re-time only their partial-prefix shapes")
    ap.add_argument("--batched-only", default=None, metavar="FIELDS",
                    help="comma-separated field names: re-time only their batched full-transform slots")
    a = ap.parse_args()
    if a.partial_only:
        retune(a.out, a.partial_only.split(","), "partial", a.reps)
    elif a.batched_only:
        retune(a.out, a.batched_only.split(","), "batched", a.reps, a.max_log)
    else:
        build(a.out, a.max_log, a.reps, not a.no_partial)
