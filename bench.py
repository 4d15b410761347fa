"""Benchmark of the NTT engine on one MI355X.

Headline workload: Goldilocks (p = 2^64 - 2^32 + 1) forward NTT of
N = 2^24 elements, natural order in and out, inputs uniform in [0, p) and
already resident in HBM.  One step = one transform.  Inputs and outputs
rotate over 3 buffer pairs (3 x 2 x 128 MiB) so that each step reads its
input from HBM rather than from the 256 MB Infinity Cache.

Prints ONE JSON line (rank 0):
  value        elements/s of the 2^24 forward transform
  roofline     algorithmic bytes (16 N per transform: read N + write N words)
               per pass launch / average pass-launch duration, vs 8 TB/s
  cpu_baseline the CPU restatement of the reference's WinterfellFourStep at
               2^24 (OpenMP over the host cores), timed on a bounded sample
Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "omr-ntt_amd"))

HBM_PEAK_GBS = 8000.0
VALU_LANE_OPS_PER_S = 256 * 64 * 2.4e9  # MI355X: 256 CUs x 4 SIMD16 x 2.4 GHz


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def prewarm(fn, torch, seconds):
    """Untimed calls of fn for a fixed wall time: the GPU clock ramps up over
    ~100 ms of work after an idle period, so a short --warmup alone would time
    the ramp.  Returns the number of calls made."""
    calls = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        for _ in range(8):
            fn(calls)
            calls += 1
        torch.cuda.synchronize()
    return calls


def time_device(fn, steps, warmup, torch):
    """Times `steps` calls of fn() with events on torch's current stream
    (the stream the engine launches on); returns seconds per call."""
    for i in range(warmup):
        fn(i)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record()
    for i in range(steps):
        fn(i)
    end.record()
    torch.cuda.synchronize()
    return start.elapsed_time(end) / 1e3 / steps


def _oracle_native():
    """The oracle module on its -march=native build, compiled on this host
    (the reference builds with target-cpu=native, .cargo/config.toml:2);
    returns (module, build description)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    native = O.use_native_build()
    return O, ("gcc -O3 -march=native -fopenmp (built on this host)" if native
               else "gcc -O3 -march=x86-64-v3 -fopenmp (native build failed)")


def cpu_baseline(x, n, budget_s=20.0):
    """WinterfellFourStep restated in C (oracle/): transposes tiled and
    spread over the host's threads like transpose_par, rows in parallel,
    derange in parallel (winterfell_four_step.rs:31-81, utils.rs:85-101)."""
    O, build = _oracle_native()
    f = O.GOLDILOCKS
    d = O.Domain(f, n)
    times = []
    out = None
    t_start = time.time()
    while len(times) < 10 and time.time() - t_start < budget_s:
        t0 = time.perf_counter()
        out = O.encoder("WinterfellFourStep", f, x, dom=d)
        times.append(time.perf_counter() - t0)
    times.sort()
    med = times[len(times) // 2]
    return {"value": n / med, "unit": "elems/s", "cores": O.num_threads(), "kind": "port",
            "sample": f"WinterfellFourStep (C restatement) Goldilocks 2^{n.bit_length() - 1}, "
                      f"median of {len(times)} runs", "build": build}, out


def cpu_extras(np, budget_s=4.0):
    """CPU baselines of the other workloads (BabyBear 2^16, Goldilocks 2^20 /
    2^22 / 2^24, 2^22 prefixes): the C restatements of the reference
    encoders that run them, OpenMP over the host cores, median of up
    to 10 runs within a time budget per item (inputs built outside the
    timing, like criterion's iter_batched)."""
    O, _ = _oracle_native()
    items = [
        ("babybear_2p16_radix2_dit", O.BABYBEAR, 16, "ArkRadix2", None),
        ("goldilocks_2p20_winterfell", O.GOLDILOCKS, 20, "WinterfellFourStep", None),
        ("goldilocks_2p24_winterfell", O.GOLDILOCKS, 24, "WinterfellFourStep", None),
        ("goldilocks_2p22_winterfell", O.GOLDILOCKS, 22, "WinterfellFourStep", None),
        ("goldilocks_2p24_stockham8", O.GOLDILOCKS, 24, "TfheStockhamRadix8", None),
        ("goldilocks_2p24_plonky3", O.GOLDILOCKS, 24, "Plonky3Radix2LayerSplit", None),
        ("babybear_2p22_td_out2p8", O.BABYBEAR, 22, "TransformDecomposition", 1 << 7),
        ("goldilocks_2p22_td_out2p8", O.GOLDILOCKS, 22, "TransformDecomposition", 1 << 7),
        ("babybear_2p22_td_out2p12", O.BABYBEAR, 22, "TransformDecomposition", 1 << 11),
    ]
    res = {}
    for key, f, logn, name, s in items:
        n = 1 << logn
        x = O.uniform(f, n, 7)
        d = O.Domain(f, n)
        times = []
        t_start = time.time()
        while len(times) < 10 and (len(times) < 1 or time.time() - t_start < budget_s):
            t0 = time.perf_counter()
            O.encoder(name, f, x, dom=d, s=s)
            times.append(time.perf_counter() - t0)
        times.sort()
        med = times[len(times) // 2]
        res[key] = {"us": med * 1e6, "elems_per_s": n / med, "encoder": f"{name} (C restatement)",
                    "runs": len(times), "threads": O.num_threads()}
        log(f"bench: cpu {key}: {med * 1e3:.2f} ms ({len(times)} runs)")
    return res


def other_sizes(E, F, torch, np):
    """Goldilocks 2^20 / 2^22 forward and inverse (default schedules), inputs
    rotating over buffer sets larger than the 256 MB Infinity Cache; GB/s =
    16 N algorithmic bytes (read + write N words) per transform."""
    out = {}
    for ln in (20, 22):
        n = 1 << ln
        nb = max(3, (1 << 26) // n)  # 2 x nb x 8 n bytes >= 1 GiB
        xs = [E.to_device(np.random.default_rng(ln + i).integers(0, F.p, size=n, dtype=np.uint64), F)
              for i in range(nb)]
        ys = [E.empty_device(n, F) for _ in range(nb)]
        fwd = time_device(lambda i: E.ntt_device(xs[i % nb], ys[i % nb], F), 200, 200, torch)
        inv = time_device(lambda i: E.intt_device(ys[i % nb], xs[i % nb], F), 200, 200, torch)
        gbs = 16 * n / fwd / 1e9
        out[f"2^{ln}"] = {"forward_us": fwd * 1e6, "inverse_us": inv * 1e6, "schedule": E.default_schedule(F, n),
                          "elems_per_s": n / fwd, "alg_GBps": gbs, "hbm_frac": gbs / HBM_PEAK_GBS}
        log(f"bench: 2^{ln} forward {fwd * 1e6:.1f} us ({gbs:.0f} GB/s, {gbs / HBM_PEAK_GBS:.1%} of peak) "
            f"inverse {inv * 1e6:.1f} us")
    return out


def variants_2p24(E, F, torch, np):
    """Goldilocks 2^24 under the reference variants that name a structure:
    WinterfellFourStep (n1 = n2 = 2^12: two passes of radix 4096) and
    TfheStockhamRadix8 (radix-8 digits), three rotating HBM-resident inputs."""
    n = 1 << 24
    xs = [E.to_device(np.random.default_rng(240 + i).integers(0, F.p, size=n, dtype=np.uint64), F)
          for i in range(3)]
    ys = [E.empty_device(n, F) for _ in range(3)]
    out = {}
    for name, v in (("WinterfellFourStep", E.WINTERFELL_FOUR_STEP), ("TfheStockhamRadix8", E.TFHE_STOCKHAM_RADIX8)):
        t = time_device(lambda i: E.ntt_device(xs[i % 3], ys[i % 3], F, v), 100, 50, torch)
        gbs = 16 * n / t / 1e9
        out[name] = {"forward_us": t * 1e6, "elems_per_s": n / t, "alg_GBps": gbs, "hbm_frac": gbs / HBM_PEAK_GBS}
        log(f"bench: Goldilocks 2^24 {name} {t * 1e6:.1f} us ({gbs:.0f} GB/s)")
    del xs, ys
    torch.cuda.empty_cache()
    return out


def host_path(E, F, n, host_x, np):
    """The trait drop-in path (NttEncoder::ntt on a host slice, omr_ntt):
    copy in over PCIe, transform, copy out.  Timed for a pageable buffer
    (staged through the library's pinned ring, and with the ring off) and
    for a buffer from omr_host_alloc (direct DMA).  PCIe-inclusive: never the headline value."""
    dom = E.NttDomain(n, F)
    enc = E.ArkRadix2()
    out = {}
    for kind in ("pageable", "pageable_unstaged", "pinned"):
        buf = E.host_empty(n, F) if kind == "pinned" else np.empty(n, dtype=F.np_dtype)
        times = []
        if kind == "pageable_unstaged":
            os.environ["OMR_HOST_STAGING"] = "0"  # plain hipMemcpy of the pageable buffer
        for _ in range(6):
            buf[:] = host_x
            t0 = time.perf_counter()
            enc.ntt(buf, dom)
            times.append(time.perf_counter() - t0)
        os.environ.pop("OMR_HOST_STAGING", None)
        med = float(np.median(times[1:]))
        out[kind] = {"ms": med * 1e3, "elems_per_s": n / med,
                     "pcie_gb_per_s": 2 * n * F.bytes / med / 1e9}
        log(f"bench: host path ({kind}) {med * 1e3:.2f} ms per call")
        del buf
    return out


def other_fields(E, torch, np):
    """Forward and inverse 2^24 transforms of the three 32-bit fields
    (BabyBear, KoalaBear, TeddyBear; src/fields.rs:4-17), three rotating
    HBM-resident inputs, tuned schedules."""
    out = {}
    n = 1 << 24
    for F in (E.BabyBear, E.KoalaBear, E.TeddyBear):
        xs = [E.to_device(np.random.default_rng(24 + i).integers(0, F.p, size=n, dtype=np.uint64), F)
              for i in range(3)]
        ys = [E.empty_device(n, F) for _ in range(3)]
        fwd = time_device(lambda i: E.ntt_device(xs[i % 3], ys[i % 3], F), 100, 200, torch)
        inv = time_device(lambda i: E.intt_device(ys[i % 3], xs[i % 3], F), 100, 200, torch)
        out[F.name] = {"forward_us": fwd * 1e6, "inverse_us": inv * 1e6, "elems_per_s": n / fwd,
                       "schedule": E.default_schedule(F, n)}
        log(f"bench: {F.name} 2^24 forward {fwd * 1e6:.1f} us inverse {inv * 1e6:.1f} us")
        del xs, ys
    torch.cuda.empty_cache()
    return out


def small_babybear(E, torch, np):
    """BabyBear 2^16 forward + inverse, latency per transform."""
    F = E.BabyBear
    n = 1 << 16
    x = E.to_device(np.random.default_rng(16).integers(0, F.p, size=n, dtype=np.uint64), F)
    y = E.empty_device(n, F)
    fwd = time_device(lambda i: E.ntt_device(x, y, F), 200, 200, torch)
    inv = time_device(lambda i: E.intt_device(y, x, F), 200, 200, torch)
    log(f"bench: BabyBear 2^16 forward {fwd * 1e6:.1f} us inverse {inv * 1e6:.1f} us")
    return {"forward_us": fwd * 1e6, "inverse_us": inv * 1e6, "schedule": E.default_schedule(F, n),
            "elems_per_s": n / fwd}


def partial_2p22(E, torch, np):
    """Partial outputs of N = 2^22: prefix (out_len) and stride subset.
    Inputs rotate over 32 vectors (512 MiB BabyBear, 1 GiB Goldilocks), more
    than the 256 MB Infinity Cache holds, so each call reads from HBM."""
    out = {}
    n = 1 << 22
    nb = 32
    for F in (E.BabyBear, E.Goldilocks):
        base = E.to_device(np.random.default_rng(22).integers(0, F.p, size=n, dtype=np.uint64), F)
        xs = [base] + [torch.roll(base, k) for k in range(1, nb)]
        y = E.empty_device(n, F)
        full = time_device(lambda i: E.ntt_device(xs[i % nb], y, F), 100, 100, torch)
        rows = {"full_us": full * 1e6, "rotating_inputs": nb, "rotating_bytes": nb * n * F.bytes}
        for lo in (8, 12, 16, 20):
            m = 1 << lo
            o = E.empty_device(m, F)
            pre = time_device(lambda i: E.partial_device(xs[i % nb], o, F, m, mode=E.PREFIX), 100, 100, torch)
            strd = time_device(lambda i: E.partial_device(xs[i % nb], o, F, m, mode=E.STRIDE), 100, 100, torch)
            rows[f"2^{lo}"] = {"prefix_us": pre * 1e6, "stride_us": strd * 1e6,
                               "prefix_in_GBps": n * F.bytes / pre / 1e9}
        out[F.name] = rows
        del xs
        log(f"bench: partial {F.name} 2^22: " + ", ".join(
            f"{k} {v['prefix_us']:.1f}/{v['stride_us']:.1f} us" for k, v in rows.items() if isinstance(v, dict))
            + f" (full {full * 1e6:.1f} us)")
    return out


def encoder(E, torch, np, with_model=True):
    """256 BabyBear columns x 2^20 coefficients -> blowup 4 coset LDE
    (n = 2^22) -> decode, with the schedule chosen by the learned cost model
    (trained live on autotuner timings; held-out regret reported), or with
    the shipped tuned table when with_model is False."""
    F = E.BabyBear
    n, cols, blow = 1 << 20, 256, 4
    m = n * blow
    res, cls, timed, sched = {}, None, 0, None
    if with_model:
        from omr_ntt import costmodel as CM
        # trained on BabyBear + Goldilocks full transforms, log n 11..22, at batches 1, 16 and 256;
        # regret on held-out shapes of that grid, on KoalaBear / TeddyBear
        # (fields it never saw), on batch 64 (a batch it never saw) and on the
        # encoder's own shape (2^22, 256 columns), each against exhaustive timing
        enc_shape = (F, m, m, cols)
        res, model = CM.experiment_extrapolation(reps=3, seed=0, steps=3000, extra_shapes=[enc_shape])
        # model-pruned search: the fastest of the model's 3 best-predicted
        # classes (3 timing runs instead of the exhaustive ~13)
        cls, timed = E.autotune_shape(F, m, m, cols, model=model, use_model=True, top_k=3)
        sched = E.schedule_class(F, m, m, cls)
        ex = res["extra"]["per_shape"][0]
        times = E.tune_classes(F, m, m, cols, reps=3)
        best = times[times > 0].min()
        res["encoder_shape"] = {"model_argmin_regret": ex["regret_model"], "default_regret": ex["regret_default"],
                                "top3_pick_class": cls, "top3_pick_regret": float(times[cls] / best - 1.0),
                                "best_class": int(np.argmin(np.where(times > 0, times, np.inf)))}
    coeffs = torch.randint(0, F.p, (n * cols,), dtype=torch.int32, device="cuda")
    evals = torch.empty(m * cols, dtype=torch.int32, device="cuda")
    back = torch.empty_like(coeffs)
    nz = torch.zeros(1, dtype=torch.int32, device="cuda")
    enc = time_device(lambda i: E.coset_lde_device(coeffs, evals, F, n, cols, blow), 10, 5, torch)
    dec = time_device(lambda i: E.coset_decode_device(evals, back, F, n, cols, blow, nonzero=nz), 10, 5, torch)
    ok = bool(torch.equal(back, coeffs)) and int(nz.item()) == 0
    if with_model:
        E.tune_clear()
    enc_bytes = cols * (n + m) * 4
    log(f"bench: encoder 256x2^20 BabyBear blowup 4: encode {enc * 1e3:.2f} ms decode {dec * 1e3:.2f} ms "
        f"round trip ok={ok}; class {cls} {sched} ({timed} timed)"
        + (f"; regret interp {res['interpolation']['mean_regret_model']:.3f}, unseen fields "
           f"{res['unseen_fields']['mean_regret_model']:.3f}, unseen batch {res['unseen_batch_64']['mean_regret_model']:.3f}, "
           f"encoder shape {res['encoder_shape']['top3_pick_regret']:.3f}" if with_model else ""))
    out = {"encode_ms": enc * 1e3, "decode_ms": dec * 1e3, "round_trip_ok": ok,
           "encode_alg_GBps": enc_bytes / enc / 1e9}
    if with_model:
        out.update({"schedule_class": cls, "schedule": sched,
                    "schedule_search": f"fastest of the cost model's top 3 ({timed} classes timed)",
                    "cost_model": {k: ({kk: vv for kk, vv in v.items() if kk != "per_shape"} if isinstance(v, dict) else v)
                                   for k, v in res.items() if k != "extra"}})
    return out


def batched(E, F, torch, np):
    """4096 Goldilocks vectors, log2 n ~ U{10..20} (seed 42), mixed forward /
    inverse: graph-replayed persistent plan vs one launch sequence per vector."""
    rng = np.random.default_rng(42)
    logs = rng.integers(10, 21, size=4096)
    dirs = rng.integers(0, 2, size=4096)
    items, off = [], 0
    for ln, d in zip(logs, dirs):
        items.append((off, off, int(ln), int(d)))
        off += 1 << int(ln)
    x = torch.randint(0, 2**62, (off,), dtype=torch.int64, device="cuda")
    y_graph = torch.empty_like(x)
    y_plain = torch.empty_like(x)
    plan = E.BatchPlan(F, items)
    info = plan.info()
    y_grouped = torch.empty_like(x)
    t_graph = time_device(lambda i: plan.execute(x, y_graph, "graph"), 10, 3, torch)
    t_pers = time_device(lambda i: plan.execute(x, y_graph, "persistent"), 10, 3, torch)
    t_grouped = time_device(lambda i: plan.execute(x, y_grouped, "grouped_graph"), 10, 5, torch)
    t_plain = time_device(lambda i: plan.execute(x, y_plain, "plain"), 3, 2, torch)
    same = bool(torch.equal(y_graph, y_plain)) and bool(torch.equal(y_grouped, y_plain))
    plan.close()
    best = min(t_graph, t_grouped)
    res = {"vectors": 4096, "total_elems": info["total_elems"], "launches": info["launches"],
           "tasks": info["tasks"], "graph_ms": t_graph * 1e3, "persistent_ms": t_pers * 1e3,
           "grouped_graph_ms": t_grouped * 1e3, "plain_ms": t_plain * 1e3,
           "batched_form": "grouped_graph" if t_grouped < t_graph else "graph",
           "speedup_vs_plain": t_plain / best, "elems_per_s": info["total_elems"] / best,
           "batched_equals_plain": same}
    log(f"bench: batched persistent graph {t_graph * 1e3:.2f} ms, grouped graph {t_grouped * 1e3:.2f} ms, "
        f"plain {t_plain * 1e3:.2f} ms: speedup {t_plain / best:.2f}x equal={same}")
    return res


def _pmc_per_dispatch(run_dir, counter, match):
    """{dispatch_id: summed counter value} for the kernels whose name passes
    `match`, from one rocprofv3 --pmc run directory."""
    import glob
    import sqlite3
    dbs = glob.glob(os.path.join(run_dir, "**", "*results.db"), recursive=True)
    if not dbs:
        raise FileNotFoundError(f"no rocprofv3 results database under {run_dir}")
    out = {}
    with sqlite3.connect(dbs[0]) as db:
        rows = db.execute("select dispatch_id, name, counter_name, counter_value from pmc_events "
                          "where counter_name = ?", (counter,))
        for did, name, _, val in rows:
            if match(name):
                out[did] = out.get(did, 0.0) + float(val)
    return out


def _dispatch_names(run_dir, match):
    """{dispatch_id: kernel name} for the kernels whose name passes `match`."""
    import glob
    import sqlite3
    dbs = glob.glob(os.path.join(run_dir, "**", "*results.db"), recursive=True)
    if not dbs:
        raise FileNotFoundError(f"no rocprofv3 results database under {run_dir}")
    with sqlite3.connect(dbs[0]) as db:
        return {did: name for did, name in db.execute("select distinct dispatch_id, name from pmc_events")
                if match(name)}


# pass-kernel name prefixes of the forward transform per field (Goldilocks:
# the forward instantiations, register-DFT unit 39; the 31/32-bit fields use
# one set of kernels for both directions)
FWD_KERNEL = {"Goldilocks": "k_stockham_pass<omr::Goldilocks,", "BabyBear": "k_stockham_pass<omr::Mont32<2013265921u",
              "KoalaBear": "k_stockham_pass<omr::Mont32<2130706433u", "TeddyBear": "k_stockham_pass<omr::Mont32<3221225473u"}


def is_fwd_pass(name, field="Goldilocks"):
    if FWD_KERNEL[field] not in name:
        return False
    return field != "Goldilocks" or ", 39," in name or ", 39>" in name


def summarize_profiles(prof_dir, out_path, logn=24, field="Goldilocks"):
    """Per-pass HBM traffic and VALU work of the 2^N forward transform of
    `field` (Goldilocks by default) from five rocprofv3 --pmc runs under prof_dir (copy_fetch,
    copy_write: `bench.py --copy-calibration`; fetch, write, valu: the bench
    itself), scaled by the copy of known size.  Writes the JSON bench.py
    reads for roofline.traffic and valu, and returns it."""
    n = 1 << logn
    copy_bytes = (1 << 24) * 8  # --copy-calibration: 128 MiB read + 128 MiB written per copy
    def is_copy(name):
        return "copy" in name.lower()

    def fwd(name):  # the forward transform's pass kernels
        return is_fwd_pass(name, field)
    cf = _pmc_per_dispatch(os.path.join(prof_dir, "copy_fetch"), "FETCH_SIZE", is_copy)
    cw = _pmc_per_dispatch(os.path.join(prof_dir, "copy_write"), "WRITE_SIZE", is_copy)
    fetch_scale = copy_bytes / 1024 / (sum(cf.values()) / len(cf))
    write_scale = copy_bytes / 1024 / (sum(cw.values()) / len(cw))
    f = _pmc_per_dispatch(os.path.join(prof_dir, "fetch"), "FETCH_SIZE", fwd)
    w = _pmc_per_dispatch(os.path.join(prof_dir, "write"), "WRITE_SIZE", fwd)
    v = _pmc_per_dispatch(os.path.join(prof_dir, "valu"), "SQ_INSTS_VALU", fwd)
    sys.path.insert(0, os.path.join(ROOT, "omr-ntt_amd"))
    import omr_ntt as E
    F = getattr(E, field)
    sched = E.default_schedule(F, n)
    npass = len(sched)

    def per_pass(d):  # skip the first transform (warm-up), average each pass index
        vals = [d[k] for k in sorted(d)][npass:]
        vals = vals[:len(vals) // npass * npass]
        return [sum(vals[i::npass]) / len(vals[i::npass]) for i in range(npass)]
    fk, wk, vi = per_pass(f), per_pass(w), per_pass(v)
    passes = [{"pass": i, "fetch_KiB_raw": fk[i], "write_KiB_raw": wk[i],
               "fetch_bytes": fk[i] * 1024 * fetch_scale, "write_bytes": wk[i] * 1024 * write_scale,
               "valu_wave_instr": vi[i], "valu_lane_ops_per_elem": vi[i] * 64 / n} for i in range(npass)]
    res = {"kernel": "k_stockham_pass", "workload": f"{field} forward NTT 2^{logn}",
           "schedule_log_radix": sched,
           "traffic_bytes_per_launch": sum(p["fetch_bytes"] + p["write_bytes"] for p in passes) / npass,
           "fetch_scale": fetch_scale, "write_scale": write_scale, "per_pass": passes}
    with open(out_path, "w") as fh:
        json.dump(res, fh, indent=1)
    print("| pass | FETCH_SIZE (KiB, raw) | HBM read (MiB) | WRITE_SIZE (KiB, raw) | HBM write (MiB) "
          "| VALU lane-ops / element |")
    print("|---|---|---|---|---|---|")
    for p in passes:
        print(f"| {p['pass']} | {p['fetch_KiB_raw']:,.0f} | {p['fetch_bytes'] / 2**20:.1f} | "
              f"{p['write_KiB_raw']:,.0f} | {p['write_bytes'] / 2**20:.1f} | {p['valu_lane_ops_per_elem']:.1f} |")
    print(f"scale factors: FETCH_SIZE x{fetch_scale:.3f}, WRITE_SIZE x{write_scale:.3f}; "
          f"traffic per launch {res['traffic_bytes_per_launch'] / 1e6:.1f} MB; "
          f"VALU lane-ops per element {sum(p['valu_lane_ops_per_elem'] for p in passes):.1f}")
    return res


ENC_N, ENC_COLS, ENC_BLOWUP = 1 << 20, 256, 4
ENC_REPS = 4  # encodes of an --encoder-only run


def encoder_only(E, torch, reps):
    """`reps` encodes of the 256 x 2^20 BabyBear blowup-4 column batch on the
    shipped tuned schedule, nothing else: the run a counter profile of the
    encoder's kernels is taken from (--summarize-encoder)."""
    F = E.BabyBear
    m = ENC_N * ENC_BLOWUP
    coeffs = torch.randint(0, F.p, (ENC_N * ENC_COLS,), dtype=torch.int32, device="cuda")
    evals = torch.empty(m * ENC_COLS, dtype=torch.int32, device="cuda")
    for _ in range(reps):
        E.coset_lde_device(coeffs, evals, F, ENC_N, ENC_COLS, ENC_BLOWUP)
    torch.cuda.synchronize()
    print(json.dumps({"encoder_only": {"encodes": reps, "columns": ENC_COLS, "n": ENC_N, "blowup": ENC_BLOWUP}}),
          flush=True)


def summarize_encoder(prof_dir):
    """HBM bytes per column of the encoder (`bench.py --encoder-only`) from
    rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs under prof_dir (fetch_enc,
    write_enc), scaled by the copy calibration runs (copy_fetch, copy_write),
    per pass kernel; the first encode is skipped.  Markdown on stdout."""
    copy_bytes = (1 << 24) * 8

    def is_copy(name):
        return "copy" in name.lower()

    def bb_pass(name):  # the encoder's pass kernels: BabyBear instantiations
        return "k_stockham_pass<omr::Mont32<2013265921u" in name
    cf = _pmc_per_dispatch(os.path.join(prof_dir, "copy_fetch"), "FETCH_SIZE", is_copy)
    cw = _pmc_per_dispatch(os.path.join(prof_dir, "copy_write"), "WRITE_SIZE", is_copy)
    fetch_scale = copy_bytes / 1024 / (sum(cf.values()) / len(cf))
    write_scale = copy_bytes / 1024 / (sum(cw.values()) / len(cw))
    f = _pmc_per_dispatch(os.path.join(prof_dir, "fetch_enc"), "FETCH_SIZE", bb_pass)
    w = _pmc_per_dispatch(os.path.join(prof_dir, "write_enc"), "WRITE_SIZE", bb_pass)
    names = _dispatch_names(os.path.join(prof_dir, "fetch_enc"), bb_pass)
    ids = sorted(f)
    npe = len(ids) // ENC_REPS  # pass launches per encode (every column group's passes)
    ids_w = sorted(w)
    first = [names[i] for i in ids[npe:2 * npe]]
    npass = next((k for k in range(1, npe) if first[k] == first[0]), npe)  # passes per column group
    rows = [{"name": first[k], "fetch": 0.0, "write": 0.0} for k in range(npass)]
    for k, (i, iw) in enumerate(zip(ids[npe:], ids_w[npe:])):
        r = rows[(k % npe) % npass]
        r["fetch"] += f[i] * 1024 * fetch_scale
        r["write"] += w[iw] * 1024 * write_scale
    encodes = ENC_REPS - 1
    alg = (ENC_N + ENC_N * ENC_BLOWUP) * 4
    print(f"{npe // npass} column groups of {npass} pass launches per encode\n")
    print("| pass | kernel | HBM read per column (MiB) | HBM write per column (MiB) |")
    print("|---|---|---|---|")
    tot_r = tot_w = 0.0
    for k, r in enumerate(rows):
        rd = r["fetch"] / encodes / ENC_COLS
        wr = r["write"] / encodes / ENC_COLS
        tot_r += rd
        tot_w += wr
        nm = r["name"].split("(")[0].replace("omr::Mont32<2013265921u, 31u, 27, 0>", "BabyBear")
        print(f"| {k} | `{nm}` | {rd / 2**20:.2f} | {wr / 2**20:.2f} |")
    print(f"\nper column: read {tot_r / 2**20:.2f} MiB + write {tot_w / 2**20:.2f} MiB = "
          f"{(tot_r + tot_w) / 2**20:.2f} MiB against {alg / 2**20:.0f} MiB algorithmic "
          f"(read 2^20 + write 2^22 words of 4 bytes): {(tot_r + tot_w) / alg:.2f}x; "
          f"scale factors FETCH_SIZE x{fetch_scale:.3f}, WRITE_SIZE x{write_scale:.3f}")


def summarize_sq(sq_dir, clk_dir, logn=24, field="Goldilocks"):
    """Per-pass SQ counters of the forward transform (Goldilocks, U = 39
    instantiation) from two rocprofv3 --pmc runs of this bench: the wait / LDS
    counters, and GRBM_COUNT with kernel durations (--kernel-trace) for the
    clock and the VALU utilisation (4 cycles per wave64 VALU instruction on a
    16-lane SIMD, 1024 SIMDs).  Markdown on stdout."""
    import glob
    import sqlite3

    def load(d, with_dur=False):
        db = glob.glob(os.path.join(d, "**", "*results.db"), recursive=True)[0]
        with sqlite3.connect(db) as c:
            pm = c.execute("select dispatch_id, name, counter_name, counter_value from pmc_events").fetchall()
            dur = dict(c.execute("select dispatch_id, duration from kernels").fetchall()) if with_dur else {}
        per = {}
        for did, name, cn, v in pm:
            if is_fwd_pass(name, field):
                per.setdefault(did, {})
                per[did][cn] = per[did].get(cn, 0.0) + float(v)
        return per, dur

    sys.path.insert(0, os.path.join(ROOT, "omr-ntt_amd"))
    import omr_ntt as E
    npass = len(E.default_schedule(getattr(E, field), 1 << logn))
    sq, _ = load(sq_dir)
    ck, dur = load(clk_dir, True)

    def by_pass(per):  # skip the first transform, average each pass index
        ids = sorted(per)[npass:]
        ids = ids[:len(ids) // npass * npass]
        return [{k: sum(per[i][k] for i in ids[p::npass]) / len(ids[p::npass]) for k in per[ids[p]]}
                for p in range(npass)], [ids[p::npass] for p in range(npass)]

    sqp, _ = by_pass(sq)
    ckp, ckids = by_pass(ck)
    print("| pass | µs | clock (GHz) | VALU instr (wave) | VALU util | wait / wave-cycle | LDS active / wave-cycle | "
          "bank-conflict / LDS cycle |")
    print("|---|---|---|---|---|---|---|---|")
    for p in range(npass):
        us = sum(dur[i] for i in ckids[p]) / len(ckids[p]) / 1e3
        ghz = ckp[p]["GRBM_COUNT"] / 8 / us / 1e3  # one GRBM instance per XCD
        util = 4 * ckp[p]["SQ_INSTS_VALU"] / (1024 * us * 1e3 * ghz)
        wc = sqp[p]["SQ_WAVE_CYCLES"]
        print(f"| {p} | {us:.1f} | {ghz:.2f} | {ckp[p]['SQ_INSTS_VALU']:.3g} | {util:.2f} | "
              f"{sqp[p]['SQ_WAIT_INST_ANY'] / wc:.2f} | {sqp[p]['SQ_ACTIVE_INST_LDS'] / wc:.3f} | "
              f"{sqp[p]['SQ_LDS_BANK_CONFLICT'] / max(sqp[p]['SQ_ACTIVE_INST_LDS'], 1):.2f} |")


def summarize_trace(trace_dir, logn=24, field="Goldilocks"):
    """Kernel-duration table of a rocprofv3 --kernel-trace run of this bench
    (calls, mean/min/max microseconds per kernel) and the forward
    transform's mean per pass in launch order, as markdown."""
    import glob
    import sqlite3
    dbs = glob.glob(os.path.join(trace_dir, "**", "*results.db"), recursive=True)
    if not dbs:
        raise FileNotFoundError(f"no rocprofv3 results database under {trace_dir}")
    with sqlite3.connect(dbs[0]) as db:
        rows = db.execute("select name, duration, dispatch_id, vgpr_count, lds_size from kernels "
                          "order by dispatch_id").fetchall()
    stats = {}
    for name, dur, _, vg, lds in rows:
        st = stats.setdefault(name, [0, 0.0, float("inf"), 0.0, vg, lds])
        st[0] += 1
        st[1] += dur / 1e3
        st[2] = min(st[2], dur / 1e3)
        st[3] = max(st[3], dur / 1e3)
    print("| kernel | calls | mean (µs) | min | max | VGPRs | LDS (B) |")
    print("|---|---|---|---|---|---|---|")
    for name, (k, tot, lo, hi, vg, lds) in sorted(stats.items(), key=lambda kv: -kv[1][1]):
        print(f"| `{name[:90]}` | {k} | {tot / k:.1f} | {lo:.1f} | {hi:.1f} | {vg} | {lds} |")
    sys.path.insert(0, os.path.join(ROOT, "omr-ntt_amd"))
    import omr_ntt as E
    npass = len(E.default_schedule(getattr(E, field), 1 << logn))
    fwd = [d / 1e3 for name, d, _, _, _ in rows if is_fwd_pass(name, field)]
    fwd = fwd[npass:len(fwd) // npass * npass]  # skip the first transform
    if fwd:
        per = [sum(fwd[i::npass]) / len(fwd[i::npass]) for i in range(npass)]
        print(f"forward transform per pass (launch order): {' / '.join(f'{x:.1f}' for x in per)} µs; "
              f"sum {sum(per):.1f} µs under the profiler")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--logn", type=int, default=24)
    ap.add_argument("--field", default="Goldilocks", choices=sorted(FWD_KERNEL),
                    help="field of the headline transform (for counter profiles of the other fields' pass "
                         "kernels; the headline metric is Goldilocks)")
    ap.add_argument("--schedule", default=None,
                    help="comma-separated log2 pass radices for the headline transform (e.g. 12,12 for the "
                         "four-step 2^12 x 2^12), to profile a schedule; default: the engine's tuned one")
    ap.add_argument("--prewarm-s", type=float, default=0.5,
                    help="untimed transforms for this many seconds before --warmup (clock ramp-up)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 128 MiB device copy (torch copy_) 10 times, to calibrate "
                         "rocprof FETCH_SIZE / WRITE_SIZE against a kernel of known traffic")
    ap.add_argument("--extras", default="host,sizes,fields,batched,small,partial,encoder,costmodel",
                    help="comma-separated legs after the headline (costmodel: train the cost model "
                         "and pick the 256 x 2^20 encoder's schedule with it)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the 2^20/2^22 sizes and the batched uneven-size workload")
    ap.add_argument("--summarize-profiles", metavar="DIR",
                    help="no GPU work: summarise the rocprofv3 --pmc runs under DIR (copy_fetch, copy_write, "
                         "fetch, write, valu) into per-pass HBM traffic and VALU work (see --profile-out)")
    ap.add_argument("--profile-out", default=os.path.join(ROOT, "profiles", "traffic_k_stockham_pass_goldilocks_2p24.json"))
    ap.add_argument("--summarize-sq", nargs=2, metavar=("SQ_DIR", "CLOCK_DIR"),
                    help="print per-pass SQ counters and clock from two rocprofv3 --pmc runs of this bench")
    ap.add_argument("--encoder-only", action="store_true",
                    help=f"only run {ENC_REPS} encodes of the 256 x 2^20 BabyBear blowup-4 columns (tuned schedule), "
                         "for a counter profile of the encoder")
    ap.add_argument("--summarize-encoder", metavar="DIR",
                    help="no GPU work: HBM bytes per column of the encoder from rocprofv3 --pmc runs under DIR "
                         "(copy_fetch, copy_write, fetch_enc, write_enc)")
    ap.add_argument("--summarize-trace", metavar="DIR",
                    help="no GPU work: print the kernel-duration table of a rocprofv3 --kernel-trace run under DIR")
    args = ap.parse_args()
    unknown = set(args.extras.split(",")) - {"host", "sizes", "fields", "batched", "small", "partial",
                                              "encoder", "costmodel"}
    if unknown and not args.no_extras:
        print(f"bench.py: unknown --extras legs {sorted(unknown)}", file=sys.stderr)
        sys.exit(2)
    if args.summarize_profiles:
        summarize_profiles(args.summarize_profiles, args.profile_out, args.logn, args.field)
        return
    if args.summarize_sq:
        summarize_sq(*args.summarize_sq, args.logn, args.field)
        return
    if args.summarize_trace:
        summarize_trace(args.summarize_trace, args.logn, args.field)
        return
    if args.summarize_encoder:
        summarize_encoder(args.summarize_encoder)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the engine has no multi-GPU path "
              "(each GPU would run its own share of the work). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    if args.copy_calibration:
        a = torch.randint(0, 2**62, (1 << 24,), dtype=torch.int64, device="cuda")
        bufs = [torch.empty_like(a) for _ in range(3)]
        for i in range(10):
            bufs[i % 3].copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration": {"bytes_read": a.numel() * 8, "bytes_written": a.numel() * 8,
                                               "launches": 10}}), flush=True)
        return

    import omr_ntt as E

    if args.encoder_only:
        encoder_only(E, torch, ENC_REPS)
        return

    F = getattr(E, args.field)
    n = 1 << args.logn
    sched = [int(v) for v in args.schedule.split(",")] if args.schedule else E.default_schedule(F, n)
    npass = len(sched)
    log(f"bench: {F.name} 2^{args.logn} forward NTT, schedule {sched}, steps={args.steps} warmup={args.warmup}")

    rng = np.random.default_rng(42)
    host_x = rng.integers(0, F.p, size=n, dtype=np.uint64)
    nbuf = 3
    xs = [E.to_device(host_x if i == 0 else rng.integers(0, F.p, size=n, dtype=np.uint64), F)
          for i in range(nbuf)]
    ys = [E.empty_device(n, F) for _ in range(nbuf)]

    def fwd(i):
        if args.schedule:
            E.ntt_schedule_device(xs[i % nbuf], ys[i % nbuf], F, sched)
        else:
            E.ntt_device(xs[i % nbuf], ys[i % nbuf], F)

    prewarm_calls = prewarm(fwd, torch, args.prewarm_s)
    sec = time_device(fwd, args.steps, args.warmup, torch)
    log(f"bench: forward {sec * 1e6:.1f} us/transform")

    # bit-exact check on the GPU: inverse round trip of every rotated buffer
    back = E.empty_device(n, F)
    round_trip = True
    for i in range(nbuf):
        E.ntt_device(xs[i], ys[i], F)
        E.intt_device(ys[i], back, F)
        round_trip &= bool(torch.equal(back, xs[i]))

    def inv(i):
        if args.schedule:
            E.ntt_schedule_device(ys[i % nbuf], xs[i % nbuf], F, sched, inverse=True)
        else:
            E.intt_device(ys[i % nbuf], xs[i % nbuf], F)

    # intt(ntt(x)) == x, so the inverse steps rewrite the inputs unchanged
    prewarm(inv, torch, min(args.prewarm_s, 0.2))
    sec_inv = time_device(inv, args.steps, args.warmup, torch)
    log(f"bench: inverse {sec_inv * 1e6:.1f} us/transform, round trip ok={round_trip}")

    alg_bytes = 2 * F.bytes * n
    per_launch = alg_bytes / npass
    achieved = per_launch / (sec / npass) / 1e9
    # HBM bytes per launch measured by rocprofv3 FETCH_SIZE / WRITE_SIZE runs
    # (scaled by a 128 MiB copy); used only when it covers this schedule
    traffic = None
    valu_ops = None
    tpath = os.path.join(ROOT, "profiles", "traffic_k_stockham_pass_goldilocks_2p24.json")
    if args.logn == 24 and F is E.Goldilocks and os.path.exists(tpath):
        with open(tpath) as fh:
            t = json.load(fh)
        if t.get("schedule_log_radix") == sched:
            traffic = t["traffic_bytes_per_launch"]
            valu_ops = sum(x["valu_lane_ops_per_elem"] for x in t["per_pass"])
    result = {
        "metric": ("Goldilocks 2^24 NTT elems/s & %HBM peak; batched-vs-plain speedup, bit-exact"
                   if F is E.Goldilocks else f"{F.name} 2^{args.logn} NTT elems/s & %HBM peak"),
        "value": n / sec,
        "unit": "elems/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        # untimed transforms before the --warmup ones, for a fixed wall time
        "prewarm": {"seconds": args.prewarm_s, "transforms": prewarm_calls},
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 (Goldilocks)" if F.bytes == 8 else f"u32 ({F.name})",
        "data": "synthetic: uniform in [0, p), numpy PCG64 seed 42, 3 rotating HBM-resident buffers",
        "config": {"workload": f"{F.name} forward NTT N=2^{args.logn}, natural order in/out",
                   "schedule_log_radix": sched, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": "k_stockham_pass (one launch per pass)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": "profiles/traffic_k_stockham_pass_goldilocks_2p24.json" if traffic else None,
                     "alg_bytes_per_launch": per_launch, "launches_per_step": npass},
        "inverse_us": sec_inv * 1e6,
        # VALU issue limit: measured lane-ops per element (rocprofv3 SQ_INSTS_VALU)
        # over 256 CUs x 64 lanes x 2.4 GHz
        "valu": None if valu_ops is None else {
            "lane_ops_per_elem": valu_ops, "issue_limit_us": valu_ops * n / VALU_LANE_OPS_PER_S * 1e6,
            "utilization": valu_ops * n / VALU_LANE_OPS_PER_S / sec},
        "bit_exact": {"gpu_round_trip": round_trip},
    }
    legs = set() if args.no_extras else set(args.extras.split(","))
    if legs:
        del xs, ys, back
        torch.cuda.empty_cache()
    if "host" in legs:
        result["host_path"] = host_path(E, F, n, host_x, np)
    if "sizes" in legs:
        result["sizes_us"] = other_sizes(E, F, torch, np)
        result["variants_2p24"] = variants_2p24(E, F, torch, np)
    if "fields" in legs:
        result["fields_2p24"] = other_fields(E, torch, np)
    if "batched" in legs:
        result["batched"] = batched(E, F, torch, np)
        torch.cuda.empty_cache()
    if "small" in legs:
        result["babybear_2p16"] = small_babybear(E, torch, np)
    if "partial" in legs:
        result["partial_2p22"] = partial_2p22(E, torch, np)
        torch.cuda.empty_cache()
    if "encoder" in legs:
        result["encoder_256x2p20"] = encoder(E, torch, np, "costmodel" in legs)
    if not args.no_cpu and F is E.Goldilocks:  # the CPU restatement timed is the Goldilocks one
        cb, cpu_out = cpu_baseline(host_x, n)
        result["cpu_baseline"] = cb
        if not args.no_extras:
            result["cpu_baselines_other_workloads"] = cpu_extras(np)
        x0 = E.to_device(host_x, F)
        y0 = E.empty_device(n, F)
        E.ntt_device(x0, y0, F)
        result["bit_exact"]["gpu_vs_cpu_restatement"] = bool(
            np.array_equal(E.from_device(y0, F), cpu_out))
        log(f"bench: cpu baseline {cb['value']:.3e} elems/s on {cb['cores']} threads")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
