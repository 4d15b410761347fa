"""Training data, training and regret evaluation of the learned schedule
cost model (the engine's HIP MLP, omr_cost_model_*).

A shape is (field, N, out_len, batch).  For every shape the timing autotuner
measures all valid schedule classes; the model learns per-shape-centred log
times from the engine's own 4 features, and picks argmin over the valid
classes.  Regret of a choice = time(choice) / time(exhaustive best) - 1.
The method without the model is the engine's untuned default class.
"""
import numpy as np

from . import (CostModel, FIELDS, Goldilocks, BabyBear, KoalaBear, TeddyBear, default_class, features,
               num_schedule_classes, tune_classes)


def shape_grid(fields=None, log_ns=range(10, 23), batches=(1, 16), partial=True, max_elems=1 << 26):
    fields = fields or FIELDS
    shapes = []
    for F in fields:
        for ln in log_ns:
            if ln > F.TWO_ADICITY:
                continue
            for b in batches:
                if (1 << ln) * b <= max_elems:
                    shapes.append((F, 1 << ln, 1 << ln, b))
            if partial and ln >= 14:
                for lo in (2, 6, 10, ln - 4):
                    shapes.append((F, 1 << ln, 1 << lo, 1))
    return shapes


def collect(shapes, reps=3):
    """Times every class of every shape: returns X (S x 4), times (S x K)."""
    K = num_schedule_classes()
    X = np.zeros((len(shapes), 4), dtype=np.float32)
    times = np.full((len(shapes), K), -1.0)
    for i, (F, n, out_len, b) in enumerate(shapes):
        X[i] = features(F, n, out_len, b)
        times[i] = tune_classes(F, n, out_len, b, reps)
    return X, times


def targets(times):
    """Per-shape centred log times and the validity mask."""
    M = (times > 0).astype(np.float32)
    lt = np.where(times > 0, np.log(np.maximum(times, 1e-3)), 0.0)
    mean = (lt * M).sum(1, keepdims=True) / np.maximum(M.sum(1, keepdims=True), 1)
    T = np.where(M > 0, lt - mean, 0.0).astype(np.float32)
    return T, M


def train(X, times, seed=0, steps=3000, lr=3e-3):
    T, M = targets(times)
    m = CostModel(seed)
    loss = m.train(X, T, M, steps=steps, lr=lr)
    return m, loss


def regrets(model, shapes, times, top_k=3):
    """Per-shape regret of the model's choice, of the fastest of its top_k
    predicted classes (what autotune_shape(top_k=...) installs after timing
    those k), and of the default class."""
    out = []
    for (F, n, out_len, b), t in zip(shapes, times):
        best = t[t > 0].min()
        c_model = model.choose(F, n, out_len, b)
        c_def = default_class(F, n, out_len)
        r_model = t[c_model] / best - 1.0
        r_def = (t[c_def] / best - 1.0) if c_def >= 0 and t[c_def] > 0 else float("nan")
        pred = model.forward(features(F, n, out_len, b))[0]
        valid = np.flatnonzero(t > 0)
        topk = valid[np.argsort(pred[valid])[:top_k]]
        r_topk = t[topk].min() / best - 1.0
        out.append({"field": F.name, "n": n, "out_len": out_len, "batch": b, "model_class": c_model,
                    "default_class": c_def, "best_class": int(np.argmax(np.where(t > 0, -t, -np.inf))),
                    "regret_model": float(r_model), "regret_default": float(r_def),
                    f"regret_top{top_k}": float(r_topk)})
    return out


def split(shapes, holdout_every=4):
    train_idx = [i for i in range(len(shapes)) if i % holdout_every != 1]
    test_idx = [i for i in range(len(shapes)) if i % holdout_every == 1]
    return train_idx, test_idx


def experiment(shapes=None, reps=3, seed=0, steps=3000):
    shapes = shapes or shape_grid()
    X, times = collect(shapes, reps)
    tr, te = split(shapes)
    model, loss = train(X[tr], times[tr], seed=seed, steps=steps)
    held = regrets(model, [shapes[i] for i in te], times[te])
    rm = np.array([r["regret_model"] for r in held])
    rd = np.array([r["regret_default"] for r in held])
    return {
        "shapes_train": len(tr), "shapes_heldout": len(te), "final_loss": loss,
        "heldout_mean_regret_model": float(np.mean(rm)),
        "heldout_mean_regret_default": float(np.nanmean(rd)),
        "heldout_frac_within_10pct_model": float(np.mean(rm <= 0.10)),
        "heldout_frac_within_10pct_default": float(np.nanmean(rd <= 0.10)),
        "timing_runs_exhaustive_per_shape": float(np.mean((times[te] > 0).sum(1))),
        "timing_runs_model_per_shape": 0.0,
        "per_shape": held,
    }, model


def _summary(held):
    rm = np.array([r["regret_model"] for r in held])
    rd = np.array([r["regret_default"] for r in held])
    r3 = np.array([r["regret_top3"] for r in held])
    return {"shapes": len(held), "mean_regret_model": float(np.mean(rm)),
            "mean_regret_default": float(np.nanmean(rd)),
            "mean_regret_top3": float(np.mean(r3)),
            "max_regret_model": float(np.max(rm)),
            "frac_within_10pct_model": float(np.mean(rm <= 0.10)),
            "frac_within_10pct_default": float(np.nanmean(rd <= 0.10))}


def experiment_extrapolation(reps=3, seed=0, steps=3000, log_ns=range(11, 23), max_elems=1 << 28,
                             train_fields=(BabyBear, Goldilocks), held_fields=(KoalaBear, TeddyBear),
                             train_batches=(1, 16, 256), held_batch=64, extra_shapes=()):
    """Regret where the model has to extrapolate: trained on full transforms
    of `train_fields` at `train_batches`, evaluated on (a) the same sizes in
    fields it never saw (their field feature is new), (b) a batch size it
    never saw, and (c) `extra_shapes` (e.g. the column encoder's shape),
    each against exhaustive timing of every class.  Every log n of the range
    is a training size: with every other one (12, 14, .., 22) the held-out
    regret was 6.9 % interpolated and 5.2 % at batch 64, against 1.1 % and
    2.1 % on the dense grid (MI355X; repeat timings differ by 0.4 %)."""
    grid = shape_grid(list(train_fields), log_ns, train_batches, partial=False, max_elems=max_elems)
    tr, te = split(grid)  # every 4th shape of the training grid held out: interpolation
    tr_shapes = [grid[i] for i in tr]
    groups = {
        "interpolation": [grid[i] for i in te],
        "unseen_fields": shape_grid(list(held_fields), log_ns, train_batches, partial=False, max_elems=max_elems),
        f"unseen_batch_{held_batch}": shape_grid(list(train_fields), log_ns, (held_batch,), partial=False,
                                                 max_elems=max_elems),
    }
    if extra_shapes:
        groups["extra"] = list(extra_shapes)
    X, times = collect(tr_shapes, reps)
    model, loss = train(X, times, seed=seed, steps=steps)
    res = {"train_shapes": len(tr_shapes), "train_fields": [F.name for F in train_fields],
           "train_log_n": [min(log_ns), max(log_ns)], "train_batches": list(train_batches), "final_loss": loss}
    for name, shapes in groups.items():
        _, t = collect(shapes, reps)
        held = regrets(model, shapes, t)
        res[name] = _summary(held)
        res[name]["per_shape"] = held
    return res, model
