"""CPU checks of the schedule-class vocabulary, the cost-model features and
the float64 MLP oracle (pinned by central finite differences)."""
import numpy as np
import pytest

import mlp_oracle as MO
import omr_ntt as E


def test_mlp_oracle_matches_finite_differences():
    rng = np.random.default_rng(0)
    K = 36
    P = rng.normal(0, 0.3, size=64 * 4 + 64 + 64 * 64 + 64 + K * 64 + K)
    X = rng.uniform(0, 1, size=(7, 4))
    T = rng.normal(0, 1, size=(7, K))
    M = (rng.uniform(0, 1, size=(7, K)) > 0.3).astype(np.float64)
    g, _ = MO.gradients(P, X, T, M, K)
    h = 1e-6
    for idx in rng.choice(P.size, size=60, replace=False):
        e = np.zeros_like(P)
        e[idx] = h
        fd = (MO.loss(P + e, X, T, M, K) - MO.loss(P - e, X, T, M, K)) / (2 * h)
        assert abs(fd - g[idx]) <= 1e-6 + 1e-5 * abs(fd), idx


def test_schedule_classes_are_valid_decompositions():
    assert E.num_schedule_classes() == 36
    for F in E.FIELDS:
        maxlr = 12 if F.bytes == 8 else 13
        for ln in range(4, min(F.TWO_ADICITY, 30) + 1):
            n = 1 << ln
            valid = 0
            for c in range(30):
                s = E.schedule_class(F, n, n, c)
                if s is None:
                    continue
                valid += 1
                assert sum(s["log_radix"]) == ln
                assert all(1 <= r <= maxlr for r in s["log_radix"])
                assert 2 <= s["log_e"] <= 4
            assert valid >= 1, (F, ln)
            d = E.default_class(F, n, n)
            assert d >= 0, (F, ln)
            assert E.schedule_class(F, n, n, d)["log_radix"] == E.default_schedule(F, n)
            for lo in range(0, ln):
                pc = [c for c in range(30, 36) if E.schedule_class(F, n, 1 << lo, c) is not None]
                assert 35 in pc  # the pruned full transform is always available
                assert E.default_class(F, n, 1 << lo) in pc


def test_features():
    """The field enters as its element width, the batch as its total
    elements: fields of one width share a feature value, and so do batches
    of the same total size."""
    x = E.features(E.Goldilocks, 1 << 24, 1 << 8, 16)
    assert np.allclose(x, [1.0, 24 / 32, 8 / 32, 28 / 40])
    for F in (E.BabyBear, E.KoalaBear, E.TeddyBear):
        assert np.allclose(E.features(F, 1 << 20, 1 << 20, 64), [0.0, 20 / 32, 20 / 32, 26 / 40])


def test_multipass_prefix_class_rule():
    """Class 34: the decomposition with n1 = 2^min(ceil(log out_len) + 2,
    log N - 6), fused over several passes; valid only when that n1 is at
    least out_len and above 2^10.  The default class follows the auto rule."""
    G = E.Goldilocks
    n = 1 << 22
    assert E.schedule_class(G, n, 1 << 12, 34) == {"method": 2, "log_n1": 14}
    assert E.schedule_class(G, n, 1 << 16, 34) == {"method": 2, "log_n1": 16}
    assert E.schedule_class(G, n, 1 << 9, 34) == {"method": 2, "log_n1": 11}
    assert E.schedule_class(G, n, 1 << 17, 34) is None  # n1 would fall below out_len
    assert E.schedule_class(G, 1 << 16, 1 << 10, 34) is None  # n1 = 2^10: the one-pass classes
    assert E.default_class(G, n, 1 << 12) == 34
    assert E.default_class(G, n, 1 << 18) == 35
    assert E.default_class(G, n, 1 << 8) == 32  # one fused pass, n1 = 2^8
    assert E.default_class(G, n, 4) == 30


def test_tuning_batched_slot_save_load(tmp_path):
    """Batches of 32 MB or more have a table slot of their own: installing a
    class for one leaves the single-vector entry alone, and both survive a
    save / load round trip ('F' and 'B' lines)."""
    import os
    G = E.Goldilocks
    n = 1 << 16
    valid = [c for c in range(30) if E.schedule_class(G, n, n, c)]
    c1, c2 = valid[0], valid[-1]
    assert E.schedule_class(G, n, n, c1)["log_radix"] != E.schedule_class(G, n, n, c2)["log_radix"]
    shipped = os.path.join(os.path.dirname(E.LIB_PATH), "tuning", "mi355x.tune")
    try:
        E.tune_clear()
        E.tune_set_class(G, n, n, c1)
        E.tune_set_class(G, n, n, c2, batch=64)  # 64 x 512 KiB = 32 MiB: the batched slot
        assert E.default_schedule(G, n) == E.schedule_class(G, n, n, c1)["log_radix"]
        p1, p2 = str(tmp_path / "a.tune"), str(tmp_path / "b.tune")
        E.tune_save(p1)
        lines = open(p1).read().split("\n")
        assert any(l.startswith("F 2 16 ") for l in lines) and any(l.startswith("B 2 16 ") for l in lines)
        E.tune_clear()
        E.tune_load(p1)
        E.tune_save(p2)
        assert open(p1).read() == open(p2).read()
        E.tune_set_class(G, n, n, c1, batch=8)  # 4 MiB: below the threshold, the single slot
        assert E.default_schedule(G, n) == E.schedule_class(G, n, n, c1)["log_radix"]
    finally:
        E.tune_clear()
        E.tune_load(shipped)
