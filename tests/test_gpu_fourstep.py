"""GPU parity of the Goldilocks radix-4096 passes (two 64-point register
DFTs, the passes of the four-step 2^12 x 2^12 transform of config 2 and
winterfell_four_step.rs:31-81) against the oracle: every schedule position
(first pass with stride s >= 8, last pass with s = 1, a middle pass),
forward and inverse, batches, in place, the grouped batch form, and the
fallback to the general pass kernel for a root whose 64th root is not a
power of 2."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import omr_ntt as E  # noqa: E402
import oracle as O  # noqa: E402

from _gpu import inputs  # noqa: E402

F = E.Goldilocks

SCHEDULES = [
    (24, [12, 12]),
    (15, [12, 3]),
    (15, [3, 12]),
    (20, [12, 8]),
    (20, [8, 12]),
    (22, [4, 12, 6]),
    (18, [6, 12]),
    (16, [12, 4]),
]


@pytest.mark.parametrize("logn,sched", SCHEDULES, ids=[f"2^{l}-{'x'.join(map(str, s))}" for l, s in SCHEDULES])
def test_radix4096_schedules_forward_inverse(logn, sched):
    n = 1 << logn
    print(f"case: four-step Goldilocks 2^{logn} schedule {sched}", flush=True)
    batch = 2 if logn <= 20 else 1
    x = inputs(F, n * batch, 900 + logn)
    exp = np.concatenate([O.ntt(F.id, x[b * n:(b + 1) * n]) for b in range(batch)])
    t = E.to_device(x, F)
    out = torch.empty_like(t)
    E.ntt_schedule_device(t, out, F, sched, batch=batch, n=n)
    assert np.array_equal(E.from_device(out, F), exp), "forward"
    back = torch.empty_like(t)
    E.ntt_schedule_device(out, back, F, sched, inverse=True, batch=batch, n=n)
    assert torch.equal(back, t), "inverse"
    # in place, both directions
    buf = t.clone()
    E.ntt_schedule_device(buf, buf, F, sched, batch=batch, n=n)
    assert torch.equal(buf, out), "forward in place"
    E.ntt_schedule_device(buf, buf, F, sched, inverse=True, batch=batch, n=n)
    assert torch.equal(buf, t), "inverse in place"


def test_winterfell_four_step_variant_at_2p24():
    """The WinterfellFourStep encoder at 2^24 is the 2^12 x 2^12 schedule."""
    n = 1 << 24
    x = inputs(F, n, 4242)
    t = E.to_device(x, F)
    E.ntt_device(t, t, F, E.WINTERFELL_FOUR_STEP)
    assert np.array_equal(E.from_device(t, F), O.ntt(F.id, x))


def test_radix4096_custom_root_falls_back():
    """A primitive root other than the ark one (w^3) has no power-of-2 64th
    root; its radix-4096 passes run on the general pass kernel."""
    n = 1 << 16
    w = pow(E.root_of_unity(F, n), 3, F.p)
    x = inputs(F, n, 17)
    t = E.to_device(x, F)
    out = torch.empty_like(t)
    E.ntt_schedule_device(t, out, F, [12, 4], omega=w)
    assert np.array_equal(E.from_device(out, F), O.ntt(F.id, x, omega=w))


def test_radix4096_in_grouped_batches():
    """Grouped batch form with 2^16 vectors under a [12, 4] tuned choice: the
    passes read and write through per-vector offset tables."""
    rng = np.random.default_rng(5)
    items, off = [], 0
    for k in range(7):
        ln = int(rng.choice([15, 16, 17]))
        items.append((off, off, ln, k % 2))
        off += (1 << ln) + 5
    x = inputs(F, off, 6)
    exp = O.ntt_batch(F.id, x, [i[0] for i in items], [i[2] for i in items], [i[3] for i in items])
    try:
        for ln, sched in ((15, [12, 3]), (16, [12, 4]), (17, [5, 12])):
            cls = _class_with(ln, sched)
            if cls is not None:
                E.tune_set_class(F, 1 << ln, 1 << ln, cls)
        plan = E.BatchPlan(F, items)
        t = E.to_device(x, F)
        for mode in ("grouped", "grouped_graph"):
            out = t.clone()
            plan.execute(t, out, mode)
            assert np.array_equal(E.from_device(out, F), exp), mode
        plan.close()
    finally:
        E.tune_clear()
        E.tune_load(os.path.join(os.path.dirname(E.LIB_PATH), "tuning", "mi355x.tune"))


def _class_with(ln, sched):
    for cls in range(E.num_schedule_classes()):
        sc = E.schedule_class(F, 1 << ln, 1 << ln, cls)
        if sc and sc.get("log_radix") == sched:
            return cls
    return None
