"""GPU parity at the reference's own large partial-output matrix and at the
full configuration sizes.

* TransformDecomposition(V2) at s in {1, 2, 3, 50, 101, 16383, 16384} and
  N in {2^14, 2^15, 2^22, 2^23, 2^24} (src/lib.rs:167-180: odd and even
  log N, s above and below sqrt N, out_len up to N), through the host
  drop-in (omr_ntt: copy in, run, copy out) and the device form
  (omr_ntt_device, out of place and in place), BabyBear and Goldilocks.
  Each prefix out[:min(2s, N)] is compared with the oracle's full transform
  restricted to it (the reference's own check, src/lib.rs:56-81), and the
  tail must keep the input (transform_decomposition.rs:96).
* DftPartial run in place on the device (d_in == d_out).
* An uneven batch planned under a non-default radix cap.
* The column encoder at full width: 256 x 2^20 BabyBear, blowup 4.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import omr_ntt as E  # noqa: E402
import oracle as O  # noqa: E402

from _gpu import inputs  # noqa: E402

S_VALUES = [1, 2, 3, 50, 101, 16383, 16384]  # src/lib.rs:170
LOG_NS = [14, 15, 22, 23, 24]  # src/lib.rs:171


@pytest.mark.parametrize("logn", LOG_NS)
@pytest.mark.parametrize("fname", ["BabyBear", "Goldilocks"])
def test_td_reference_matrix(fname, logn):
    F = {"BabyBear": E.BabyBear, "Goldilocks": E.Goldilocks}[fname]
    n = 1 << logn
    print(f"case: TD matrix {fname} 2^{logn}", flush=True)
    x = inputs(F, n, 1000 + logn)
    full = O.ntt(F.id, x)
    dom = E.NttDomain(n, F)
    xd = E.to_device(x, F)
    out = E.empty_device(n, F)
    for s in S_VALUES:
        ol = min(2 * s, n)
        for enc in (E.TransformDecomposition(s), E.TransformDecompositionV2(s)):
            # host drop-in: in place on a numpy buffer
            buf = x.astype(F.np_dtype)
            enc.ntt(buf, dom)
            got = buf.astype(np.uint64)
            assert np.array_equal(got[:ol], full[:ol]), (enc.name(), "host", s)
            assert np.array_equal(got[ol:], x[ol:]), (enc.name(), "host tail", s)
            # device form, out of place: out_len elements written
            out.fill_(0)
            E.ntt_device(xd, out, F, enc.variant, s=s)
            got = E.from_device(out, F)
            assert np.array_equal(got[:ol], full[:ol]), (enc.name(), "device", s)
            assert not got[ol:].any(), (enc.name(), "device wrote past out_len", s)
            # device form, in place: the tail keeps the input
            t = xd.clone()
            E.ntt_device(t, t, F, enc.variant, s=s)
            got = E.from_device(t, F)
            assert np.array_equal(got[:ol], full[:ol]), (enc.name(), "device in place", s)
            assert np.array_equal(got[ol:], x[ol:]), (enc.name(), "device in place tail", s)
    del xd, out
    torch.cuda.empty_cache()


def test_td_huge_s_clamps():
    """s beyond N/2 (even 2^63 and above, where 2s wraps a 64-bit word) is
    out_len = N: the whole transform."""
    F = E.Goldilocks
    n = 1 << 12
    x = inputs(F, n, 77)
    full = O.ntt(F.id, x)
    for s in (n // 2, n // 2 + 1, 1 << 63, (1 << 64) - 1):
        for v in (E.TRANSFORM_DECOMPOSITION, E.DFT_PARTIAL):
            t = E.to_device(x, F)
            E.ntt_device(t, t, F, v, s=s)
            assert np.array_equal(E.from_device(t, F), full), (v, s)
        buf = x.copy()
        E.TransformDecomposition(s).ntt(buf, E.NttDomain(n, F))
        assert np.array_equal(buf, full), s


@pytest.mark.parametrize("fname", ["BabyBear", "Goldilocks"])
def test_dft_partial_in_place(fname):
    """DftPartial with d_in == d_out: every output reads all N inputs, so no
    output may be stored over an input still to be read (one chunk of 2^14
    or fewer inputs)."""
    F = {"BabyBear": E.BabyBear, "Goldilocks": E.Goldilocks}[fname]
    for logn in (6, 8, 12, 14, 15):
        n = 1 << logn
        x = inputs(F, n, 300 + logn)
        full = O.ntt(F.id, x)
        for s in (1, 3, 50):
            ol = min(2 * s, n)
            t = E.to_device(x, F)
            E.ntt_device(t, t, F, E.DFT_PARTIAL, s=s)
            got = E.from_device(t, F)
            assert np.array_equal(got[:ol], full[:ol]), (logn, s)
            assert np.array_equal(got[ol:], x[ol:]), (logn, s, "tail")
            # the partial-output entry point, direct method, in place too
            t = E.to_device(x, F)
            E.partial_device(t, t, F, ol, method=1)
            got = E.from_device(t, F)
            assert np.array_equal(got[:ol], full[:ol]), (logn, s, "omr_partial_device")


def test_batch_plan_at_radix_cap_5():
    """The smallest radix cap the planner accepts (2^5) splits log n = 13..21
    into passes of 2^4 / 2^5; every execution mode matches the oracle.  Caps
    below 2^5 are refused."""
    F = E.Goldilocks
    with pytest.raises(E.OmrError):
        E.batch_set_max_log_radix(4)
    E.batch_set_max_log_radix(5)
    try:
        items, off = [], 0
        for k, ln in enumerate([13, 14, 15, 16, 17, 19, 21, 12, 4]):
            items.append((off, off, ln, k % 2))
            off += 1 << ln
        assert E.batch_vector_schedule(13) == [5, 4, 4]
        x = inputs(F, off, 41)
        plan = E.BatchPlan(F, items)
        exp = O.ntt_batch(F.id, x, [it[0] for it in items], [it[2] for it in items], [it[3] for it in items])
        t_in = E.to_device(x, F)
        for mode in ("persistent", "graph", "grouped", "plain"):
            out = torch.zeros_like(t_in)
            plan.execute(t_in, out, mode)
            assert np.array_equal(E.from_device(out, F), exp), mode
        plan.close()
    finally:
        E.batch_set_max_log_radix(0)  # automatic


@pytest.mark.slow
def test_encoder_full_width_256_columns():
    """The column encoder at its full configuration: 256 BabyBear columns of
    2^20 coefficients, blowup 4 (n = 2^22 per column), with the shipped tuned
    schedule.  decode(encode(c)) == c with no nonzero tail coefficient, and
    four columns (first, last, two inside) equal the oracle's coset LDE."""
    F = E.BabyBear
    n, cols, blow = 1 << 20, 256, 4
    m = n * blow
    print("case: encoder 256 x 2^20 BabyBear blowup 4", flush=True)
    gen = torch.Generator(device="cuda").manual_seed(5)
    coeffs = torch.randint(0, F.p, (n * cols,), dtype=torch.int64, device="cuda", generator=gen).to(torch.int32)
    evals = torch.empty(m * cols, dtype=torch.int32, device="cuda")
    E.coset_lde_device(coeffs, evals, F, n, cols, blow)
    back = torch.empty_like(coeffs)
    nz = torch.zeros(1, dtype=torch.int32, device="cuda")
    E.coset_decode_device(evals, back, F, n, cols, blow, nonzero=nz)
    assert torch.equal(back, coeffs)
    assert int(nz.item()) == 0
    for c in (0, 1, 131, cols - 1):
        cc = E.from_device(coeffs[c * n:(c + 1) * n], F)
        got = E.from_device(evals[c * m:(c + 1) * m], F)
        assert np.array_equal(got, O.coset_lde(F.id, cc, blow)), c
    del coeffs, evals, back
    torch.cuda.empty_cache()
