"""GPU parity: the HIP engine through the C ABI against the CPU oracle and
the pure-Python golden vectors.  Mirrors the reference's relational matrix
(src/lib.rs:24-180) for all four fields and extends it to the config sizes."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import omr_ntt as E  # noqa: E402
import oracle as O  # noqa: E402

from _gpu import inputs, gpu_ntt  # noqa: E402

GOLD = os.path.join(os.path.dirname(__file__), "golden")
POWERS_OF_TWO = [1 << k for k in range(1, 17)]
EVEN_POWERS_OF_TWO = [1 << k for k in range(2, 17, 2)]
POWERS_OF_TWO_DIV_BY_3 = [8, 64, 512, 4096, 32768]
FIELD_IDS = [f.name for f in E.FIELDS]


def field(name):
    return {f.name: f for f in E.FIELDS}[name]


def load(name):
    with open(os.path.join(GOLD, name)) as fh:
        return json.load(fh)


@pytest.mark.parametrize("fname", FIELD_IDS)
def test_device_field_ops_golden(fname):
    F = field(fname)
    g = load("fieldops.json")[fname.lower()]
    a = E.to_device(np.array(g["a"], dtype=np.uint64), F)
    b = E.to_device(np.array(g["b"], dtype=np.uint64), F)
    out = E.empty_device(a.numel(), F)
    for op, key in ((0, "add"), (1, "sub"), (2, "mul"), (4, "mul")):
        E.field_op_device(F, op, a, b, out)
        assert [int(v) for v in E.from_device(out, F)] == g[key], (fname, key, op)
    E.field_op_device(F, 3, a, b, out)
    assert [int(v) for v in E.from_device(out, F)] == g["a"]
    # random bulk vs oracle
    x, y = inputs(F, 1 << 16, 1), inputs(F, 1 << 16, 2)
    a, b = E.to_device(x, F), E.to_device(y, F)
    out = E.empty_device(x.size, F)
    for op in (0, 1, 2):
        E.field_op_device(F, op, a, b, out)
        assert np.array_equal(E.from_device(out, F), O.vec_op(F.id, op, x, y))


@pytest.mark.parametrize("fname", FIELD_IDS)
def test_dft_golden_vectors(fname):
    F = field(fname)
    for case in load("dft.json")[fname.lower()]:
        x = np.array(case["x"], dtype=np.uint64)
        for variant in (E.ARK_RADIX2, E.DFT):
            got = gpu_ntt(F, x, variant)
            assert [int(v) for v in got] == case["X"], (fname, case["n"], variant)


@pytest.mark.parametrize("fname", FIELD_IDS)
@pytest.mark.parametrize("enc", ["ArkRadix2", "LambdaBowers", "WinterfellFourStep", "Plonky3Radix2LayerSplit"])
def test_full_encoders_agree_with_oracle(fname, enc):  # src/lib.rs:114-127
    F = field(fname)
    for n in POWERS_OF_TWO:
        x = inputs(F, n, 1000 + n)
        exp = O.encoder("ArkRadix2", F.id, x)
        dom = E.NttDomain(n, F)
        buf = x.astype(F.np_dtype)
        getattr(E, enc)().ntt(buf, dom)  # host path (the trait drop-in)
        assert np.array_equal(buf.astype(np.uint64), exp), (enc, n)
        t = E.to_device(x, F)
        getattr(E, enc)().ntt(t, dom)  # device path
        assert np.array_equal(E.from_device(t, F), exp), (enc, n)


@pytest.mark.parametrize("fname", FIELD_IDS)
def test_radix4_stockham8_and_constraints(fname):  # src/lib.rs:130-139
    F = field(fname)
    for n in EVEN_POWERS_OF_TWO:
        x = inputs(F, n, n)
        assert np.array_equal(gpu_ntt(F, x, E.LAMBDA_RADIX4), O.ntt(F.id, x)), n
    for n in POWERS_OF_TWO_DIV_BY_3:
        x = inputs(F, n, n)
        assert np.array_equal(gpu_ntt(F, x, E.TFHE_STOCKHAM_RADIX8), O.ntt(F.id, x)), n
    with pytest.raises(E.OmrError):
        E.LambdaRadix4().ntt(inputs(F, 8, 1).astype(F.np_dtype), E.NttDomain(8, F))
    with pytest.raises(E.OmrError):
        E.TfheStockhamRadix8().ntt(inputs(F, 16, 1).astype(F.np_dtype), E.NttDomain(16, F))


@pytest.mark.parametrize("fname", FIELD_IDS)
def test_smallest_domains(fname):
    """N = 1 (omega = 1: identity; the reference accepts 1 as a power of two,
    src/encoder.rs:24) and N = 2 (omega = p - 1), through every full and
    partial encoder on the host path, the device forward and inverse."""
    F = field(fname)
    assert E.root_of_unity(F, 1) == 1 and E.root_of_unity(F, 2) == F.p - 1
    for n in (1, 2):
        x = inputs(F, n, 90 + n)
        want = O.ntt(F.id, x)
        a = [int(v) for v in x]
        assert [int(v) for v in want] == (a if n == 1 else [(a[0] + a[1]) % F.p, (a[0] - a[1]) % F.p])
        for enc in E.encoders.all(n.bit_length() - 1) + E.encoders.all_partial(1):
            buf = x.astype(F.np_dtype)
            enc.ntt(buf, E.NttDomain(n, F))
            m = n if enc.s == 0 else min(2 * enc.s, n)
            assert np.array_equal(buf[:m].astype(np.uint64), want[:m]), (enc.name(), n)
        t = E.to_device(x, F)
        E.ntt_device(t, t, F)
        assert np.array_equal(E.from_device(t, F), want), n
        E.intt_device(t, t, F)
        assert np.array_equal(E.from_device(t, F), x), n
        xb = inputs(F, 3 * n, 95 + n)  # three receivers' vectors, their whole prefix
        out = E.empty_device(3 * n, F)
        E.partial_batch_device(E.to_device(xb, F), out, F, n, 3, n)
        got = E.from_device(out, F)
        for b in range(3):
            assert np.array_equal(got[b * n:(b + 1) * n], O.ntt(F.id, xb[b * n:(b + 1) * n])), (n, b)


@pytest.mark.parametrize("fname", FIELD_IDS)
def test_dft_variant_agrees(fname):  # src/lib.rs:85-88
    F = field(fname)
    for n in (64, 128, 256, 4096):
        x = inputs(F, n, n)
        assert np.array_equal(gpu_ntt(F, x, E.DFT), O.ntt(F.id, x))


@pytest.mark.parametrize("fname", FIELD_IDS)
def test_partial_encoders_prefix(fname):  # src/lib.rs:141-165
    F = field(fname)
    cases = [("DftPartial", [64, 128, 256]), ("TransformDecomposition", POWERS_OF_TWO),
             ("TransformDecompositionV2", [1024, 4096, 32768, 65536])]
    for enc, sizes in cases:
        for s in (2, 4, 8, 16, 50):
            for n in sizes:
                x = inputs(F, n, s * 7 + n)
                m = min(2 * s, n)
                exp = O.ntt(F.id, x)
                buf = x.astype(F.np_dtype)
                getattr(E, enc)(s).ntt(buf, E.NttDomain(n, F))
                got = buf.astype(np.uint64)
                assert np.array_equal(got[:m], exp[:m]), (enc, s, n)
                assert np.array_equal(got[m:], x[m:]), (enc, s, n)  # tail keeps the input


@pytest.mark.parametrize("fname", FIELD_IDS)
def test_inverse_round_trip(fname):
    F = field(fname)
    for n in [2, 16, 1 << 10, 1 << 13, 1 << 16, 1 << 18]:
        x = inputs(F, n, 77 + n)
        t = E.to_device(x, F)
        E.ntt_device(t, t, F)
        X = E.from_device(t, F)
        assert np.array_equal(X, O.ntt(F.id, x))
        E.intt_device(t, t, F)
        assert np.array_equal(E.from_device(t, F), x), n
        if n <= 1 << 13:
            assert np.array_equal(O.intt(F.id, X), x)


@pytest.mark.parametrize("fname", FIELD_IDS)
def test_explicit_schedules_agree(fname):
    F = field(fname)
    maxlr = 12 if F.bytes == 8 else 13
    n = 1 << 16
    x = inputs(F, n, 5)
    exp = O.ntt(F.id, x)
    t_in = E.to_device(x, F)
    scheds = [[16], [8, 8], [4, 12], [12, 4], [6, 5, 5], [2, 3, 4, 7], [1, 15], [3] * 5 + [1], [2] * 8]
    for sched in scheds:
        if any(r > maxlr for r in sched):
            continue
        for log_e in (0, 2, 3, 4):
            out = E.empty_device(n, F)
            E.ntt_schedule_device(t_in, out, F, sched, log_e=log_e)
            assert np.array_equal(E.from_device(out, F), exp), (sched, log_e)
            back = E.empty_device(n, F)
            E.ntt_schedule_device(out, back, F, sched, log_e=log_e, inverse=True)
            assert np.array_equal(E.from_device(back, F), x), (sched, log_e)


@pytest.mark.parametrize("fname", FIELD_IDS)
def test_batched_same_size(fname):
    F = field(fname)
    n, batch = 1 << 12, 37
    x = inputs(F, n * batch, 11)
    t = E.to_device(x, F)
    out = E.empty_device(n * batch, F)
    E.ntt_device(t, out, F, batch=batch, n=n)
    got = E.from_device(out, F).reshape(batch, n)
    for b in (0, 1, 17, 36):
        assert np.array_equal(got[b], O.ntt(F.id, x[b * n:(b + 1) * n]))


def test_custom_omega_is_honoured():
    # any primitive N-th root (not only the ark-convention one) is accepted
    F = E.Goldilocks
    n = 1024
    w = E.root_of_unity(F, n)
    w3 = pow(w, 3, F.p)  # also primitive (3 is odd)
    x = inputs(F, n, 3)
    t = E.to_device(x, F)
    E.ntt_device(t, t, F, omega=w3)
    assert np.array_equal(E.from_device(t, F), O.ntt(F.id, x, omega=w3))
    with pytest.raises(E.OmrError):
        E.ntt_device(t, t, F, omega=pow(w, 2, F.p))  # order N/2: not primitive


@pytest.mark.parametrize("logn", [20, 22, 24])
def test_goldilocks_large_forward_inverse(logn):  # Goldilocks 2^20 / 2^22 / 2^24
    F = E.Goldilocks
    n = 1 << logn
    x = inputs(F, n, 42)
    t = E.to_device(x, F)
    out = E.empty_device(n, F)
    E.ntt_device(t, out, F)
    X = E.from_device(out, F)
    assert np.array_equal(X, O.ntt(F.id, x))
    back = E.empty_device(n, F)
    E.intt_device(out, back, F)
    assert torch.equal(back, t)
    # four-step 2^12 x 2^12 at 2^24 and the radix-8 Stockham agree too
    if logn == 24:
        for sched, le in (([12, 12], 0), ([8, 8, 8], 0), ([9, 9, 6], 3), ([6, 6, 6, 6], 0)):
            o2 = E.empty_device(n, F)
            E.ntt_schedule_device(t, o2, F, sched, log_e=le)
            assert torch.equal(o2, out), sched


@pytest.mark.parametrize("fname", ["BabyBear", "Goldilocks"])
def test_partial_large_prefix_and_stride(fname):  # N = 2^22, out_len 2^8 .. 2^20
    F = field(fname)
    n = 1 << 22
    x = inputs(F, n, 42)
    X = O.ntt(F.id, x)
    t = E.to_device(x, F)
    for logs in (8, 12, 16, 20):
        m = 1 << logs
        out = E.empty_device(m, F)
        for method in (0, 2, 3):
            if method == 2 and m > (1 << 14):
                continue
            E.partial_device(t, out, F, m, mode=E.PREFIX, method=method)
            assert np.array_equal(E.from_device(out, F), X[:m]), (logs, method)
        E.partial_device(t, out, F, m, mode=E.STRIDE)
        assert np.array_equal(E.from_device(out, F), X[:: n // m]), logs
    # reference quirk sizes (src/lib.rs:167-180): s in {1, 3, 101}
    for s in (1, 3, 101):
        m = 2 * s
        out = E.empty_device(m, F)
        for method in (1, 2, 3):
            E.partial_device(t, out, F, m, method=method)
            assert np.array_equal(E.from_device(out, F), X[:m]), (s, method)


@pytest.mark.parametrize("fname", FIELD_IDS)
def test_partial_prefix_fused(fname):
    """Decomposition with the recombine fused into the first pass (out_len <=
    n1 <= 2^10): every sub-NTT size, odd and power-of-two out_len, against
    the oracle's full transform; ragged out_len at the smallest sizes."""
    F = field(fname)
    for logn in (6, 10, 16, 20):
        n = 1 << logn
        x = inputs(F, n, logn)
        X = O.ntt(F.id, x)
        t = E.to_device(x, F)
        for out_len in (1, 2, 5, 64, 100, 256, 1000, 1024):
            if out_len > n:
                continue
            out = E.empty_device(out_len, F)
            for log_n1 in (0, 6, 7, 8, 10):
                if log_n1 > logn or (log_n1 and (1 << log_n1) < out_len):
                    continue
                E.partial_device(t, out, F, out_len, mode=E.PREFIX, method=2 if log_n1 else 0, log_n1=log_n1)
                assert np.array_equal(E.from_device(out, F), X[:out_len]), (logn, out_len, log_n1)


@pytest.mark.parametrize("fname", FIELD_IDS)
def test_partial_prefix_fused_multipass(fname):
    """Sub-NTTs of size n1 > 2^10 in two or three passes, the recombine fused
    into the last one (its tiles carry one value p of the earlier digits,
    k' = p + P k).  out_len below P (tiles with nothing to add), ragged and
    power-of-two out_len, s = N/n1 down to 4 columns."""
    F = field(fname)
    for logn, cases in ((16, [(11, (50, 64, 2047, 2048)), (12, (50, 100, 4096)), (14, (4095, 16384))]),
                        (20, [(12, (4096,)), (16, (1000, 65536)), (18, (262144,))]),
                        (22, [(20, (5000,))])):
        n = 1 << logn
        x = inputs(F, n, logn + 3)
        X = O.ntt(F.id, x)
        t = E.to_device(x, F)
        for log_n1, lens in cases:
            for out_len in lens:
                out = E.empty_device(out_len, F)
                E.partial_device(t, out, F, out_len, mode=E.PREFIX, method=2, log_n1=log_n1)
                assert np.array_equal(E.from_device(out, F), X[:out_len]), (logn, log_n1, out_len)


@pytest.mark.parametrize("fname", FIELD_IDS)
def test_partial_batch(fname):
    """The OMR epoch: a batch of receivers, each needing k = 2s outputs of its
    own vector (benches/common/mod.rs:10-12; s = 1000 is the reference's
    default): fused in one pass (k <= 1024), in two passes (k = 2000 at
    N = 2^18) and the full transform with a pruned last pass (N = 2^12)."""
    F = field(fname)
    for logn, batch, ks in ((12, 5, (1, 7, 256, 1024, 2000, 4096)), (18, 3, (2000, 200)), (20, 2, (1024,))):
        n = 1 << logn
        x = inputs(F, n * batch, logn + batch)
        t = E.to_device(x, F)
        want = [O.ntt(F.id, x[b * n:(b + 1) * n]) for b in range(batch)]
        for k in ks:
            out = E.empty_device(k * batch, F)
            E.partial_batch_device(t, out, F, n, batch, k)
            got = E.from_device(out, F)
            for b in range(batch):
                assert np.array_equal(got[b * k:(b + 1) * k], want[b][:k]), (logn, batch, k, b)


@pytest.mark.parametrize("fname", FIELD_IDS)
def test_host_forms(fname):
    """Host-buffer forms (copy in, run, copy out): inverse, batched forward,
    coset LDE and decode, against the oracle."""
    F = field(fname)
    n = 1 << 12
    x = inputs(F, n, 3)
    X = O.ntt(F.id, x)
    y = X.copy()
    E.intt(y, F)
    assert np.array_equal(y, x)
    b = 3
    xs = inputs(F, n * b, 4)
    ys = xs.astype(F.np_dtype)
    E.ntt_batched(ys, F, b)
    for i in range(b):
        assert np.array_equal(ys[i * n:(i + 1) * n], O.ntt(F.id, xs[i * n:(i + 1) * n]))
    m, cols = 256, 2
    c = inputs(F, m * cols, 5)
    ev = E.coset_lde(c, F, m, cols, 4)
    for j in range(cols):
        assert np.array_equal(ev[j * 4 * m:(j + 1) * 4 * m], O.coset_lde(F.id, c[j * m:(j + 1) * m], 4))
    back, nz = E.coset_decode(ev, F, m, cols, 4)
    assert np.array_equal(back, c) and nz == 0
    ev[7] = (int(ev[7]) + 1) % F.p  # a corrupted codeword has a nonzero tail
    _, nz = E.coset_decode(ev, F, m, cols, 4)
    assert nz > 0


@pytest.mark.parametrize("fname", FIELD_IDS)
def test_montgomery_words_pass_through(fname):
    """Every operation is F-linear, so words in Montgomery form (x R mod p,
    R = 2^64 for ark-ff's one-limb BigInt, 2^32 for a u32 form) go in and
    come out in that form: a caller holding Montgomery limbs needs no
    conversion of the buffer, only of the scalar omega."""
    F = field(fname)
    n = 1 << 12
    x = inputs(F, n, 21)
    for R in (1 << 64, 1 << 32):
        r = np.full(n, R % F.p, dtype=np.uint64)
        mont = lambda v: O.vec_op(F.id, 2, v, r[:v.size])  # noqa: E731
        assert np.array_equal(gpu_ntt(F, mont(x)), mont(O.ntt(F.id, x)))
        t = E.to_device(mont(x), F)
        E.intt_device(t, t, F)
        assert np.array_equal(E.from_device(t, F), mont(O.intt(F.id, x)))
        m = 64
        c = E.to_device(mont(x[:m]), F)
        ev = E.empty_device(4 * m, F)
        E.coset_lde_device(c, ev, F, m, 1, 4)
        assert np.array_equal(E.from_device(ev, F), mont(O.coset_lde(F.id, x[:m], 4)))
        out = E.empty_device(32, F)
        E.partial_device(E.to_device(mont(x), F), out, F, 32)
        assert np.array_equal(E.from_device(out, F), mont(O.ntt(F.id, x)[:32]))


@pytest.mark.parametrize("fname", ["Goldilocks", "BabyBear"])
def test_host_staging_large(fname):
    """Host buffers of 8 MiB and more go through the pinned staging ring
    (16 MiB slots, the last one ragged); omr_host_alloc buffers go direct.
    Every form must give the device path's result bit for bit."""
    F = field(fname)
    n = 1 << 22  # 32 MiB (Goldilocks) / 16 MiB (BabyBear)
    x = inputs(F, n, 11)
    t = E.to_device(x, F)
    E.ntt_device(t, t, F)
    want = E.from_device(t, F)
    y = x.astype(F.np_dtype)
    E.ArkRadix2().ntt(y, E.NttDomain(n, F))  # pageable: staged
    assert np.array_equal(y, want)
    E.intt(y, F)
    assert np.array_equal(y, x)
    pinned = E.host_empty(n, F)
    pinned[:] = x
    E.ArkRadix2().ntt(pinned, E.NttDomain(n, F))  # pinned: direct
    assert np.array_equal(pinned, want)
    # ragged chunking: 5 vectors of 2^20 (40 MiB / 20 MiB)
    nb, b = 1 << 20, 5
    xs = inputs(F, nb * b, 12)
    ts = E.to_device(xs, F)
    E.ntt_device(ts, ts, F, batch=b, n=nb)
    ys = xs.copy()
    E.ntt_batched(ys, F, b)
    assert np.array_equal(ys, E.from_device(ts, F))
    # coset LDE of 4 columns of 2^20 (out 4x larger) and its decode
    m, cols = 1 << 20, 4
    c = inputs(F, m * cols, 13)
    ev = E.coset_lde(c, F, m, cols, 4)
    dc, dev = E.to_device(c, F), E.empty_device(m * cols * 4, F)
    E.coset_lde_device(dc, dev, F, m, cols, 4)
    assert np.array_equal(ev, E.from_device(dev, F))
    back, nz = E.coset_decode(ev, F, m, cols, 4)
    assert np.array_equal(back, c) and nz == 0


def test_encoder_registry_names():
    names = [e.name() for e in E.encoders.all(12)]
    assert names == ["ArkRadix2", "LambdaBowers", "WinterfellFourStep", "Plonky3Radix2LayerSplit",
                     "LambdaRadix4", "TfheStockhamRadix8"]
    assert [e.name() for e in E.encoders.all_partial(50)] == [
        "DftPartial", "TransformDecomposition", "TransformDecompositionV2"]


@pytest.mark.parametrize("fname", FIELD_IDS)
def test_coset_lde_golden_and_decode(fname):  # the column encoder's semantics at golden sizes
    F = field(fname)
    for case in load("lde.json")[fname.lower()]:
        n, blow = case["n"], case["blowup"]
        c = E.to_device(np.array(case["coeffs"], dtype=np.uint64), F)
        ev = E.empty_device(n * blow, F)
        E.coset_lde_device(c, ev, F, n, 1, blow, case["shift"])
        assert [int(v) for v in E.from_device(ev, F)] == case["evals"]
        back = E.empty_device(n, F)
        nz = torch.zeros(1, dtype=torch.int32, device="cuda")
        E.coset_decode_device(ev, back, F, n, 1, blow, case["shift"], nonzero=nz)
        assert [int(v) for v in E.from_device(back, F)] == case["coeffs"]
        assert int(nz.item()) == 0


@pytest.mark.parametrize("fname", FIELD_IDS)
def test_coset_lde_columns(fname):
    F = field(fname)
    n, ncols, blow = 1 << 12, 8, 4
    x = inputs(F, n * ncols, 21)
    c = E.to_device(x, F)
    ev = E.empty_device(n * blow * ncols, F)
    E.coset_lde_device(c, ev, F, n, ncols, blow)
    got = E.from_device(ev, F).reshape(ncols, n * blow)
    for col in (0, 3, 7):
        assert np.array_equal(got[col], O.coset_lde(F.id, x[col * n:(col + 1) * n], blow))
    back = E.empty_device(n * ncols, F)
    nz = torch.zeros(1, dtype=torch.int32, device="cuda")
    E.coset_decode_device(ev, back, F, n, ncols, blow, nonzero=nz)
    assert torch.equal(back, c) and int(nz.item()) == 0
    # a corrupted evaluation is detected (its decode has nonzero high coefficients)
    bad = ev.clone()
    bad[5] = (bad[5] + 1) % F.p if F.bytes == 4 else bad[5] ^ 1
    E.coset_decode_device(bad, back, F, n, ncols, blow, nonzero=nz)
    assert int(nz.item()) > 0


@pytest.mark.parametrize("fname", FIELD_IDS)
def test_coset_lde_blowups(fname):
    """Blowups 2 to 64 and n from 1 up, several columns, default and custom
    shifts.  The first pass (radix = blowup) is skipped: the second reads the
    replicated input c[pos mod n] h^(pos mod n)."""
    F = field(fname)
    for i, (n, blow, ncols) in enumerate(((1, 4, 3), (2, 2, 2), (8, 8, 2), (256, 16, 3), (1 << 14, 2, 2),
                                          (1 << 12, 64, 1), (1 << 16, 4, 2))):
        shift = 0 if i % 2 == 0 else 11
        x = inputs(F, n * ncols, n + blow)
        c = E.to_device(x, F)
        ev = E.empty_device(n * blow * ncols, F)
        E.coset_lde_device(c, ev, F, n, ncols, blow, shift)
        got = E.from_device(ev, F).reshape(ncols, n * blow)
        for col in range(ncols):
            want = O.coset_lde(F.id, x[col * n:(col + 1) * n], blow, shift if shift else None)
            assert np.array_equal(got[col], want), (n, blow, ncols, shift, col)
        back = E.empty_device(n * ncols, F)
        nz = torch.zeros(1, dtype=torch.int32, device="cuda")
        E.coset_decode_device(ev, back, F, n, ncols, blow, shift, nonzero=nz)
        assert torch.equal(back, c) and int(nz.item()) == 0, (n, blow, shift)


def test_cpp_reference_test_suite():
    """tests/cpp/test_encoders.cpp: the reference's src/lib.rs tests in C++
    through include/omrntt.hpp (built by __graft_entry__.build)."""
    import subprocess
    exe = os.path.join(os.path.dirname(__file__), "cpp", "test_encoders")
    if not os.path.exists(exe):
        pytest.fail("tests/cpp/test_encoders not built (run __graft_entry__.build())")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]


def _batch_items(rng, count, lo, hi, layout_gap=0):
    items, off = [], 0
    for _ in range(count):
        ln = int(rng.integers(lo, hi + 1))
        items.append((off, off, ln, int(rng.integers(0, 2))))
        off += (1 << ln) + layout_gap
    return items, off


@pytest.mark.parametrize("fname", ["Goldilocks", "BabyBear", "KoalaBear", "TeddyBear"])
def test_batched_uneven_sizes_all_modes(fname):
    """Every execution mode of a batch plan against the oracle, in each field
    (8192-element batch tiles for Goldilocks / BabyBear / KoalaBear, 4096 for
    TeddyBear); the 2^19 / 2^20 vectors take the two-pass splits of the
    default radix cap ([10, 9] / [10, 10] for Goldilocks)."""
    F = field(fname)
    rng = np.random.default_rng(7)
    items, total = _batch_items(rng, 96, 4, 16, layout_gap=3)
    items += [(total, total, 4, 0), (total + 16, total + 16, 13, 1)]
    total += 16 + (1 << 13)
    items += [(total, total, 19, 0), (total + (1 << 19), total + (1 << 19), 20, 1)]
    total += (1 << 19) + (1 << 20)
    x = inputs(F, total, 12)
    t_in = E.to_device(x, F)
    plan = E.BatchPlan(F, items)
    info = plan.info()
    assert info["total_elems"] == sum(1 << it[2] for it in items)
    outs = {}
    for mode in ("persistent", "graph", "plain", "grouped", "grouped_graph"):
        out = torch.zeros_like(t_in)
        plan.execute(t_in, out, mode)
        outs[mode] = E.from_device(out, F)
    for (io, oo, ln, inv) in items:
        n = 1 << ln
        seg = x[io:io + n]
        exp = O.intt(F.id, seg) if inv else O.ntt(F.id, seg)
        for mode, got in outs.items():
            assert np.array_equal(got[oo:oo + n], exp), (mode, ln, inv)
    # graph replays with eager work in between stay bit-exact
    for mode in ("graph", "grouped_graph"):
        for rep in range(3):
            out = torch.zeros_like(t_in)
            plan.execute(t_in, out, mode)
            torch.cuda.synchronize()
            _ = (t_in.float().sum() if rep == 1 else None)
            assert np.array_equal(E.from_device(out, F), outs["plain"]), (mode, rep)
    plan.close()


@pytest.mark.parametrize("lanes", ["1", "3", "4"])
def test_grouped_lanes(lanes, monkeypatch):
    """The grouped form runs its (size, direction) groups on 1..4 streams
    (OMR_BATCH_LANES, read when the plan is made): every lane count gives
    the oracle's outputs, eagerly and over graph replays, in and out of place."""
    monkeypatch.setenv("OMR_BATCH_LANES", lanes)
    F = field("Goldilocks")
    rng = np.random.default_rng(int(lanes))
    items, total = _batch_items(rng, 60, 4, 15, layout_gap=1)
    x = inputs(F, total, 21)
    t_in = E.to_device(x, F)
    plan = E.BatchPlan(F, items)
    exp = x.copy()
    for (io, oo, ln, inv) in items:
        seg = x[io:io + (1 << ln)]
        exp[oo:oo + (1 << ln)] = O.intt(F.id, seg) if inv else O.ntt(F.id, seg)
    for rep in range(3):
        for mode in ("grouped", "grouped_graph"):
            out = t_in.clone()  # the gaps between vectors keep the input
            plan.execute(t_in, out, mode)
            torch.cuda.synchronize()
            assert np.array_equal(E.from_device(out, F), exp), (lanes, mode, rep)
    buf = t_in.clone()
    plan.execute(buf, buf, "grouped")
    torch.cuda.synchronize()
    assert np.array_equal(E.from_device(buf, F), exp), (lanes, "in place")
    plan.close()


def test_batch_plan_survives_twiddle_cache_eviction():
    """A plan bakes twiddle-table pointers into its device task tables; the
    context's table cache (bounded at 128 entries) must not free them while
    the plan lives.  Fill the cache with 140 other roots, then replay."""
    F = E.Goldilocks
    rng = np.random.default_rng(3)
    items, total = _batch_items(rng, 12, 4, 14)
    x = inputs(F, total, 5)
    t_in = E.to_device(x, F)
    plan = E.BatchPlan(F, items)
    first = torch.zeros_like(t_in)
    plan.execute(t_in, first, "graph")
    want = E.from_device(first, F)
    n = 1 << 10
    w = E.root_of_unity(F, n)
    y = E.to_device(inputs(F, n, 6), F)
    out = E.empty_device(n, F)
    for k in range(1, 2 * 140, 2):  # 140 distinct primitive roots w^k, k odd
        E.ntt_device(y, out, F, omega=pow(w, k, F.p))
    for mode in ("graph", "persistent"):
        again = torch.zeros_like(t_in)
        plan.execute(t_in, again, mode)
        assert np.array_equal(E.from_device(again, F), want), mode
    plan.close()


def test_batched_in_place():
    F = E.Goldilocks
    rng = np.random.default_rng(8)
    items, total = _batch_items(rng, 40, 8, 15)
    x = inputs(F, total, 13)
    t = E.to_device(x, F)
    plan = E.BatchPlan(F, items)
    for mode in ("persistent", "grouped"):
        t = E.to_device(x, F)
        plan.execute(t, t, mode)
        got = E.from_device(t, F)
        for (io, oo, ln, inv) in items:
            n = 1 << ln
            seg = x[io:io + n]
            exp = O.intt(F.id, seg) if inv else O.ntt(F.id, seg)
            assert np.array_equal(got[oo:oo + n], exp), (mode, ln, inv)
    plan.close()


@pytest.mark.slow
def test_batched_4096_uneven_full_size():
    """4096 Goldilocks vectors, log2 n ~ U{10..20} (seed 42), mixed directions:
    every vector of the batched forms matches the oracle (all 4096, on the
    host's threads), graph == plain bit for bit, and the forward-then-inverse
    round trip restores the input."""
    F = E.Goldilocks
    rng = np.random.default_rng(42)
    logs = rng.integers(10, 21, size=4096)
    dirs = rng.integers(0, 2, size=4096)
    items, off = [], 0
    for ln, d in zip(logs, dirs):
        items.append((off, off, int(ln), int(d)))
        off += 1 << int(ln)
    x = torch.randint(0, 2**62, (off,), dtype=torch.int64, device="cuda")  # < p
    plan = E.BatchPlan(F, items)
    y_graph = torch.empty_like(x)
    y_plain = torch.empty_like(x)
    plan.execute(x, y_graph, "graph")
    plan.execute(x, y_plain, "plain")
    assert torch.equal(y_graph, y_plain)
    y_grouped = torch.empty_like(x)
    for _ in range(2):  # the capturing call, then a replay
        y_grouped.zero_()
        plan.execute(x, y_grouped, "grouped_graph")
        assert torch.equal(y_grouped, y_plain)
    del y_grouped
    inv_items = [(o, o, ln, 1 - d) for (o, _, ln, d) in items]
    plan_inv = E.BatchPlan(F, inv_items)
    back = torch.empty_like(x)
    plan_inv.execute(y_graph, back, "graph")
    assert torch.equal(back, x)
    del back
    xs = E.from_device(x, F)
    exp = O.ntt_batch(F.id, xs, [it[0] for it in items], [it[2] for it in items], [it[3] for it in items])
    del xs
    ys = E.from_device(y_graph, F)
    bad = [i for i, (o, _, ln, _) in enumerate(items) if not np.array_equal(ys[o:o + (1 << ln)], exp[o:o + (1 << ln)])]
    assert not bad, bad[:10]
    plan.close()
    plan_inv.close()


MAX_LOG = {"BabyBear": 27, "KoalaBear": 24, "TeddyBear": 30, "Goldilocks": 32}


def _sum_mod_p(t, p, sign=False):
    """sum_j (+-1)^j t[j] mod p, exact: 32-bit halves summed in int64 chunks."""
    total = 0
    chunk = 1 << 26
    for s0 in range(0, t.numel(), chunk):
        v = t[s0:s0 + chunk].to(torch.int64)
        lo, hi = v & 0xFFFFFFFF, (v >> 32) & 0xFFFFFFFF
        if sign:  # chunk starts at an even index
            total += int(lo[0::2].sum()) + (int(hi[0::2].sum()) << 32)
            total -= int(lo[1::2].sum()) + (int(hi[1::2].sum()) << 32)
        else:
            total += int(lo.sum()) + (int(hi.sum()) << 32)
    return total % p


@pytest.mark.parametrize("fname", FIELD_IDS)
def test_maximum_sizes(fname):
    """Each field at its two-adicity limit (the largest N NttDomain::new
    accepts, src/encoder.rs:28-32): forward then inverse gives the input
    back, and X[0] = sum x, X[N/2] = sum (-1)^j x_j (closed forms, summed on
    the device in exact integer chunks)."""
    F = field(fname)
    n = 1 << MAX_LOG[fname]
    if F.bytes == 8:
        x = torch.randint(0, 2**62, (n,), dtype=torch.int64, device="cuda")
    else:  # int32 words < min(p, 2^31)
        x = torch.randint(0, min(F.p, 2**31 - 1), (n,), dtype=torch.int32, device="cuda")
    X = torch.empty_like(x)
    E.ntt_device(x, X, F)
    as_u = (lambda v: v.to(torch.int64) & 0xFFFFFFFF) if F.bytes == 4 else (lambda v: v)
    got0 = int(as_u(X[0:1]).item()) % (1 << 64)
    gotn2 = int(as_u(X[n // 2:n // 2 + 1]).item()) % (1 << 64)
    assert got0 == _sum_mod_p(x, F.p), fname
    assert gotn2 == _sum_mod_p(x, F.p, sign=True), fname
    back = torch.empty_like(x)
    E.intt_device(X, back, F)
    assert torch.equal(back, x), fname
    del x, X, back
    torch.cuda.empty_cache()


def _random_sweep(seed, cases, max_ln, max_batch):
    """Seeded sweep over shapes the fixed tests do not enumerate: every field,
    sizes 2^1..2^max_ln, every full encoder where it applies, batches, prefix
    and stride-subset lengths (ragged prefixes included), explicit methods and
    sub-NTT sizes, explicit schedules and register radices, inverse."""
    rng = np.random.default_rng(seed)
    full = [E.ARK_RADIX2, E.DFT, E.LAMBDA_BOWERS, E.LAMBDA_RADIX4, E.PLONKY3_LAYER_SPLIT,
            E.WINTERFELL_FOUR_STEP, E.TFHE_STOCKHAM_RADIX8]
    for case in range(cases):
        F = E.FIELDS[int(rng.integers(0, 4))]
        ln = int(rng.integers(1, max_ln + 1))
        n = 1 << ln
        x = inputs(F, n, 5000 + case)
        X = O.ntt(F.id, x)
        t = E.to_device(x, F)
        var = full[int(rng.integers(0, len(full)))]
        if lib_supports(var, ln) and not (var == E.DFT and ln > 12):
            out = E.empty_device(n, F)
            E.ntt_device(t, out, F, var)
            assert np.array_equal(E.from_device(out, F), X), (case, F.name, ln, var)
        b = int(rng.integers(1, max_batch + 1))
        xb = inputs(F, n * b, 6000 + case)
        tb = E.to_device(xb, F)
        ob = E.empty_device(n * b, F)
        E.ntt_device(tb, ob, F, batch=b, n=n)
        gb = E.from_device(ob, F)
        for i in range(b):
            assert np.array_equal(gb[i * n:(i + 1) * n], O.ntt(F.id, xb[i * n:(i + 1) * n])), (case, i)
        out_len = int(rng.integers(1, n + 1))
        meth = int(rng.integers(0, 4))
        if ln > 14 and meth in (1, 2):  # direct sums and tiny sub-NTTs cost out_len x N
            meth = 0
        l1 = int(rng.integers(0, ln + 1)) if meth == 2 else 0
        po = E.empty_device(out_len, F)
        try:
            E.partial_device(t, po, F, out_len, mode=E.PREFIX, method=meth, log_n1=l1)
        except E.OmrError as err:  # an explicit sub-NTT size that cannot serve this prefix
            assert err.status == 11 and meth == 2, (case, ln, out_len, meth, l1)
        else:
            assert np.array_equal(E.from_device(po, F), X[:out_len]), (case, F.name, ln, out_len, meth, l1)
        m = 1 << int(rng.integers(0, ln + 1))
        so = E.empty_device(m, F)
        E.partial_device(t, so, F, m, mode=E.STRIDE)
        assert np.array_equal(E.from_device(so, F), X[::n // m]), (case, F.name, ln, m)
        maxlr = 12 if F.bytes == 8 else 13
        parts, left = [], ln
        while left:
            r = int(rng.integers(1, min(left, maxlr) + 1))
            parts.append(r)
            left -= r
        if len(parts) <= 8:
            le = int(rng.integers(0, 5))
            so2 = E.empty_device(n, F)
            E.ntt_schedule_device(t, so2, F, parts, log_e=le if le != 1 else 0)
            assert np.array_equal(E.from_device(so2, F), X), (case, F.name, parts, le)
            E.ntt_schedule_device(so2, so2, F, parts, log_e=le if le != 1 else 0, inverse=True)
            assert torch.equal(so2, t), (case, F.name, parts, le)


def test_random_shapes_against_oracle():
    _random_sweep(2024, 250, 14, 4)


def test_random_large_shapes_against_oracle():
    """Sizes up to 2^20, where every radix runs its full-column kernels."""
    _random_sweep(7777, 80, 20, 2)


def lib_supports(variant, log_n):
    return bool(E.lib().omr_variant_supports(variant, log_n))


def test_random_structured_against_oracle():
    """Seeded sweep over the larger and structured paths: prefixes of
    N = 2^15..2^20 (one- and multi-pass fused, pruned full transform), the
    batched prefix, coset LDE / decode with random blowup, shift and columns,
    and uneven batch plans in every execution mode."""
    rng = np.random.default_rng(77)
    for case in range(24):
        F = E.FIELDS[case % 4]
        ln = int(rng.integers(15, 21))
        n = 1 << ln
        x = inputs(F, n, 7000 + case)
        X = O.ntt(F.id, x)
        t = E.to_device(x, F)
        out_len = int(min(n, 2 ** rng.uniform(1, ln)))
        po = E.empty_device(out_len, F)
        E.partial_device(t, po, F, out_len, mode=E.PREFIX)
        assert np.array_equal(E.from_device(po, F), X[:out_len]), (case, F.name, ln, out_len)
        b, k = int(rng.integers(1, 4)), int(min(n, 2 ** rng.uniform(0, ln - 2)))
        xb = inputs(F, n * b, 7100 + case)
        ob = E.empty_device(k * b, F)
        E.partial_batch_device(E.to_device(xb, F), ob, F, n, b, k)
        gb = E.from_device(ob, F)
        for i in range(b):
            assert np.array_equal(gb[i * k:(i + 1) * k], O.ntt(F.id, xb[i * n:(i + 1) * n])[:k]), (case, i, k)
        blow = 1 << int(rng.integers(1, 5))
        lm = int(rng.integers(0, min(14, F.TWO_ADICITY - 4) + 1))
        m, cols = 1 << lm, int(rng.integers(1, 4))
        shift = int(rng.integers(2, 1000)) if case % 2 else 0
        c = inputs(F, m * cols, 7200 + case)
        ev = E.empty_device(m * blow * cols, F)
        E.coset_lde_device(E.to_device(c, F), ev, F, m, cols, blow, shift)
        got = E.from_device(ev, F)
        for j in range(cols):
            want = O.coset_lde(F.id, c[j * m:(j + 1) * m], blow, shift if shift else None)
            assert np.array_equal(got[j * m * blow:(j + 1) * m * blow], want), (case, m, blow, shift, j)
        back = E.empty_device(m * cols, F)
        nz = torch.zeros(1, dtype=torch.int32, device="cuda")
        E.coset_decode_device(ev, back, F, m, cols, blow, shift, nonzero=nz)
        assert np.array_equal(E.from_device(back, F), c) and int(nz.item()) == 0, (case, m, blow)
    for case in range(6):
        F = E.FIELDS[case % 4]
        items, total = _batch_items(rng, int(rng.integers(1, 40)), 4, 15, layout_gap=int(rng.integers(0, 5)))
        x = inputs(F, total, 7300 + case)
        t_in = E.to_device(x, F)
        plan = E.BatchPlan(F, items)
        for mode in ("persistent", "graph", "plain", "grouped", "grouped_graph"):
            out = torch.zeros_like(t_in)
            plan.execute(t_in, out, mode)
            got = E.from_device(out, F)
            for (io, oo, ln, inv) in items:
                seg = x[io:io + (1 << ln)]
                want = O.intt(F.id, seg) if inv else O.ntt(F.id, seg)
                assert np.array_equal(got[oo:oo + (1 << ln)], want), (case, mode, ln, inv)
        plan.close()


def test_concurrent_host_threads_share_one_context():
    """The reference trait is Send + Sync (src/encoder.rs:70): concurrent
    calls on one engine must be safe.  Four host threads share the process's
    context (ctypes drops the GIL during each call), each on its own stream,
    mixing host-buffer calls (staging ring, table cache) with device calls
    (forward, inverse, prefix) at different sizes; every result is checked
    against the oracle."""
    import threading
    errors = []

    def worker(k):
        try:
            rng = np.random.default_rng(100 + k)
            stream = torch.cuda.Stream()
            for it in range(6):
                F = (E.Goldilocks, E.BabyBear)[(k + it) % 2]
                ln = int(rng.integers(8, 19))
                n = 1 << ln
                x = inputs(F, n, 1000 * k + it)
                if it % 3 == 0:  # host buffer: the trait path
                    buf = x.astype(F.np_dtype)  # the field's word type
                    E.ArkRadix2().ntt(buf, E.NttDomain(n, F))
                    assert np.array_equal(buf.astype(np.uint64), O.ntt(F.id, x)), ("host", k, it, ln)
                    continue
                with torch.cuda.stream(stream):
                    t = E.to_device(x, F)
                    o = torch.empty_like(t)
                    if it % 3 == 1:
                        E.ntt_device(t, o, F)
                        E.intt_device(o, t, F)
                        stream.synchronize()
                        assert np.array_equal(E.from_device(o, F), O.ntt(F.id, x)), ("fwd", k, it, ln)
                        assert np.array_equal(E.from_device(t, F), x), ("inv", k, it, ln)
                    else:
                        m = 1 << int(rng.integers(2, ln))
                        E.partial_device(t, o, F, m)
                        stream.synchronize()
                        assert np.array_equal(E.from_device(o, F)[:m], O.ntt(F.id, x)[:m]), ("prefix", k, it, ln, m)
        except Exception as e:  # reported by the main thread
            errors.append(repr(e))

    E.context()
    threads = [threading.Thread(target=worker, args=(k,)) for k in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=120)
    assert not any(th.is_alive() for th in threads), "a worker did not finish"
    assert not errors, errors


@pytest.mark.parametrize("fname", FIELD_IDS)
def test_stride_subset_one_launch(fname):
    """Stride subsets of 16 <= m <= 512 outputs run as one kernel: folds over
    up to 512 workgroups finished by two levels of last-workgroup-done
    counters and an m-point NTT in LDS.  Every m around that range (8 and
    1024 take the multi-launch path), vectors too short for a full row subset
    (fallback), calls back to back with different m (the counters are cleared
    by each launch), and in place (d_in == d_out)."""
    F = field(fname)
    for logn in (8, 10, 14, 18, 22):
        n = 1 << logn
        x = inputs(F, n, 700 + logn)
        X = O.ntt(F.id, x)
        t = E.to_device(x, F)
        for logm in range(3, 11):
            m = 1 << logm
            if m > n:
                continue
            print(f"case: stride subset {F} 2^{logn} m={m}", flush=True)
            out = E.empty_device(m, F)
            for rep in range(2):
                E.partial_device(t, out, F, m, mode=E.STRIDE)
                assert np.array_equal(E.from_device(out, F), X[:: n // m]), (logn, m, rep)
        if n >= 1 << 10:
            buf = t.clone()
            E.partial_device(buf, buf, F, 256, mode=E.STRIDE)
            assert np.array_equal(E.from_device(buf, F)[:256], X[:: n // 256]), (logn, "in place")


@pytest.mark.parametrize("fname", FIELD_IDS)
def test_coset_lde_columns_two_streams(fname):
    """Column batches larger than one group of ~128 MiB of intermediates run
    their groups on two streams (the caller's and a second one with its own
    intermediates).  2^20 -> 2^22 columns, two more columns than one group
    holds: sampled columns of both streams against the oracle, the round
    trip, and a corrupted evaluation on either stream caught by decode's
    nonzero count."""
    F = field(fname)
    n, blow = 1 << 20, 4
    group = (128 << 20) // (n * blow * F.bytes)
    ncols = group + 2
    print(f"case: two-stream LDE {F} {ncols} columns of 2^20 -> 2^22 (group {group})", flush=True)
    x = inputs(F, n * ncols, 77)
    c = E.to_device(x, F)
    ev = E.empty_device(n * blow * ncols, F)
    E.coset_lde_device(c, ev, F, n, ncols, blow)
    for col in (0, group - 1, group, ncols - 1):  # first stream, then the second
        got = E.from_device(ev[col * n * blow:(col + 1) * n * blow], F)
        assert np.array_equal(got, O.coset_lde(F.id, x[col * n:(col + 1) * n], blow)), col
    back = E.empty_device(n * ncols, F)
    nz = torch.zeros(1, dtype=torch.int32, device="cuda")
    E.coset_decode_device(ev, back, F, n, ncols, blow, nonzero=nz)
    assert torch.equal(back, c) and int(nz.item()) == 0
    for col in (1, ncols - 1):
        bad = ev.clone()
        i = col * n * blow + 12345
        bad[i] = (bad[i] + 1) % F.p if F.bytes == 4 else bad[i] ^ 1
        E.coset_decode_device(bad, back, F, n, ncols, blow, nonzero=nz)
        assert int(nz.item()) > 0, col
        del bad
    torch.cuda.empty_cache()


@pytest.mark.parametrize("fname", FIELD_IDS)
def test_domain_tables(fname):
    """NttDomain's public twiddles / bitrev_twiddles (src/encoder.rs:14-18)
    built on the device, against the oracle's restatement of NttDomain::new
    (itself pinned to Python big-int powers), N = 1 .. 2^20, the ark root
    and a custom root; the device form and the host form agree."""
    F = field(fname)
    for logn in (0, 1, 2, 3, 10, 16, 20):
        n = 1 << logn
        for omega in (None, pow(E.root_of_unity(F, n), 3, F.p) if n > 2 else None):
            d = E.NttDomain(n, F, omega)
            ref = O.Domain(F.id, n, d.omega)
            assert np.array_equal(d.twiddles, ref.twiddles), (logn, omega)
            assert np.array_equal(d.bitrev_twiddles, ref.bitrev_twiddles[:n // 2]), (logn, omega)
            tw, br = E.domain_tables_device(F, n, d.omega)
            assert np.array_equal(E.from_device(tw, F), ref.twiddles), (logn, omega, "device")
            assert np.array_equal(E.from_device(br, F), ref.bitrev_twiddles[:n // 2]), (logn, omega, "device")


@pytest.mark.parametrize("fname", ["Goldilocks", "BabyBear", "TeddyBear"])
def test_post_twiddle_batches_and_odd_sizes(fname):
    """Vectors of 2^22 and more take each next pass's twiddle one pass early
    (the first passes' outputs carry it; an inverse's N^-1 rides in the last
    such chain; 31/32-bit fields at register radix 16 and pass radix
    2^6..2^11 only, other passes keep their own twiddle): batches of 2^22,
    2^23 (odd log N) and explicit schedules mixing passes that can and
    cannot carry it, forward and inverse, in place and out of place, against
    the oracle."""
    F = field(fname)
    for logn, batch in ((22, 3), (23, 1)):
        n = 1 << logn
        print(f"case: post-twiddle {fname} 2^{logn} x {batch}", flush=True)
        x = inputs(F, n * batch, 3300 + logn)
        t = E.to_device(x, F)
        out = torch.empty_like(t)
        E.ntt_device(t, out, F, batch=batch, n=n)
        got = E.from_device(out, F)
        for b in range(batch):
            assert np.array_equal(got[b * n:(b + 1) * n], O.ntt(F.id, x[b * n:(b + 1) * n])), (logn, b)
        back = torch.empty_like(t)
        E.intt_device(out, back, F, batch=batch, n=n)
        assert torch.equal(back, t), (logn, "inverse")
        buf = t.clone()
        E.ntt_device(buf, buf, F, batch=batch, n=n)
        assert torch.equal(buf, out), (logn, "in place")
        E.intt_device(buf, buf, F, batch=batch, n=n)
        assert torch.equal(buf, t), (logn, "inverse in place")
        del out, back, buf
    n = 1 << 22
    x = inputs(F, n, 3399)
    X = O.ntt(F.id, x)
    t = E.to_device(x, F)
    scheds = [[6, 8, 8], [11, 11], [4, 6, 6, 6], [12, 10], [10, 12]]
    if F.bytes == 4:
        scheds += [[5, 9, 8], [8, 8, 6], [13, 9], [9, 13], [7, 7, 8]]
    for sched in scheds:
        print(f"case: post-twiddle {fname} 2^22 schedule {sched}", flush=True)
        o2 = torch.empty_like(t)
        E.ntt_schedule_device(t, o2, F, sched)
        assert np.array_equal(E.from_device(o2, F), X), sched
        E.ntt_schedule_device(o2, o2, F, sched, inverse=True)
        assert torch.equal(o2, t), (sched, "inverse")
    torch.cuda.empty_cache()
