"""Pins the CPU oracle (oracle/omr_oracle.c) against pure-Python big-int
golden vectors, closed forms and cross-agreement of the restated encoders.
Mirrors the reference's relational test matrix (src/lib.rs:24-180) for all
four fields."""
import json
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
NAMES = ["babybear", "koalabear", "goldilocks", "teddybear"]
POWERS_OF_TWO = [1 << k for k in range(1, 17)]          # src/lib.rs:24-26
EVEN_POWERS_OF_TWO = [1 << k for k in range(2, 17, 2)]   # src/lib.rs:27
POWERS_OF_TWO_DIV_BY_3 = [8, 64, 512, 4096, 32768]       # src/lib.rs:28


def load(name):
    with open(os.path.join(GOLD, name)) as fh:
        return json.load(fh)


@pytest.mark.parametrize("fname", NAMES)
def test_field_ops_vs_bigint(fname):
    f = O.FIELDS[fname]
    g = load("fieldops.json")[fname]
    for op, key in ((0, "add"), (1, "sub"), (2, "mul")):
        out = O.vec_op(f, op, g["a"], g["b"])
        assert [int(v) for v in out] == g[key], (fname, key)
    rng = np.random.default_rng(7)
    for _ in range(2000):
        a, b = (int(v) for v in rng.integers(0, O.P[f], 2, dtype=np.uint64))
        assert O.mul(f, a, b) == O.mul_textbook(f, a, b) == (a * b) % O.P[f]


@pytest.mark.parametrize("fname", NAMES)
def test_roots_of_unity_closed_form(fname):
    f = O.FIELDS[fname]
    g = load("omega.json")[fname]
    assert g["two_adicity"] == O.TWO_ADICITY[f]
    for k, w in g["omega"].items():
        assert O.root_of_unity(f, 1 << int(k)) == w
    assert O.root_of_unity(f, 1 << (O.TWO_ADICITY[f] + 1)) == 0
    assert O.root_of_unity(f, 12) == 0


def test_goldilocks_small_roots_are_powers_of_two():
    p = O.P[2]
    for n, k in ((2, 96), (4, 48), (8, 120), (16, 156), (32, 78), (64, 39)):
        assert O.root_of_unity(2, n) == pow(2, k, p)


@pytest.mark.parametrize("fname", NAMES)
def test_dft_golden(fname):
    f = O.FIELDS[fname]
    for case in load("dft.json")[fname]:
        x = np.array(case["x"], dtype=np.uint64)
        assert [int(v) for v in O.dft(f, x)] == case["X"]
        assert [int(v) for v in O.ntt(f, x)] == case["X"]
        assert [int(v) for v in O.intt(f, np.array(case["X"], dtype=np.uint64))] == case["x"]


@pytest.mark.parametrize("fname", NAMES)
def test_domain_tables(fname):
    f = O.FIELDS[fname]
    n = 64
    d = O.Domain(f, n)
    p = O.P[f]
    assert [int(v) for v in d.twiddles] == [pow(d.omega, k, p) for k in range(n)]
    for i in range(n // 2):
        assert int(d.bitrev_twiddles[i]) == pow(d.omega, O.bitrev(i, 5), p)


FULL = ["ArkRadix2", "LambdaBowers", "WinterfellFourStep", "Plonky3Radix2LayerSplit"]


@pytest.mark.parametrize("fname", NAMES)
def test_dft_agrees_with_ark(fname):  # src/lib.rs:85-88
    f = O.FIELDS[fname]
    for n in (64, 128, 256):
        x = O.uniform(f, n, n)
        assert np.array_equal(O.encoder("Dft", f, x), O.encoder("ArkRadix2", f, x))


@pytest.mark.parametrize("fname", NAMES)
@pytest.mark.parametrize("name", FULL)
def test_full_encoders_agree(fname, name):  # src/lib.rs:114-127
    f = O.FIELDS[fname]
    for n in POWERS_OF_TWO:
        x = O.uniform(f, n, 1000 + n)
        exp = O.encoder("ArkRadix2", f, x)
        assert np.array_equal(O.encoder(name, f, x), exp), (name, n)


@pytest.mark.parametrize("fname", NAMES)
def test_radix4_and_stockham8(fname):  # src/lib.rs:130-139
    f = O.FIELDS[fname]
    for n in EVEN_POWERS_OF_TWO:
        x = O.uniform(f, n, n)
        assert np.array_equal(O.encoder("LambdaRadix4", f, x), O.encoder("ArkRadix2", f, x))
    for n in POWERS_OF_TWO_DIV_BY_3:
        x = O.uniform(f, n, n)
        assert np.array_equal(O.encoder("TfheStockhamRadix8", f, x), O.encoder("ArkRadix2", f, x))
    with pytest.raises(ValueError):
        O.encoder("LambdaRadix4", f, O.uniform(f, 8, 1))
    with pytest.raises(ValueError):
        O.encoder("TfheStockhamRadix8", f, O.uniform(f, 16, 1))


@pytest.mark.parametrize("fname", NAMES)
def test_partial_encoders(fname):  # src/lib.rs:141-165
    f = O.FIELDS[fname]
    for s in (2, 4, 8, 16, 50):
        for n in (64, 128, 256):
            x = O.uniform(f, n, s * n)
            m = min(2 * s, n)
            exp = O.encoder("ArkRadix2", f, x)
            got = O.encoder("DftPartial", f, x, s=s)
            assert np.array_equal(got[:m], exp[:m])
            assert np.array_equal(got[m:], x[m:])  # tail keeps the input (dft.rs:88)
        for n in POWERS_OF_TWO:
            x = O.uniform(f, n, s + n)
            m = min(2 * s, n)
            exp = O.encoder("ArkRadix2", f, x)
            assert np.array_equal(O.encoder("TransformDecomposition", f, x, s=s)[:m], exp[:m])
        for n in (1024, 4096, 32768, 65536):
            x = O.uniform(f, n, s + 3 * n)
            m = min(2 * s, n)
            exp = O.encoder("ArkRadix2", f, x)
            assert np.array_equal(O.encoder("TransformDecompositionV2", f, x, s=s)[:m], exp[:m])


def test_linearity():  # src/lib.rs:90-112, all fields
    for f in range(4):
        p = O.P[f]
        n = 32
        x, y = O.uniform(f, n, 1), O.uniform(f, n, 2)
        a, b = 12345 % p, (p - 7)
        xy = O.vec_op(f, 0, O.vec_op(f, 2, np.full(n, a, np.uint64), x),
                      O.vec_op(f, 2, np.full(n, b, np.uint64), y))
        lhs = O.dft(f, xy)
        rhs = O.vec_op(f, 0, O.vec_op(f, 2, np.full(n, a, np.uint64), O.dft(f, x)),
                       O.vec_op(f, 2, np.full(n, b, np.uint64), O.dft(f, y)))
        assert np.array_equal(lhs, rhs)


def test_convolution_theorem():
    # NTT(x) * NTT(y) = NTT(cyclic convolution of x and y): catches a wrong
    # root or a transposed index that linearity alone would pass.
    for f in range(4):
        p = O.P[f]
        n = 16
        x, y = O.uniform(f, n, 3), O.uniform(f, n, 4)
        conv = [sum(int(x[i]) * int(y[(k - i) % n]) for i in range(n)) % p for k in range(n)]
        lhs = O.vec_op(f, 2, O.ntt(f, x), O.ntt(f, y))
        assert [int(v) for v in lhs] == [int(v) for v in O.ntt(f, np.array(conv, np.uint64))]


@pytest.mark.parametrize("fname", NAMES)
def test_lde_golden_and_decode(fname):
    f = O.FIELDS[fname]
    for case in load("lde.json")[fname]:
        ev = O.coset_lde(f, case["coeffs"], case["blowup"], case["shift"])
        assert [int(v) for v in ev] == case["evals"]
        c, bad = O.coset_decode(f, ev, case["n"], case["blowup"], case["shift"])
        assert bad == 0 and [int(v) for v in c] == case["coeffs"]
        ev2 = ev.copy()
        ev2[3] = (int(ev2[3]) + 1) % O.P[f]
        assert O.coset_decode(f, ev2, case["n"], case["blowup"], case["shift"])[1] > 0


def test_stride_subset_fold_identity():
    for f in range(4):
        n = 1024
        x = O.uniform(f, n, 9)
        full = O.ntt(f, x)
        for m in (1, 2, 16, 256, 1024):
            assert np.array_equal(O.stride_subset(f, x, m), full[:: n // m])


def test_large_partial_matrix():  # src/lib.rs:167-180 at sizes that run fast here
    for f in (0, 2):
        for n in (1 << 14, 1 << 15):
            x = O.uniform(f, n, n)
            exp = O.ntt(f, x)
            for s in (1, 3, 101, 16383, 16384):
                m = min(2 * s, n)
                assert np.array_equal(O.encoder("TransformDecomposition", f, x, s=s)[:m], exp[:m])


def test_ntt_batch_matches_per_vector_calls():
    """orc_ntt_batch (uneven vectors on the host's threads) equals one
    orc_ntt / orc_intt call per vector, and leaves the gaps untouched."""
    rng = np.random.default_rng(9)
    for f in range(4):
        items, off = [], 0
        for _ in range(40):
            ln = int(rng.integers(1, 12))
            items.append((off, ln, int(rng.integers(0, 2))))
            off += (1 << ln) + 2
        x = O.uniform(f, off, 3)
        got = O.ntt_batch(f, x, [i[0] for i in items], [i[1] for i in items], [i[2] for i in items])
        exp = x.copy()
        for o, ln, inv in items:
            seg = x[o:o + (1 << ln)]
            exp[o:o + (1 << ln)] = O.intt(f, seg) if inv else O.ntt(f, seg)
        assert np.array_equal(got, exp), f
