"""GPU parity of the word-width entry points (omr_ntt_device_words,
omr_ntt_words): BabyBear, KoalaBear and TeddyBear vectors held one value per
64-bit word -- the one-u64-limb storage of the crate's define_field! types
(src/fields.rs:1-17) -- give the same values as the 32-bit path and the
oracle, with zero upper halves; partial variants keep the tail; ark-ff-style
Montgomery words (x 2^64 mod p) come back as the Montgomery words of the
transform."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import omr_ntt as E  # noqa: E402
import oracle as O  # noqa: E402

from _gpu import inputs  # noqa: E402

FIELDS32 = ["BabyBear", "KoalaBear", "TeddyBear"]


def field(name):
    return {f.name: f for f in E.FIELDS}[name]


def words_dev(x):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.uint64).view(np.int64).copy()).cuda()


def words_host(t):
    return t.cpu().numpy().view(np.uint64).copy()


@pytest.mark.parametrize("fname", FIELDS32)
def test_words_full_variants(fname):
    F = field(fname)
    for logn, batch in ((1, 3), (4, 1), (10, 2), (13, 1), (16, 2), (20, 1), (22, 1)):
        n = 1 << logn
        print(f"case: words {fname} 2^{logn} batch {batch}", flush=True)
        x = inputs(F, n * batch, 500 + logn)
        exp = np.concatenate([O.ntt(F.id, x[b * n:(b + 1) * n]) for b in range(batch)])
        xw = words_dev(x)
        out = torch.full_like(xw, -1)
        E.ntt_device_words(xw, out, F, batch=batch, n=n)
        assert np.array_equal(words_host(out), exp), (logn, "out of place")
        buf = xw.clone()
        E.ntt_device_words(buf, buf, F, batch=batch, n=n)
        assert np.array_equal(words_host(buf), exp), (logn, "in place")
    n = 1 << 12
    x = inputs(F, n, 91)
    exp = O.ntt(F.id, x)
    for v in (E.LAMBDA_BOWERS, E.LAMBDA_RADIX4, E.WINTERFELL_FOUR_STEP, E.PLONKY3_LAYER_SPLIT,
              E.TFHE_STOCKHAM_RADIX8, E.DFT):
        out = torch.empty(n, dtype=torch.int64, device="cuda")
        E.ntt_device_words(words_dev(x), out, F, v)
        assert np.array_equal(words_host(out), exp), v


@pytest.mark.parametrize("fname", FIELDS32)
def test_words_partial_variants_keep_the_tail(fname):
    F = field(fname)
    for logn in (6, 14, 20):
        n = 1 << logn
        x = inputs(F, n, 600 + logn)
        full = O.ntt(F.id, x)
        for s in (1, 3, 50, 16384):
            ol = min(2 * s, n)
            for v in (E.DFT_PARTIAL, E.TRANSFORM_DECOMPOSITION, E.TRANSFORM_DECOMPOSITION_V2):
                if v == E.DFT_PARTIAL and logn > 14:
                    continue
                buf = words_dev(x)
                E.ntt_device_words(buf, buf, F, v, s=s)
                got = words_host(buf)
                assert np.array_equal(got[:ol], full[:ol]), (logn, s, v)
                assert np.array_equal(got[ol:], x[ol:]), (logn, s, v, "tail")
                h = x.copy()  # host drop-in on 64-bit words
                E.ntt_words(h, F, v, s=s)
                assert np.array_equal(h[:ol], full[:ol]), (logn, s, v, "host")
                assert np.array_equal(h[ol:], x[ol:]), (logn, s, v, "host tail")


@pytest.mark.parametrize("fname", FIELDS32)
def test_words_montgomery_pass_through(fname):
    """ark-ff stores a one-limb element as x 2^64 mod p; the transform of
    those words is the Montgomery form of the transform (linearity)."""
    F = field(fname)
    n = 1 << 18
    x = inputs(F, n, 7)
    R = pow(2, 64, F.p)
    mont = np.array([(int(v) * R) % F.p for v in x[:4096]], dtype=np.uint64)
    mont = np.resize(mont, n)  # tile the converted block (the values only need to be < p)
    xs = np.resize(x[:4096], n)
    buf = mont.copy()
    E.ntt_words(buf, F)
    exp = O.ntt(F.id, xs)
    want = np.array([(int(v) * R) % F.p for v in exp[:2048]], dtype=np.uint64)
    assert np.array_equal(buf[:2048], want)
    assert not (buf >> np.uint64(32)).any()


def test_words_width_rules():
    gl = E.Goldilocks
    n = 1 << 10
    x = inputs(gl, n, 3)
    t = E.to_device(x, gl)
    out = torch.empty_like(t)
    L = E.lib()
    w = E.root_of_unity(gl, n)
    # Goldilocks words are already 8 bytes: the plain path
    assert L.omr_ntt_device_words(E.context(), gl.id, 0, E._dptr(t), E._dptr(out), n, 1, w, 0, 8, None) == 0
    torch.cuda.synchronize()
    assert np.array_equal(E.from_device(out, gl), O.ntt(gl.id, x))
    assert L.omr_ntt_device_words(E.context(), gl.id, 0, E._dptr(t), E._dptr(out), n, 1, w, 0, 4, None) == 12
    bb = E.BabyBear
    wb = E.root_of_unity(bb, n)
    assert L.omr_ntt_device_words(E.context(), bb.id, 0, E._dptr(t), E._dptr(out), n, 1, wb, 0, 2, None) == 12
    assert L.omr_ntt_device_words(E.context(), bb.id, 0, E._dptr(t), E._dptr(out), 1000, 1, wb, 0, 8, None) == 2
