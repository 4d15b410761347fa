"""omr_ntt_device_noalloc (intermediates and twiddle tables in the library's
fixed device array, no allocation) against omr_ntt_device / omr_intt_device
through the C ABI: bit-exact on two calls in a row, the second with a smaller
input (other tables, other schedule), forward and inverse, in and out of
place, batches, and the size limit refused before any GPU work."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import omr_ntt as E  # noqa: E402
import oracle as O  # noqa: E402

from _gpu import inputs  # noqa: E402

MAXB = 32 << 20  # OMR_NOALLOC_MAX_BYTES


def _noalloc(x, out, F, n, batch=1, inverse=False, variant=E.ARK_RADIX2, omega=None):
    omega = E.root_of_unity(F, n) if omega is None else omega
    return E.lib().omr_ntt_device_noalloc(E.context(), F.id, variant, ctypes.c_void_p(x.data_ptr()),
                                          ctypes.c_void_p(out.data_ptr()), n, batch, omega, 1 if inverse else 0,
                                          E._stream_ptr())


def _reference(x, F, n, batch, inverse, variant=E.ARK_RADIX2):
    out = torch.empty_like(x)
    if inverse:
        E.intt_device(x, out, F, batch=batch, n=n)
    else:
        E.ntt_device(x, out, F, variant, batch=batch, n=n)
    return out


@pytest.mark.parametrize("fname", ["Goldilocks", "BabyBear", "TeddyBear"])
def test_noalloc_matches_allocating_entry_points(fname):
    F = {f.name: f for f in E.FIELDS}[fname]
    big = MAXB // F.bytes  # the largest single vector: 2^22 / 2^23 elements
    cases = [(big, 1, False), (1 << 14, 1, False),  # two calls in a row, the second smaller
             (big, 1, True), (1 << 10, 3, True),
             (big // 4, 4, False), (1 << 16, 2, True), (2, 5, False), (1, 3, True)]
    for n, batch, inv in cases:
        print(f"case: noalloc {fname} n={n} batch={batch} inverse={inv}", flush=True)
        x = E.to_device(inputs(F, n * batch, n + batch), F)
        want = _reference(x, F, n, batch, inv)
        got = torch.empty_like(x)
        assert _noalloc(x, got, F, n, batch, inv) == 0
        torch.cuda.synchronize()
        assert torch.equal(got, want), (n, batch, inv, "out of place")
        buf = x.clone()
        assert _noalloc(buf, buf, F, n, batch, inv) == 0
        torch.cuda.synchronize()
        assert torch.equal(buf, want), (n, batch, inv, "in place")
    # against the oracle too, once per field, after a larger call left other tables behind
    n = 1 << 12
    xh = inputs(F, n, 5)
    x = E.to_device(xh, F)
    got = torch.empty_like(x)
    assert _noalloc(x, got, F, n) == 0
    assert np.array_equal(E.from_device(got, F), O.ntt(F.id, xh))


def test_noalloc_variants_and_custom_omega():
    F = E.Goldilocks
    n = 1 << 20
    x = E.to_device(inputs(F, n, 31), F)
    want = _reference(x, F, n, 1, False)
    for v in (E.LAMBDA_BOWERS, E.LAMBDA_RADIX4, E.WINTERFELL_FOUR_STEP, E.PLONKY3_LAYER_SPLIT):
        got = torch.empty_like(x)
        assert _noalloc(x, got, F, n, variant=v) == 0
        assert torch.equal(got, want), v
    w = pow(E.root_of_unity(F, n), 3, F.p)  # another primitive root
    got = torch.empty_like(x)
    assert _noalloc(x, got, F, n, omega=w) == 0
    ref = torch.empty_like(x)
    E.ntt_device(x, ref, F, omega=w)
    assert torch.equal(got, ref)


def test_noalloc_refuses_before_gpu_work():
    F = E.Goldilocks
    x = E.empty_device(8, F)
    n_over = MAXB // F.bytes * 2
    w = E.root_of_unity(F, n_over)
    L = E.lib()
    p = ctypes.c_void_p(x.data_ptr())
    assert L.omr_ntt_device_noalloc(E.context(), F.id, 0, p, p, n_over, 1, w, 0, None) == 13
    w2 = E.root_of_unity(F, 1 << 20)
    assert L.omr_ntt_device_noalloc(E.context(), F.id, 0, p, p, 1 << 20, 5, w2, 0, None) == 13
    assert L.omr_ntt_device_noalloc(E.context(), F.id, E.DFT, p, p, 8, 1, E.root_of_unity(F, 8), 0, None) == 5
    assert L.omr_ntt_device_noalloc(E.context(), F.id, E.TRANSFORM_DECOMPOSITION, p, p, 8, 1,
                                    E.root_of_unity(F, 8), 0, None) == 5
    assert L.omr_ntt_device_noalloc(None, F.id, 0, p, p, 8, 1, E.root_of_unity(F, 8), 0, None) == 1
