"""CPU-side checks of the C ABI: the library loads, exports every function
include/omrntt.h declares, and rejects bad arguments before any GPU work
(the reference's panics, src/encoder.rs:24-32, lambdaworks_radix4.rs:17-21,
tfhe_stockham_radix8.rs:17-21, become status codes)."""
import ctypes
import os
import re

import numpy as np
import pytest

import omr_ntt as E

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    inc = os.path.join(ROOT, "include")
    for fn in sorted(os.listdir(inc)):
        if fn.endswith(".h"):
            txt = open(os.path.join(inc, fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names += re.findall(r"\b(omr_[a-z0-9_]+)\s*\(", txt)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = E.lib()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n


def test_field_info_and_roots():
    for F in E.FIELDS:
        p, g, s = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_int()
        assert E.lib().omr_field_info(F.id, ctypes.byref(p), ctypes.byref(g), ctypes.byref(s)) == 0
        assert (p.value, g.value, s.value) == (F.p, F.g, F.TWO_ADICITY)
        for k in range(0, F.TWO_ADICITY + 1, 3):
            assert E.root_of_unity(F, 1 << k) == pow(F.g, (F.p - 1) >> k, F.p)
    assert E.lib().omr_field_info(7, None, None, None) == 4


def test_domain_errors_mirror_reference_panics():
    with pytest.raises(E.OmrError) as e:
        E.NttDomain(12, E.BabyBear)
    assert e.value.status == 2
    with pytest.raises(E.OmrError) as e:
        E.NttDomain(1 << 28, E.BabyBear)
    assert e.value.status == 3
    with pytest.raises(E.OmrError):
        E.root_of_unity(E.KoalaBear, 1 << 25)


def _call_ntt(field, variant, n, omega, s=0, buf=True):
    b = np.zeros(max(n, 1), dtype=np.uint64)
    ptr = b.ctypes.data_as(ctypes.c_void_p) if buf else None
    return E.lib().omr_ntt(None, field, variant, ptr, n, omega, s)


def test_validation_before_device_work():
    L = E.lib()
    gl = E.Goldilocks
    w = E.root_of_unity(gl, 1024)
    assert _call_ntt(2, 0, 1000, w) == 2          # not a power of two
    assert _call_ntt(2, 0, 0, w) == 2             # empty: 0 is not a power of two (src/encoder.rs:24)
    assert _call_ntt(3, 0, 1 << 31, 1) == 3       # beyond TeddyBear's two-adicity (30)
    assert _call_ntt(0, 0, 1 << 28, 1) == 3       # beyond BabyBear two-adicity
    assert _call_ntt(9, 0, 1024, w) == 4          # unknown field
    assert _call_ntt(2, 42, 1024, w) == 5         # unknown variant
    assert _call_ntt(2, 3, 512, E.root_of_unity(gl, 512)) == 6   # radix-4 needs even log N
    assert _call_ntt(2, 6, 1024, w) == 6          # radix-8 needs log N % 3 == 0
    assert _call_ntt(2, 0, 1024, pow(w, 2, gl.p)) == 7   # omega of order N/2
    assert _call_ntt(2, 0, 1024, gl.p + 5) == 7          # omega out of range
    assert _call_ntt(2, 0, 1024, w) == 1                 # valid args, null context
    lr = (ctypes.c_int * 2)(10, 10)
    assert L.omr_ntt_schedule_device(None, 2, None, None, 1024, 1, w, 0, lr, 2, 0, None) == 11
    lr = (ctypes.c_int * 2)(13, 0)
    assert L.omr_ntt_schedule_device(None, 2, None, None, 1 << 13, 1, E.root_of_unity(gl, 1 << 13), 0, lr, 1, 0, None) == 11
    assert L.omr_partial_device(None, 2, None, None, 1024, w, 2048, 0, 0, 0, None) == 8
    assert L.omr_partial_device(None, 2, None, None, 1024, w, 24, 1, 0, 0, None) == 8
    # batched prefix: out_len and batch checked before pointers
    assert L.omr_partial_batch_device(None, 2, None, None, 1024, 4, w, 0, None) == 8
    assert L.omr_partial_batch_device(None, 2, None, None, 1024, 4, w, 2048, None) == 8
    assert L.omr_partial_batch_device(None, 2, None, None, 1024, 0, w, 16, None) == 12
    assert L.omr_partial_batch_device(None, 2, None, None, 1000, 4, w, 16, None) == 2
    assert L.omr_partial_batch_device(None, 2, None, None, 1024, 4, w, 16, None) == 1
    # host forms: validated like their device forms, before any pointer use
    assert L.omr_intt(None, 2, None, 1000, w) == 2
    assert L.omr_intt(None, 2, None, 1024, w) == 1
    assert L.omr_ntt_batched(None, 2, None, 1024, 0, w) == 12
    assert L.omr_ntt_batched(None, 2, None, 1024, 3, w) == 1
    # dense batches: at most 65535 vectors per call, checked before pointers
    # (n * batch * 8 bytes would overflow 64 bits for batch >= 2^51 at n = 1024)
    big = 1 << 60
    assert L.omr_ntt_batched(None, 2, None, 1024, 65536, w) == 12
    assert L.omr_ntt_batched(None, 2, None, 1024, big, w) == 12
    assert L.omr_ntt_device(None, 2, 1, None, None, 1024, big, w, 0, None) == 12   # Dft variant
    assert L.omr_ntt_device(None, 2, 8, None, None, 1024, 65536, w, 8, None) == 12  # partial variant
    assert L.omr_ntt_device(None, 2, 0, None, None, 1024, 65535, w, 0, None) == 1
    assert L.omr_intt_device(None, 2, None, None, 1024, 65536, w, None) == 12
    lr = (ctypes.c_int * 2)(5, 5)
    assert L.omr_ntt_schedule_device(None, 2, None, None, 1024, 65536, w, 0, lr, 2, 0, None) == 12
    assert L.omr_coset_lde(None, 0, None, None, 1 << 26, 1, 4, 0) == 3   # 2^28 > BabyBear 2-adicity
    assert L.omr_coset_lde(None, 9, None, None, 1024, 1, 4, 0) == 4
    assert L.omr_coset_lde(None, 2, None, None, 1024, 1, 3, 0) == 2
    assert L.omr_coset_lde(None, 2, None, None, 1024, 1, 4, 0) == 1
    assert L.omr_coset_decode(None, 2, None, None, 1024, 1, 4, 0, None) == 1
    # pinned host blocks: argument checks, and only blocks of omr_host_alloc are freed
    p = ctypes.c_void_p()
    assert L.omr_host_alloc(0, ctypes.byref(p)) == 12
    assert L.omr_host_alloc(64, None) == 1
    assert L.omr_host_free(None) == 1
    stray = np.zeros(8, dtype=np.uint64)
    assert L.omr_host_free(stray.ctypes.data) == 12
    assert L.omr_variant_name(99) is None
    assert [L.omr_variant_name(i).decode() for i in range(10)] == [
        "ArkRadix2", "Dft", "LambdaBowers", "LambdaRadix4", "Plonky3Radix2LayerSplit",
        "WinterfellFourStep", "TfheStockhamRadix8", "DftPartial", "TransformDecomposition",
        "TransformDecompositionV2"]


def test_default_schedules():
    assert E.default_schedule(E.Goldilocks, 1 << 24) == [8, 8, 8]
    assert E.default_schedule(E.BabyBear, 1 << 16) == [8, 8]
    assert sum(E.default_schedule(E.Goldilocks, 1 << 20)) == 20
    assert E.default_schedule(E.Goldilocks, 1 << 12) == [12]


def test_missing_library_fails_loudly():
    """No CPU fallback: without the HIP library every entry point raises."""
    import subprocess
    import sys
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "import numpy as np, omr_ntt as E\n"
        "E.LIB_PATH = '/nonexistent/libomrntt.so'\n"
        "for f in (lambda: E.lib(), lambda: E.ArkRadix2().ntt(np.zeros(16, np.uint32), E.NttDomain(16, E.BabyBear))):\n"
        "    try:\n"
        "        f()\n"
        "    except ImportError as e:\n"
        "        assert 'missing' in str(e), e\n"
        "    else:\n"
        "        raise SystemExit('no error')\n"
        "print('raised')\n" % os.path.join(ROOT, "omr-ntt_amd"))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "raised" in r.stdout, r.stderr


def test_batch_vector_schedules_fit_the_tile_kernel():
    """Every (radix cap, log n) the batch planner accepts splits log n into
    passes of 2^4 .. 2^12 (what the batch tile kernel runs, register radix
    16) that sum to log n; caps below 2^5 or above 2^12 are refused."""
    for cap in range(5, 13):
        for ln in range(4, 37):
            sched = E.batch_vector_schedule(ln, cap)
            assert sum(sched) == ln, (cap, ln, sched)
            assert all(4 <= r <= 12 for r in sched), (cap, ln, sched)
            if ln > 12:
                assert max(sched) <= cap and max(sched) - min(sched) <= 1, (cap, ln, sched)
    for cap in (1, 2, 3, 4, 13):
        assert E.lib().omr_batch_set_max_log_radix(cap) == 11
        with pytest.raises(E.OmrError):
            E.batch_vector_schedule(16, cap)
    with pytest.raises(E.OmrError):
        E.batch_vector_schedule(3, 9)  # vectors start at 2^4
    # the automatic cap (0, the default): two passes up to 2^20 for Goldilocks
    assert E.lib().omr_batch_set_max_log_radix(0) == 0
    assert E.batch_vector_schedule(20) == [10, 10]
    assert E.batch_vector_schedule(19, 9) == [7, 6, 6]
