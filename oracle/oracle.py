"""ctypes wrapper over oracle/liborc.so, the CPU restatement of the reference.

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke entry point and by
the CPU-baseline leg of bench.py, always as the checker.  The product library
never routes through this module.

All arrays are numpy uint64 holding canonical field elements.  Field ids:
0 BabyBear, 1 KoalaBear, 2 Goldilocks, 3 TeddyBear (src/fields.rs:3-17).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")
NATIVE_PATH = os.path.join(HERE, "liborc_native.so")

BABYBEAR, KOALABEAR, GOLDILOCKS, TEDDYBEAR = 0, 1, 2, 3
FIELDS = {"babybear": 0, "koalabear": 1, "goldilocks": 2, "teddybear": 3}
P = [2013265921, 2130706433, 18446744069414584321, 3221225473]
G = [31, 3, 7, 5]
TWO_ADICITY = [27, 24, 32, 30]

_lib = None
_U64P = ctypes.POINTER(ctypes.c_uint64)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def use_native_build():
    """Switches this module to the -march=native build of the oracle, compiled
    for the host running it (the CPU baseline's build, mirroring the
    reference's target-cpu=native, .cargo/config.toml:2).  Call before the
    first use of lib().  Returns False (portable build kept) if the compile
    fails."""
    global _lib, LIB_PATH
    if _lib is not None and LIB_PATH == NATIVE_PATH:
        return True
    r = subprocess.run(["make", "-s", "-B", "-C", HERE, "native"], capture_output=True, text=True)
    if r.returncode != 0:
        return False
    LIB_PATH, _lib = NATIVE_PATH, None
    return True


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        u64, i32, vp = ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p
        sig = {
            "orc_add": (u64, [i32, u64, u64]),
            "orc_sub": (u64, [i32, u64, u64]),
            "orc_mul": (u64, [i32, u64, u64]),
            "orc_mul_textbook": (u64, [i32, u64, u64]),
            "orc_pow": (u64, [i32, u64, u64]),
            "orc_inv": (u64, [i32, u64]),
            "orc_root_of_unity": (u64, [i32, u64]),
            "orc_bitrev": (u64, [u64, i32]),
            "orc_domain": (None, [i32, u64, u64, vp, vp]),
            "orc_derange": (None, [vp, u64, i32]),
            "orc_transpose": (None, [vp, vp, u64, u64]),
            "orc_radix2": (None, [i32, vp, u64, vp]),
            "orc_dft": (None, [i32, vp, u64, u64]),
            "orc_dft_partial": (None, [i32, vp, u64, u64, u64]),
            "orc_lambda_radix4": (i32, [i32, vp, u64, vp]),
            "orc_lambda_bowers": (None, [i32, vp, u64, vp]),
            "orc_plonky3_layer_split": (None, [i32, vp, u64, vp, vp]),
            "orc_winterfell_four_step": (None, [i32, vp, u64, vp]),
            "orc_tfhe_stockham_radix8": (i32, [i32, vp, u64, vp]),
            "orc_transform_decomposition": (None, [i32, vp, u64, u64, vp, u64, i32]),
            "orc_intt": (None, [i32, vp, u64, u64]),
            "orc_ntt_batch": (None, [i32, vp, vp, vp, vp, u64]),
            "orc_ntt": (None, [i32, vp, u64, u64]),
            "orc_coset_lde": (i32, [i32, vp, u64, u64, u64, vp]),
            "orc_coset_decode": (ctypes.c_int64, [i32, vp, u64, u64, u64, vp]),
            "orc_stride_subset": (None, [i32, vp, u64, u64, u64, vp]),
            "orc_vec_op": (None, [i32, i32, vp, vp, vp, u64]),
            "orc_num_threads": (i32, []),
            "orc_set_num_threads": (None, [i32]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


def _u64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint64))


# ---- scalar field ops -------------------------------------------------------
def add(f, a, b):
    return lib().orc_add(f, a, b)


def sub(f, a, b):
    return lib().orc_sub(f, a, b)


def mul(f, a, b):
    return lib().orc_mul(f, a, b)


def mul_textbook(f, a, b):
    return lib().orc_mul_textbook(f, a, b)


def pow_(f, a, e):
    return lib().orc_pow(f, a, e)


def inv(f, a):
    return lib().orc_inv(f, a)


def root_of_unity(f, n):
    return lib().orc_root_of_unity(f, n)


def bitrev(a, log_len):
    return lib().orc_bitrev(a, log_len)


def vec_op(f, op, a, b):
    a, b = _u64(a), _u64(b)
    out = np.empty_like(a)
    lib().orc_vec_op(f, op, _p(a), _p(b), _p(out), a.size)
    return out


# ---- NttDomain (src/encoder.rs:23-52) ---------------------------------------
class Domain:
    def __init__(self, f, n, omega=None):
        self.f, self.N = f, n
        self.log_N = n.bit_length() - 1
        self.omega = root_of_unity(f, n) if omega is None else omega
        self.twiddles = np.empty(n, dtype=np.uint64)
        self.bitrev_twiddles = np.empty(max(n // 2, 1), dtype=np.uint64)
        lib().orc_domain(f, n, self.omega, _p(self.twiddles), _p(self.bitrev_twiddles))


# ---- encoders: each returns a new array -----------------------------------
def ntt(f, x, omega=None):
    """Textbook radix-2 forward NTT (the ArkRadix2 role)."""
    x = _u64(x).copy()
    w = root_of_unity(f, x.size) if omega is None else omega
    lib().orc_ntt(f, _p(x), x.size, w)
    return x


def intt(f, x, omega=None):
    x = _u64(x).copy()
    w = root_of_unity(f, x.size) if omega is None else omega
    lib().orc_intt(f, _p(x), x.size, w)
    return x


def dft(f, x, omega=None):
    x = _u64(x).copy()
    w = root_of_unity(f, x.size) if omega is None else omega
    lib().orc_dft(f, _p(x), x.size, w)
    return x


def encoder(name, f, x, dom=None, s=None):
    """Run the restatement of reference encoder `name` on a copy of x."""
    x = _u64(x).copy()
    n = x.size
    d = dom if dom is not None else Domain(f, n)
    L = lib()
    if name == "ArkRadix2":
        L.orc_radix2(f, _p(x), n, _p(d.twiddles))
    elif name == "Dft":
        L.orc_dft(f, _p(x), n, d.omega)
    elif name == "DftPartial":
        L.orc_dft_partial(f, _p(x), n, d.omega, s)
    elif name == "LambdaRadix4":
        if L.orc_lambda_radix4(f, _p(x), n, _p(d.bitrev_twiddles)) != 0:
            raise ValueError("LambdaRadix4 requires log2 N even")
    elif name == "LambdaBowers":
        L.orc_lambda_bowers(f, _p(x), n, _p(d.twiddles))
    elif name == "Plonky3Radix2LayerSplit":
        L.orc_plonky3_layer_split(f, _p(x), n, _p(d.twiddles), _p(d.bitrev_twiddles))
    elif name == "WinterfellFourStep":
        L.orc_winterfell_four_step(f, _p(x), n, _p(d.bitrev_twiddles))
    elif name == "TfheStockhamRadix8":
        if L.orc_tfhe_stockham_radix8(f, _p(x), n, _p(d.twiddles)) != 0:
            raise ValueError("TfheStockhamRadix8 requires log2 N divisible by 3")
    elif name == "TransformDecomposition":
        L.orc_transform_decomposition(f, _p(x), n, d.omega, _p(d.twiddles), s, 0)
    elif name == "TransformDecompositionV2":
        L.orc_transform_decomposition(f, _p(x), n, d.omega, _p(d.twiddles), s, 1)
    else:
        raise KeyError(name)
    return x


def ntt_batch(f, buf, offsets, log_ns, inverse):
    """Vectors of uneven sizes in one buffer: vector i = buf[off[i] : off[i] +
    2^log_ns[i]], forward (inverse[i] == 0) or inverse; returns a new array."""
    b = _u64(buf).copy()
    off = _u64(offsets)
    ln = np.ascontiguousarray(np.asarray(log_ns, dtype=np.uint32))
    iv = np.ascontiguousarray(np.asarray(inverse, dtype=np.uint32))
    assert off.size == ln.size == iv.size
    assert all(int(o) + (1 << int(l)) <= b.size for o, l in zip(off, ln))
    lib().orc_ntt_batch(f, _p(b), _p(off), ln.ctypes.data_as(ctypes.c_void_p), iv.ctypes.data_as(ctypes.c_void_p),
                        off.size)
    return b


def coset_lde(f, coeffs, blowup, shift=None):
    c = _u64(coeffs)
    out = np.empty(c.size * blowup, dtype=np.uint64)
    sh = G[f] if shift is None else shift
    if lib().orc_coset_lde(f, _p(c), c.size, blowup, sh, _p(out)) != 0:
        raise ValueError("size exceeds two-adicity")
    return out


def coset_decode(f, evals, n, blowup, shift=None):
    e = _u64(evals)
    out = np.empty(n, dtype=np.uint64)
    sh = G[f] if shift is None else shift
    bad = lib().orc_coset_decode(f, _p(e), n, blowup, sh, _p(out))
    return out, bad


def stride_subset(f, x, m, omega=None):
    x = _u64(x)
    w = root_of_unity(f, x.size) if omega is None else omega
    out = np.empty(m, dtype=np.uint64)
    lib().orc_stride_subset(f, _p(x), x.size, w, m, _p(out))
    return out


def transpose(x, w, h):
    x = _u64(x)
    out = np.empty_like(x)
    lib().orc_transpose(_p(x), _p(out), w, h)
    return out


def num_threads():
    return lib().orc_num_threads()


def set_num_threads(t):
    lib().orc_set_num_threads(t)


def uniform(f, n, seed):
    """Seeded uniform field elements in [0, p) (numpy PCG64)."""
    rng = np.random.default_rng(seed)
    return rng.integers(0, P[f], size=n, dtype=np.uint64)
