"""omr_ntt -- Python host mirror of the omr-ntt operator API over the gfx950
engine's C ABI (include/omrntt.h, libomrntt.so).

Mirrors the reference crate's surface:
  * fields (src/fields.rs:3-19): BabyBear, KoalaBear, Goldilocks, TeddyBear,
    DefaultField = BabyBear;
  * NttDomain(N) (src/encoder.rs:9-52): N, log_N, omega and the same panics
    (here: OmrError) for non-power-of-two N and N beyond the two-adicity;
  * NttEncoder.ntt(buf, domain) / .name() (src/encoder.rs:70-74) with the
    twelve variant names of src/encoders/mod.rs:15-23;
  * encoders.all(log_n) / encoders.all_partial(s) (src/encoders/mod.rs:32-54).

buf is a numpy array of canonical values (host path, synchronous, in place)
or a torch tensor on the GPU (device path, in place, on torch's current
stream).  There is no CPU fallback: without libomrntt.so every call raises.
"""
import ctypes
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "libomrntt.so")

# ---- fields ----------------------------------------------------------------


class Field:
    def __init__(self, fid, name, p, g, two_adicity):
        self.id, self.name, self.p, self.g, self.TWO_ADICITY = fid, name, p, g, two_adicity
        self.bytes = 8 if p > 2**32 else 4
        self.np_dtype = np.uint64 if self.bytes == 8 else np.uint32

    def __repr__(self):
        return self.name


BabyBear = Field(0, "BabyBear", 2013265921, 31, 27)
KoalaBear = Field(1, "KoalaBear", 2130706433, 3, 24)
Goldilocks = Field(2, "Goldilocks", 18446744069414584321, 7, 32)
TeddyBear = Field(3, "TeddyBear", 3221225473, 5, 30)
DefaultField = BabyBear
FIELDS = [BabyBear, KoalaBear, Goldilocks, TeddyBear]

# ---- variants (src/encoders/mod.rs:15-23) -----------------------------------
ARK_RADIX2, DFT, LAMBDA_BOWERS, LAMBDA_RADIX4, PLONKY3_LAYER_SPLIT, WINTERFELL_FOUR_STEP, \
    TFHE_STOCKHAM_RADIX8, DFT_PARTIAL, TRANSFORM_DECOMPOSITION, TRANSFORM_DECOMPOSITION_V2 = range(10)

STATUS = {0: "ok", 1: "E_NULL", 2: "E_SIZE", 3: "E_TWO_ADICITY", 4: "E_FIELD", 5: "E_VARIANT",
          6: "E_RADIX", 7: "E_OMEGA", 8: "E_OUTLEN", 9: "E_ALLOC", 10: "E_DEVICE",
          11: "E_SCHEDULE", 12: "E_ARG", 13: "E_SCRATCH"}


class OmrError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        msg = lib().omr_status_str(status).decode()
        super().__init__(f"{what}: {STATUS.get(status, status)} ({msg})" if what else msg)


_lib = None
_ctx = None
_VP = ctypes.c_void_p


def lib():
    """Loads libomrntt.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with python omr-ntt_amd/build.py")
        # The PyTorch wheel carries its own HIP runtime (device tensors handed
        # to the engine come from it).  It must initialise the GPU before this
        # library's runtime does: in the other order torch finds no device.
        try:
            import torch
            if torch.cuda.device_count() > 0:
                torch.cuda.init()
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        i32, u64, u32 = ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32
        P64, Pi = ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_int)
        sig = {
            "omr_status_str": (ctypes.c_char_p, [i32]),
            "omr_version": (ctypes.c_char_p, []),
            "omr_ctx_create": (i32, [i32, ctypes.POINTER(_VP)]),
            "omr_ctx_destroy": (i32, [_VP]),
            "omr_ctx_synchronize": (i32, [_VP, _VP]),
            "omr_field_info": (i32, [i32, P64, P64, Pi]),
            "omr_field_bytes": (i32, [i32]),
            "omr_root_of_unity": (i32, [i32, u64, P64]),
            "omr_domain_tables": (i32, [_VP, i32, u64, u64, _VP, _VP]),
            "omr_domain_tables_device": (i32, [_VP, i32, u64, u64, _VP, _VP, _VP]),
            "omr_variant_name": (ctypes.c_char_p, [i32]),
            "omr_variant_supports": (i32, [i32, u32]),
            "omr_ntt": (i32, [_VP, i32, i32, _VP, u64, u64, u64]),
            "omr_ntt_words": (i32, [_VP, i32, i32, _VP, u64, u64, u64, u32]),
            "omr_ntt_device_words": (i32, [_VP, i32, i32, _VP, _VP, u64, u64, u64, u64, u32, _VP]),
            "omr_ntt_device": (i32, [_VP, i32, i32, _VP, _VP, u64, u64, u64, u64, _VP]),
            "omr_intt_device": (i32, [_VP, i32, _VP, _VP, u64, u64, u64, _VP]),
            "omr_ntt_device_noalloc": (i32, [_VP, i32, i32, _VP, _VP, u64, u64, u64, i32, _VP]),
            "omr_ntt_schedule_device": (i32, [_VP, i32, _VP, _VP, u64, u64, u64, i32, Pi, i32, i32, _VP]),
            "omr_default_schedule": (i32, [i32, u64, Pi, Pi]),
            "omr_partial_device": (i32, [_VP, i32, _VP, _VP, u64, u64, u64, i32, i32, u32, _VP]),
            "omr_field_op_device": (i32, [i32, i32, _VP, _VP, _VP, u64, _VP]),
            "omr_partial_batch_device": (i32, [_VP, i32, _VP, _VP, u64, u64, u64, u64, _VP]),
            "omr_intt": (i32, [_VP, i32, _VP, u64, u64]),
            "omr_ntt_batched": (i32, [_VP, i32, _VP, u64, u64, u64]),
            "omr_coset_lde": (i32, [_VP, i32, _VP, _VP, u64, u64, u32, u64]),
            "omr_coset_decode": (i32, [_VP, i32, _VP, _VP, u64, u64, u32, u64, ctypes.POINTER(u32)]),
            "omr_batch_plan_create": (i32, [_VP, i32, _VP, u32, ctypes.POINTER(_VP)]),
            "omr_batch_plan_destroy": (i32, [_VP]),
            "omr_batch_plan_info": (i32, [_VP, P64, ctypes.POINTER(u32), ctypes.POINTER(u32)]),
            "omr_batch_execute": (i32, [_VP, _VP, _VP, _VP]),
            "omr_batch_execute_graph": (i32, [_VP, _VP, _VP, _VP]),
            "omr_batch_execute_plain": (i32, [_VP, _VP, _VP, _VP]),
            "omr_batch_execute_grouped": (i32, [_VP, _VP, _VP, _VP]),
            "omr_batch_execute_grouped_graph": (i32, [_VP, _VP, _VP, _VP]),
            "omr_batch_set_max_log_radix": (i32, [i32]),
            "omr_batch_vector_schedule": (i32, [u32, i32, Pi, Pi]),
            "omr_num_schedule_classes": (i32, []),
            "omr_schedule_class": (i32, [i32, u64, u64, i32, Pi, Pi, Pi, Pi, ctypes.POINTER(u32)]),
            "omr_default_class": (i32, [i32, u64, u64]),
            "omr_tune_classes": (i32, [_VP, i32, u64, u64, u64, i32, ctypes.POINTER(ctypes.c_float)]),
            "omr_tune_set_class": (i32, [i32, u64, u64, i32]),
            "omr_tune_set_class_batch": (i32, [i32, u64, u64, u64, i32]),
            "omr_tune_clear": (i32, []),
            "omr_tune_save": (i32, [ctypes.c_char_p]),
            "omr_tune_load": (i32, [ctypes.c_char_p]),
            "omr_cost_model_create": (i32, [_VP, u32, ctypes.POINTER(_VP)]),
            "omr_cost_model_destroy": (i32, [_VP]),
            "omr_cost_model_num_params": (i32, []),
            "omr_cost_model_num_inputs": (i32, []),
            "omr_cost_model_num_outputs": (i32, []),
            "omr_cost_model_get_params": (i32, [_VP, _VP]),
            "omr_cost_model_set_params": (i32, [_VP, _VP]),
            "omr_cost_model_features": (i32, [i32, u64, u64, u64, _VP]),
            "omr_cost_model_forward": (i32, [_VP, _VP, u32, _VP]),
            "omr_cost_model_gradients": (i32, [_VP, _VP, _VP, _VP, u32, _VP, _VP]),
            "omr_cost_model_train": (i32, [_VP, _VP, _VP, _VP, u32, u32, ctypes.c_float, _VP]),
            "omr_cost_model_choose": (i32, [_VP, i32, u64, u64, u64, Pi]),
            "omr_cost_model_save": (i32, [_VP, ctypes.c_char_p]),
            "omr_cost_model_load": (i32, [_VP, ctypes.c_char_p]),
            "omr_autotune_shape": (i32, [_VP, _VP, i32, i32, u64, u64, u64, i32, Pi, Pi]),
            "omr_coset_lde_device": (i32, [_VP, i32, _VP, _VP, u64, u64, u32, u64, _VP]),
            "omr_coset_decode_device": (i32, [_VP, i32, _VP, _VP, u64, u64, u32, u64, _VP, _VP]),
            "omr_host_alloc": (i32, [ctypes.c_size_t, ctypes.POINTER(_VP)]),
            "omr_host_free": (i32, [_VP]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        _lib = L
    return _lib


def _check(status, what=""):
    if status != 0:
        raise OmrError(status, what)


def context(device=0):
    """Process-wide engine context on HIP device `device`."""
    global _ctx
    if _ctx is None:
        c = _VP()
        _check(lib().omr_ctx_create(device, ctypes.byref(c)), "omr_ctx_create")
        _ctx = c
    return _ctx


_roots = {}


def root_of_unity(field, n):
    """F::get_root_of_unity(N) (the ark-ff convention g^((p-1)/N)), cached."""
    key = (field.id, n)
    w = _roots.get(key)
    if w is None:
        c = ctypes.c_uint64()
        _check(lib().omr_root_of_unity(field.id, n, ctypes.byref(c)), f"root_of_unity(N={n})")
        w = _roots[key] = c.value
    return w


def default_schedule(field, n):
    lr = (ctypes.c_int * 8)()
    npass = ctypes.c_int()
    _check(lib().omr_default_schedule(field.id, n, lr, ctypes.byref(npass)))
    return [lr[i] for i in range(npass.value)]


# ---- NttDomain (src/encoder.rs:9-52) ---------------------------------------


class NttDomain:
    """Reusable NTT parameters: N, log_N and the primitive N-th root omega
    (F::get_root_of_unity(N)), plus the reference's public tables twiddles
    (omega^k, k < N) and bitrev_twiddles (omega^bitrev(i, log_N - 1),
    i < N/2).  The engine derives its own device tables from omega, so the
    two tables are built (on the GPU) only when read."""

    def __init__(self, N, field=DefaultField, omega=None):
        if N <= 0 or N & (N - 1):
            raise OmrError(2, f"NTT size N={N} must be a power of 2")
        log_N = N.bit_length() - 1
        if log_N > field.TWO_ADICITY:
            raise OmrError(3, f"N={N} requires log2(N)={log_N} <= TWO_ADICITY={field.TWO_ADICITY}")
        self.field, self.N, self.log_N = field, N, log_N
        self.omega = root_of_unity(field, N) if omega is None else int(omega)
        self._tables = None

    def _host_tables(self):
        if self._tables is None:
            tw = np.empty(self.N, dtype=self.field.np_dtype)
            br = np.empty(self.N // 2, dtype=self.field.np_dtype)
            _check(lib().omr_domain_tables(context(), self.field.id, self.N, self.omega, tw.ctypes.data,
                                           br.ctypes.data if br.size else None), "omr_domain_tables")
            self._tables = (tw.astype(np.uint64), br.astype(np.uint64))
        return self._tables

    @property
    def twiddles(self):
        """twiddles[k] = omega^k for k < N (src/encoder.rs:16)."""
        return self._host_tables()[0]

    @property
    def bitrev_twiddles(self):
        """bitrev_twiddles[i] = omega^bitrev(i, log_N - 1) for i < N/2 (src/encoder.rs:18)."""
        return self._host_tables()[1]


def domain_tables_device(field, n, omega=None):
    """NttDomain's twiddles and bitrev_twiddles as device tensors (n and n/2
    words), built on torch's current stream."""
    import torch
    omega = root_of_unity(field, n) if omega is None else omega
    dt = torch.int64 if field.bytes == 8 else torch.int32
    tw = torch.empty(n, dtype=dt, device="cuda")
    br = torch.empty(n // 2, dtype=dt, device="cuda")
    _check(lib().omr_domain_tables_device(context(), field.id, n, omega, _dptr(tw),
                                          _dptr(br) if n >= 2 else None, _stream_ptr()),
           "omr_domain_tables_device")
    return tw, br


# ---- device-pointer helpers --------------------------------------------------


def _is_torch(x):
    return type(x).__module__.startswith("torch")


def _stream_ptr():
    """torch's current stream on the current device, as a raw HIP stream
    handle.  torch._C._cuda_getCurrentRawStream returns the handle without
    building a Stream object (0.2 against 2.5 us a call, which small
    transforms pay on every call); current_stream() where it is missing."""
    import torch
    raw = getattr(torch._C, "_cuda_getCurrentRawStream", None)
    if raw is not None:
        return _VP(raw(torch.cuda.current_device()))
    return _VP(torch.cuda.current_stream().cuda_stream)


def _dptr(t):
    return _VP(t.data_ptr())


def _check_dev_tensor(t, field, n_elems=None):
    if not t.is_cuda or not t.is_contiguous():
        raise ValueError("device buffers must be contiguous GPU tensors")
    if t.element_size() != field.bytes:
        raise ValueError(f"{field} elements are {field.bytes} bytes, tensor has {t.element_size()}")
    if n_elems is not None and t.numel() < n_elems:
        raise ValueError("tensor too small")


# ---- device entry points -----------------------------------------------------


def ntt_device(x, out, field, variant=ARK_RADIX2, omega=None, s=0, batch=1, n=None):
    n = n or x.numel() // batch
    omega = root_of_unity(field, n) if omega is None else omega
    _check_dev_tensor(x, field, n * batch)
    _check_dev_tensor(out, field)
    _check(lib().omr_ntt_device(context(), field.id, variant, _dptr(x), _dptr(out), n, batch, omega, s,
                                _stream_ptr()), "omr_ntt_device")
    return out


def ntt_device_words(x, out, field, variant=ARK_RADIX2, omega=None, s=0, batch=1, n=None):
    """ntt_device on int64 tensors holding one value of a 31/32-bit field per
    64-bit word (the crate's one-limb storage)."""
    n = n or x.numel() // batch
    omega = root_of_unity(field, n) if omega is None else omega
    for t in (x, out):
        if not t.is_cuda or not t.is_contiguous() or t.element_size() != 8:
            raise ValueError("word buffers must be contiguous 64-bit GPU tensors")
    if x.numel() < n * batch:
        raise ValueError("tensor too small")
    _check(lib().omr_ntt_device_words(context(), field.id, variant, _dptr(x), _dptr(out), n, batch, omega, s, 8,
                                      _stream_ptr()), "omr_ntt_device_words")
    return out


def ntt_words(buf, field, variant=ARK_RADIX2, omega=None, s=0):
    """Host form: NttEncoder::ntt on a numpy uint64 array of 64-bit words, in place."""
    if not isinstance(buf, np.ndarray) or buf.dtype != np.uint64 or not buf.flags["C_CONTIGUOUS"]:
        raise TypeError("buf must be a contiguous numpy uint64 array")
    n = buf.size
    omega = root_of_unity(field, n) if omega is None else omega
    _check(lib().omr_ntt_words(context(), field.id, variant, buf.ctypes.data_as(_VP), n, omega, s, 8),
           "omr_ntt_words")
    return buf


def intt_device(x, out, field, omega=None, batch=1, n=None):
    n = n or x.numel() // batch
    omega = root_of_unity(field, n) if omega is None else omega
    _check_dev_tensor(x, field, n * batch)
    _check_dev_tensor(out, field, n * batch)
    _check(lib().omr_intt_device(context(), field.id, _dptr(x), _dptr(out), n, batch, omega,
                                 _stream_ptr()), "omr_intt_device")
    return out


def ntt_schedule_device(x, out, field, log_radix, log_e=0, inverse=False, omega=None, batch=1, n=None):
    n = n or x.numel() // batch
    omega = root_of_unity(field, n) if omega is None else omega
    _check_dev_tensor(x, field, n * batch)
    _check_dev_tensor(out, field, n * batch)
    lr = (ctypes.c_int * len(log_radix))(*log_radix)
    _check(lib().omr_ntt_schedule_device(context(), field.id, _dptr(x), _dptr(out), n, batch, omega,
                                         1 if inverse else 0, lr, len(log_radix), log_e,
                                         _stream_ptr()), "omr_ntt_schedule_device")
    return out


PREFIX, STRIDE = 0, 1


def partial_device(x, out, field, out_len, mode=PREFIX, method=0, log_n1=0, omega=None):
    n = x.numel()
    omega = root_of_unity(field, n) if omega is None else omega
    _check_dev_tensor(x, field, n)
    _check_dev_tensor(out, field, out_len)
    _check(lib().omr_partial_device(context(), field.id, _dptr(x), _dptr(out), n, omega, out_len,
                                    mode, method, log_n1, _stream_ptr()), "omr_partial_device")
    return out


def host_empty(n, field):
    """A numpy array of n field words in pinned host memory (omr_host_alloc):
    the host-buffer entry points copy it directly instead of through their
    staging ring.  The block is freed when the last view of it goes away."""
    nbytes = int(n) * field.bytes
    p = _VP()
    _check(lib().omr_host_alloc(max(nbytes, 1), ctypes.byref(p)), "omr_host_alloc")
    block = (ctypes.c_char * max(nbytes, 1)).from_address(p.value)
    weakref.finalize(block, lib().omr_host_free, p.value)
    return np.frombuffer(block, dtype=field.np_dtype, count=int(n))


def _host_view(buf, field):
    """(buf, a contiguous array of the field's word type holding the same values)."""
    if isinstance(buf, np.ndarray) and buf.dtype == field.np_dtype and buf.flags["C_CONTIGUOUS"]:
        return buf, buf
    return buf, np.ascontiguousarray(np.asarray(buf).astype(field.np_dtype))


def _host_writeback(buf, view, field):
    if view is not buf:
        buf[...] = view


def intt(buf, field, omega=None):
    """Host form: inverse NTT of a numpy array in place (omega = the forward root)."""
    a, view = _host_view(buf, field)
    n = a.size
    omega = root_of_unity(field, n) if omega is None else omega
    _check(lib().omr_intt(context(), field.id, view.ctypes.data, n, omega), "omr_intt")
    _host_writeback(buf, view, field)
    return buf


def ntt_batched(buf, field, batch, omega=None):
    """Host form: `batch` consecutive vectors forward-transformed in place."""
    a, view = _host_view(buf, field)
    n = a.size // batch
    omega = root_of_unity(field, n) if omega is None else omega
    _check(lib().omr_ntt_batched(context(), field.id, view.ctypes.data, n, batch, omega), "omr_ntt_batched")
    _host_writeback(buf, view, field)
    return buf


def coset_lde(coeffs, field, n, ncols=1, blowup=4, shift=0):
    """Host form of coset_lde_device: returns the evaluations (numpy)."""
    _, cv = _host_view(coeffs, field)
    ev = np.empty(n * ncols * blowup, dtype=field.np_dtype)
    _check(lib().omr_coset_lde(context(), field.id, cv.ctypes.data, ev.ctypes.data, n, ncols, blowup, shift),
           "omr_coset_lde")
    return ev.astype(np.uint64)


def coset_decode(evals, field, n, ncols=1, blowup=4, shift=0):
    """Host form of coset_decode_device: returns (coefficients, nonzero tail count)."""
    _, ev = _host_view(evals, field)
    co = np.empty(n * ncols, dtype=field.np_dtype)
    nz = ctypes.c_uint32(0)
    _check(lib().omr_coset_decode(context(), field.id, ev.ctypes.data, co.ctypes.data, n, ncols, blowup, shift,
                                  ctypes.byref(nz)), "omr_coset_decode")
    return co.astype(np.uint64), nz.value


def partial_batch_device(x, out, field, n, batch, out_len, omega=None):
    """Prefix outputs of `batch` vectors of n (the OMR epoch: each receiver's
    k = 2s outputs): out[b*out_len + k] = NTT(x[b*n:(b+1)*n])[k]."""
    omega = root_of_unity(field, n) if omega is None else omega
    _check_dev_tensor(x, field, n * batch)
    _check_dev_tensor(out, field, out_len * batch)
    _check(lib().omr_partial_batch_device(context(), field.id, _dptr(x), _dptr(out), n, batch, omega, out_len,
                                          _stream_ptr()), "omr_partial_batch_device")
    return out


def coset_lde_device(coeffs, evals, field, n, ncols=1, blowup=4, shift=0):
    """evals[c][i] = sum_j coeffs[c][j] (shift w_m^i)^j, m = n * blowup (shift 0 = g)."""
    _check_dev_tensor(coeffs, field, n * ncols)
    _check_dev_tensor(evals, field, n * blowup * ncols)
    _check(lib().omr_coset_lde_device(context(), field.id, _dptr(coeffs), _dptr(evals), n, ncols,
                                      blowup, shift, _stream_ptr()), "omr_coset_lde_device")
    return evals


def coset_decode_device(evals, coeffs, field, n, ncols=1, blowup=4, shift=0, nonzero=None):
    """Inverse of coset_lde_device; nonzero (int32 GPU tensor of 1) receives the
    count of nonzero coefficients at j >= n."""
    _check_dev_tensor(evals, field, n * blowup * ncols)
    _check_dev_tensor(coeffs, field, n * ncols)
    nz = _dptr(nonzero) if nonzero is not None else None
    _check(lib().omr_coset_decode_device(context(), field.id, _dptr(evals), _dptr(coeffs), n, ncols,
                                         blowup, shift, nz, _stream_ptr()), "omr_coset_decode_device")
    return coeffs


class BatchItem(ctypes.Structure):
    _fields_ = [("in_offset", ctypes.c_uint64), ("out_offset", ctypes.c_uint64),
                ("log_n", ctypes.c_uint32), ("inverse", ctypes.c_uint32)]


class BatchPlan:
    """Batch of vectors of uneven power-of-two sizes (omr_batch_plan_*).
    items: list of (in_offset, out_offset, log_n, inverse) in elements."""

    def __init__(self, field, items):
        self.field = field
        arr = (BatchItem * len(items))(*[BatchItem(*it) for it in items])
        self._p = _VP()
        _check(lib().omr_batch_plan_create(context(), field.id, ctypes.cast(arr, _VP), len(items),
                                           ctypes.byref(self._p)), "omr_batch_plan_create")

    def info(self):
        tot, nl, nt = ctypes.c_uint64(), ctypes.c_uint32(), ctypes.c_uint32()
        _check(lib().omr_batch_plan_info(self._p, ctypes.byref(tot), ctypes.byref(nl), ctypes.byref(nt)))
        return {"total_elems": tot.value, "launches": nl.value, "tasks": nt.value}

    def execute(self, x, out, mode="graph"):
        fn = {"graph": lib().omr_batch_execute_graph, "persistent": lib().omr_batch_execute,
              "plain": lib().omr_batch_execute_plain, "grouped": lib().omr_batch_execute_grouped,
              "grouped_graph": lib().omr_batch_execute_grouped_graph}[mode]
        _check(fn(self._p, _dptr(x), _dptr(out), _stream_ptr()), f"omr_batch_execute({mode})")
        return out

    def close(self):
        if self._p:
            lib().omr_batch_plan_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_set_max_log_radix(r):
    """Radix cap 2^r (5..12) for later batch plans; 0 = automatic (2^10
    Goldilocks, 2^9 for the 4-byte fields)."""
    _check(lib().omr_batch_set_max_log_radix(r))


def batch_vector_schedule(log_n, max_log_radix=0):
    """Pass radices (log2) the batch planner gives one vector of 2^log_n."""
    lr = (ctypes.c_int * 9)()
    npass = ctypes.c_int()
    _check(lib().omr_batch_vector_schedule(log_n, max_log_radix, lr, ctypes.byref(npass)),
           "omr_batch_vector_schedule")
    return [lr[i] for i in range(npass.value)]


# ---- autotuner and learned cost model ---------------------------------------


def num_schedule_classes():
    return lib().omr_num_schedule_classes()


def schedule_class(field, n, out_len, cls):
    """Concrete schedule of class `cls` for a shape, or None if invalid:
    {"log_radix": [...], "log_e": e} for full transforms, {"method": m,
    "log_n1": l} for partial ones."""
    lr = (ctypes.c_int * 8)()
    npass, le, meth = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    l1 = ctypes.c_uint32()
    st = lib().omr_schedule_class(field.id, n, out_len, cls, lr, ctypes.byref(npass), ctypes.byref(le),
                                  ctypes.byref(meth), ctypes.byref(l1))
    if st == 11:
        return None
    _check(st, "omr_schedule_class")
    if out_len == n:
        return {"log_radix": [lr[i] for i in range(npass.value)], "log_e": le.value}
    return {"method": meth.value, "log_n1": l1.value}


def default_class(field, n, out_len):
    return lib().omr_default_class(field.id, n, out_len)


def tune_classes(field, n, out_len=None, batch=1, reps=5):
    """Median microseconds of every schedule class (-1 where invalid)."""
    out_len = n if out_len is None else out_len
    t = (ctypes.c_float * num_schedule_classes())()
    _check(lib().omr_tune_classes(context(), field.id, n, out_len, batch, reps, t), "omr_tune_classes")
    return np.array(t[:], dtype=np.float64)


def tune_set_class(field, n, out_len, cls, batch=1):
    """Installs class `cls` for `batch` vectors of the shape (batches of 32 MB
    or more have a table slot of their own)."""
    _check(lib().omr_tune_set_class_batch(field.id, n, out_len, batch, cls), "omr_tune_set_class_batch")


def tune_clear():
    _check(lib().omr_tune_clear())


def tune_save(path):
    _check(lib().omr_tune_save(path.encode()), "omr_tune_save")


def tune_load(path):
    _check(lib().omr_tune_load(path.encode()), "omr_tune_load")


def features(field, n, out_len, batch):
    x = np.zeros(4, dtype=np.float32)
    _check(lib().omr_cost_model_features(field.id, n, out_len, batch, x.ctypes.data_as(_VP)))
    return x


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


class CostModel:
    """The engine's 4 -> 64 -> 64 -> 36 MLP (HIP forward/backward kernels)."""

    def __init__(self, seed=0):
        self._p = _VP()
        _check(lib().omr_cost_model_create(context(), seed, ctypes.byref(self._p)), "omr_cost_model_create")
        self.n_params = lib().omr_cost_model_num_params()
        self.n_out = lib().omr_cost_model_num_outputs()

    def params(self):
        out = np.empty(self.n_params, dtype=np.float32)
        _check(lib().omr_cost_model_get_params(self._p, out.ctypes.data_as(_VP)))
        return out

    def set_params(self, p):
        p = _f32(p)
        assert p.size == self.n_params
        _check(lib().omr_cost_model_set_params(self._p, p.ctypes.data_as(_VP)))

    def forward(self, X):
        X = _f32(X).reshape(-1, 4)
        Y = np.empty((X.shape[0], self.n_out), dtype=np.float32)
        _check(lib().omr_cost_model_forward(self._p, X.ctypes.data_as(_VP), X.shape[0], Y.ctypes.data_as(_VP)))
        return Y

    def gradients(self, X, T, M):
        X, T, M = _f32(X).reshape(-1, 4), _f32(T), _f32(M)
        g = np.empty(self.n_params, dtype=np.float32)
        loss = ctypes.c_float()
        _check(lib().omr_cost_model_gradients(self._p, X.ctypes.data_as(_VP), T.ctypes.data_as(_VP),
                                              M.ctypes.data_as(_VP), X.shape[0], g.ctypes.data_as(_VP),
                                              ctypes.byref(loss)), "omr_cost_model_gradients")
        return g, loss.value

    def train(self, X, T, M, steps=2000, lr=3e-3):
        X, T, M = _f32(X).reshape(-1, 4), _f32(T), _f32(M)
        loss = ctypes.c_float()
        _check(lib().omr_cost_model_train(self._p, X.ctypes.data_as(_VP), T.ctypes.data_as(_VP),
                                          M.ctypes.data_as(_VP), X.shape[0], steps, lr, ctypes.byref(loss)),
               "omr_cost_model_train")
        return loss.value

    def choose(self, field, n, out_len=None, batch=1):
        c = ctypes.c_int()
        _check(lib().omr_cost_model_choose(self._p, field.id, n, n if out_len is None else out_len, batch,
                                           ctypes.byref(c)), "omr_cost_model_choose")
        return c.value

    def save(self, path):
        _check(lib().omr_cost_model_save(self._p, path.encode()), "omr_cost_model_save")

    def load(self, path):
        _check(lib().omr_cost_model_load(self._p, path.encode()), "omr_cost_model_load")

    def close(self):
        if self._p:
            lib().omr_cost_model_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune_shape(field, n, out_len=None, batch=1, model=None, use_model=True, reps=5, top_k=1):
    """Installs a schedule for the shape: the model's choice (top_k = 1, no
    timing), the fastest of the model's top_k predicted classes (top_k >= 2,
    top_k timing runs), or exhaustive timing of every class (use_model
    False).  Returns (class, classes_timed)."""
    c, timed = ctypes.c_int(), ctypes.c_int()
    mode = max(1, int(top_k)) if use_model else 0
    _check(lib().omr_autotune_shape(context(), model._p if model else None, mode, field.id, n,
                                    n if out_len is None else out_len, batch, reps, ctypes.byref(c),
                                    ctypes.byref(timed)), "omr_autotune_shape")
    return c.value, timed.value


def field_op_device(field, op, a, b, out):
    _check(lib().omr_field_op_device(field.id, op, _dptr(a), _dptr(b), _dptr(out), a.numel(),
                                     _stream_ptr()), "omr_field_op_device")
    return out


def synchronize():
    _check(lib().omr_ctx_synchronize(context(), _stream_ptr()))


# ---- tensors <-> numpy -----------------------------------------------------


def torch_dtype(field):
    import torch
    return torch.int64 if field.bytes == 8 else torch.int32


def to_device(arr, field):
    """numpy canonical values -> contiguous GPU tensor (bit-identical)."""
    import torch
    a = np.ascontiguousarray(np.asarray(arr).astype(field.np_dtype))
    view = a.view(np.int64 if field.bytes == 8 else np.int32)
    return torch.from_numpy(view.copy()).cuda()


def from_device(t, field):
    a = t.detach().cpu().numpy()
    return a.view(np.uint64 if field.bytes == 8 else np.uint32).astype(np.uint64)


def empty_device(n, field):
    import torch
    return torch.empty(n, dtype=torch_dtype(field), device="cuda")


# ---- the operator API (src/encoder.rs:70-83) ---------------------------------


class NttEncoder:
    variant = None
    s = 0

    def ntt(self, buf, domain):
        field = domain.field
        if _is_torch(buf):  # device-resident, in place
            out_len = buf.numel() if self.variant < DFT_PARTIAL else min(2 * self.s, domain.N)
            if self.variant >= DFT_PARTIAL:
                tmp = empty_device(max(out_len, 1), field)
                ntt_device(buf, tmp, field, self.variant, domain.omega, self.s, n=domain.N)
                buf[:out_len].copy_(tmp[:out_len])
            else:
                ntt_device(buf, buf, field, self.variant, domain.omega, self.s, n=domain.N)
            return
        if not isinstance(buf, np.ndarray) or buf.dtype != field.np_dtype or not buf.flags["C_CONTIGUOUS"]:
            raise TypeError(f"buf must be a contiguous numpy {field.np_dtype.__name__} array")
        if buf.size != domain.N:
            raise OmrError(12, f"buf.len()={buf.size} != N={domain.N}")
        _check(lib().omr_ntt(context(), field.id, self.variant, buf.ctypes.data_as(_VP), domain.N,
                             domain.omega, self.s), self.name())

    def name(self):
        return lib().omr_variant_name(self.variant).decode()


def _full(variant_id, cls_name):
    return type(cls_name, (NttEncoder,), {"variant": variant_id})


def _partial(variant_id, cls_name):
    def __init__(self, s):
        self.s = s
    return type(cls_name, (NttEncoder,), {"variant": variant_id, "__init__": __init__})


ArkRadix2 = _full(ARK_RADIX2, "ArkRadix2")
Dft = _full(DFT, "Dft")
LambdaBowers = _full(LAMBDA_BOWERS, "LambdaBowers")
LambdaRadix4 = _full(LAMBDA_RADIX4, "LambdaRadix4")
Plonky3Radix2LayerSplit = _full(PLONKY3_LAYER_SPLIT, "Plonky3Radix2LayerSplit")
WinterfellFourStep = _full(WINTERFELL_FOUR_STEP, "WinterfellFourStep")
TfheStockhamRadix8 = _full(TFHE_STOCKHAM_RADIX8, "TfheStockhamRadix8")
DftPartial = _partial(DFT_PARTIAL, "DftPartial")
TransformDecomposition = _partial(TRANSFORM_DECOMPOSITION, "TransformDecomposition")
TransformDecompositionV2 = _partial(TRANSFORM_DECOMPOSITION_V2, "TransformDecompositionV2")


class encoders:
    """Registry (src/encoders/mod.rs:32-54)."""

    @staticmethod
    def all(log_n):
        v = [ArkRadix2(), LambdaBowers(), WinterfellFourStep(), Plonky3Radix2LayerSplit()]
        if log_n % 2 == 0:
            v.append(LambdaRadix4())
        if log_n % 3 == 0:
            v.append(TfheStockhamRadix8())
        return v

    @staticmethod
    def all_partial(s):
        return [DftPartial(s), TransformDecomposition(s), TransformDecompositionV2(s)]
