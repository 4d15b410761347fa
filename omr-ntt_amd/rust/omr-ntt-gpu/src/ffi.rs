//! The C ABI of include/omrntt.h used by the encoders (status codes,
//! variants and fields are the enums declared there).
use std::os::raw::{c_char, c_int, c_void};

#[repr(C)]
pub struct OmrCtx {
    _p: [u8; 0],
}

pub const OMR_OK: c_int = 0;

extern "C" {
    pub fn omr_status_str(s: c_int) -> *const c_char;
    pub fn omr_ctx_create(device: c_int, out: *mut *mut OmrCtx) -> c_int;
    pub fn omr_ctx_destroy(ctx: *mut OmrCtx) -> c_int;
    pub fn omr_variant_name(variant: c_int) -> *const c_char;
    pub fn omr_variant_supports(variant: c_int, log_n: u32) -> c_int;
    pub fn omr_field_info(field: c_int, modulus: *mut u64, generator: *mut u64, two_adicity: *mut c_int) -> c_int;
    /// NttEncoder::ntt on a host slice of `word_bytes`-wide words.
    pub fn omr_ntt_words(ctx: *mut OmrCtx, field: c_int, variant: c_int, buf: *mut c_void, n: u64, omega: u64,
                         s: u64, word_bytes: u32) -> c_int;
    /// Inverse on a host slice of canonical words of the field's width.
    pub fn omr_intt(ctx: *mut OmrCtx, field: c_int, buf: *mut c_void, n: u64, omega: u64) -> c_int;
    /// NttDomain's public tables (src/encoder.rs:14-18) as canonical words of
    /// the field's width: twiddles (n words) and bitrev_twiddles (n/2 words);
    /// either pointer may be null.
    pub fn omr_domain_tables(ctx: *mut OmrCtx, field: c_int, n: u64, omega: u64, twiddles: *mut c_void,
                             bitrev_twiddles: *mut c_void) -> c_int;
    pub fn omr_host_alloc(bytes: usize, out: *mut *mut c_void) -> c_int;
    pub fn omr_host_free(p: *mut c_void) -> c_int;
}
