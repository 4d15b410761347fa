//! `GpuEncoder`: the crate's `NttEncoder<F>` (src/encoder.rs:70-74) on the
//! MI355X engine.  The slice is handed to the engine as it lies in memory:
//! every field of src/fields.rs is an ark-ff one-limb type (one u64 word per
//! element, Montgomery form), and every transform is F-linear, so the words
//! x R mod p come back as the words of the transform -- no conversion pass.
//! Only `domain.omega` crosses canonical.  Status codes become panics with
//! the library's message, as the reference's `assert!`s panic.
pub mod ffi;

use ark_ff::{BigInteger, FftField, PrimeField};
use ntt::{NttDomain, NttEncoder};
use std::ffi::CStr;

/// Variant ids of include/omrntt.h (the reference registry's names).
pub mod variant {
    pub const ARK_RADIX2: i32 = 0;
    pub const DFT: i32 = 1;
    pub const LAMBDA_BOWERS: i32 = 2;
    pub const LAMBDA_RADIX4: i32 = 3;
    pub const PLONKY3_RADIX2_LAYER_SPLIT: i32 = 4;
    pub const WINTERFELL_FOUR_STEP: i32 = 5;
    pub const TFHE_STOCKHAM_RADIX8: i32 = 6;
    pub const DFT_PARTIAL: i32 = 7;
    pub const TRANSFORM_DECOMPOSITION: i32 = 8;
    pub const TRANSFORM_DECOMPOSITION_V2: i32 = 9;
}

/// One engine context (one GPU) shared by the encoders built on it.
pub struct GpuContext {
    ctx: *mut ffi::OmrCtx,
}
unsafe impl Send for GpuContext {}
unsafe impl Sync for GpuContext {} // the library locks the context per call

impl GpuContext {
    pub fn new(device: i32) -> Self {
        let mut ctx = std::ptr::null_mut();
        check(unsafe { ffi::omr_ctx_create(device, &mut ctx) }, "omr_ctx_create");
        GpuContext { ctx }
    }
}

impl Drop for GpuContext {
    fn drop(&mut self) {
        unsafe { ffi::omr_ctx_destroy(self.ctx) };
    }
}

fn check(status: i32, what: &str) {
    if status != ffi::OMR_OK {
        let msg = unsafe { CStr::from_ptr(ffi::omr_status_str(status)) };
        panic!("{what}: {}", msg.to_string_lossy());
    }
}

/// The engine's field id for F (by modulus), if it has one.
fn field_id<F: PrimeField>() -> Option<i32> {
    let m = F::MODULUS.as_ref();
    if m.len() != 1 && m[1..].iter().any(|&w| w != 0) {
        return None;
    }
    (0..4).find(|&f| {
        let mut p = 0u64;
        unsafe { ffi::omr_field_info(f, &mut p, std::ptr::null_mut(), std::ptr::null_mut()) == ffi::OMR_OK }
            && p == m[0]
    })
}

/// A reference encoder variant (full, or partial with its `s`) on the GPU.
pub struct GpuEncoder<'a> {
    ctx: &'a GpuContext,
    variant: i32,
    s: u64,
}

impl<'a> GpuEncoder<'a> {
    pub fn new(ctx: &'a GpuContext, variant: i32, s: u64) -> Self {
        GpuEncoder { ctx, variant, s }
    }
}

impl<'a, F: FftField + PrimeField> NttEncoder<F> for GpuEncoder<'a> {
    fn ntt(&self, buf: &mut [F], domain: &NttDomain<F>) {
        assert!(buf.len() == domain.N, "buf.len() != N");
        let fid = field_id::<F>().expect("field not supported by the GPU engine");
        // one u64 word per element (ark-ff one-limb storage)
        assert_eq!(std::mem::size_of::<F>(), 8, "expected one 64-bit limb per element");
        let omega = domain.omega.into_bigint().as_ref()[0]; // canonical
        let status = unsafe {
            ffi::omr_ntt_words(self.ctx.ctx, fid, self.variant, buf.as_mut_ptr().cast(), domain.N as u64, omega,
                               self.s, 8)
        };
        check(status, self.name());
    }

    fn name(&self) -> &str {
        unsafe { CStr::from_ptr(ffi::omr_variant_name(self.variant)) }.to_str().unwrap()
    }
}

/// The GPU counterparts of `encoders::all(log_n)` (src/encoders/mod.rs:32-46).
pub fn all<'a, F: FftField + PrimeField>(ctx: &'a GpuContext, log_n: u32) -> Vec<Box<dyn NttEncoder<F> + 'a>> {
    let mut v: Vec<Box<dyn NttEncoder<F> + 'a>> = Vec::new();
    for id in [variant::ARK_RADIX2, variant::LAMBDA_BOWERS, variant::WINTERFELL_FOUR_STEP,
               variant::PLONKY3_RADIX2_LAYER_SPLIT, variant::LAMBDA_RADIX4, variant::TFHE_STOCKHAM_RADIX8] {
        if unsafe { ffi::omr_variant_supports(id, log_n) } != 0 {
            v.push(Box::new(GpuEncoder::new(ctx, id, 0)));
        }
    }
    v
}

/// The GPU counterparts of `encoders::all_partial(s)` (src/encoders/mod.rs:48-54).
pub fn all_partial<'a, F: FftField + PrimeField>(ctx: &'a GpuContext, s: u64) -> Vec<Box<dyn NttEncoder<F> + 'a>> {
    vec![
        Box::new(GpuEncoder::new(ctx, variant::DFT_PARTIAL, s)),
        Box::new(GpuEncoder::new(ctx, variant::TRANSFORM_DECOMPOSITION, s)),
        Box::new(GpuEncoder::new(ctx, variant::TRANSFORM_DECOMPOSITION_V2, s)),
    ]
}
