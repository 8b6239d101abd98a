//! Declarations of the C ABI in `include/particular_amd.h` that this crate calls.

use std::os::raw::{c_char, c_double, c_float, c_int, c_uint};

/// Opaque context (`pnb_ctx`): one GPU, one stream, grow-only device buffers.
#[repr(C)]
pub struct PnbCtx {
    _private: [u8; 0],
}

pub const PNB_OK: c_int = 0;
pub const PNB_BH_MLP: c_uint = 1;
pub const PNB_BH_REF_COM: c_uint = 2;
pub const PNB_BH_RELATIVE: c_uint = 4;

extern "C" {
    pub fn pnb_create(device: c_int, out: *mut *mut PnbCtx) -> c_int;
    pub fn pnb_destroy(ctx: *mut PnbCtx);
    pub fn pnb_last_error(ctx: *const PnbCtx) -> *const c_char;
    pub fn pnb_status_string(status: c_int) -> *const c_char;

    pub fn pnb_bf_accel_f32(ctx: *mut PnbCtx, affected: *const c_float, n: usize,
                            affecting: *const c_float, m: usize, softening: c_float,
                            checked: c_int, out: *mut c_float) -> c_int;
    pub fn pnb_reordered_bf_accel_f32(ctx: *mut PnbCtx, particles: *const c_float, n: usize,
                                      softening: c_float, checked: c_int,
                                      out: *mut c_float) -> c_int;
    pub fn pnb_bf_pairs_accel_f32(ctx: *mut PnbCtx, particles: *const c_float, n: usize,
                                  softening: c_float, checked: c_int,
                                  out: *mut c_float) -> c_int;
    pub fn pnb_bh_accel_f32(ctx: *mut PnbCtx, affected: *const c_float, n: usize,
                            affecting: *const c_float, m: usize, theta: c_float,
                            softening: c_float, checked: c_int, flags: c_uint,
                            out: *mut c_float) -> c_int;
    pub fn pnb_reordered_bh_accel_f32(ctx: *mut PnbCtx, particles: *const c_float, n: usize,
                                      theta: c_float, softening: c_float, checked: c_int,
                                      flags: c_uint, out: *mut c_float) -> c_int;
    pub fn pnb_bh_set_relative(ctx: *mut PnbCtx, alpha: c_float, estimate_theta: c_float,
                               per_particle: c_int, bucket: c_int) -> c_int;
    pub fn pnb_bh_set_capturable(ctx: *mut PnbCtx, enable: c_int) -> c_int;
    pub fn pnb_bh_check(ctx: *mut PnbCtx) -> c_int;

    pub fn pnb_bf_accel_f32_2d(ctx: *mut PnbCtx, affected: *const c_float, n: usize,
                               affecting: *const c_float, m: usize, softening: c_float,
                               checked: c_int, out: *mut c_float) -> c_int;
    pub fn pnb_bh_accel_f32_2d(ctx: *mut PnbCtx, affected: *const c_float, n: usize,
                               affecting: *const c_float, m: usize, theta: c_float,
                               softening: c_float, checked: c_int, flags: c_uint,
                               out: *mut c_float) -> c_int;

    pub fn pnb_bf_accel_f64(ctx: *mut PnbCtx, affected: *const c_double, n: usize,
                            affecting: *const c_double, m: usize, softening: c_double,
                            checked: c_int, out: *mut c_double) -> c_int;

    pub fn pnb_mlp_num_params() -> c_int;
    pub fn pnb_mlp_set_params(ctx: *mut PnbCtx, params: *const c_float, n: usize) -> c_int;
}
