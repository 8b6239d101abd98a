//! MI355X (gfx950) backend for particular's `Interaction` API.
//!
//! The algorithms here replace `gpu::BruteForce` (particular/src/gpu/mod.rs:149-208) and
//! `parallel::BarnesHut` (particular/src/parallel.rs:298-367) with hand-written HIP kernels
//! behind the C ABI of `include/particular_amd.h`. Each implements
//! `Interaction<Between<&[P1], &[P2]>>`, so particular's blanket impls
//! (particular/src/storage.rs:207-241) give `&[P]`, `&Ordered<P>` and `&Reordered<P, F>` for free,
//! which is how the massive/massless split keeps working unchanged. `HipCompute` is the
//! `GpuCompute` sugar (particular/src/gpu/mod.rs:13-37).
//!
//! ```ignore
//! use particular::prelude::*;
//! use particular::gravity::newtonian::AccelerationSoftened;
//! use particular_hip::{HipCompute, HipResources};
//!
//! let mut res = HipResources::new(0);
//! let reordered = Reordered::new(&particles, |p| p.mu() != 0.0);
//! let accelerations: Vec<glam::Vec3> = reordered
//!     .hip_brute_force(&mut res, AccelerationSoftened::checked(1e-3))
//!     .collect();
//! let tree: Vec<glam::Vec3> = particles.as_slice()
//!     .hip_barnes_hut(&mut res, 0.5, Acceleration::checked())
//!     .collect();
//! ```
//!
//! Supported particles: `glam::Vec3` / `f32` (brute force and Barnes-Hut), `glam::Vec2` / `f32`
//! (both), `glam::DVec3` / `f64` (brute force; the reference computes these on the CPU only).

pub mod ffi;

use std::ffi::CStr;
use std::fmt;
use std::os::raw::{c_int, c_uint};

use particular::gravity::newtonian::{Acceleration, AccelerationSoftened};
use particular::gravity::{Mass, Position};
use particular::{Between, Interaction, Ordered, Reordered};

/// A failed call: the library's status code and message.
#[derive(Clone, Debug)]
pub struct HipError {
    pub status: i32,
    pub message: String,
}

impl fmt::Display for HipError {
    fn fmt(&self, f: &mut fmt::Formatter<'_>) -> fmt::Result {
        write!(f, "particular_amd status {}: {}", self.status, self.message)
    }
}

impl std::error::Error for HipError {}

/// The `GpuResources` analogue (particular/src/gpu/mod.rs:84-143): one context per GPU that owns
/// its stream and grow-only device buffers, reused across calls. Like the reference's
/// `&mut GpuResources`, a context serves one call at a time.
pub struct HipResources {
    ctx: *mut ffi::PnbCtx,
}

// A context may move between threads; it is used by one thread at a time (&mut borrows).
unsafe impl Send for HipResources {}

impl HipResources {
    /// Opens a context on `device`. Panics when no usable GPU exists, as the reference panics
    /// when wgpu finds no adapter.
    pub fn new(device: i32) -> Self {
        Self::try_new(device).unwrap_or_else(|e| panic!("{e}"))
    }

    pub fn try_new(device: i32) -> Result<Self, HipError> {
        let mut ctx = std::ptr::null_mut();
        let rc = unsafe { ffi::pnb_create(device, &mut ctx) };
        if rc != ffi::PNB_OK {
            let msg = unsafe { CStr::from_ptr(ffi::pnb_status_string(rc)) };
            return Err(HipError { status: rc, message: msg.to_string_lossy().into_owned() });
        }
        Ok(Self { ctx })
    }

    /// The raw context, for the entry points this crate does not wrap.
    pub fn as_ptr(&self) -> *mut ffi::PnbCtx {
        self.ctx
    }

    fn status(&self, rc: c_int) -> Result<(), HipError> {
        if rc == ffi::PNB_OK {
            return Ok(());
        }
        let msg = unsafe { CStr::from_ptr(ffi::pnb_last_error(self.ctx)) };
        Err(HipError { status: rc, message: msg.to_string_lossy().into_owned() })
    }

    /// The reference panics on device failures (particular/src/gpu/resources.rs:24,246,345); the
    /// `Interaction` impls do the same, with the library's message.
    fn expect(&self, rc: c_int) {
        if let Err(e) = self.status(rc) {
            panic!("{e}");
        }
    }

    /// Parameters of the relative opening criterion (`BarnesHut::relative`) for later calls.
    pub fn set_relative(&mut self, alpha: f32, estimate_theta: f32, per_particle: bool,
                        bucket: i32) -> Result<(), HipError> {
        let rc = unsafe {
            ffi::pnb_bh_set_relative(self.ctx, alpha, estimate_theta, per_particle as c_int, bucket)
        };
        self.status(rc)
    }

    /// Loads the learned correction's network (`BarnesHut::learned`): `params` holds the
    /// library's parameter vector (`pnb_mlp_num_params` values, e.g. a trained file's contents).
    pub fn set_network(&mut self, params: &[f32]) -> Result<(), HipError> {
        let want = unsafe { ffi::pnb_mlp_num_params() } as usize;
        if params.len() != want {
            return Err(HipError {
                status: 1,
                message: format!("network has {} parameters, expected {want}", params.len()),
            });
        }
        let rc = unsafe { ffi::pnb_mlp_set_params(self.ctx, params.as_ptr(), params.len()) };
        self.status(rc)
    }
}

impl Drop for HipResources {
    fn drop(&mut self) {
        unsafe { ffi::pnb_destroy(self.ctx) }
    }
}

/// What the kernels need from particular's interaction types: the softening and whether pairs
/// at zero distance are checked (particular/src/gravity/newtonian/acceleration.rs:46-58,
/// acceleration_softened.rs:51-63).
pub trait HipInteraction<S> {
    const CHECKED: bool;
    fn softening(&self) -> S;
}

impl<const C: bool, S: Default> HipInteraction<S> for Acceleration<C> {
    const CHECKED: bool = C;
    fn softening(&self) -> S {
        S::default()
    }
}

impl<const C: bool, S: Copy> HipInteraction<S> for AccelerationSoftened<S, C> {
    const CHECKED: bool = C;
    fn softening(&self) -> S {
        self.softening
    }
}

/// A vector type the library takes: its record layout and its entry points.
pub trait HipVector: Sized + Copy {
    type Scalar: Copy + Default;
    /// Scalars per particle record ({x, y, z, mu} or {x, y, mu}).
    const RECORD: usize;
    /// Scalars per output record.
    const OUTPUT: usize;
    fn write_record(&self, mu: Self::Scalar, out: &mut Vec<Self::Scalar>);
    fn read_output(o: &[Self::Scalar]) -> Self;
    /// # Safety
    /// `a`, `b` and `out` hold `n`, `m` and `n` records.
    unsafe fn brute_force(ctx: *mut ffi::PnbCtx, a: *const Self::Scalar, n: usize,
                          b: *const Self::Scalar, m: usize, softening: Self::Scalar,
                          checked: c_int, out: *mut Self::Scalar) -> c_int;
}

/// Vector types with a GPU Barnes-Hut (f32, two and three dimensions).
pub trait HipTreeVector: HipVector<Scalar = f32> {
    /// # Safety
    /// As `HipVector::brute_force`.
    unsafe fn barnes_hut(ctx: *mut ffi::PnbCtx, a: *const f32, n: usize, b: *const f32, m: usize,
                         theta: f32, softening: f32, checked: c_int, flags: c_uint,
                         out: *mut f32) -> c_int;
}

impl HipVector for glam::Vec3 {
    type Scalar = f32;
    const RECORD: usize = 4;
    const OUTPUT: usize = 4;
    fn write_record(&self, mu: f32, out: &mut Vec<f32>) {
        out.extend_from_slice(&[self.x, self.y, self.z, mu]);
    }
    fn read_output(o: &[f32]) -> Self {
        glam::Vec3::new(o[0], o[1], o[2])
    }
    unsafe fn brute_force(ctx: *mut ffi::PnbCtx, a: *const f32, n: usize, b: *const f32, m: usize,
                          softening: f32, checked: c_int, out: *mut f32) -> c_int {
        ffi::pnb_bf_accel_f32(ctx, a, n, b, m, softening, checked, out)
    }
}

impl HipTreeVector for glam::Vec3 {
    unsafe fn barnes_hut(ctx: *mut ffi::PnbCtx, a: *const f32, n: usize, b: *const f32, m: usize,
                         theta: f32, softening: f32, checked: c_int, flags: c_uint,
                         out: *mut f32) -> c_int {
        ffi::pnb_bh_accel_f32(ctx, a, n, b, m, theta, softening, checked, flags, out)
    }
}

impl HipVector for glam::Vec2 {
    type Scalar = f32;
    const RECORD: usize = 3;
    const OUTPUT: usize = 2;
    fn write_record(&self, mu: f32, out: &mut Vec<f32>) {
        out.extend_from_slice(&[self.x, self.y, mu]);
    }
    fn read_output(o: &[f32]) -> Self {
        glam::Vec2::new(o[0], o[1])
    }
    unsafe fn brute_force(ctx: *mut ffi::PnbCtx, a: *const f32, n: usize, b: *const f32, m: usize,
                          softening: f32, checked: c_int, out: *mut f32) -> c_int {
        ffi::pnb_bf_accel_f32_2d(ctx, a, n, b, m, softening, checked, out)
    }
}

impl HipTreeVector for glam::Vec2 {
    unsafe fn barnes_hut(ctx: *mut ffi::PnbCtx, a: *const f32, n: usize, b: *const f32, m: usize,
                         theta: f32, softening: f32, checked: c_int, flags: c_uint,
                         out: *mut f32) -> c_int {
        ffi::pnb_bh_accel_f32_2d(ctx, a, n, b, m, theta, softening, checked, flags, out)
    }
}

impl HipVector for glam::DVec3 {
    type Scalar = f64;
    const RECORD: usize = 4;
    const OUTPUT: usize = 4;
    fn write_record(&self, mu: f64, out: &mut Vec<f64>) {
        out.extend_from_slice(&[self.x, self.y, self.z, mu]);
    }
    fn read_output(o: &[f64]) -> Self {
        glam::DVec3::new(o[0], o[1], o[2])
    }
    unsafe fn brute_force(ctx: *mut ffi::PnbCtx, a: *const f64, n: usize, b: *const f64, m: usize,
                          softening: f64, checked: c_int, out: *mut f64) -> c_int {
        ffi::pnb_bf_accel_f64(ctx, a, n, b, m, softening, checked, out)
    }
}

/// Particle records in the library's layout (`GravitationalField::from`, particular/src/gravity/
/// mod.rs:150-161: position and mu).
fn pack<P, V>(ps: &[P]) -> Vec<V::Scalar>
where
    V: HipVector,
    P: Position<Vector = V> + Mass<Scalar = V::Scalar>,
{
    let mut v = Vec::with_capacity(ps.len() * V::RECORD);
    for p in ps {
        p.position().write_record(p.mu(), &mut v);
    }
    v
}

fn unpack<V: HipVector>(out: &[V::Scalar]) -> std::vec::IntoIter<V> {
    out.chunks_exact(V::OUTPUT).map(V::read_output).collect::<Vec<_>>().into_iter()
}

/// Brute force on the GPU: replaces `gpu::BruteForce` (particular/src/gpu/mod.rs:149-208) with the
/// results of `sequential::BruteForce` (particular/src/sequential.rs:181-194), compensated sums.
pub struct BruteForce<'a, T> {
    /// Context reused across calls (like the reference's `GpuResources`, it should not be
    /// recreated for every call).
    pub resources: &'a mut HipResources,
    /// Interaction to compute between two particles.
    pub interaction: T,
}

impl<'a, T> BruteForce<'a, T> {
    pub fn new(resources: &'a mut HipResources, interaction: T) -> Self {
        Self { resources, interaction }
    }
}

impl<P1, P2, V, T> Interaction<Between<&[P1], &[P2]>> for BruteForce<'_, T>
where
    V: HipVector,
    P1: Position<Vector = V> + Mass<Scalar = V::Scalar>,
    P2: Position<Vector = V> + Mass<Scalar = V::Scalar>,
    T: HipInteraction<V::Scalar>,
{
    type Output = std::vec::IntoIter<V>;

    fn compute(&mut self, Between(affected, affecting): Between<&[P1], &[P2]>) -> Self::Output {
        let (a, b) = (pack(affected), pack(affecting));
        let mut out = vec![V::Scalar::default(); affected.len() * V::OUTPUT];
        let rc = unsafe {
            V::brute_force(self.resources.ctx, a.as_ptr(), affected.len(), b.as_ptr(),
                           affecting.len(), self.interaction.softening(), T::CHECKED as c_int,
                           out.as_mut_ptr())
        };
        self.resources.expect(rc);
        unpack::<V>(&out)
    }
}

/// Particles whose massive/massless split the GPU performs (`Reordered::new(p, |p| p.mu() != 0)`
/// without the host copy, particular/src/storage.rs:85-95, 219-229): every particle is affected,
/// the affecting set is those with mu != 0 in their original order, outputs follow the input.
#[derive(Clone, Copy)]
pub struct SplitOnDevice<'p, P>(pub &'p [P]);

impl<'p, P, T> Interaction<SplitOnDevice<'p, P>> for BruteForce<'_, T>
where
    P: Position<Vector = glam::Vec3> + Mass<Scalar = f32>,
    T: HipInteraction<f32>,
{
    type Output = std::vec::IntoIter<glam::Vec3>;

    fn compute(&mut self, SplitOnDevice(ps): SplitOnDevice<'p, P>) -> Self::Output {
        let a = pack(ps);
        let mut out = vec![0f32; ps.len() * 4];
        let rc = unsafe {
            ffi::pnb_reordered_bf_accel_f32(self.resources.ctx, a.as_ptr(), ps.len(),
                                            self.interaction.softening(), T::CHECKED as c_int,
                                            out.as_mut_ptr())
        };
        self.resources.expect(rc);
        unpack::<glam::Vec3>(&out)
    }
}

/// Barnes-Hut on the GPU: replaces `parallel::BarnesHut` (particular/src/parallel.rs:298-367).
/// The octree is the reference's (particular/src/tree/mod.rs:73-139), node for node; every
/// affected particle applies the reference's traversal rule (particular/src/sequential.rs:456-506)
/// with `theta`. Centres of mass are reduced in FP64 unless `reference_com` asks for the
/// reference's own f32 folds.
pub struct BarnesHut<'a, T> {
    pub resources: &'a mut HipResources,
    pub theta: f32,
    pub interaction: T,
    /// `ffi::PNB_BH_*` bits.
    pub flags: u32,
}

impl<'a, T> BarnesHut<'a, T> {
    pub fn new(resources: &'a mut HipResources, theta: f32, interaction: T) -> Self {
        Self { resources, theta, interaction, flags: 0 }
    }

    /// Node data exactly as the reference computes it (f32 sequential folds); for parity checks.
    pub fn reference_com(mut self) -> Self {
        self.flags |= ffi::PNB_BH_REF_COM;
        self
    }

    /// Adds the learned correction of each accepted internal node (a network loaded with
    /// `HipResources::set_network`). No reference counterpart.
    pub fn learned(mut self) -> Self {
        self.flags |= ffi::PNB_BH_MLP;
        self
    }

    /// Relative opening criterion with the given alpha (no reference counterpart): a node also
    /// opens when its first neglected term exceeds alpha times the particle's estimated
    /// acceleration. The estimate's parameters keep the context's values (defaults 1.2, one
    /// estimate per wavefront, buckets of 24).
    pub fn relative(mut self, alpha: f32) -> Self {
        self.resources.set_relative(alpha, 1.2, false, 24).unwrap_or_else(|e| panic!("{e}"));
        self.flags |= ffi::PNB_BH_RELATIVE;
        self
    }
}

impl<P1, P2, V, T> Interaction<Between<&[P1], &[P2]>> for BarnesHut<'_, T>
where
    V: HipTreeVector,
    P1: Position<Vector = V> + Mass<Scalar = f32>,
    P2: Position<Vector = V> + Mass<Scalar = f32>,
    T: HipInteraction<f32>,
{
    type Output = std::vec::IntoIter<V>;

    fn compute(&mut self, Between(affected, affecting): Between<&[P1], &[P2]>) -> Self::Output {
        let (a, b) = (pack(affected), pack(affecting));
        let mut out = vec![0f32; affected.len() * V::OUTPUT];
        let rc = unsafe {
            V::barnes_hut(self.resources.ctx, a.as_ptr(), affected.len(), b.as_ptr(),
                          affecting.len(), self.theta, self.interaction.softening(),
                          T::CHECKED as c_int, self.flags as c_uint, out.as_mut_ptr())
        };
        self.resources.expect(rc);
        unpack::<V>(&out)
    }
}

impl<'p, P, T> Interaction<SplitOnDevice<'p, P>> for BarnesHut<'_, T>
where
    P: Position<Vector = glam::Vec3> + Mass<Scalar = f32>,
    T: HipInteraction<f32>,
{
    type Output = std::vec::IntoIter<glam::Vec3>;

    fn compute(&mut self, SplitOnDevice(ps): SplitOnDevice<'p, P>) -> Self::Output {
        let a = pack(ps);
        let mut out = vec![0f32; ps.len() * 4];
        let rc = unsafe {
            ffi::pnb_reordered_bh_accel_f32(self.resources.ctx, a.as_ptr(), ps.len(), self.theta,
                                            self.interaction.softening(), T::CHECKED as c_int,
                                            self.flags as c_uint, out.as_mut_ptr())
        };
        self.resources.expect(rc);
        unpack::<glam::Vec3>(&out)
    }
}

/// `GpuCompute` sugar (particular/src/gpu/mod.rs:13-37): `storage.hip_brute_force(&mut res, i)`.
pub trait HipCompute<T>: Sized {
    fn hip_brute_force<'a>(self, resources: &'a mut HipResources, interaction: T)
        -> <BruteForce<'a, T> as Interaction<Self>>::Output
    where
        BruteForce<'a, T>: Interaction<Self>,
    {
        BruteForce::new(resources, interaction).compute(self)
    }

    fn hip_barnes_hut<'a>(self, resources: &'a mut HipResources, theta: f32, interaction: T)
        -> <BarnesHut<'a, T> as Interaction<Self>>::Output
    where
        BarnesHut<'a, T>: Interaction<Self>,
    {
        BarnesHut::new(resources, theta, interaction).compute(self)
    }
}

// The storages the reference's GpuCompute covers (particular/src/gpu/mod.rs:33-37), plus the
// device-side split.
impl<T, P> HipCompute<T> for &[P] {}
impl<T, P> HipCompute<T> for &Ordered<P> {}
impl<T, P, F> HipCompute<T> for &Reordered<'_, P, F> {}
impl<T, S1, S2> HipCompute<T> for Between<S1, S2> {}
impl<T, P> HipCompute<T> for SplitOnDevice<'_, P> {}

#[cfg(test)]
mod tests {
    //! Need an MI355X: `cargo test -- --ignored`. The data are those of the reference's
    //! `acceleration_error!` test (particular/src/gravity/newtonian/mod.rs:230-262).
    use super::*;
    use particular::gravity::GravitationalField;
    use particular::prelude::SequentialCompute;

    fn bodies() -> Vec<GravitationalField<glam::Vec3, f32>> {
        [(10.0, 0.0), (0.0, 20.0), (1.0, 30.0), (-3.0, 40.0), (30.0, 0.0), (-45.0, 0.0)]
            .into_iter()
            .map(|(x, m)| GravitationalField::new(glam::Vec3::splat(x), m))
            .collect()
    }

    fn close(a: &[glam::Vec3], b: &[glam::Vec3], tol: f32) {
        for (x, y) in a.iter().zip(b) {
            assert!((*x - *y).length() <= tol * y.length().max(f32::MIN_POSITIVE), "{x} vs {y}");
        }
    }

    #[test]
    #[ignore = "needs an MI355X"]
    fn brute_force_matches_sequential() {
        let ps = bodies();
        let mut res = HipResources::new(0);
        let want: Vec<_> = ps.as_slice().brute_force(Acceleration::checked()).collect();
        let got: Vec<_> = ps.as_slice().hip_brute_force(&mut res, Acceleration::checked()).collect();
        close(&got, &want, 1e-6);
        let reordered = Reordered::new(&ps, |p| p.m != 0.0);
        let got: Vec<_> = reordered.hip_brute_force(&mut res, Acceleration::checked()).collect();
        close(&got, &want, 1e-6);
        let got: Vec<_> =
            SplitOnDevice(ps.as_slice()).hip_brute_force(&mut res, Acceleration::checked()).collect();
        close(&got, &want, 1e-6);
    }

    #[test]
    #[ignore = "needs an MI355X"]
    fn barnes_hut_at_theta_zero_is_brute_force() {
        let ps = bodies();
        let mut res = HipResources::new(0);
        let want: Vec<_> = ps.as_slice().brute_force(Acceleration::checked()).collect();
        let got: Vec<_> =
            ps.as_slice().hip_barnes_hut(&mut res, 0.0, Acceleration::checked()).collect();
        close(&got, &want, 1e-6);
    }
}
