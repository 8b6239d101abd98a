"""particular_amd: MI355X (gfx950) backend for particular's gravitational acceleration path.

Host-side mirror of the reference interface, over the C ABI of ``libparticular_amd.so``:

==============================  ==============================================================
reference (particular 0.8)       here
==============================  ==============================================================
``Between(affected, affecting)``  :class:`Between`            (particular/src/lib.rs:299-300)
``Interaction::compute``          ``.compute(storage)``       (particular/src/lib.rs:364-370)
``Acceleration::checked()``       :meth:`Acceleration.checked` (newtonian/acceleration.rs:27-44)
``AccelerationSoftened``          :class:`AccelerationSoftened` (acceleration_softened.rs:17-49)
``Ordered`` / ``Reordered``       :class:`Ordered` / :class:`Reordered` (storage.rs:51-205)
``GpuResources``                  :class:`HipResources`       (particular/src/gpu/mod.rs:84-143)
``gpu::BruteForce``               :class:`BruteForce`         (particular/src/gpu/mod.rs:149-208)
``sequential::BruteForcePairs``   :class:`BruteForcePairs`    (particular/src/sequential.rs:275-305)
``parallel::BarnesHut``           :class:`BarnesHut`          (particular/src/parallel.rs:298-367)
``GpuCompute`` sugar              :func:`hip_brute_force`, :func:`hip_barnes_hut` (gpu/mod.rs:13-37)
==============================  ==============================================================

Storages lower to ``Between`` exactly as the reference's blanket impls do (storage.rs:207-241):
a particle array ``p`` -> ``Between(p, p)``; ``Ordered`` -> ``Between(particles, affecting)``;
``Reordered`` -> ``Between(unordered, affecting)`` with outputs in the original order.

Particles are ``(n, 4)`` float32 arrays ``{x, y, z, mu}`` (numpy on the host, or torch tensors on
the GPU, which stay resident in HBM). Results are ``(n, 3)`` float32 numpy arrays for host input and
``(n, 4)`` float32 torch tensors (w = 0) for device input.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Any, Callable, NamedTuple

import numpy as np

from . import _lib
from ._lib import PnbError, check

__all__ = [
    "Between", "Acceleration", "AccelerationSoftened", "Ordered", "Reordered", "HipResources",
    "BruteForce", "BruteForcePairs", "BarnesHut", "hip_brute_force", "hip_barnes_hut", "PnbError",
    "device_count",
]


class Between(NamedTuple):
    """A pair of storages: the first is affected by the second (particular/src/lib.rs:299-300)."""

    affected: Any
    affecting: Any


@dataclass(frozen=True)
class Acceleration:
    """Newtonian acceleration between point masses (newtonian/acceleration.rs:17-44)."""

    is_checked: bool = True
    softening: float = 0.0

    @staticmethod
    def checked() -> "Acceleration":
        return Acceleration(True)

    @staticmethod
    def unchecked() -> "Acceleration":
        return Acceleration(False)


@dataclass(frozen=True)
class AccelerationSoftened:
    """Newtonian acceleration with Plummer softening eps (acceleration_softened.rs:17-49)."""

    softening: float
    is_checked: bool = True

    @staticmethod
    def checked(softening: float) -> "AccelerationSoftened":
        return AccelerationSoftened(float(softening), True)

    @staticmethod
    def unchecked(softening: float) -> "AccelerationSoftened":
        return AccelerationSoftened(float(softening), False)


def _is_affecting(particles) -> np.ndarray:
    """GravitationalField::is_affecting: mu != 0 (particular/src/gravity/mod.rs:27-34)."""
    return np.asarray(particles)[:, -1] != 0  # mu is the last field (3D {x,y,z,mu}, 2D {x,y,mu})


class Ordered:
    """Affecting particles first, each group in its original order (storage.rs:51-138)."""

    def __init__(self, particles: np.ndarray, affecting_len: int):
        self._particles = particles
        self._affecting_len = int(affecting_len)

    @staticmethod
    def new(unordered, is_affecting_fn: Callable | None = None) -> "Ordered":
        unordered = _as_host(unordered, allow_2d=True)
        mask = _is_affecting(unordered) if is_affecting_fn is None else np.asarray(
            [bool(is_affecting_fn(p)) for p in unordered])
        parts = np.concatenate([unordered[mask], unordered[~mask]])
        return Ordered(np.ascontiguousarray(parts), int(mask.sum()))

    def affecting_len(self) -> int:
        return self._affecting_len

    def affecting(self) -> np.ndarray:
        return self._particles[: self._affecting_len]

    def non_affecting(self) -> np.ndarray:
        return self._particles[self._affecting_len:]

    def particles(self) -> np.ndarray:
        return self._particles


class Reordered:
    """The original particles plus an :class:`Ordered` copy (storage.rs:140-205).

    With the default affecting test (mu != 0) the split of a device tensor is done on the GPU by
    stream compaction instead of a host copy.
    """

    def __init__(self, unordered, is_affecting_fn: Callable | None = None):
        self.unordered = unordered
        self._fn = is_affecting_fn
        self._ordered = None

    @staticmethod
    def new(unordered, is_affecting_fn: Callable | None = None) -> "Reordered":
        return Reordered(unordered, is_affecting_fn)

    def ordered(self) -> Ordered:
        if self._ordered is None:
            self._ordered = Ordered.new(self.unordered, self._fn)
        return self._ordered

    def affecting_len(self) -> int:
        return self.ordered().affecting_len()

    def affecting(self) -> np.ndarray:
        return self.ordered().affecting()

    def non_affecting(self) -> np.ndarray:
        return self.ordered().non_affecting()

    def reordered(self) -> np.ndarray:
        return self.ordered().particles()


def device_count() -> int:
    n = ctypes.c_int(0)
    rc = _lib.lib().pnb_device_count(ctypes.byref(n))
    return n.value if rc == _lib.PNB_OK else 0


class HipResources:
    """Device resources of one GPU: a context with its stream and grow-only buffers.

    The GpuResources analogue (particular/src/gpu/mod.rs:84-143): create it once and reuse it;
    one host thread at a time.
    """

    def __init__(self, device: int = 0):
        L = _lib.lib()
        h = ctypes.c_void_p()
        check(L.pnb_create(int(device), ctypes.byref(h)))
        self.handle = h
        self.device = int(device)

    def close(self) -> None:
        if getattr(self, "handle", None):
            _lib.lib().pnb_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def check(self, status: int) -> None:
        check(status, self.handle)

    def set_bh_capturable(self, enable: bool = True) -> None:
        """Barnes-Hut without host round trips (pnb_bh_set_capturable): after one ordinary call
        at given sizes (and with the same mlp flag), device-pointer calls, plain or with the
        learned correction, can be captured in a graph and replayed."""
        self.check(_lib.lib().pnb_bh_set_capturable(self.handle, int(enable)))

    def bh_check(self) -> None:
        """Raises when the last capturable Barnes-Hut build did not fit its sizes (pnb_bh_check)."""
        self.check(_lib.lib().pnb_bh_check(self.handle))

    def last_error(self) -> str:
        """Message of the last failure on this context (pnb_last_error)."""
        return _lib.lib().pnb_last_error(self.handle).decode()

    def synchronize(self) -> None:
        self.check(_lib.lib().pnb_synchronize(self.handle))

    def set_stream(self, stream_handle: int | None) -> None:
        """Launch on the given hipStream_t (None or 0: HIP's null stream)."""
        self.check(_lib.lib().pnb_set_stream(self.handle, stream_handle or None))

    def reset_stream(self) -> None:
        """Launch on the context's own stream again."""
        self.check(_lib.lib().pnb_reset_stream(self.handle))

    def use_torch_stream(self) -> None:
        """Launch on torch's current stream, so device tensors need no extra synchronisation."""
        import torch

        self.set_stream(torch.cuda.current_stream(self.device).cuda_stream)

    def profile(self, enable: bool = True) -> None:
        self.check(_lib.lib().pnb_profile_enable(self.handle, int(enable)))

    def profile_read(self, kernel_id: int) -> tuple[float, int]:
        ms = ctypes.c_double()
        cnt = ctypes.c_uint64()
        self.check(_lib.lib().pnb_profile_read(self.handle, kernel_id, ctypes.byref(ms),
                                               ctypes.byref(cnt)))
        return ms.value, cnt.value

    def set_bf_tuning(self, n_min: int, variant: int, per_lane: int, j_split: int) -> None:
        self.check(_lib.lib().pnb_bf_set_tuning(self.handle, n_min, variant, per_lane, j_split))

    def clear_bf_tuning(self) -> None:
        self.check(_lib.lib().pnb_bf_clear_tuning(self.handle))


# ---- storage handling --------------------------------------------------------------------------

def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


def _as_host(x, allow_2d: bool = False) -> np.ndarray:
    if _is_torch(x):
        x = x.detach().cpu().numpy()
    a = np.ascontiguousarray(x, dtype=np.float32)
    if a.ndim != 2 or not (a.shape[1] == 4 or (allow_2d and a.shape[1] == 3)):
        raise ValueError("particles must be an (n, 4) float32 array {x, y, z, mu}"
                         + (" or (n, 3) {x, y, mu}" if allow_2d else ""))
    return a


def _check_dev(t):
    import torch

    if t.dtype != torch.float32 or t.dim() != 2 or t.shape[1] != 4 or not t.is_cuda:
        raise ValueError("device particles must be a CUDA float32 tensor of shape (n, 4)")
    return t.contiguous()


def _lower(storage):
    """Lower a storage to (kind, affected, affecting) following storage.rs:207-241."""
    if isinstance(storage, Between):
        return "between", storage.affected, storage.affecting
    if isinstance(storage, Reordered):
        if storage._fn is None:
            return "reordered", storage.unordered, None
        return "between", storage.unordered, storage.affecting()
    if isinstance(storage, Ordered):
        return "between", storage.particles(), storage.affecting()
    return "between", storage, storage


class _Algorithm:
    def __init__(self, resources: HipResources, interaction):
        if not isinstance(interaction, (Acceleration, AccelerationSoftened)):
            raise TypeError("interaction must be Acceleration or AccelerationSoftened")
        self.resources = resources
        self.interaction = interaction

    def _params(self):
        return float(self.interaction.softening), int(self.interaction.is_checked)

    def compute(self, storage):
        kind, affected, affecting = _lower(storage)
        if affected.shape[1] == 3:  # {x, y, mu}: two-dimensional particles
            return self._compute_2d(kind, affected, affecting)
        dev = _is_torch(affected) and affected.is_cuda
        if dev:
            import torch

            affected = _check_dev(affected)
            out = torch.empty((affected.shape[0], 4), dtype=torch.float32, device=affected.device)
            self.resources.use_torch_stream()
            if kind == "reordered":
                self._reordered_dev(affected, out)
            else:
                affecting = _check_dev(affecting)
                self._between_dev(affected, affecting, out)
            return out
        affected = _as_host(affected)
        out = np.zeros((affected.shape[0], 4), np.float32)
        if kind == "reordered":
            self._reordered_host(affected, out)
        else:
            self._between_host(affected, _as_host(affecting), out)
        return out[:, :3]

    def _compute_2d(self, kind, affected, affecting):
        """Two-dimensional particles {x, y, mu} (n x 3): accelerations (n x 2). The Reordered
        split (mu != 0) is taken here, then Between(affected, affecting)."""
        if _is_torch(affected) and affected.is_cuda:
            import torch

            a = affected.to(torch.float32).contiguous()
            b = a[a[:, 2] != 0].contiguous() if kind == "reordered" else \
                affecting.to(torch.float32).contiguous()
            out = torch.empty((a.shape[0], 2), dtype=torch.float32, device=a.device)
            self.resources.use_torch_stream()
            self._between_2d(a.data_ptr(), a.shape[0], b.data_ptr(), b.shape[0], out.data_ptr(),
                             dev=True)
            return out
        a = np.ascontiguousarray(affected, np.float32)
        if kind == "reordered":
            b = np.ascontiguousarray(a[a[:, 2] != 0])
        elif affecting is affected:
            b = a
        else:
            b = np.ascontiguousarray(affecting, np.float32)
        out = np.zeros((a.shape[0], 2), np.float32)
        self._between_2d(a.ctypes.data, a.shape[0], b.ctypes.data, b.shape[0], out.ctypes.data,
                         dev=False)
        return out


class BruteForce(_Algorithm):
    """Brute force on the GPU: replaces gpu::BruteForce (particular/src/gpu/mod.rs:149-208).

    Per affected particle, the sum over every affecting particle of the pair interaction, as
    sequential::BruteForce computes it (particular/src/sequential.rs:181-194). float64 host arrays
    take the FP64 kernel (the reference's DVec3 / f64 scalar path).
    """

    def compute(self, storage):
        kind, affected, affecting = _lower(storage)
        if ((not _is_torch(affected)) and np.asarray(affected).dtype == np.float64
                and np.asarray(affected).shape[1] == 4):
            a = np.ascontiguousarray(affected, np.float64)
            if kind == "reordered":
                b = np.ascontiguousarray(a[a[:, 3] != 0])
            else:
                b = np.ascontiguousarray(affecting, np.float64)
            out = np.zeros((a.shape[0], 4), np.float64)
            eps, chk = self._params()
            self.resources.check(_lib.lib().pnb_bf_accel_f64(
                self.resources.handle, a.ctypes.data, a.shape[0], b.ctypes.data, b.shape[0], eps,
                chk, out.ctypes.data))
            return out[:, :3]
        return super().compute(storage)

    def _between_host(self, a, b, out):
        eps, chk = self._params()
        self.resources.check(_lib.lib().pnb_bf_accel_f32(
            self.resources.handle, a.ctypes.data, a.shape[0], b.ctypes.data, b.shape[0], eps,
            chk, out.ctypes.data))

    def _reordered_host(self, a, out):
        eps, chk = self._params()
        self.resources.check(_lib.lib().pnb_reordered_bf_accel_f32(
            self.resources.handle, a.ctypes.data, a.shape[0], eps, chk, out.ctypes.data))

    def _between_dev(self, a, b, out):
        eps, chk = self._params()
        self.resources.check(_lib.lib().pnb_bf_accel_f32_dev(
            self.resources.handle, a.data_ptr(), a.shape[0], b.data_ptr(), b.shape[0], eps, chk,
            out.data_ptr()))

    def _reordered_dev(self, a, out):
        eps, chk = self._params()
        self.resources.check(_lib.lib().pnb_reordered_bf_accel_f32_dev(
            self.resources.handle, a.data_ptr(), a.shape[0], eps, chk, out.data_ptr()))

    def _between_2d(self, a, n, b, m, out, dev):
        eps, chk = self._params()
        fn = _lib.lib().pnb_bf_accel_f32_2d_dev if dev else _lib.lib().pnb_bf_accel_f32_2d
        self.resources.check(fn(self.resources.handle, a, n, b, m, eps, chk, out))


class BruteForcePairs(_Algorithm):
    """Brute force over pairs on the GPU: sequential::BruteForcePairs
    (particular/src/sequential.rs:275-305). One particle set, every pair evaluated once for both
    accelerations (Newton's third law). Storages as the reference takes them: a slice (an (n, 4)
    array or CUDA tensor), Ordered and Reordered, whose sums are those over the whole set (pairs
    of two massless particles add zero). Between has no pair form in the reference."""

    def compute(self, storage):
        if isinstance(storage, Between):
            raise TypeError("BruteForcePairs computes one particle set (a slice, Ordered or "
                            "Reordered), not Between")
        if isinstance(storage, Reordered):
            particles = storage.unordered
        elif isinstance(storage, Ordered):
            particles = storage.particles()
        else:
            particles = storage
        eps, chk = self._params()
        if _is_torch(particles) and particles.is_cuda:
            import torch

            a = _check_dev(particles)
            out = torch.empty((a.shape[0], 4), dtype=torch.float32, device=a.device)
            self.resources.use_torch_stream()
            self.resources.check(_lib.lib().pnb_bf_pairs_accel_f32_dev(
                self.resources.handle, a.data_ptr(), a.shape[0], eps, chk, out.data_ptr()))
            return out
        a = _as_host(particles)
        out = np.zeros((a.shape[0], 4), np.float32)
        self.resources.check(_lib.lib().pnb_bf_pairs_accel_f32(
            self.resources.handle, a.ctypes.data, a.shape[0], eps, chk, out.ctypes.data))
        return out[:, :3]


class BarnesHut(_Algorithm):
    """Barnes-Hut on the GPU: replaces parallel::BarnesHut (particular/src/parallel.rs:298-367).

    The octree is built from the affecting particles node for node as the reference builds it
    (FP64 centres of mass); each affected particle applies the reference's traversal rule
    (particular/src/sequential.rs:465-505). ``theta = 0`` degenerates to brute force.
    """

    def __init__(self, resources: HipResources, theta: float, interaction, mlp: bool = False,
                 reference_com: bool = False, relative: float | None = None,
                 estimate_theta: float = 1.2, estimate_per_particle: bool = False,
                 bucket: int = 24):
        """``relative=alpha`` adds the relative opening criterion (no reference counterpart): a
        node is also opened when its first neglected term, M w^p / r^(p+2) (p = 2; with ``mlp``
        3, or 4 when the network has octupole outputs), exceeds alpha times the acceleration from an estimate pass at
        ``estimate_theta`` (at the centroid of the particle's wavefront, or at the particle with
        ``estimate_per_particle``); ``theta`` then acts as a geometric floor, and an opened node
        of at most ``bucket`` bodies is summed body by body."""
        super().__init__(resources, interaction)
        self.theta = float(theta)
        self.mlp = bool(mlp)
        self.reference_com = bool(reference_com)
        self.relative = None if relative is None else float(relative)
        self.estimate_theta = float(estimate_theta)
        self.estimate_per_particle = bool(estimate_per_particle)
        self.bucket = int(bucket)

    def _flags(self) -> int:
        f = (_lib.BH_MLP if self.mlp else 0) | (_lib.BH_REF_COM if self.reference_com else 0)
        if self.relative is not None:
            self.resources.check(_lib.lib().pnb_bh_set_relative(
                self.resources.handle, self.relative, self.estimate_theta,
                int(self.estimate_per_particle), self.bucket))
            f |= _lib.BH_RELATIVE
        return f

    def _between_host(self, a, b, out):
        # Between(p, p) over one array passes one pointer twice; the library then uploads one
        # copy and reuses the tree order for the affected particles.
        eps, chk = self._params()
        self.resources.check(_lib.lib().pnb_bh_accel_f32(
            self.resources.handle, a.ctypes.data, a.shape[0], b.ctypes.data, b.shape[0],
            self.theta, eps, chk, self._flags(), out.ctypes.data))

    def _reordered_host(self, a, out):
        eps, chk = self._params()
        self.resources.check(_lib.lib().pnb_reordered_bh_accel_f32(
            self.resources.handle, a.ctypes.data, a.shape[0], self.theta, eps, chk, self._flags(),
            out.ctypes.data))

    def _between_dev(self, a, b, out):
        eps, chk = self._params()
        self.resources.check(_lib.lib().pnb_bh_accel_f32_dev(
            self.resources.handle, a.data_ptr(), a.shape[0], b.data_ptr(), b.shape[0],
            self.theta, eps, chk, self._flags(), out.data_ptr()))

    def _reordered_dev(self, a, out):
        eps, chk = self._params()
        self.resources.check(_lib.lib().pnb_reordered_bh_accel_f32_dev(
            self.resources.handle, a.data_ptr(), a.shape[0], self.theta, eps, chk,
            self._flags(), out.data_ptr()))

    def _between_2d(self, a, n, b, m, out, dev):
        eps, chk = self._params()
        fn = _lib.lib().pnb_bh_accel_f32_2d_dev if dev else _lib.lib().pnb_bh_accel_f32_2d
        self.resources.check(fn(self.resources.handle, a, n, b, m, self.theta, eps, chk,
                                self._flags(), out))


def tree_export(resources: HipResources, m: int | None = None):
    """Nodes of the last octree built on `resources` (pre-order) and the tree particle order."""
    L = _lib.lib()
    nodes = ctypes.c_size_t()
    lvl = ctypes.c_int()
    resources.check(L.pnb_bh_tree_size(resources.handle, ctypes.byref(nodes), ctypes.byref(lvl)))
    arr = np.zeros(nodes.value, _lib.tree_node_dtype())
    order = np.zeros(m if m is not None else 0, np.uint32)
    resources.check(L.pnb_bh_tree_export(resources.handle, arr.ctypes.data, arr.shape[0],
                                         order.ctypes.data if order.size else None, order.size))
    return arr, order


def bh_stats(resources: HipResources) -> tuple[int, int]:
    """(wavefront node steps, accepted pair terms) of the walks since the last build."""
    a = ctypes.c_uint64()
    b = ctypes.c_uint64()
    resources.check(_lib.lib().pnb_bh_stats(resources.handle, ctypes.byref(a), ctypes.byref(b)))
    return a.value, b.value


def _bh_available() -> bool:
    return hasattr(_lib.lib(), "pnb_bh_accel_f32")


def hip_brute_force(storage, resources: HipResources, interaction):
    """GpuCompute::gpu_brute_force sugar (particular/src/gpu/mod.rs:19-30)."""
    return BruteForce(resources, interaction).compute(storage)


def hip_barnes_hut(storage, resources: HipResources, theta: float, interaction):
    return BarnesHut(resources, theta, interaction).compute(storage)
