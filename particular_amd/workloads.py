"""Seeded synthetic inputs for the acceleration path.

Every generator draws from numpy's Philox counter-based generator, so a given seed produces
bit-identical float32 particle arrays on any machine; the GPU path and the CPU baseline always see
the same inputs. Particles are returned as ``(n, 4)`` float32 arrays ``{x, y, z, mu}`` -- the
16-byte ``GravitationalField<Vec3, f32>`` record of particular/src/gravity/mod.rs:12-18 -- and,
for the step loops, FP64 positions and velocities.

Plummer spheres follow Aarseth, Henon & Wielen (1974, A&A 37, 183): radius from the inverted
cumulative mass profile, isotropic directions, speeds by von Neumann rejection from
g(q) = q^2 (1 - q^2)^{7/2}, in units G = 1.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

DEFAULT_SEED = 1808  # the reference benchmark's seed (particular/benches/benchmark.rs:101)
MASS_FRACTION_CAP = 0.999  # truncates the Plummer sphere at r = 38.7 a


def _rng(seed: int, stream: int = 0) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(key=[seed & 0xFFFFFFFFFFFFFFFF, stream]))


def _isotropic(rng: np.random.Generator, n: int) -> np.ndarray:
    cos_t = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * np.pi, n)
    sin_t = np.sqrt(np.maximum(0.0, 1.0 - cos_t * cos_t))
    return np.stack([sin_t * np.cos(phi), sin_t * np.sin(phi), cos_t], axis=1)


def plummer(n: int, seed: int = DEFAULT_SEED, scale: float = 1.0, total_mass: float = 1.0,
            stream: int = 0, cap: float = MASS_FRACTION_CAP):
    """Plummer sphere of ``n`` equal-mass bodies centred at the origin.

    Returns ``(pos, vel)`` as FP64 ``(n, 3)`` arrays (G = 1, total mass ``total_mass``, scale
    radius ``scale``).
    """
    rng = _rng(seed, stream)
    x1 = rng.uniform(0.0, cap, n)
    x1 = np.maximum(x1, 1e-300)
    r = scale / np.sqrt(x1 ** (-2.0 / 3.0) - 1.0)
    pos = _isotropic(rng, n) * r[:, None]
    q = np.empty(n)
    todo = np.arange(n)
    while todo.size:
        qq = rng.uniform(0.0, 1.0, todo.size)
        yy = rng.uniform(0.0, 0.1, todo.size)
        ok = yy < qq * qq * (1.0 - qq * qq) ** 3.5
        q[todo[ok]] = qq[ok]
        todo = todo[~ok]
    v_esc = np.sqrt(2.0 * total_mass / scale) * (1.0 + (r / scale) ** 2) ** -0.25
    vel = _isotropic(rng, n) * (q * v_esc)[:, None]
    return pos, vel


def uniform_cube(n: int, side: float, seed: int = DEFAULT_SEED, stream: int = 1) -> np.ndarray:
    rng = _rng(seed, stream)
    return rng.uniform(-0.5 * side, 0.5 * side, (n, 3))


def pack(pos, mu) -> np.ndarray:
    """(n, 3) positions + per-body mu -> (n, 4) float32 {x, y, z, mu}."""
    pos = np.asarray(pos)
    out = np.empty((pos.shape[0], 4), np.float32)
    out[:, :3] = pos.astype(np.float32)
    out[:, 3] = np.broadcast_to(np.asarray(mu, np.float32), (pos.shape[0],))
    return out


@dataclass
class System:
    """A particle set: float32 records plus the FP64 state the step loops integrate."""

    particles: np.ndarray  # (n, 4) float32 {x, y, z, mu}
    pos: np.ndarray  # (n, 3) float64
    vel: np.ndarray  # (n, 3) float64
    softening: float = 0.0

    @property
    def n(self) -> int:
        return self.particles.shape[0]


def direct_sum_case(n: int = 4096, seed: int = DEFAULT_SEED) -> System:
    """Plain direct sum: n Plummer bodies (scale 1), mu = 1/n, all massive, no softening."""
    pos, vel = plummer(n, seed)
    return System(pack(pos, 1.0 / n), pos, vel, 0.0)


def massive_massless_case(n_massive: int = 65536, n_total: int = 1048576,
                          seed: int = DEFAULT_SEED, side: float = 4.0,
                          softening: float = 1e-3) -> System:
    """Direct sum at scale: a Plummer sphere of ``n_massive`` bodies (mu = 1/n_massive) and
    ``n_total - n_massive`` massless bodies uniform in a cube of side ``side``, interleaved by a
    seeded permutation so the massive/massless split has to be done by the storage."""
    n_light = n_total - n_massive
    pm, vm = plummer(n_massive, seed)
    pl = uniform_cube(n_light, side, seed)
    pos = np.concatenate([pm, pl])
    vel = np.concatenate([vm, np.zeros((n_light, 3))])
    mu = np.concatenate([np.full(n_massive, 1.0 / n_massive), np.zeros(n_light)])
    perm = _rng(seed, 7).permutation(n_total)
    pos, vel, mu = pos[perm], vel[perm], mu[perm]
    return System(pack(pos, mu), pos, vel, softening)


def two_plummer_case(n: int = 4194304, seed: int = DEFAULT_SEED, separation: float = 10.0,
                     bulk_speed: float = 0.3) -> System:
    """Two Plummer spheres of n/2 bodies each (scale 1, total mass 1/2 each, mu = 1/n), centred
    at x = -separation/2 and +separation/2, approaching along x at +-bulk_speed."""
    h = n // 2
    pa, va = plummer(h, seed, total_mass=0.5, stream=0)
    pb, vb = plummer(n - h, seed, total_mass=0.5, stream=2)
    pa[:, 0] -= 0.5 * separation
    pb[:, 0] += 0.5 * separation
    va[:, 0] += bulk_speed
    vb[:, 0] -= bulk_speed
    pos = np.concatenate([pa, pb])
    vel = np.concatenate([va, vb])
    return System(pack(pos, 1.0 / n), pos, vel, 0.0)


@dataclass
class Ensemble:
    """Independent systems stored back to back; system s owns [offsets[s], offsets[s+1])."""

    particles: np.ndarray
    pos: np.ndarray
    vel: np.ndarray
    offsets: np.ndarray  # int64, num_systems + 1
    softening: float

    @property
    def num_systems(self) -> int:
        return self.offsets.shape[0] - 1


def ensemble_case(num_systems: int = 2048, n_min: int = 64, n_max: int = 65536,
                  seed: int = DEFAULT_SEED, softening: float = 1e-3) -> Ensemble:
    """Batched ensemble: sizes log-uniform in [n_min, n_max], each a Plummer sphere with scale
    radius uniform in [0.5, 2] and mu = 1/N_s."""
    rng = _rng(seed, 11)
    sizes = np.exp(rng.uniform(np.log(n_min), np.log(n_max), num_systems)).astype(np.int64)
    sizes = np.clip(sizes, n_min, n_max)
    scales = rng.uniform(0.5, 2.0, num_systems)
    offsets = np.zeros(num_systems + 1, np.int64)
    offsets[1:] = np.cumsum(sizes)
    total = int(offsets[-1])
    pos = np.empty((total, 3))
    vel = np.empty((total, 3))
    mu = np.empty(total)
    for s in range(num_systems):
        a, b = offsets[s], offsets[s + 1]
        p, v = plummer(int(b - a), seed, scale=float(scales[s]), stream=1000 + s)
        pos[a:b], vel[a:b], mu[a:b] = p, v, 1.0 / float(b - a)
    return Ensemble(pack(pos, mu), pos, vel, offsets, softening)
