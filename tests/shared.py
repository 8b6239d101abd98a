"""Helpers shared by the test modules (not collected)."""
import numpy as np

def kat_particles():
    """The `acceleration_error!` data of particular/src/gravity/newtonian/mod.rs:230-242."""
    return np.array([[10, 10, 10, 0], [0, 0, 0, 20], [1, 1, 1, 30], [-3, -3, -3, 40],
                     [30, 30, 30, 0], [-45, -45, -45, 0]], np.float32)


def circular_orbit(step_accel, orbits, dt=1.0 / 60.0):
    """circular_orbit! (particular/src/gravity/newtonian/mod.rs:281-347): mu = 1e6 at the origin,
    a massless body at x = 100 with v_y = 100; symplectic Euler; returns (distance, energy) errors."""
    p = np.array([[0, 0, 0, 1e6], [100, 0, 0, 0]], np.float32)
    v = np.array([[0, 0, 0], [0, 100, 0]], np.float32)

    def state(p):
        r = float(np.linalg.norm((p[0, :3] - p[1, :3]).astype(np.float32)))
        return r, -(p[0, 3] + p[1, 3]) / (r + r)

    r0, e0 = state(p)
    steps = int(round(2 * np.pi * np.sqrt(r0**3 / 1e6) / dt))
    for _ in range(steps * orbits):
        a = step_accel(p)
        v = v + a[:, :3] * np.float32(dt)
        p[:, :3] = p[:, :3] + v * np.float32(dt)
    r1, e1 = state(p)
    return abs(1 - r0 / r1), abs(1 - e0 / e1)
