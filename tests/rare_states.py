"""States that drive the face kernel's rare branches (include/mhd/kernels/muscl_hll.hpp):

* negative_pressure_patch: the Orszag-Tang state with a block of cells whose total energy is
  below their kinetic plus magnetic energy (p < 0). Every face touching the block has a
  non-positive reconstructed pressure on some side, so both sides revert to the cell values
  (the positivity / finiteness fallback, :145-168).
* cold_slow_flow: a smooth flow with speeds of ~1e-16 (p ~ 1e-32, B = 0): the global signal
  speed c_h and with it every face's wave-speed span sR - sL stay below 1e-14, with sL < 0 < sR,
  so every face takes the degenerate-span average (FL + FR) / 2 (:196-205). The flow is a
  slowed-down sound wave, finite and smooth for many steps.
Both are returned as full (9, ny + 4, nx + 4) float64 states with the ghosts filled by the
given boundary filler (oracle.apply_bc)."""
import math

import numpy as np

GAMMA = 5.0 / 3.0


def _conserved(rho, vx, vy, vz, p, bx, by, bz, gamma=GAMMA):
    e_kin = 0.5 * rho * (vx * vx + vy * vy + vz * vz)
    e_mag = 0.5 * (bx * bx + by * by + bz * bz)
    return np.stack([rho, rho * vx, rho * vy, rho * vz, p / (gamma - 1.0) + e_kin + e_mag,
                     bx, by, bz, np.zeros_like(rho)])


def negative_pressure_patch(oracle, g, rows=slice(15, 22), cols=slice(18, 27)):
    U = oracle.init_ic("orszag_tang", g)
    u = U[:, 2:-2, 2:-2]
    rho = u[0]
    e_kin = 0.5 * (u[1] ** 2 + u[2] ** 2 + u[3] ** 2) / rho
    e_mag = 0.5 * (u[5] ** 2 + u[6] ** 2 + u[7] ** 2)
    p = (GAMMA - 1.0) * (u[4] - e_kin - e_mag)
    # p -> -0.3 p inside the block
    u[4][rows, cols] = (-0.3 * p[rows, cols]) / (GAMMA - 1.0) + e_kin[rows, cols] + \
        e_mag[rows, cols]
    oracle.apply_bc(g, U, (0, 0))
    return U


def cold_slow_flow(oracle, g):
    x = (np.arange(g.nx) + 0.5) / g.nx
    y = (np.arange(g.ny) + 0.5) / g.ny
    X, Y = np.meshgrid(x, y)
    s = np.sin(2 * math.pi * X) * np.cos(2 * math.pi * Y)
    rho = 1.0 + 0.1 * s
    vx = 1e-18 * np.sin(2 * math.pi * Y)
    vy = -1e-18 * np.cos(2 * math.pi * X)
    p = 1e-32 * (1.0 + 0.5 * np.cos(2 * math.pi * (X + Y)))
    z = np.zeros_like(rho)
    U = np.zeros((9, g.ny + 4, g.nx + 4))
    U[:, 2:-2, 2:-2] = _conserved(rho, vx, vy, z, p, z, z, z)
    oracle.apply_bc(g, U, (0, 0))
    return U
