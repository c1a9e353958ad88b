"""Seeded synthetic workloads: the mixed-size ensemble and the request stream.

All randomness comes from numpy's PCG64 generator (numpy.random.default_rng(seed)), so a seed
reproduces the same instances on any machine.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from .solver import init_ic

B0_STD = 1.0 / math.sqrt(4.0 * math.pi)  # Orszag-Tang field normalisation B0 = 1/sqrt(4 pi)
PERIODIC, OUTFLOW = 0, 2


@dataclasses.dataclass
class Spec:
    """One solve: initial condition, grid, gamma, field scale, boundaries, end time (and the
    step limit past which the solve fails, as the reference's solve() does)."""
    ic: str
    n: int
    gamma: float
    b0: float
    bc: tuple
    t_end: float = math.inf
    seed: int = 0
    arrival: float = 0.0
    max_steps: int = 1_000_000
    limiter: object = None   # SlopeLimiter; None: the engine's
    nonideal: object = None  # NonIdealConfig; None: the engine's
    mask: object = None      # (n, n) uint8 FluidMask (1 = fluid); None: all fluid
    walls: bool = True       # reflective wall fill of the mask's solid cells

    def state(self):
        return init_ic(self.ic, self.n, self.n, gamma=self.gamma, b0=self.b0, seed=self.seed)


def ensemble(count=512, seed=2024, sizes=(64, 96, 128, 192, 256)):
    """Independent Orszag-Tang / rotor instances on square grids drawn uniformly from `sizes`;
    gamma ~ U[1.4, 5/3]; the field scale ~ U[0.5, 2] multiplies the Orszag-Tang B0 = 1/sqrt(4 pi)
    or the rotor's Bx = 5/sqrt(4 pi). Orszag-Tang runs periodic, the rotor with outflow."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        ot = bool(rng.integers(0, 2) == 0)
        n = int(rng.choice(sizes))
        gamma = float(rng.uniform(1.4, 5.0 / 3.0))
        scale = float(rng.uniform(0.5, 2.0))
        if ot:
            out.append(Spec("orszag_tang_b0", n, gamma, scale * B0_STD, (PERIODIC, PERIODIC)))
        else:
            out.append(Spec("rotor_b0", n, gamma, scale, (OUTFLOW, OUTFLOW)))
    return out


def request_stream(count, rate=200.0, seed=7, n_min=64, n_max=1024, t_min=0.05, t_max=0.5):
    """Poisson arrivals (mean `rate` per second) of solves mixing the oblique circularly
    polarised Alfven wave, the tanh Kelvin-Helmholtz layer (seeded vy noise) and the rotor;
    square grids N ~ U{n_min..n_max}; t_end ~ U[t_min, t_max]."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(count):
        t += float(rng.exponential(1.0 / rate))
        kind = int(rng.integers(0, 3))
        n = int(rng.integers(n_min, n_max + 1))
        t_end = float(rng.uniform(t_min, t_max))
        if kind == 0:
            s = Spec("alfven_oblique", n, 5.0 / 3.0, 1.0, (PERIODIC, PERIODIC), t_end)
        elif kind == 1:
            s = Spec("kh_tanh", n, 5.0 / 3.0, 1.0, (PERIODIC, PERIODIC), t_end,
                     seed=int(rng.integers(0, 2 ** 31)))
        else:
            s = Spec("rotor", n, 5.0 / 3.0, 1.0, (OUTFLOW, OUTFLOW), t_end)
        s.arrival = t
        out.append(s)
    return out


def circles_mask(n, count=8, seed=11, r_min=0.02, r_max=0.06):
    """uint8 (n, n) interior mask on [0,1]^2, 1 = fluid, with `count` solid circles of radius
    ~ U[r_min, r_max] and seeded centres, rasterised at cell centres with
    (x - cx)^2 + (y - cy)^2 <= r^2 (the reference GUI's rule, gui/geometry.py:29-38)."""
    rng = np.random.default_rng(seed)
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x)
    mask = np.ones((n, n), dtype=np.uint8)
    for _ in range(count):
        r = rng.uniform(r_min, r_max)
        cx, cy = rng.uniform(r, 1.0 - r, size=2)
        mask[(X - cx) ** 2 + (Y - cy) ** 2 <= r * r] = 0
    return mask


def rotor_circles(n, seed=11, t_end=math.inf, eta_ohm=1e-3, eta_hall=1e-2, eta_ambipolar=1e-3):
    """The non-ideal + geometry solve as one request: the rotor with Ohmic, Hall and ambipolar
    terms, outflow boundaries, minmod, and a mask of 8 seeded solid circles acting as
    reflective walls."""
    from .solver import NonIdealConfig, SlopeLimiter
    ni = NonIdealConfig(ohmic=True, hall=True, ambipolar=True, eta_ohm=eta_ohm,
                        eta_hall=eta_hall, eta_ambipolar=eta_ambipolar)
    return Spec("rotor", n, 5.0 / 3.0, 1.0, (OUTFLOW, OUTFLOW), t_end, limiter=SlopeLimiter.Minmod,
                nonideal=ni, mask=circles_mask(n, 8, seed=seed), walls=True)
