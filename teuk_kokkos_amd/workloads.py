"""Synthetic workloads as schema-v1 configuration texts.

The five standard cases of this project:
  1. linear Schwarzschild, nr=257 D8-4, m=+-2, ell<=10
  2. linear Kerr a=0.7 ringdown, nr=1025 D10-5, ell<=16
  3. batched sweep of 512 independent Kerr evolutions (uneven nr, ellmax,
     spin, m), D10-5, CFL 0.5
  4. near-extremal a=0.999, nr=4097 D10-5, ell<=32
  5. second order: 64 instances, parents m=+-2 (ell<=8), targets {-4,0,4}
     (ell<=12), nr=1025 D10-5, all 13 fields live, unrestricted source
Values the case descriptions leave open come from the reference templates
(configs/*.cfg: L, reduction damping, free-damped reduction).  Draws use a
SplitMix64 stream with Box-Muller normals so every host produces identical
inputs.
"""
from __future__ import annotations

import math

import numpy as np

MASK = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014) with uniform and Box-Muller draws."""

    def __init__(self, seed: int):
        self.state = seed & MASK

    def next(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next() >> 11) * (1.0 / (1 << 53))

    def choice(self, values):
        return values[min(len(values) - 1, int(self.uniform() * len(values)))]

    def normal(self) -> float:
        u1 = max(self.uniform(), 1e-300)
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def seed_for(case: int, instance: int) -> int:
    return 0x7E0C0000 + case * 1000 + instance


def minimum_ntheta(ell_first: int, ell_second: int) -> int:
    """Smallest ntheta the schema accepts (config.hpp:457-473)."""
    return max(ell_first + 1, ell_second + 1, (2 * ell_first + ell_second + 2) // 2)


def _gauss_legendre_x(n: int) -> np.ndarray:
    x, _ = np.polynomial.legendre.leggauss(n)
    return x


def maximum_coordinate_speed(mass, spin, length, nr, ntheta) -> float:
    """max |coordinate velocity| of the radial principal part over the grid
    (solver_driver.hpp:36-57, boundary.hpp:23-44)."""
    rp = mass + math.sqrt(max(mass * mass - spin * spin, 0.0))
    horizon = length * length / rp
    R = np.arange(nr) * (horizon / (nr - 1))
    x = _gauss_legendre_x(ntheta)
    sin2 = 1.0 - x * x
    L2 = length * length
    L4 = L2 * L2
    a2 = spin * spin
    mrl = mass * R / L2
    time_r = 8.0 * mass * (2.0 * mass - a2 * R / L2) * (1.0 + 2.0 * mrl)
    adv = L2 - (8.0 * mass * mass - a2) * R * R / L2 + 4.0 * a2 * mass * R ** 3 / L4
    principal = R * R / L4 * (L4 - 2.0 * L2 * mass * R + a2 * R * R)
    ct = time_r[:, None] - a2 * sin2[None, :]
    root = np.sqrt(adv[:, None] ** 2 + ct * principal[:, None])
    lm = (adv[:, None] - root) / ct
    lp = (adv[:, None] + root) / ct
    return float(max(np.abs(lm).max(), np.abs(lp).max()))


def cfl_time_step(mass, spin, length, nr, ntheta, cfl) -> float:
    rp = mass + math.sqrt(max(mass * mass - spin * spin, 0.0))
    h = length * length / rp / (nr - 1)
    return cfl * h / maximum_coordinate_speed(mass, spin, length, nr, ntheta)


def config_text(**kw) -> str:
    """Schema-v1 text from keyword values; `seeds` is a list of
    (ell, m, complex amplitude); dotted keys use double underscores."""
    seeds = kw.pop("seeds", None)
    lines = ["config_version = 1"]
    for key, value in kw.items():
        key = key.replace("__", ".")
        if isinstance(value, (list, tuple)):
            value = ",".join(str(v) for v in value)
        elif isinstance(value, bool):
            value = "true" if value else "false"
        elif isinstance(value, float):
            value = repr(value)
        lines.append(f"{key} = {value}")
    if seeds:
        for k, (ell, m, amp) in enumerate(seeds):
            amp = complex(amp)
            lines += [f"initial_data.mode.{k}.ell = {ell}", f"initial_data.mode.{k}.m = {m}",
                      f"initial_data.mode.{k}.amplitude_real = {amp.real!r}",
                      f"initial_data.mode.{k}.amplitude_imag = {amp.imag!r}"]
    return "\n".join(lines) + "\n"


def _linear(mass, spin, length, nr, ell, scheme, modes, seeds, steps, cfl, dissipation,
            center, width, add_sharp=False, damping=0.1):
    nt = minimum_ntheta(ell, ell)
    dt = cfl_time_step(mass, spin, length, nr, nt, cfl) * (1.0 - 1e-9)
    return dict(mass=mass, spin=spin, compactification_length=length, nr=nr, ntheta=nt,
                ellmax_first=ell, ellmax_second=ell, first_order_modes=modes,
                second_order_modes=modes, final_time=dt * steps, steps=steps, cfl=cfl,
                reduction_damping=damping, dissipation=dissipation,
                radial_discretization=scheme, reduction_mode="free_damped",
                initial_data__center=center, initial_data__width=width,
                initial_data__add_sharp_partner=add_sharp, seeds=seeds)


def case1(cfl=0.02, dissipation=0.0) -> dict:
    """Linear Schwarzschild, nr=257 D8-4, m=+-2 band ell=2..10, pulse in (2,2).

    The D8-4 closure's large boundary eigenvalue makes RK4 blow up at CFL 0.5
    (and 0.05); CFL 0.02 is the default that stays bounded."""
    return _linear(1.0, 0.0, 2.0, 257, 10, "d8-4", [-2, 2], [(2, 2, 1.0e-3)], 2000, cfl,
                   dissipation, 0.5, 0.05)


def case2(cfl=0.5, dissipation=0.0) -> dict:
    """Kerr a=0.7 ringdown, nr=1025 D10-5, ell=2..16, pulses in (2,2) and (3,2)."""
    rng = SplitMix64(seed_for(2, 0))
    seeds = [(ell, 2, complex(1e-3 * rng.normal(), 1e-3 * rng.normal())) for ell in (2, 3)]
    return _linear(1.0, 0.7, 2.0, 1025, 16, "d10-5", [-2, 2], seeds, 20000, cfl, dissipation,
                   0.5, 0.05)


def case4(cfl=0.5, dissipation=0.0) -> dict:
    """Near-extremal a=0.999, L=1, nr=4097 D10-5, ell=2..32."""
    return _linear(1.0, 0.999, 1.0, 4097, 32, "d10-5", [-2, 2], [(2, 2, 1.0e-3)], 100000, cfl,
                   dissipation, 0.5, 0.05)


def sweep_instance(k: int, cfl=0.5) -> dict:
    """Instance k of the 512-evolution Kerr sweep."""
    rng = SplitMix64(seed_for(3, k))
    spin = 0.99 * rng.uniform()
    m = rng.choice(list(range(-4, 5)))
    ell = rng.choice([8, 12, 16, 24])
    nr = rng.choice([257, 513, 1025, 2049])
    amp = complex(1e-3 * rng.normal(), 1e-3 * rng.normal())
    center = 0.2 + 0.6 * rng.uniform()
    width = 0.03 + 0.07 * rng.uniform()
    modes = [0] if m == 0 else [-abs(m), abs(m)]
    seed_ell = max(2, abs(m))
    return _linear(1.0, spin, 2.0, nr, ell, "d10-5", modes, [(seed_ell, m, amp)], 20000, cfl, 0.0,
                   center, width, add_sharp=(m != 0))


def second_order_instance(k: int, cfl=0.5, steps=200) -> dict:
    """Instance k of the 64-instance second-order self-coupling batch:
    a ~ U[0, 0.95], L = 1.5, parents m = +-2 (ell <= 8), targets
    {-4, 0, 4} (ell <= 12), N_theta = 15, nr = 1025 D10-5, (2, +-2) pulses
    with sharp partners, the unrestricted source switched on at 10 dt
    (configs/second_order_22_self_coupling.cfg for L, damping, center and
    width)."""
    rng = SplitMix64(seed_for(5, k))
    spin = 0.95 * rng.uniform()
    nt = minimum_ntheta(8, 12)
    nr = 1025
    dt = cfl_time_step(1.0, spin, 1.5, nr, nt, cfl) * (1.0 - 1e-9)
    return dict(mass=1.0, spin=spin, compactification_length=1.5, nr=nr, ntheta=nt,
                ellmax_first=8, ellmax_second=12, first_order_modes=[-2, 2],
                second_order_modes=[-4, 0, 4], final_time=dt * steps, steps=steps, cfl=cfl,
                reduction_damping=0.1, dissipation=0.0, radial_discretization="d10-5",
                reduction_mode="free_damped", initial_data__center=0.45,
                initial_data__width=0.12, initial_data__add_sharp_partner=True,
                second_order__enabled=True, second_order__source_mode="unrestricted",
                second_order__source_start_time=10.0 * dt, seeds=[(2, 2, 1.0e-3)])


def second_order_batch(n: int = 64, cfl=0.5):
    return [second_order_instance(k, cfl) for k in range(n)]


def sweep(n: int = 512, cfl=0.5):
    return [sweep_instance(k, cfl) for k in range(n)]


def texts(cases):
    return [config_text(**dict(c)) for c in cases]
