"""Host-side pins of the sampled diagnostics: the horizon stencil the GPU
diagnostics are compared against differentiates every polynomial of degree
<= 8 exactly (horizon_diagnostics.hpp:18-79)."""
from __future__ import annotations

import math

import helpers as H


def test_horizon_weights_are_exact_on_degree_eight_polynomials():
    h = 0.013
    R = [1.7 - h * j for j in range(9)]  # R_H, R_H - h, ...
    for d in range(5):
        w = H.horizon_weights(d, h)
        for k in range(9):
            exact = math.factorial(k) / math.factorial(k - d) * 1.7 ** (k - d) if k >= d else 0.0
            got = sum(w[j] * R[j] ** k for j in range(9))
            # rounding bound: the terms' magnitudes, not the result, set it
            mag = sum(abs(w[j]) * abs(R[j]) ** k for j in range(9))
            assert abs(got - exact) <= 1e-12 * mag, (d, k)
