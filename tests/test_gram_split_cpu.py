"""Error bound of the split Gram projection the stage kernels use for bands
with ell_max > 16: k = G X + M Psi (G = A S, M = G Lambda G) formed as
[X + Lambda Psi] in FP64 plus [E X + N Psi] (E = G - I, N = M - Lambda) with
FP32 operands, coefficients and sums.  Checked here in numpy against the
product in long double, on the oracle's band matrices and random rows whose
magnitudes span many decades: the split form must be as accurate as the plain
FP64 product, and dropping the correction must not be (so a kernel that lost
it would show)."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from teuk_kokkos_amd import workloads as W

EPS = 2.0 ** -53


def _bands(m, ell_max):
    nt = W.minimum_ntheta(ell_max, ell_max)
    S, A = oracle.transform_matrices(-2, m, ell_max, nt)
    ell_min = max(2, abs(m))
    lam = np.array([-(l + 2) * (l - 1) for l in range(ell_min, ell_max + 1)], dtype=float)
    return S, A, lam


def _rows(rng, rows, nl):
    scale = 10.0 ** rng.uniform(-12, 3, size=(rows, 1))
    return scale * (rng.standard_normal((rows, nl)) + 1j * rng.standard_normal((rows, nl)))


def _f32_product(mat32, x):
    """mat32 @ x with x rounded to FP32 and the sum accumulated in FP32."""
    re = (x.real.astype(np.float32) @ mat32.T).astype(np.float64)
    im = (x.imag.astype(np.float32) @ mat32.T).astype(np.float64)
    return re + 1j * im


@pytest.mark.parametrize("m,ell_max", [(2, 24), (0, 24), (-3, 24), (2, 32)])
def test_split_gram_projection_is_as_accurate_as_fp64(m, ell_max):
    S, A, lam = _bands(m, ell_max)
    nl = lam.size
    G = A @ S
    gl = G.astype(np.longdouble)
    Ml = gl @ np.diag(lam.astype(np.longdouble)) @ gl
    M = Ml.astype(np.float64)
    E32 = (gl - np.eye(nl, dtype=np.longdouble)).astype(np.float32)
    N32 = (Ml - np.diag(lam.astype(np.longdouble))).astype(np.float32)
    rng = np.random.default_rng(1000 * ell_max + m + 7)
    X = _rows(rng, 64, nl)
    Psi = _rows(rng, 64, nl)

    def exact():
        xr, xi = X.real.astype(np.longdouble), X.imag.astype(np.longdouble)
        pr, pi = Psi.real.astype(np.longdouble), Psi.imag.astype(np.longdouble)
        re = xr @ gl.T + pr @ Ml.T
        im = xi @ gl.T + pi @ Ml.T
        return re, im

    ere, eim = exact()
    k_exact = ere.astype(np.float64) + 1j * eim.astype(np.float64)
    # per-row scale: the size of the terms, which sets every form's rounding
    row_scale = np.abs(X).max(axis=1) + np.abs(lam)[None, :].max() * np.abs(Psi).max(axis=1)

    def err(k):
        return float((np.abs(k - k_exact).max(axis=1) / row_scale).max())

    k64 = X @ G.T + Psi @ M.T
    identity = X + Psi * lam[None, :]
    k_split = identity + _f32_product(E32, X) + _f32_product(N32, Psi)
    e64, esplit, enone = err(k64), err(k_split), err(identity)
    print(f"[cpu] ell_max {ell_max} m {m}: FP64 product {e64:.2e}, split {esplit:.2e}, "
          f"identity only {enone:.2e} (units of the row's term size)")
    assert esplit <= 8 * EPS
    assert esplit <= 2.0 * max(e64, EPS)
    # the correction is what the reference's A S carries: without it the
    # projection is off by |A S - I|
    assert enone >= 100 * esplit


def _bf16(x):
    """FP32 values rounded to BF16 (to nearest, ties to even), as FP32."""
    b = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32)


def _bf16_product(mat32, x):
    """mat32 @ x with both operands rounded FP32 -> BF16 (the FP32 value
    first, as the kernels round) and FP32 sums: v_mfma_f32_16x16x32_bf16."""
    m = _bf16(mat32)
    re = (_bf16(x.real.astype(np.float32)) @ m.T).astype(np.float64)
    im = (_bf16(x.imag.astype(np.float32)) @ m.T).astype(np.float64)
    return re + 1j * im


# the plan's limit for the BF16 correction (plan.hip, kGramBf16Limit)
BF16_LIMIT = 2e-11


@pytest.mark.parametrize("m,ell_max", [(2, 24), (0, 24), (-3, 24), (4, 24), (2, 32)])
def test_bf16_gram_correction_error_and_eligibility(m, ell_max):
    """The Kerr-Gram stage kernels take the correction E X + N Psi on BF16
    operands for bands whose E and N are below 2e-11 (of 1 and of max|Lambda|):
    the ell-24 bands, not the ell-32 ones.  Where taken, the error against
    the long-double product stays below 1e-13 of the terms' size; where not,
    BF16 would not be accurate enough."""
    S, A, lam = _bands(m, ell_max)
    nl = lam.size
    G = A @ S
    gl = G.astype(np.longdouble)
    Ml = gl @ np.diag(lam.astype(np.longdouble)) @ gl
    El = gl - np.eye(nl, dtype=np.longdouble)
    Nl = Ml - np.diag(lam.astype(np.longdouble))
    eligible = (float(np.abs(El).max()) <= BF16_LIMIT
                and float(np.abs(Nl).max()) <= BF16_LIMIT * float(np.abs(lam).max()))
    assert eligible == (ell_max == 24)
    rng = np.random.default_rng(1000 * ell_max + m + 11)
    X = _rows(rng, 64, nl)
    Psi = _rows(rng, 64, nl)
    re = X.real.astype(np.longdouble) @ gl.T + Psi.real.astype(np.longdouble) @ Ml.T
    im = X.imag.astype(np.longdouble) @ gl.T + Psi.imag.astype(np.longdouble) @ Ml.T
    k_exact = re.astype(np.float64) + 1j * im.astype(np.float64)
    row_scale = np.abs(X).max(axis=1) + np.abs(lam).max() * np.abs(Psi).max(axis=1)
    identity = X + Psi * lam[None, :]
    k_bf16 = (identity + _bf16_product(El.astype(np.float32), X)
              + _bf16_product(Nl.astype(np.float32), Psi))
    err = float((np.abs(k_bf16 - k_exact).max(axis=1) / row_scale).max())
    enone = float((np.abs(identity - k_exact).max(axis=1) / row_scale).max())
    print(f"[cpu] ell_max {ell_max} m {m}: BF16 correction {err:.2e}, identity only "
          f"{enone:.2e}, eligible {eligible}")
    if eligible:
        assert err <= 1e-13
        assert enone >= 10 * err
    else:
        assert err > 1e-13
