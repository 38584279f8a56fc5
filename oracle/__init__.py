"""ctypes binding of the CPU oracle (oracle/libteuk_oracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
The product package never imports it.

The oracle is a nodal FP64 restatement of the reference's linear stage
operator (include/teuk/full_spatial.hpp:99-398,
include/teuk/spatial_pipeline.hpp:458-620) and classical RK4
(include/teuk/device_rk4.hpp:84-128).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libteuk_oracle.so")
_lib = None

SCHEMES = {"d4-2": 0, "d8-4": 1, "d10-5": 2}
REDUCTIONS = {"free_damped": 0, "stage_constrained": 1}


class OracleParams(ctypes.Structure):
    _fields_ = [
        ("mass", ctypes.c_double), ("spin", ctypes.c_double), ("length", ctypes.c_double),
        ("nr", ctypes.c_int), ("ntheta", ctypes.c_int),
        ("ell_max_first", ctypes.c_int), ("ell_max_second", ctypes.c_int),
        ("n_modes", ctypes.c_int), ("modes", ctypes.c_int * 64),
        ("n_parents", ctypes.c_int), ("parents", ctypes.c_int * 64),
        ("n_targets", ctypes.c_int), ("targets", ctypes.c_int * 64),
        ("reduction_damping", ctypes.c_double), ("dissipation", ctypes.c_double),
        ("reduction", ctypes.c_int), ("scheme", ctypes.c_int), ("threads", ctypes.c_int),
        ("source_mode", ctypes.c_int), ("source_start_time", ctypes.c_double),
        ("constraint_tolerance", ctypes.c_double), ("required_passes", ctypes.c_int),
    ]


class OracleActivation(ctypes.Structure):
    _fields_ = [("active", ctypes.c_int), ("activation_time", ctypes.c_double),
                ("consecutive_passes", ctypes.c_int),
                ("last_eligibility_time", ctypes.c_double)]


class OracleConstraintNorms(ctypes.Structure):
    _fields_ = [("absolute_maximum", ctypes.c_double * 3),
                ("normalized_maximum", ctypes.c_double * 3),
                ("weighted_rms", ctypes.c_double * 3),
                ("normalized_weighted", ctypes.c_double * 3),
                ("controlling_normalized", ctypes.c_double), ("all_finite", ctypes.c_int)]


SOURCE_MODES = {"disabled": 0, "constraint_aware": 1, "unrestricted": 2}


def build() -> None:
    """Compile the oracle with its Makefile (g++, OpenMP)."""
    import subprocess
    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        d, i, p = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.oracle_create.restype = p
        L.oracle_create.argtypes = [ctypes.POINTER(OracleParams)]
        L.oracle_destroy.argtypes = [p]
        L.oracle_last_error.restype = ctypes.c_char_p
        L.oracle_value_count.restype = ctypes.c_longlong
        L.oracle_value_count.argtypes = [p]
        L.oracle_get_state.argtypes = [p, dp]
        L.oracle_set_state.argtypes = [p, dp]
        L.oracle_init_gaussian.restype = d
        L.oracle_init_gaussian.argtypes = [p, d, d, i, i, ip, ip, dp, dp]
        L.oracle_evaluate_rhs.argtypes = [p, dp, dp]
        L.oracle_step.restype = i
        L.oracle_step.argtypes = [p, d, d]
        L.oracle_advance.restype = i
        L.oracle_advance.argtypes = [p, i, d, d]
        L.oracle_get_activation.argtypes = [p, ctypes.POINTER(OracleActivation)]
        L.oracle_set_activation.argtypes = [p, ctypes.POINTER(OracleActivation)]
        L.oracle_constraint_norms.argtypes = [p, ctypes.POINTER(OracleConstraintNorms)]
        L.oracle_consider_activation.restype = i
        L.oracle_consider_activation.argtypes = [p, d]
        L.oracle_last_stage_sources.restype = i
        L.oracle_last_stage_sources.argtypes = [p, ip]
        L.oracle_intermediate.restype = ctypes.c_longlong
        L.oracle_intermediate.argtypes = [p, i, dp]
        L.oracle_background.argtypes = [d, d, d, d, d, d, dp]
        L.oracle_pair_source.argtypes = [d, dp, dp, dp, dp]
        L.oracle_ghp_row.restype = i
        L.oracle_ghp_row.argtypes = [i, i, i, i, i, i, d, d, d, dp, dp, dp]
        L.oracle_delta_n.argtypes = [dp, dp, dp, i, d, d, d, dp]
        L.oracle_recon_dt.argtypes = [dp, dp, dp, i, d, d, d, dp]
        L.oracle_field_weights.argtypes = [ip]
        L.oracle_eth_point.argtypes = [i, i, i, d, d, d, d, d, dp, dp, dp, dp]
        L.oracle_recon_rhs.argtypes = [d, dp, dp, dp]
        L.oracle_outer_source_point.argtypes = [d, dp, dp, dp]
        L.oracle_coordinate_forcing.argtypes = [d, d, d, d, dp, dp]
        L.oracle_constraint_terms.argtypes = [d, dp, dp, dp]
        L.oracle_thorn.argtypes = [dp, dp, dp, i, i, i, i, d, d, d, d, d, dp, dp]
        L.oracle_endpoint_modal.restype = i
        L.oracle_endpoint_modal.argtypes = [p, i, dp, dp]
        L.oracle_gauss_legendre.argtypes = [i, dp, dp]
        L.oracle_swsh_theta.restype = d
        L.oracle_swsh_theta.argtypes = [i, i, i, d]
        L.oracle_transform_matrices.restype = i
        L.oracle_transform_matrices.argtypes = [i, i, i, i, dp, dp]
        L.oracle_teukolsky_coefficients.argtypes = [d, d, d, i, i, d, d, dp]
        L.oracle_sbp_derivative_entry.restype = d
        L.oracle_sbp_derivative_entry.argtypes = [i, i, i, i]
        L.oracle_sbp_norm_weight.restype = d
        L.oracle_sbp_norm_weight.argtypes = [i, i, i]
        L.oracle_sbp_minimum_points.restype = i
        L.oracle_sbp_minimum_points.argtypes = [i]
        L.oracle_sbp_derivative.argtypes = [i, i, d, dp, dp]
        L.oracle_sbp_dissipation.argtypes = [i, i, d, d, dp, dp]
        L.oracle_dissipation_spectral_radius_bound.restype = d
        L.oracle_dissipation_spectral_radius_bound.argtypes = [i]
        L.oracle_dissipation_step_admissible.restype = i
        L.oracle_dissipation_step_admissible.argtypes = [i, d, d]
        L.oracle_maximum_coordinate_speed.restype = d
        L.oracle_maximum_coordinate_speed.argtypes = [d, d, d, i, i]
        L.oracle_dense_solve.restype = d
        L.oracle_dense_solve.argtypes = [i, dp, dp, dp]
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _cview(a: np.ndarray) -> np.ndarray:
    """complex128 array -> float64 view (interleaved)."""
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a.view(np.float64)


# ----------------------------------------------------------------- primitives
def gauss_legendre(n: int):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().oracle_gauss_legendre(n, _dp(x), _dp(w))
    return x, w


def swsh_theta(ell: int, m: int, spin: int, theta: float) -> float:
    return lib().oracle_swsh_theta(ell, m, spin, theta)


def transform_matrices(spin: int, m: int, ell_max: int, ntheta: int):
    ell_min = max(abs(spin), abs(m))
    nm = ell_max - ell_min + 1
    S = np.zeros((ntheta, nm))
    A = np.zeros((nm, ntheta))
    got = lib().oracle_transform_matrices(spin, m, ell_max, ntheta, _dp(S), _dp(A))
    if got != nm:
        raise ValueError(lib().oracle_last_error().decode())
    return S, A


def teukolsky_coefficients(mass, spin, length, spin_weight, m, radius, theta):
    out = np.zeros(9)
    lib().oracle_teukolsky_coefficients(mass, spin, length, spin_weight, m, radius,
                                        theta, _dp(out))
    return {"time": out[0], "radial_advection": out[1], "radial_principal": out[2],
            "definition": complex(out[3], out[4]), "q": complex(out[5], out[6]),
            "psi": complex(out[7], out[8])}


def sbp_matrix(scheme: int, n: int, h: float = 1.0):
    """Dense D (derivative) and H (norm) matrices for an n-point grid."""
    L = lib()
    D = np.array([[L.oracle_sbp_derivative_entry(scheme, n, r, c) for c in range(n)]
                  for r in range(n)]) / h
    H = np.diag([h * L.oracle_sbp_norm_weight(scheme, n, i) for i in range(n)])
    return D, H


def sbp_derivative(scheme: int, values: np.ndarray, h: float) -> np.ndarray:
    v = _cview(values)
    out = np.zeros_like(v)
    lib().oracle_sbp_derivative(scheme, len(values), h, _dp(v), _dp(out))
    return out.view(np.complex128)


def sbp_dissipation(scheme: int, values: np.ndarray, h: float, strength: float):
    v = _cview(values)
    out = np.zeros_like(v)
    lib().oracle_sbp_dissipation(scheme, len(values), h, strength, _dp(v), _dp(out))
    return out.view(np.complex128)


def norm_weight(scheme: int, n: int, i: int) -> float:
    return lib().oracle_sbp_norm_weight(scheme, n, i)


def maximum_coordinate_speed(mass, spin, length, nr, ntheta):
    return lib().oracle_maximum_coordinate_speed(mass, spin, length, nr, ntheta)


def dense_solve(matrix: np.ndarray, rhs: np.ndarray):
    n = len(rhs)
    m = _cview(matrix)
    b = _cview(rhs)
    x = np.zeros(2 * n)
    res = lib().oracle_dense_solve(n, _dp(m), _dp(b), _dp(x))
    return x.view(np.complex128), res


# ---------------------------------------------------- second-order primitives
BACKGROUND_NAMES = ("mu0", "tau0", "pi0", "psi20", "rho0", "epsilon0")


def background(mass, spin, length, radius, cos_theta, sin_theta) -> dict:
    out = np.zeros(12)
    lib().oracle_background(mass, spin, length, radius, cos_theta, sin_theta, _dp(out))
    v = out.view(np.complex128)
    return dict(zip(BACKGROUND_NAMES, v))


def pair_source(radius, bg: dict, fields, derivs):
    """fields: 19 (value, tangent) complex pairs, derivs: 11 pairs.
    Returns ((D, D_t), (T, T_t))."""
    b = np.array([bg["mu0"], bg["tau0"], bg["pi0"]], dtype=np.complex128).view(np.float64).copy()
    f = np.array(fields, dtype=np.complex128).reshape(19, 2).view(np.float64).copy()
    g = np.array(derivs, dtype=np.complex128).reshape(11, 2).view(np.float64).copy()
    out = np.zeros(8)
    lib().oracle_pair_source(radius, _dp(b), _dp(f.reshape(-1)), _dp(g.reshape(-1)), _dp(out))
    v = out.view(np.complex128)
    return (v[0], v[1]), (v[2], v[3])


def ghp_row(spin, boost, m, ell_max, x, dt, *, prime=False, kerr_spin=0.0, length=1.0,
            radius=0.5):
    x = _cview(x)
    dt = _cview(dt)
    out = np.zeros_like(x)
    rc = lib().oracle_ghp_row(spin, boost, m, ell_max, len(x) // 2, int(prime), kerr_spin,
                              length, radius, _dp(x), _dp(dt), _dp(out))
    if rc < 0:
        raise ValueError(lib().oracle_last_error().decode())
    return out.view(np.complex128)


def _c1(z) -> np.ndarray:
    return np.array([complex(z)], dtype=np.complex128).view(np.float64).copy()


def _cn(values) -> np.ndarray:
    return np.array([complex(z) for z in values], dtype=np.complex128).view(np.float64).copy()


WEIGHT_FIELDS = ("F", "G", "H", "Lambda", "Pi", "B", "C", "U", "D", "T")


def field_weights() -> dict:
    """{field: (spin, boost, falloff)} as the oracle's pipeline uses them."""
    out = (ctypes.c_int * 30)()
    lib().oracle_field_weights(out)
    return {f: (out[3 * k], out[3 * k + 1], out[3 * k + 2]) for k, f in enumerate(WEIGHT_FIELDS)}


def eth_point(prime, spin, boost, kerr_spin, length, radius, cos_theta, sin_theta, x, dt,
              ang) -> complex:
    out = np.zeros(2)
    lib().oracle_eth_point(int(prime), spin, boost, kerr_spin, length, radius, cos_theta,
                           sin_theta, _dp(_c1(x)), _dp(_c1(dt)), _dp(_c1(ang)), _dp(out))
    return complex(out[0], out[1])


RECON_INPUTS = ("F", "G", "Lambda", "H", "B", "Pi", "C", "U", "Pi_s", "B_s", "C_s", "eth1F",
                "eth2G", "eth2C", "eth2Pi", "ethp1B_s", "ethp2C_s")
RECON_OUTPUTS = ("G", "Lambda", "H", "B", "Pi", "C", "U")


def recon_rhs(radius, bg: dict, inputs: dict) -> dict:
    """R_X of Delta_n X = R_X for the seven reconstruction fields."""
    b = _cn([bg["mu0"], bg["tau0"], bg["pi0"]])
    x = _cn([inputs.get(k, 0j) for k in RECON_INPUTS])
    out = np.zeros(14)
    lib().oracle_recon_rhs(radius, _dp(b), _dp(x), _dp(out))
    return dict(zip(RECON_OUTPUTS, out.view(np.complex128)))


def outer_source_point(radius, bg: dict, delta3_d, d, ethp3_t, t) -> complex:
    b = _cn([bg["mu0"], bg["tau0"], bg["pi0"]])
    x = _cn([delta3_d, d, ethp3_t, t])
    out = np.zeros(2)
    lib().oracle_outer_source_point(radius, _dp(b), _dp(x), _dp(out))
    return complex(out[0], out[1])


def coordinate_forcing(length, kerr_spin, radius, cos_theta, source) -> complex:
    out = np.zeros(2)
    lib().oracle_coordinate_forcing(length, kerr_spin, radius, cos_theta, _dp(_c1(source)),
                                    _dp(out))
    return complex(out[0], out[1])


CONSTRAINT_INPUTS = ("F", "G", "H", "Lambda", "Pi", "B", "C", "U", "U_s", "thorn1F", "thorn2G",
                     "ethp2G", "ethp3H")


def constraint_terms(radius, bg: dict, inputs: dict):
    """(left, right) pairs of the psi3 Bianchi, psi2 Bianchi and h_ll
    reality constraints."""
    b = _cn([bg[k] for k in BACKGROUND_NAMES])
    x = _cn([inputs.get(k, 0j) for k in CONSTRAINT_INPUTS])
    out = np.zeros(12)
    lib().oracle_constraint_terms(radius, _dp(b), _dp(x), _dp(out))
    v = out.view(np.complex128)
    return [(v[0], v[1]), (v[2], v[3]), (v[4], v[5])]


def thorn(value, dt, dr, n, spin, boost, m, radius, cos_theta, mass, kerr_spin, length,
          epsilon0) -> complex:
    out = np.zeros(2)
    lib().oracle_thorn(_dp(_c1(value)), _dp(_c1(dt)), _dp(_c1(dr)), n, spin, boost, m, radius,
                       cos_theta, mass, kerr_spin, length, _dp(_c1(epsilon0)), _dp(out))
    return complex(out[0], out[1])


def delta_n(value, dt, dr, n, radius, mass, length) -> complex:
    a, b, c = (_cview(np.array([z])) for z in (value, dt, dr))
    out = np.zeros(2)
    lib().oracle_delta_n(_dp(a), _dp(b), _dp(c), n, radius, mass, length, _dp(out))
    return complex(out[0], out[1])


def recon_dt(value, dr, rhs, n, radius, mass, length) -> complex:
    a, b, c = (_cview(np.array([z])) for z in (value, dr, rhs))
    out = np.zeros(2)
    lib().oracle_recon_dt(_dp(a), _dp(b), _dp(c), n, radius, mass, length, _dp(out))
    return complex(out[0], out[1])


# ------------------------------------------------------------------- pipeline
@dataclass
class Setup:
    """Plain parameters of one linear run (the config-schema subset the
    oracle needs)."""
    mass: float = 1.0
    spin: float = 0.0
    length: float = 1.0
    nr: int = 65
    ntheta: int = 7
    ell_max_first: int = 4
    ell_max_second: int = 4
    first_order_modes: list = field(default_factory=lambda: [-2, 2])
    second_order_modes: list = field(default_factory=lambda: [-2, 2])
    reduction_damping: float = 0.1
    dissipation: float = 0.0
    reduction: str = "free_damped"
    scheme: str = "d4-2"
    source_mode: str = "disabled"
    source_start_time: float = 0.0
    constraint_tolerance: float = 1.0e-10
    required_passes: int = 1

    @property
    def horizon(self) -> float:
        rp = self.mass + np.sqrt(max(self.mass ** 2 - self.spin ** 2, 0.0))
        return self.length ** 2 / rp

    @property
    def spacing(self) -> float:
        return self.horizon / (self.nr - 1)

    @property
    def modes(self):
        return sorted(set(self.first_order_modes) | set(self.second_order_modes))


class Pipeline:
    """Nodal oracle pipeline for one linear run."""

    def __init__(self, setup: Setup, threads: int = 0):
        self.setup = setup
        p = OracleParams()
        p.mass, p.spin, p.length = setup.mass, setup.spin, setup.length
        p.nr, p.ntheta = setup.nr, setup.ntheta
        p.ell_max_first, p.ell_max_second = setup.ell_max_first, setup.ell_max_second
        modes = setup.modes
        p.n_modes = len(modes)
        for k, m in enumerate(modes):
            p.modes[k] = m
        p.n_parents = len(setup.first_order_modes)
        for k, m in enumerate(sorted(setup.first_order_modes)):
            p.parents[k] = m
        p.n_targets = len(setup.second_order_modes)
        for k, m in enumerate(sorted(setup.second_order_modes)):
            p.targets[k] = m
        p.reduction_damping = setup.reduction_damping
        p.dissipation = setup.dissipation
        p.reduction = REDUCTIONS[setup.reduction]
        p.scheme = SCHEMES[setup.scheme]
        p.threads = threads
        p.source_mode = SOURCE_MODES[setup.source_mode]
        p.source_start_time = setup.source_start_time
        p.constraint_tolerance = setup.constraint_tolerance
        p.required_passes = setup.required_passes
        self._h = lib().oracle_create(ctypes.byref(p))
        if not self._h:
            raise ValueError(lib().oracle_last_error().decode())
        self.shape = (len(modes), 13, setup.nr, setup.ntheta)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().oracle_destroy(h)
            self._h = None

    def get_state(self) -> np.ndarray:
        out = np.zeros(2 * int(np.prod(self.shape)))
        lib().oracle_get_state(self._h, _dp(out))
        return out.view(np.complex128).reshape(self.shape)

    def set_state(self, state: np.ndarray) -> None:
        v = _cview(state.reshape(-1))
        lib().oracle_set_state(self._h, _dp(v))

    def init_gaussian(self, center: float, width: float, seeds, compact_support=False):
        """seeds: list of (ell, m, complex amplitude), already expanded."""
        seeds = sorted(seeds, key=lambda s: (s[1], s[0]))
        n = len(seeds)
        ells = (ctypes.c_int * n)(*[s[0] for s in seeds])
        ms = (ctypes.c_int * n)(*[s[1] for s in seeds])
        re = np.array([complex(s[2]).real for s in seeds])
        im = np.array([complex(s[2]).imag for s in seeds])
        res = lib().oracle_init_gaussian(self._h, center, width, int(compact_support), n,
                                         ells, ms, _dp(re), _dp(im))
        if res < 0:
            raise RuntimeError(lib().oracle_last_error().decode())
        return res

    def evaluate_rhs(self, stage: np.ndarray) -> np.ndarray:
        v = _cview(stage.reshape(-1))
        out = np.zeros_like(v)
        lib().oracle_evaluate_rhs(self._h, _dp(v), _dp(out))
        return out.view(np.complex128).reshape(self.shape)

    def _ok(self, rc: int) -> None:
        if rc < 0:
            raise RuntimeError(lib().oracle_last_error().decode())

    def step(self, time: float, dt: float) -> None:
        self._ok(lib().oracle_step(self._h, time, dt))

    def advance(self, steps: int, dt: float, time0: float = 0.0) -> None:
        self._ok(lib().oracle_advance(self._h, steps, time0, dt))

    def activation(self) -> dict:
        a = OracleActivation()
        lib().oracle_get_activation(self._h, ctypes.byref(a))
        return {"active": bool(a.active), "activation_time": a.activation_time,
                "consecutive_passes": a.consecutive_passes,
                "last_eligibility_time": a.last_eligibility_time}

    def set_activation(self, active, activation_time, consecutive_passes,
                       last_eligibility_time) -> None:
        a = OracleActivation(int(active), activation_time, consecutive_passes,
                             last_eligibility_time)
        lib().oracle_set_activation(self._h, ctypes.byref(a))

    def consider_activation(self, time: float) -> bool:
        rc = lib().oracle_consider_activation(self._h, time)
        self._ok(rc)
        return bool(rc)

    def constraint_norms(self) -> dict:
        n = OracleConstraintNorms()
        lib().oracle_constraint_norms(self._h, ctypes.byref(n))
        return {"absolute_maximum": list(n.absolute_maximum),
                "normalized_maximum": list(n.normalized_maximum),
                "weighted_rms": list(n.weighted_rms),
                "normalized_weighted": list(n.normalized_weighted),
                "controlling_normalized": n.controlling_normalized,
                "all_finite": bool(n.all_finite)}

    def last_stage_sources(self):
        flags = (ctypes.c_int * 8)()
        n = lib().oracle_last_stage_sources(self._h, flags)
        return [int(flags[k]) for k in range(n)]

    INTERMEDIATES = {"forcing": (0, 1), "source_over_r3": (1, 1), "inner": (2, 2),
                     "inner_tangent": (3, 2), "constraints": (4, 3), "constraint_terms": (5, 6),
                     "radial_derivatives": (6, 7), "angular_inputs": (7, 6),
                     "pair_sums": (8, 2)}

    def intermediate(self, name: str) -> np.ndarray:
        which, slots = self.INTERMEDIATES[name]
        n = lib().oracle_intermediate(self._h, which, None)
        if n < 0:
            raise ValueError("intermediates exist only with the source policy enabled")
        out = np.zeros(2 * n)
        lib().oracle_intermediate(self._h, which, _dp(out))
        M, _, N, T = self.shape
        shape = (M, N, T) if slots == 1 else (M, slots, N, T)
        return out.view(np.complex128).reshape(shape)

    def endpoint_modal(self, mode_index: int):
        nm = self.setup.ell_max_first + 1
        a = np.zeros(2 * nm)
        b = np.zeros(2 * nm)
        got = lib().oracle_endpoint_modal(self._h, mode_index, _dp(a), _dp(b))
        if got < 0:
            raise ValueError("mode is not a first-order mode")
        return a[:2 * got].view(np.complex128), b[:2 * got].view(np.complex128)
