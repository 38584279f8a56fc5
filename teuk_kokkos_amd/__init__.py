"""MI355X-native Teukolsky stage operator and RK4 stepper.

Python mirror of the reference's SpatialPipeline interface
(include/teuk/spatial_pipeline.hpp:144-620) over the C ABI of
libteuk_hip.so (include/teuk_hip.h).  The HIP library is required: importing
works without it, but creating a Plan raises if it is missing.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libteuk_hip.so")
DIAG_COUNT = 100  # TEUK_DIAG_COUNT (include/teuk_hip.h)

TEUK_OK, TEUK_EINVAL, TEUK_ERUNTIME = 0, 1, 2
SCHEMES = {"d4-2": 0, "d8-4": 1, "d10-5": 2}


class TeukError(RuntimeError):
    """TEUK_ERUNTIME (std::runtime_error in the reference)."""


class TeukInvalidArgument(ValueError):
    """TEUK_EINVAL (std::invalid_argument in the reference)."""


class WaveformRecord(ctypes.Structure):
    _fields_ = [("step", ctypes.c_uint64), ("time", ctypes.c_double),
                ("boundary", ctypes.c_int), ("order", ctypes.c_int),
                ("m", ctypes.c_int), ("ell", ctypes.c_int),
                ("rescaled_re", ctypes.c_double), ("rescaled_im", ctypes.c_double),
                ("endpoint_re", ctypes.c_double), ("endpoint_im", ctypes.c_double)]


class SourceActivation(ctypes.Structure):
    _fields_ = [("active", ctypes.c_int), ("consecutive_passes", ctypes.c_int),
                ("activation_time", ctypes.c_double),
                ("last_eligibility_time", ctypes.c_double)]


# Every symbol include/teuk_hip.h declares, with its ctypes signature.
_P = ctypes.c_void_p
_D = ctypes.c_double
_I = ctypes.c_int
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)
_SP = ctypes.POINTER(ctypes.c_char_p)
SIGNATURES = {
    "teuk_config_resolve": (_I, [ctypes.c_char_p, _SP, _I, ctypes.c_char_p, ctypes.c_size_t,
                                 ctypes.POINTER(ctypes.c_size_t)]),
    "teuk_config_last_error": (ctypes.c_char_p, []),
    "teuk_plan_create": (_I, [ctypes.c_char_p, _SP, _I, ctypes.POINTER(_P)]),
    "teuk_plan_create_batch": (_I, [_SP, _I, _SP, _I, ctypes.POINTER(_P)]),
    "teuk_plan_destroy": (None, [_P]),
    "teuk_last_error": (ctypes.c_char_p, [_P]),
    "teuk_plan_instance_count": (_I, [_P]),
    "teuk_plan_shape": (_I, [_P, _I, _IP, _IP, _IP, _IP]),
    "teuk_plan_time_step": (_D, [_P, _I]),
    "teuk_plan_progress": (_I, [_P, _I, _DP, ctypes.POINTER(ctypes.c_uint64)]),
    "teuk_plan_grid_points": (_D, [_P]),
    "teuk_plan_modal_points": (_D, [_P]),
    "teuk_initialize": (_I, [_P]),
    "teuk_state_import": (_I, [_P, _I, _DP, _D, _DP]),
    "teuk_state_export": (_I, [_P, _I, _DP]),
    "teuk_evaluate_rhs": (_I, [_P, _I, _P, _P, _D]),
    "teuk_evaluate_rhs_noalloc": (_I, [_P, _I, _P, _P, _D]),
    "teuk_step": (_I, [_P, _D, _D]),
    "teuk_advance": (_I, [_P, _I]),
    "teuk_sync": (_I, [_P]),
    "teuk_sample_waveform": (_I, [_P, _I, ctypes.POINTER(WaveformRecord), _I, _IP]),
    "teuk_get_source_activation": (_I, [_P, _I, ctypes.POINTER(SourceActivation)]),
    "teuk_set_source_activation": (_I, [_P, _I, ctypes.POINTER(SourceActivation)]),
    "teuk_get_source_constraint_norms": (_I, [_P, _I, ctypes.POINTER(ctypes.c_double)]),
    "teuk_sample_diagnostics": (_I, [_P, _I, _D, ctypes.POINTER(ctypes.c_double)]),
    "teuk_source_pairs": (_I, [_P, _I, ctypes.POINTER(ctypes.c_double), _I, _IP]),
    "teuk_checkpoint_write": (_I, [_P, _I, ctypes.c_char_p]),
    "teuk_checkpoint_load": (_I, [_P, _I, ctypes.c_char_p]),
    "teuk_plan_set_graph": (_I, [_P, _I]),
    "teuk_plan_set_lds_budget": (_I, [_P, _I]),
    "teuk_plan_set_tile_rows": (_I, [_P, _I]),
    "teuk_plan_tiles": (_I, [_P, _IP, _IP]),
    "teuk_plan_set_kernel_timing": (_I, [_P, _I]),
    "teuk_plan_kernel_timing": (_I, [_P, _DP, ctypes.POINTER(ctypes.c_longlong)]),
    "teuk_plan_step_timing": (_I, [_P, _DP, ctypes.POINTER(ctypes.c_longlong)]),
    "teuk_plan_group_count": (_I, [_P, _IP]),
    "teuk_plan_group_info": (_I, [_P, _I, _IP, _IP, _IP, _IP, _DP]),
    "teuk_plan_group_variant": (_I, [_P, _I, _IP]),
    "teuk_plan_set_specialized": (_I, [_P, _I]),
    "teuk_plan_set_mixed_gram": (_I, [_P, _I]),
    "teuk_plan_set_group_shape": (_I, [_P, _I, _I, _I, _I, _I]),
    "teuk_plan_group_shape": (_I, [_P, _I, _IP, _IP]),
    "teuk_plan_set_stream": (_I, [_P, _P]),
    "teuk_plan_graph_replays": (_I, [_P, ctypes.POINTER(ctypes.c_longlong)]),
    "teuk_allocation_count": (_I, [ctypes.POINTER(ctypes.c_ulonglong)]),
    "teuk_device_alloc": (_I, [ctypes.c_size_t, ctypes.POINTER(_P)]),
    "teuk_device_free": (_I, [_P]),
    "teuk_memcpy": (_I, [_P, _P, ctypes.c_size_t, _I]),
    "teuk_sbp_derivative_entry": (_D, [_I, _I, _I, _I]),
    "teuk_sbp_norm_weight": (_D, [_I, _I, _I]),
    "teuk_band_matrices": (_I, [_I, _I, _I, _I, _DP, _DP]),
    "teuk_config_modes": (_I, [ctypes.c_char_p, _IP, _IP, _I, _IP, _IP, _I, _IP]),
}

_lib = None


def lib():
    """Load libteuk_hip.so; raise loudly when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise TeukError(f"{LIB_PATH} is missing: build it with "
                            "`python -m teuk_kokkos_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _strings(items):
    items = [s.encode() if isinstance(s, str) else s for s in items]
    arr = (ctypes.c_char_p * max(1, len(items)))(*items)
    return arr, len(items)


def _raise(rc: int, message: str):
    if rc == TEUK_EINVAL:
        raise TeukInvalidArgument(message)
    raise TeukError(message)


def resolve_config(text: str, overrides=()) -> str:
    """parse_run_configuration_text + resolved_configuration_text."""
    L = lib()
    ov, n = _strings(overrides)
    need = ctypes.c_size_t(0)
    rc = L.teuk_config_resolve(text.encode(), ov, n, None, 0, ctypes.byref(need))
    if rc != TEUK_OK:
        _raise(rc, L.teuk_config_last_error().decode())
    buf = ctypes.create_string_buffer(need.value)
    rc = L.teuk_config_resolve(text.encode(), ov, n, buf, need.value, ctypes.byref(need))
    if rc != TEUK_OK:
        _raise(rc, L.teuk_config_last_error().decode())
    return buf.value.decode()


def config_modes(text: str):
    """ModeRegistry of a configuration: (modes, sharp indices, ordered pairs)."""
    L = lib()
    cap = 256
    modes = (ctypes.c_int * cap)()
    sharp = (ctypes.c_int * cap)()
    pairs = (ctypes.c_int * (3 * 4096))()
    count = ctypes.c_int(0)
    pcount = ctypes.c_int(0)
    rc = L.teuk_config_modes(text.encode(), modes, sharp, cap, ctypes.byref(count), pairs, 4096,
                             ctypes.byref(pcount))
    if rc != TEUK_OK:
        _raise(rc, L.teuk_config_last_error().decode())
    n, pn = count.value, pcount.value
    return (list(modes[:n]), list(sharp[:n]),
            [tuple(pairs[3 * k:3 * k + 3]) for k in range(pn)])


def band_matrices(spin: int, m: int, ell_max: int, ntheta: int):
    ell_min = max(abs(spin), abs(m))
    nl = ell_max - ell_min + 1
    S = np.zeros((ntheta, nl))
    A = np.zeros((nl, ntheta))
    got = lib().teuk_band_matrices(spin, m, ell_max, ntheta,
                                   S.ctypes.data_as(_DP), A.ctypes.data_as(_DP))
    if got != nl:
        raise TeukInvalidArgument(lib().teuk_config_last_error().decode())
    return S, A


def allocation_count() -> int:
    """Device allocations, streams and events the library has made so far."""
    n = ctypes.c_ulonglong()
    rc = lib().teuk_allocation_count(ctypes.byref(n))
    if rc != TEUK_OK:
        _raise(rc, lib().teuk_config_last_error().decode())
    return n.value


class DeviceArray:
    """A device buffer holding one complex128 numpy-shaped array."""

    def __init__(self, host: np.ndarray):
        self.shape = host.shape
        self.nbytes = host.nbytes
        ptr = ctypes.c_void_p()
        rc = lib().teuk_device_alloc(self.nbytes, ctypes.byref(ptr))
        if rc != TEUK_OK:
            _raise(rc, lib().teuk_config_last_error().decode())
        self.ptr = ptr.value
        self.upload(host)

    def upload(self, host: np.ndarray):
        arr = np.ascontiguousarray(host, dtype=np.complex128)
        assert arr.nbytes == self.nbytes
        rc = lib().teuk_memcpy(ctypes.c_void_p(self.ptr), arr.ctypes.data_as(ctypes.c_void_p),
                               self.nbytes, 0)
        if rc != TEUK_OK:
            _raise(rc, lib().teuk_config_last_error().decode())

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.complex128)
        rc = lib().teuk_memcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr),
                               self.nbytes, 1)
        if rc != TEUK_OK:
            _raise(rc, lib().teuk_config_last_error().decode())
        return out

    def free(self):
        if getattr(self, "ptr", None):
            lib().teuk_device_free(ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        self.free()


class Plan:
    """One or more independent linear Teukolsky evolutions on the GPU.

    Mirrors SpatialPipeline: construction from a schema-v1 configuration,
    evaluate_rhs, step, state import/export in the reference's
    (mode, field=13, radial, theta) layout, and endpoint waveforms.
    """

    def __init__(self, config, overrides=()):
        L = lib()
        ov, n = _strings(overrides)
        handle = ctypes.c_void_p()
        if isinstance(config, (list, tuple)):
            cfgs, nc = _strings(list(config))
            rc = L.teuk_plan_create_batch(cfgs, nc, ov, n, ctypes.byref(handle))
        else:
            rc = L.teuk_plan_create(config.encode(), ov, n, ctypes.byref(handle))
        if rc != TEUK_OK:
            _raise(rc, L.teuk_config_last_error().decode())
        self._h = handle
        self.instances = L.teuk_plan_instance_count(self._h)

    def close(self):
        if getattr(self, "_h", None):
            lib().teuk_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc != TEUK_OK:
            _raise(rc, lib().teuk_last_error(self._h).decode())

    # ----------------------------------------------------------------- shape
    def shape(self, instance: int = 0):
        m, f, r, t = (ctypes.c_int() for _ in range(4))
        self._check(lib().teuk_plan_shape(self._h, instance, ctypes.byref(m), ctypes.byref(f),
                                          ctypes.byref(r), ctypes.byref(t)))
        return (m.value, f.value, r.value, t.value)

    def time_step(self, instance: int = 0) -> float:
        return lib().teuk_plan_time_step(self._h, instance)

    def progress(self, instance: int = 0):
        t = ctypes.c_double()
        s = ctypes.c_uint64()
        self._check(lib().teuk_plan_progress(self._h, instance, ctypes.byref(t), ctypes.byref(s)))
        return t.value, s.value

    @property
    def grid_points(self) -> float:
        return lib().teuk_plan_grid_points(self._h)

    @property
    def modal_points(self) -> float:
        return lib().teuk_plan_modal_points(self._h)

    # ----------------------------------------------------------------- state
    def initialize(self):
        self._check(lib().teuk_initialize(self._h))

    def import_state(self, state: np.ndarray, instance: int = 0, tolerance: float = 5.0e-11):
        arr = np.ascontiguousarray(state, dtype=np.complex128)
        if arr.shape != self.shape(instance):
            raise TeukInvalidArgument("imported pipeline state shape mismatch")
        off = ctypes.c_double()
        self._check(lib().teuk_state_import(self._h, instance,
                                            arr.view(np.float64).ctypes.data_as(_DP), tolerance,
                                            ctypes.byref(off)))
        return off.value

    def export_state(self, instance: int = 0) -> np.ndarray:
        out = np.zeros(self.shape(instance), dtype=np.complex128)
        self._check(lib().teuk_state_export(self._h, instance,
                                            out.view(np.float64).ctypes.data_as(_DP)))
        return out

    # -------------------------------------------------------------- operator
    def evaluate_rhs_device(self, d_stage: int, d_out: int, instance: int = 0,
                            time: float = 0.0):
        """Enqueue the RHS of device buffers in the interchange layout."""
        self._check(lib().teuk_evaluate_rhs(self._h, instance, ctypes.c_void_p(d_stage),
                                            ctypes.c_void_p(d_out), time))

    def step(self, time: float, dt: float):
        self._check(lib().teuk_step(self._h, time, dt))

    def advance(self, steps: int):
        self._check(lib().teuk_advance(self._h, steps))

    def sync(self):
        self._check(lib().teuk_sync(self._h))

    def sample_waveform(self, instance: int = 0):
        count = ctypes.c_int()
        self._check(lib().teuk_sample_waveform(self._h, instance, None, 0, ctypes.byref(count)))
        recs = (WaveformRecord * max(1, count.value))()
        self._check(lib().teuk_sample_waveform(self._h, instance, recs, count.value,
                                               ctypes.byref(count)))
        return [recs[k] for k in range(count.value)]

    def source_activation(self, instance: int = 0) -> SourceActivation:
        s = SourceActivation()
        self._check(lib().teuk_get_source_activation(self._h, instance, ctypes.byref(s)))
        return s

    def restore_source_activation(self, state: SourceActivation, instance: int = 0):
        self._check(lib().teuk_set_source_activation(self._h, instance, ctypes.byref(state)))

    def source_constraint_norms(self, instance: int = 0) -> dict:
        """Norms of the last constraint-aware activation check (second-order
        plans): per family the absolute maximum, normalized maximum, weighted
        RMS and normalized weighted norm."""
        v = (ctypes.c_double * 14)()
        self._check(lib().teuk_get_source_constraint_norms(self._h, instance, v))
        fam = lambda k: [v[k], v[4 + k], v[8 + k]]
        return {"absolute_maximum": fam(0), "normalized_maximum": fam(1),
                "weighted_rms": fam(2), "normalized_weighted": fam(3),
                "controlling_normalized": v[12], "all_finite": v[13] > 0.5}

    def sample_diagnostics(self, time: float, instance: int = 0) -> dict:
        """The reference driver's sampled diagnostics (solver_driver.hpp:
        205-242) of one instance at `time`: RMS and maximum of every state
        and RHS field, of the two reduction constraints, source_over_r3 and
        the forcing, the independent reconstruction constraints and the
        transport residuals (each a (rms, max) pair), the horizon derivative
        maxima of FirstPsi and SecondPsi, and the latch."""
        v = (ctypes.c_double * DIAG_COUNT)()
        self._check(lib().teuk_sample_diagnostics(self._h, instance, time, v))
        a = np.array(v[:])
        pair = lambda k: (a[k], a[k + 1])
        return {
            "state_rms": a[0:13], "state_max": a[13:26], "rhs_rms": a[26:39],
            "rhs_max": a[39:52], "first_constraint": pair(52), "second_constraint": pair(54),
            "source_over_r3": pair(56), "forcing": pair(58),
            "psi3_bianchi": pair(60), "psi2_bianchi": pair(62), "hll_reality": pair(64),
            "independent_combined": pair(66),
            "transport": {n: pair(68 + 2 * k)
                          for k, n in enumerate(("G", "Lambda", "H", "B", "Pi", "C", "U"))},
            "transport_combined": pair(82),
            "first_horizon": a[84:89], "second_horizon": a[89:94],
            "source_constraint_max": a[94], "source_active": a[95] > 0.5,
            "activation_time": a[96], "raw": a}

    def source_pairs(self, instance: int = 0) -> list:
        """Per-pair source report of the last sample_diagnostics call
        (second-order plans): dicts with m1, m2, target and the RMS and
        maximum of |D| and |T|."""
        n = ctypes.c_int()
        self._check(lib().teuk_source_pairs(self._h, instance, None, 0, ctypes.byref(n)))
        v = (ctypes.c_double * max(7 * n.value, 1))()
        self._check(lib().teuk_source_pairs(self._h, instance, v, n.value, ctypes.byref(n)))
        keys = ("m1", "m2", "target", "D_rms", "D_max", "T_rms", "T_max")
        out = []
        for k in range(n.value):
            rec = dict(zip(keys, v[7 * k:7 * k + 7]))
            for key in ("m1", "m2", "target"):
                rec[key] = int(rec[key])
            out.append(rec)
        return out

    def write_checkpoint(self, directory: str, instance: int = 0):
        """Checkpoint format v5 of one instance into a new directory."""
        self._check(lib().teuk_checkpoint_write(self._h, instance, str(directory).encode()))

    def load_checkpoint(self, directory: str, instance: int = 0):
        """Load a checkpoint whose registry, configuration and geometry match
        the instance: its state, latch and progress."""
        self._check(lib().teuk_checkpoint_load(self._h, instance, str(directory).encode()))

    # ---------------------------------------------------------------- tuning
    def set_graph(self, enable: bool):
        self._check(lib().teuk_plan_set_graph(self._h, int(enable)))

    def graph_replays(self) -> int:
        """Steps taken as replays of the captured step graph."""
        n = ctypes.c_longlong()
        self._check(lib().teuk_plan_graph_replays(self._h, ctypes.byref(n)))
        return n.value

    def set_lds_budget(self, nbytes: int):
        self._check(lib().teuk_plan_set_lds_budget(self._h, nbytes))

    def set_tile_rows(self, rows: int):
        self._check(lib().teuk_plan_set_tile_rows(self._h, rows))

    def set_stream(self, stream: int | None):
        """Enqueue every later call on the caller's hipStream_t (an integer
        handle, e.g. torch.cuda.Stream().cuda_stream); None returns to the
        plan's own stream."""
        self._check(lib().teuk_plan_set_stream(self._h, ctypes.c_void_p(stream or 0)))

    def set_specialized(self, enable: bool):
        """Stage kernels specialised per line variant (the default), or every
        line through the general kernel."""
        self._check(lib().teuk_plan_set_specialized(self._h, int(enable)))

    def set_mixed_gram(self, enable: bool):
        """The Gram correction of the ell-24 bands on BF16 MFMA (mixed
        precision) instead of FP32 (as accurate as the FP64 product)."""
        self._check(lib().teuk_plan_set_mixed_gram(self._h, int(enable)))

    def tiles(self):
        n = ctypes.c_int()
        lds = ctypes.c_int()
        self._check(lib().teuk_plan_tiles(self._h, ctypes.byref(n), ctypes.byref(lds)))
        return n.value, lds.value

    def set_kernel_timing(self, enable: bool):
        self._check(lib().teuk_plan_set_kernel_timing(self._h, int(enable)))

    def kernel_timing(self):
        ms = ctypes.c_double()
        n = ctypes.c_longlong()
        self._check(lib().teuk_plan_kernel_timing(self._h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def step_timing(self):
        """(GPU milliseconds, steps) of the RK4 steps timed since
        set_kernel_timing(True)."""
        ms = ctypes.c_double()
        n = ctypes.c_longlong()
        self._check(lib().teuk_plan_step_timing(self._h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def groups(self):
        """Execution groups (one stage launch each): scheme, ell_tiles, line
        variant (0 general, 1 Kerr with G = I, 2 Kerr with the Gram
        projection), tiles, lds_bytes and the group's timed kernel
        milliseconds."""
        n = ctypes.c_int()
        self._check(lib().teuk_plan_group_count(self._h, ctypes.byref(n)))
        out = []
        for g in range(n.value):
            v = [ctypes.c_int() for _ in range(4)]
            ms = ctypes.c_double()
            self._check(lib().teuk_plan_group_info(self._h, g, *[ctypes.byref(x) for x in v],
                                                   ctypes.byref(ms)))
            var = ctypes.c_int()
            self._check(lib().teuk_plan_group_variant(self._h, g, ctypes.byref(var)))
            sh = [ctypes.c_int() for _ in range(2)]
            self._check(lib().teuk_plan_group_shape(self._h, g, *[ctypes.byref(x) for x in sh]))
            out.append({"scheme": v[0].value, "ell_tiles": v[1].value, "variant": var.value,
                        "tiles": v[2].value, "lds_bytes": v[3].value, "kernel_ms": ms.value,
                        "lds_budget": sh[0].value, "tile_rows": sh[1].value})
        return out

    def set_group_shape(self, scheme: int, ell_tiles: int, variant: int, lds_budget: int = 0,
                        tile_rows: int = 0):
        """Tile shape of the execution group keyed (scheme, ell_tiles,
        variant): LDS budget and tile rows (0 = the plan's setting).
        Re-tiles; results do not depend on the shape."""
        self._check(lib().teuk_plan_set_group_shape(self._h, scheme, ell_tiles, variant,
                                                    lds_budget, tile_rows))
