"""Python host mirror of the MI355X free-streaming + Landau-matching library.

Thin ctypes layer over ``libfsm.so`` (the C ABI declared in
``include/fsm.h``).  It mirrors the reference's operator interface for this
path -- class ``FREESTREAMMILNE`` of src/FreestreamMilne.cpp:28-87 with
``initialize_from_vector``, ``run_freestream_milne``, ``output_to_vectors``,
``tau_LandauMatch`` and the ``final_*`` vectors -- and exposes the
device-resident entry points for callers that keep events in HBM.

The library is required: importing this module raises if ``libfsm.so`` has
not been built (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

try:  # load torch's HIP runtime first so one runtime serves both libraries
    import torch  # noqa: F401
except Exception:  # pragma: no cover - torch is optional for host-buffer use
    torch = None

_HERE = os.path.dirname(os.path.abspath(__file__))
_PKG = os.path.dirname(_HERE)
LIB_PATH = os.path.join(_PKG, "libfsm.so")

OUTPUT_NAMES = (
    "e", "p", "ut", "ux", "uy", "un",
    "pitt", "pitx", "pity", "pitn", "pixx", "pixy", "pixn", "piyy", "piyn", "pinn",
    "Pi",
)
NUM_OUTPUTS = len(OUTPUT_NAMES)
NUM_OUTPUTS_BARYON = NUM_OUTPUTS + 5  # + n_B, V^tau, V^x, V^y, V^eta

FSM_OK = 0
FSM_ERR_SIZE = -1
FSM_ERR_IC = -2
FSM_ERR_FILE = -3
FSM_ERR_OOM = -4
FSM_ERR_DEVICE = -5
FSM_ERR_ARG = -6
FSM_ERR_UNSUPPORTED = -7
FSM_ERR_BUSY = -8
FSM_ERR_CANCELLED = -9
FSM_ERR_TOO_LARGE = -10
FSM_ERR_NO_PLAN = -11

PATH_FP32 = 0
PATH_MIXED = 1
# 3-D rapidity quadrature (fsm_set_rapidity_rule)
RAPIDITY_TAN = 0
RAPIDITY_GAUSS_HERMITE = 1

IC_WHICH_ENERGY = 0
IC_WHICH_BARYON = 1
PROFILE_XYEV = 0
PROFILE_CPUVH = 1
PROFILE_SUPERMC_BLOCK = 2
PROFILE_TRENTO_BLOCK = 3
PROFILE_TRENTO3D_BLOCK = 4
TILES_AUTO = 0
TILES_BATCH = 1

_PARAM_FIELDS = [
    ("OUTPUTFORMAT", ctypes.c_int), ("BARYON", ctypes.c_int), ("IC_ENERGY", ctypes.c_int),
    ("IC_BARYON", ctypes.c_int), ("ETA_WIDTH", ctypes.c_float), ("ETA_FLAT", ctypes.c_float),
    ("SIGMA", ctypes.c_float), ("SIGMA_B", ctypes.c_float), ("DIM_X", ctypes.c_int),
    ("DIM_Y", ctypes.c_int), ("DIM_ETA", ctypes.c_int), ("DIM_RAP", ctypes.c_int),
    ("DIM_PHIP", ctypes.c_int), ("DX", ctypes.c_float), ("DY", ctypes.c_float),
    ("DETA", ctypes.c_float), ("DRAP", ctypes.c_float), ("DTAU", ctypes.c_float),
    ("TAU0", ctypes.c_float), ("EOS_TYPE", ctypes.c_int), ("E_FREEZE", ctypes.c_float),
    ("VISCOUS_MATCHING", ctypes.c_int), ("E_DEP_FS", ctypes.c_int), ("E_R", ctypes.c_float),
    ("TAU_R", ctypes.c_float), ("ALPHA", ctypes.c_float), ("DIM", ctypes.c_int),
    ("TAU", ctypes.c_float),
]


class Params(ctypes.Structure):
    """struct parameters of src/Parameter.h:2-33 (same fields, same order)."""

    _fields_ = _PARAM_FIELDS

    def copy(self) -> "Params":
        q = Params()
        ctypes.memmove(ctypes.byref(q), ctypes.byref(self), ctypes.sizeof(Params))
        return q

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in _PARAM_FIELDS}

    def update(self, **kw) -> "Params":
        for k, v in kw.items():
            if k not in dict(_PARAM_FIELDS):
                raise KeyError(k)
            setattr(self, k, v)
        return self

    def __repr__(self) -> str:  # pragma: no cover
        return "Params(%s)" % ", ".join("%s=%r" % kv for kv in self.as_dict().items())


class PlanInfo(ctypes.Structure):
    """fsm_plan_info: what the planner chose for an event shape."""

    _fields_ = [(n, ctypes.c_int) for n in ("dim3", "variant", "tile_w", "tile_h", "tiles", "n_phi", "n_offsets",
                                            "n_offset_groups", "n_vector_samples", "lds_bytes", "batch_variant",
                                            "batch_tiles", "tuned")] + [("tune_us", ctypes.c_float * 3)] + \
        [(n, ctypes.c_int) for n in ("mfma_batch", "mfma_one")] + \
        [(n, ctypes.c_double) for n in ("samples_kept", "samples_evaluated")] + [("est_batch_us", ctypes.c_float)]

    def as_dict(self):
        d = {n: getattr(self, n) for n, _ in self._fields_}
        d["tune_us"] = [round(float(v), 3) for v in self.tune_us]
        return d


class ServerConfig(ctypes.Structure):
    """fsm_server_config (include/fsm.h)."""

    _fields_ = [("hbm_budget_bytes", ctypes.c_size_t), ("max_batch_events", ctypes.c_int),
                ("max_inflight", ctypes.c_int), ("max_requests", ctypes.c_int), ("max_batch_us", ctypes.c_float)]


class RunReport(ctypes.Structure):
    """fsm_run_report: the values run_freestream_milne prints."""
    _fields_ = [("total_energy_before", ctypes.c_float), ("total_energy_after", ctypes.c_float),
                ("energy_inside", ctypes.c_float), ("R_Pi_Inv_center", ctypes.c_float),
                ("R_pimunu_Inv_center", ctypes.c_float), ("tau_landau_match", ctypes.c_float),
                ("dtau", ctypes.c_float), ("dim_phip", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class DiagOut(ctypes.Structure):
    """fsm_diag_out: device pointers of the diagnostics (0 = not wanted)."""
    _fields_ = [("R_Pi_Inv", ctypes.c_void_p), ("R_pimunu_Inv", ctypes.c_void_p),
                ("pi_dot_u", ctypes.c_void_p * 4), ("pi_mu_mu", ctypes.c_void_p),
                ("beta", ctypes.c_void_p * 4), ("omega", ctypes.c_void_p * 6)]


class FsmError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        super().__init__("%s: %s (status %d)" % (what, _lib().fsm_status_string(status).decode(), status))


_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                "libfsm.so not found at %s; build it with `make -C freestream-milne_amd` "
                "(or __graft_entry__.build())" % LIB_PATH)
        lib = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER(Params)
        vp = ctypes.c_void_p
        fp = ctypes.POINTER(ctypes.c_float)
        lib.fsm_version.restype = ctypes.c_char_p
        lib.fsm_status_string.restype = ctypes.c_char_p
        lib.fsm_status_string.argtypes = [ctypes.c_int]
        lib.fsm_params_default.argtypes = [P]
        lib.fsm_params_default.restype = None
        lib.fsm_params_read.argtypes = [P, ctypes.c_char_p]
        lib.fsm_params_finalize.argtypes = [P, fp, ctypes.c_size_t]
        lib.fsm_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
        lib.fsm_destroy.argtypes = [vp]
        lib.fsm_destroy.restype = None
        lib.fsm_set_path.argtypes = [vp, ctypes.c_int]
        lib.fsm_set_rapidity_rule.argtypes = [vp, ctypes.c_int]
        lib.fsm_gauss_hermite.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_double)]
        lib.fsm_plan_describe.argtypes = [vp, P, ctypes.POINTER(PlanInfo)]
        lib.fsm_set_tile_policy.argtypes = [vp, ctypes.c_int]
        lib.fsm_set_autotune.argtypes = [vp, ctypes.c_int]
        lib.fsm_autotune.argtypes = [vp, P, ctypes.POINTER(ctypes.c_int)]
        lib.fsm_run.argtypes = [vp, P, fp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p), fp]
        lib.fsm_run_baryon.argtypes = [vp, P, fp, fp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p),
                                       ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), fp]
        lib.fsm_run_device.argtypes = [vp, P, vp, vp, vp]
        lib.fsm_run_device_noalloc.argtypes = [vp, P, vp, vp, vp]
        lib.fsm_run_baryon_device.argtypes = [vp, P, vp, vp, vp, vp]
        lib.fsm_stream_tmunu_device.argtypes = [vp, P, vp, vp, vp]
        lib.fsm_landau_device.argtypes = [vp, P, vp, vp, vp, vp, vp, vp, vp]
        lib.fsm_sync.argtypes = [vp, vp]
        lib.fsm_enable_timing.argtypes = [vp, ctypes.c_int]
        lib.fsm_last_kernel_ms.argtypes = [vp, ctypes.c_int, fp]
        lib.fsm_kernel_time_totals.argtypes = [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]
        lib.fsm_batch_create.argtypes = [vp, ctypes.POINTER(vp)]
        lib.fsm_batch_destroy.argtypes = [vp]
        lib.fsm_batch_destroy.restype = None
        lib.fsm_batch_add.argtypes = [vp, P, vp, vp]
        lib.fsm_batch_add_baryon.argtypes = [vp, P, vp, vp, vp]
        lib.fsm_batch_clear.argtypes = [vp]
        lib.fsm_batch_size.argtypes = [vp]
        lib.fsm_batch_set_graph.argtypes = [vp, ctypes.c_int]
        lib.fsm_batch_run.argtypes = [vp, vp]
        lib.fsm_server_create.argtypes = [vp, ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)]
        lib.fsm_server_destroy.argtypes = [vp]
        lib.fsm_server_destroy.restype = None
        lib.fsm_server_submit.argtypes = [vp, P, fp, fp, ctypes.c_size_t]
        lib.fsm_server_submit.restype = ctypes.c_int64
        lib.fsm_server_submit_device.argtypes = [vp, P, vp, vp, ctypes.c_size_t, vp]
        lib.fsm_server_submit_device.restype = ctypes.c_int64
        lib.fsm_server_poll.argtypes = [vp, ctypes.c_int]
        lib.fsm_server_poll.restype = ctypes.c_int64
        lib.fsm_server_result.argtypes = [vp, ctypes.c_int64, ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_void_p)]
        lib.fsm_server_fetch.argtypes = [vp, ctypes.c_int64, ctypes.POINTER(ctypes.c_void_p)]
        lib.fsm_server_fetch_baryon.argtypes = [vp, ctypes.c_int64, ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p,
                                                ctypes.POINTER(ctypes.c_void_p)]
        lib.fsm_server_release.argtypes = [vp, ctypes.c_int64]
        lib.fsm_server_cancel.argtypes = [vp, ctypes.c_int64]
        lib.fsm_server_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                                         ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_size_t)]
        lib.fsm_initial_condition.argtypes = [P, ctypes.c_int, ctypes.c_char_p, fp, ctypes.c_size_t]
        lib.fsm_ic_seed.argtypes = [ctypes.c_uint]
        lib.fsm_ic_seed.restype = None
        lib.fsm_ic_rand.argtypes = []
        lib.fsm_read_profile.argtypes = [P, ctypes.c_int, ctypes.c_char_p, fp, ctypes.c_size_t]
        lib.fsm_write_scalar.argtypes = [P, ctypes.c_char_p, fp]
        lib.fsm_write_projection.argtypes = [P, ctypes.c_char_p, fp]
        lib.fsm_energy_sums.argtypes = [P, fp, fp, fp, fp]
        lib.fsm_write_energy_inside.argtypes = [ctypes.c_char_p, ctypes.c_float]
        lib.fsm_diagnostics_device.argtypes = [vp, P, vp, vp, vp, vp, vp, ctypes.POINTER(DiagOut), vp]
        lib.fsm_run_files.argtypes = [vp, P, fp, fp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p),
                                      ctypes.c_char_p, ctypes.POINTER(RunReport)]
        _LIB = lib
    return _LIB


def lib():
    """The loaded ctypes library (raises ImportError when not built)."""
    return _lib()


def gauss_hermite(n: int):
    """The library's n-point Gauss-Hermite rule (weight exp(-x^2)): (x, w)."""
    x = (ctypes.c_double * n)()
    w = (ctypes.c_double * n)()
    _check(_lib().fsm_gauss_hermite(n, x, w), "fsm_gauss_hermite")
    return np.array(x), np.array(w)


def version() -> str:
    return _lib().fsm_version().decode()


def _check(st: int, what: str):
    if st != FSM_OK:
        raise FsmError(st, what)


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def params_default() -> Params:
    """Built-in defaults, src/FreestreamMilne.cpp:149-171."""
    p = Params()
    _lib().fsm_params_default(ctypes.byref(p))
    return p


def params_read(path: str, p: Params | None = None) -> tuple[Params, int]:
    """Positional freestream_input reader (src/FileIO.cpp:166-245).

    Returns (params, status); a missing file keeps the given/default params
    and returns FSM_ERR_FILE, as the reference keeps its defaults."""
    p = params_default() if p is None else p.copy()
    st = _lib().fsm_params_read(ctypes.byref(p), path.encode())
    return p, st


def params_finalize(p: Params, e: np.ndarray | None = None) -> Params:
    """Derived fields and run-time overrides (src/FreestreamMilne.cpp:177-178,316-332)."""
    q = p.copy()
    if e is None:
        st = _lib().fsm_params_finalize(ctypes.byref(q), None, 0)
    else:
        e = np.ascontiguousarray(e, dtype=np.float32)
        st = _lib().fsm_params_finalize(ctypes.byref(q), _fptr(e), e.size)
    _check(st, "fsm_params_finalize")
    return q


def _stream_handle(stream) -> int | None:
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    h = getattr(stream, "cuda_stream", None)
    if h is not None:
        return h
    raise TypeError("stream must be None, an int handle or a torch.cuda.Stream")


def initial_condition(p: Params, which: int = IC_WHICH_ENERGY, profile_dir: str = "initial_profiles") -> np.ndarray:
    """fsm_initial_condition: the reference's IC dispatch for p.IC_ENERGY
    (which = IC_WHICH_ENERGY, fm^-4) or p.IC_BARYON (IC_WHICH_BARYON, fm^-3)."""
    n = p.DIM_X * p.DIM_Y * p.DIM_ETA
    out = np.zeros(n, dtype=np.float32)
    _check(_lib().fsm_initial_condition(ctypes.byref(p), which, profile_dir.encode(), _fptr(out), n),
           "fsm_initial_condition")
    return out


def ic_seed(seed: int = 1):
    """Restarts the rand() sequence of the fluctuating generators (srand)."""
    _lib().fsm_ic_seed(seed)


def ic_rand() -> int:
    return int(_lib().fsm_ic_rand())


def read_profile(p: Params, fmt: int, path: str) -> np.ndarray:
    """fsm_read_profile: one of the reference's profile file formats."""
    n = p.DIM_X * p.DIM_Y * p.DIM_ETA
    out = np.zeros(n, dtype=np.float32)
    _check(_lib().fsm_read_profile(ctypes.byref(p), fmt, path.encode(), _fptr(out), n), "fsm_read_profile")
    return out


def write_scalar(p: Params, path: str, var: np.ndarray):
    var = np.ascontiguousarray(var, dtype=np.float32)
    _check(_lib().fsm_write_scalar(ctypes.byref(p), path.encode(), _fptr(var)), "fsm_write_scalar")


def write_projection(p: Params, path: str, var: np.ndarray):
    var = np.ascontiguousarray(var, dtype=np.float32)
    _check(_lib().fsm_write_projection(ctypes.byref(p), path.encode(), _fptr(var)), "fsm_write_projection")


def energy_sums(p: Params, T00: np.ndarray, e: np.ndarray):
    """(total energy after streaming, energy inside the freeze-out surface)."""
    T00 = np.ascontiguousarray(T00, dtype=np.float32)
    e = np.ascontiguousarray(e, dtype=np.float32)
    tot, ins = np.zeros(1, np.float32), np.zeros(1, np.float32)
    _check(_lib().fsm_energy_sums(ctypes.byref(p), _fptr(T00), _fptr(e), _fptr(tot), _fptr(ins)), "fsm_energy_sums")
    return float(tot[0]), float(ins[0])


class Context:
    """One device context (fsm_create/fsm_destroy)."""

    def __init__(self, device: int = 0):
        self._ctx = ctypes.c_void_p()
        _check(_lib().fsm_create(device, ctypes.byref(self._ctx)), "fsm_create")
        self.device = device

    def close(self):
        if self._ctx:
            _lib().fsm_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def set_path(self, path: int):
        _check(_lib().fsm_set_path(self._ctx, path), "fsm_set_path")

    def set_rapidity_rule(self, rule: int):
        """RAPIDITY_TAN (the reference's grid) or RAPIDITY_GAUSS_HERMITE (3-D events)."""
        _check(_lib().fsm_set_rapidity_rule(self._ctx, rule), "fsm_set_rapidity_rule")

    # -- host-buffer step ------------------------------------------------
    def run(self, p: Params, e: np.ndarray):
        """Whole step on host buffers.  Returns (outputs dict, finalized params, tau)."""
        e = np.ascontiguousarray(e, dtype=np.float32).ravel()
        q = p.copy()
        n = e.size
        outs = [np.empty(n, dtype=np.float64) for _ in range(NUM_OUTPUTS)]
        arr = (ctypes.c_void_p * NUM_OUTPUTS)(*[o.ctypes.data for o in outs])
        tau = ctypes.c_float()
        st = _lib().fsm_run(self._ctx, ctypes.byref(q), _fptr(e), n, arr, ctypes.byref(tau))
        _check(st, "fsm_run")
        return dict(zip(OUTPUT_NAMES, outs)), q, tau.value

    def run_baryon(self, p: Params, e: np.ndarray, nB: np.ndarray):
        e = np.ascontiguousarray(e, dtype=np.float32).ravel()
        nB = np.ascontiguousarray(nB, dtype=np.float32).ravel()
        q = p.copy()
        n = e.size
        outs = [np.empty(n, dtype=np.float64) for _ in range(NUM_OUTPUTS)]
        nb_out = np.empty(n, dtype=np.float64)
        V = [np.empty(n, dtype=np.float64) for _ in range(4)]
        arr = (ctypes.c_void_p * NUM_OUTPUTS)(*[o.ctypes.data for o in outs])
        varr = (ctypes.c_void_p * 4)(*[v.ctypes.data for v in V])
        tau = ctypes.c_float()
        st = _lib().fsm_run_baryon(self._ctx, ctypes.byref(q), _fptr(e), _fptr(nB), n, arr,
                                   nb_out.ctypes.data, varr, ctypes.byref(tau))
        _check(st, "fsm_run_baryon")
        return dict(zip(OUTPUT_NAMES, outs)), nb_out, V, q, tau.value

    # -- device-resident entry points (raw device pointers) --------------
    def run_device(self, p: Params, d_e: int, d_out: int, stream=None):
        _check(_lib().fsm_run_device(self._ctx, ctypes.byref(p), d_e, d_out, _stream_handle(stream)),
               "fsm_run_device")

    def run_device_noalloc(self, p: Params, d_e: int, d_out: int, stream=None) -> int:
        """fsm_run_device_noalloc: returns the status (FSM_ERR_TOO_LARGE and
        FSM_ERR_NO_PLAN are not raised, so callers can test for them)."""
        return _lib().fsm_run_device_noalloc(self._ctx, ctypes.byref(p), d_e, d_out, _stream_handle(stream))

    def run_baryon_device(self, p: Params, d_e: int, d_nB: int, d_out: int, stream=None):
        """d_out: NUM_OUTPUTS_BARYON planes of DIM doubles (17 outputs, n_B, V^mu)."""
        _check(_lib().fsm_run_baryon_device(self._ctx, ctypes.byref(p), d_e, d_nB, d_out, _stream_handle(stream)),
               "fsm_run_baryon_device")

    def stream_tmunu_device(self, p: Params, d_e: int, d_T: int, stream=None):
        _check(_lib().fsm_stream_tmunu_device(self._ctx, ctypes.byref(p), d_e, d_T, _stream_handle(stream)),
               "fsm_stream_tmunu_device")

    def landau_device(self, p: Params, d_T: int, d_e=None, d_u=None, d_p=None, d_Pi=None, d_pi=None,
                      stream=None):
        _check(_lib().fsm_landau_device(self._ctx, ctypes.byref(p), d_T, d_e, d_u, d_p, d_Pi, d_pi,
                                        _stream_handle(stream)), "fsm_landau_device")

    def diagnostics_device(self, p: Params, d_e: int, d_u: int, d_p: int, d_Pi: int, d_pi: int, out: DiagOut,
                           stream=None):
        _check(_lib().fsm_diagnostics_device(self._ctx, ctypes.byref(p), d_e, d_u, d_p, d_Pi, d_pi,
                                             ctypes.byref(out), _stream_handle(stream)), "fsm_diagnostics_device")

    def run_files(self, p: Params, e: np.ndarray, nB: np.ndarray = None, output_dir: str | None = None):
        """fsm_run_files: the whole reference run (outputs, report, and the
        output/*.dat files when output_dir exists).  Returns (outputs dict,
        finalized params, report dict)."""
        e = np.ascontiguousarray(e, dtype=np.float32).ravel()
        nb = None if nB is None else np.ascontiguousarray(nB, dtype=np.float32).ravel()
        q = p.copy()
        outs = [np.empty(e.size, dtype=np.float64) for _ in range(NUM_OUTPUTS)]
        arr = (ctypes.c_void_p * NUM_OUTPUTS)(*[o.ctypes.data for o in outs])
        rep = RunReport()
        _check(_lib().fsm_run_files(self._ctx, ctypes.byref(q), _fptr(e), None if nb is None else _fptr(nb), e.size,
                                    arr, None if output_dir is None else output_dir.encode(), ctypes.byref(rep)),
               "fsm_run_files")
        return dict(zip(OUTPUT_NAMES, outs)), q, rep.as_dict()

    def sync(self, stream=None):
        _check(_lib().fsm_sync(self._ctx, _stream_handle(stream)), "fsm_sync")

    def enable_timing(self, on: bool = True):
        _check(_lib().fsm_enable_timing(self._ctx, 1 if on else 0), "fsm_enable_timing")

    def kernel_time_totals(self):
        """(ms[3] summed over calls since the last collection, number of calls)."""
        ms = (ctypes.c_double * 3)()
        n = ctypes.c_int()
        _check(_lib().fsm_kernel_time_totals(self._ctx, ms, ctypes.byref(n)), "fsm_kernel_time_totals")
        return list(ms), n.value

    def set_tile_policy(self, policy: int):
        """TILES_AUTO (latency shape for one-event calls) or TILES_BATCH."""
        _check(_lib().fsm_set_tile_policy(self._ctx, policy), "fsm_set_tile_policy")

    def set_autotune(self, on: bool):
        _check(_lib().fsm_set_autotune(self._ctx, 1 if on else 0), "fsm_set_autotune")

    def autotune(self, p: Params) -> int:
        v = ctypes.c_int()
        _check(_lib().fsm_autotune(self._ctx, ctypes.byref(p), ctypes.byref(v)), "fsm_autotune")
        return v.value

    def plan_info(self, p: Params) -> dict:
        """fsm_plan_describe for finalized parameters p (on this context's path)."""
        info = PlanInfo()
        _check(_lib().fsm_plan_describe(self._ctx, ctypes.byref(p), ctypes.byref(info)), "fsm_plan_describe")
        return info.as_dict()

    def last_kernel_ms(self, which: int) -> float:
        v = ctypes.c_float()
        _check(_lib().fsm_last_kernel_ms(self._ctx, which, ctypes.byref(v)), "fsm_last_kernel_ms")
        return v.value


class Batch:
    """Many independent events run with one launch per stage (fsm_batch_*).

    ``add(params_finalized, d_e, d_out)`` takes device pointers (input DIM
    floats in GeV/fm^3, output 17*DIM doubles); ``run(stream)`` replays a HIP
    graph of the whole batch while its composition is unchanged."""

    def __init__(self, ctx: Context):
        self._ctx = ctx
        self._b = ctypes.c_void_p()
        self._params = []  # keep the structs alive while referenced
        _check(_lib().fsm_batch_create(ctx._ctx, ctypes.byref(self._b)), "fsm_batch_create")

    def add(self, p: Params, d_e: int, d_out: int) -> int:
        q = p.copy()
        st = _lib().fsm_batch_add(self._b, ctypes.byref(q), d_e, d_out)
        if st < 0:
            raise FsmError(st, "fsm_batch_add")
        self._params.append(q)
        return st

    def add_baryon(self, p: Params, d_e: int, d_nB: int, d_out: int) -> int:
        """A BARYON = 1 event; d_out holds NUM_OUTPUTS_BARYON planes."""
        q = p.copy()
        st = _lib().fsm_batch_add_baryon(self._b, ctypes.byref(q), d_e, d_nB, d_out)
        if st < 0:
            raise FsmError(st, "fsm_batch_add_baryon")
        self._params.append(q)
        return st

    def clear(self):
        _check(_lib().fsm_batch_clear(self._b), "fsm_batch_clear")
        self._params = []

    def __len__(self):
        return _lib().fsm_batch_size(self._b)

    def set_graph(self, on: bool):
        _check(_lib().fsm_batch_set_graph(self._b, 1 if on else 0), "fsm_batch_set_graph")

    def run(self, stream=None):
        _check(_lib().fsm_batch_run(self._b, _stream_handle(stream)), "fsm_batch_run")

    def close(self):
        if self._b:
            _lib().fsm_batch_destroy(self._b)
            self._b = ctypes.c_void_p()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Request loop with continuous batching (fsm_server_*).

    ``submit(params, e)`` admits one event (host input, GeV/fm^3) and returns a
    request id, or None when the request slots / HBM budget are full
    (FSM_ERR_BUSY).  ``poll(timeout_us)`` returns the id of a completed
    request or None.  Results stay in HBM: ``result(id)`` gives (status,
    latency_us, device pointer of the 17*DIM outputs); ``fetch(id)`` copies
    them to host and releases the request."""

    def __init__(self, ctx: Context, hbm_budget_bytes: int = 0, max_batch_events: int = 0,
                 max_inflight: int = 0, max_requests: int = 0, max_batch_us: float = 0.0):
        self._ctx = ctx
        cfg = ServerConfig(hbm_budget_bytes, max_batch_events, max_inflight, max_requests, max_batch_us)
        self._s = ctypes.c_void_p()
        _check(_lib().fsm_server_create(ctx._ctx, ctypes.byref(cfg), ctypes.byref(self._s)), "fsm_server_create")
        self._dims = {}

    def submit(self, p: Params, e: np.ndarray, nB: np.ndarray = None):
        """nB (initial n_B, fm^-3) for a BARYON = 1 request."""
        e = np.ascontiguousarray(e, dtype=np.float32).ravel()
        nb = None if nB is None else np.ascontiguousarray(nB, dtype=np.float32).ravel()
        q = p.copy()
        rid = _lib().fsm_server_submit(self._s, ctypes.byref(q), _fptr(e), None if nb is None else _fptr(nb), e.size)
        if rid == FSM_ERR_BUSY:
            return None
        if rid < 0:
            raise FsmError(int(rid), "fsm_server_submit")
        self._dims[rid] = e.size
        return rid

    def submit_device(self, p: Params, d_e: int, n: int, d_nB: int = None, stream=None):
        """Same with the input (and n_B) already in device memory (pointers);
        the server copies them after the work queued on `stream`."""
        q = p.copy()
        rid = _lib().fsm_server_submit_device(self._s, ctypes.byref(q), d_e, d_nB, n, _stream_handle(stream))
        if rid == FSM_ERR_BUSY:
            return None
        if rid < 0:
            raise FsmError(int(rid), "fsm_server_submit_device")
        self._dims[rid] = n
        return rid

    def poll(self, timeout_us: int = 0):
        rid = _lib().fsm_server_poll(self._s, timeout_us)
        return None if rid < 0 else rid

    def result(self, rid: int):
        lat = ctypes.c_double()
        ptr = ctypes.c_void_p()
        st = _lib().fsm_server_result(self._s, rid, ctypes.byref(lat), ctypes.byref(ptr))
        return st, lat.value, ptr.value

    def fetch(self, rid: int):
        n = self._dims.pop(rid)
        outs = [np.empty(n, dtype=np.float64) for _ in range(NUM_OUTPUTS)]
        arr = (ctypes.c_void_p * NUM_OUTPUTS)(*[o.ctypes.data for o in outs])
        _check(_lib().fsm_server_fetch(self._s, rid, arr), "fsm_server_fetch")
        return dict(zip(OUTPUT_NAMES, outs))

    def fetch_baryon(self, rid: int):
        """(outputs dict, n_B, [V^tau, V^x, V^y, V^eta]) of a baryon request."""
        n = self._dims.pop(rid)
        outs = [np.empty(n, dtype=np.float64) for _ in range(NUM_OUTPUTS)]
        nb = np.empty(n, dtype=np.float64)
        V = [np.empty(n, dtype=np.float64) for _ in range(4)]
        arr = (ctypes.c_void_p * NUM_OUTPUTS)(*[o.ctypes.data for o in outs])
        varr = (ctypes.c_void_p * 4)(*[v.ctypes.data for v in V])
        _check(_lib().fsm_server_fetch_baryon(self._s, rid, arr, nb.ctypes.data, varr), "fsm_server_fetch_baryon")
        return dict(zip(OUTPUT_NAMES, outs)), nb, V

    def release(self, rid: int):
        self._dims.pop(rid, None)
        _check(_lib().fsm_server_release(self._s, rid), "fsm_server_release")

    def cancel(self, rid: int) -> int:
        return _lib().fsm_server_cancel(self._s, rid)

    def stats(self):
        a, c, l = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        b = ctypes.c_size_t()
        _check(_lib().fsm_server_stats(self._s, ctypes.byref(a), ctypes.byref(c), ctypes.byref(l), ctypes.byref(b)),
               "fsm_server_stats")
        return {"admitted": a.value, "completed": c.value, "launches": l.value, "bytes_in_use": b.value}

    def close(self):
        if self._s:
            _lib().fsm_server_destroy(self._s)
            self._s = ctypes.c_void_p()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


class FreestreamMilne:
    """Mirror of class FREESTREAMMILNE (src/FreestreamMilne.cpp:28-87).

    ``run_freestream_milne`` reads ``./freestream_input`` (or ``input_file``)
    positionally over the built-in defaults, exactly as the reference does,
    dispatches the initial condition on IC_ENERGY / IC_BARYON
    (src/FreestreamMilne.cpp:231-298: 5 = the vector of
    ``initialize_from_vector``, the others from the reference's generators and
    ``initial_profiles/`` files), runs the step on the GPU and writes the
    reference's ``output/*.dat`` files when ``output/`` exists.  A vector
    whose size differs from DIM raises ``FsmError(FSM_ERR_SIZE)`` where the
    reference calls exit(-1) (:259-264); an unknown initial condition raises
    ``FsmError(FSM_ERR_IC)`` where the reference prints "Not a valid initial
    Condition" and returns 0."""

    def __init__(self, device: int = 0, input_file: str = "freestream_input", context: Context | None = None,
                 output_dir: str | None = "output", profile_dir: str = "initial_profiles"):
        self._ctx = context if context is not None else Context(device)
        self.input_file = input_file
        self.output_dir = output_dir
        self.profile_dir = profile_dir
        self.gridSize = 0
        self.tau_LandauMatch = 0.0
        self.init_energy_density = np.zeros(0, dtype=np.float32)
        self.init_baryon_density = None
        self.params = None
        self.report = None
        for name in _FINAL_NAMES:
            setattr(self, name, np.zeros(0, dtype=np.float64))

    def initialize_from_vector(self, energy_density_in):
        self.init_energy_density = np.array(energy_density_in, dtype=np.float32).ravel()

    def run_freestream_milne(self, params: Params | None = None) -> int:
        if params is None:
            params, _ = params_read(self.input_file)
        params = params.copy()
        params.DIM = params.DIM_X * params.DIM_Y * params.DIM_ETA
        if params.IC_ENERGY == 5:
            e = self.init_energy_density
            if e.size != params.DIM:
                raise FsmError(FSM_ERR_SIZE, "run_freestream_milne: vector size %d, DIM %d" % (e.size, params.DIM))
        else:
            e = initial_condition(params, IC_WHICH_ENERGY, self.profile_dir)
        nb = None
        if params.BARYON:
            nb = (self.init_baryon_density if self.init_baryon_density is not None
                  else initial_condition(params, IC_WHICH_BARYON, self.profile_dir))
        outs, q, rep = self._ctx.run_files(params, e, nb, self.output_dir)
        self.params = q
        self.report = rep
        self.tau_LandauMatch = rep["tau_landau_match"]
        for name, key in zip(_FINAL_NAMES, OUTPUT_NAMES):
            setattr(self, name, outs[key])
        return 0

    def output_to_vectors(self):
        return tuple(getattr(self, n).copy() for n in _FINAL_NAMES)


_FINAL_NAMES = (
    "final_energy_density", "final_pressure", "final_ut", "final_ux", "final_uy", "final_un",
    "final_pitt", "final_pitx", "final_pity", "final_pitn", "final_pixx", "final_pixy", "final_pixn",
    "final_piyy", "final_piyn", "final_pinn", "final_Pi",
)


def exported_symbols_from_header(header_path: str) -> list[str]:
    """Function names declared in include/fsm.h (used by the ABI test)."""
    import re

    txt = open(header_path).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fsm_[a-z0-9_]+)\s*\(", txt)))


__all__ = [
    "Params", "Context", "FreestreamMilne", "FsmError", "OUTPUT_NAMES", "NUM_OUTPUTS", "NUM_OUTPUTS_BARYON",
    "params_default",
    "params_read", "params_finalize", "lib", "version", "LIB_PATH", "PATH_FP32", "PATH_MIXED", "TILES_AUTO",
    "TILES_BATCH", "RunReport", "DiagOut", "initial_condition", "ic_seed", "ic_rand", "read_profile", "write_scalar",
    "write_projection", "energy_sums", "IC_WHICH_ENERGY", "IC_WHICH_BARYON", "PROFILE_XYEV", "PROFILE_CPUVH",
    "PROFILE_SUPERMC_BLOCK", "PROFILE_TRENTO_BLOCK", "PROFILE_TRENTO3D_BLOCK",
]
