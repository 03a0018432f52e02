"""Python host binding of libsdc_rbc.so (the C-ABI in include/rbc_abi.h).

The names mirror the reference's operator classes (gpu_flow/code/operator/*.h,
timestepping/implicit_explicit_stepping.h): a `Stepper` owns one rbc context and exposes
step_time (implicit_explicit_step::step_time + make_real_signal), calculate_nonlinear
(nonlinear_operator_rayleigh_noslip::calculate_operator), calculate_inverse
(inverse_B_matrix::calculate_inverse) and calculate_implicit
(linear_implicit_operator::calculate_operator).

Fields are numpy complex128 arrays shaped (L, N, M/2+1): the reference layout
(x fastest, then y, then the vertical mode) read slowest-axis first.  F and G are (L,).

There is no CPU fallback: if the library is missing or no GPU is present the calls raise.
"""
import ctypes
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsdc_rbc.so")

F64, F32, BF16X3 = 0, 1, 2
PRECISIONS = {"f64": F64, "f32": F32, "bf16x3": BF16X3}
NSTAGES = 7


class RbcError(RuntimeError):
    pass


class RbcNumericError(RbcError):
    """RBC_ERR_NUMERIC: a step ended with a non-finite value in the state (the reference's
    find_NaN check, util/util.cu:83-105), or a factor could not be built."""


ERR_NUMERIC = -4


def _raise(rc, msg):
    raise (RbcNumericError if rc == ERR_NUMERIC else RbcError)(msg)


class rbc_params(ctypes.Structure):
    _fields_ = [("M", ctypes.c_int), ("N", ctypes.c_int), ("L", ctypes.c_int),
                ("Lx", ctypes.c_double), ("Ly", ctypes.c_double),
                ("Ra", ctypes.c_double), ("Pr", ctypes.c_double), ("dt", ctypes.c_double),
                ("precision", ctypes.c_int), ("nq", ctypes.c_int), ("device", ctypes.c_int)]


_lib = None

EXPORTS = ["rbc_create", "rbc_destroy", "rbc_last_error", "rbc_get_info", "rbc_set_state",
           "rbc_get_state", "rbc_step", "rbc_reset_history", "rbc_nonlinear", "rbc_inverse_B",
           "rbc_implicit", "rbc_nusselt", "rbc_profile_steps", "rbc_stage_name", "rbc_synchronize",
           "rbc_read_state_files", "rbc_write_state_files", "rbc_get_history", "rbc_set_history",
           "rbc_velocity_grid", "rbc_velocity_at", "rbc_tracer_create", "rbc_tracer_destroy",
           "rbc_tracer_last_error", "rbc_tracer_init", "rbc_tracer_set_positions", "rbc_tracer_step",
           "rbc_tracer_get_positions", "rbc_tracer_current", "rbc_tracer_info", "rbc_tracer_moments",
           "rbc_energy_spectrum", "rbc_energy_spectrum_ref", "rbc_ensemble_create", "rbc_ensemble_destroy",
           "rbc_ensemble_last_error", "rbc_ensemble_size", "rbc_ensemble_member", "rbc_ensemble_tune",
           "rbc_ensemble_step", "rbc_ensemble_batching", "rbc_ensemble_batched_members",
           "rbc_closure_create", "rbc_closure_destroy", "rbc_closure_last_error",
           "rbc_closure_info", "rbc_closure_restrict_state", "rbc_closure_sample", "rbc_closure_points",
           "rbc_closure_dataset", "rbc_closure_load_dataset", "rbc_closure_clear_dataset",
           "rbc_closure_fit_normalization", "rbc_closure_get_normalization", "rbc_closure_set_normalization",
           "rbc_closure_get_weights", "rbc_closure_set_weights", "rbc_closure_gradient", "rbc_closure_train",
           "rbc_closure_predict", "rbc_closure_eval", "rbc_closure_enable", "rbc_closure_interval",
           "rbc_closure_damping", "rbc_closure_step",
           "rbc_closure_status", "rbc_set_state_device", "rbc_get_state_device", "rbc_nonlinear_device",
           "rbc_inverse_B_device", "rbc_implicit_device", "rbc_stream", "rbc_velocity_at_device",
           "rbc_init_state", "rbc_init_pattern", "rbc_force_mean", "rbc_characteristic_velocity", "rbc_run_main",
           "rbc_extract_vel_main", "rbc_velocity_at_device_noalloc", "rbc_set_penalization", "rbc_freeze",
           "rbc_apply_freeze", "rbc_temp_mask", "rbc_apply_temp_mask", "rbc_mask_theta", "rbc_temperature_grid",
           "rbc_stream_packets"]

CLOSURE_NIN, CLOSURE_NOUT = 8, 3

TRACER_MODES = {"rk4_trilinear": 0, "spectral_rk2": 1}
PART_INIT = {"random": 0, "center": 1, "center_sigma": 2, "diagonal": 3, "pairs": 4, "coord_sigma": 5}


def load_library(path=LIB_PATH):
    """Loads libsdc_rbc.so and declares the C signatures.  Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RbcError(f"{path} is not built; run `python -m sdc_amd.build`")
    L = ctypes.CDLL(path)
    dp = ctypes.POINTER(ctypes.c_double)
    vp = ctypes.c_void_p
    L.rbc_create.argtypes = [ctypes.POINTER(rbc_params), ctypes.POINTER(vp)]
    L.rbc_create.restype = ctypes.c_int
    L.rbc_destroy.argtypes = [vp]
    L.rbc_destroy.restype = None
    L.rbc_last_error.argtypes = [vp]
    L.rbc_last_error.restype = ctypes.c_char_p
    ip = ctypes.POINTER(ctypes.c_int)
    L.rbc_get_info.argtypes = [vp, ip, ip, ip, ip]
    for fn in ("rbc_set_state", "rbc_get_state"):
        getattr(L, fn).argtypes = [vp] + [dp] * 5
    L.rbc_step.argtypes = [vp, ctypes.c_int]
    L.rbc_reset_history.argtypes = [vp]
    for fn in ("rbc_nonlinear", "rbc_inverse_B", "rbc_implicit"):
        getattr(L, fn).argtypes = [vp] + [dp] * 10
    L.rbc_nusselt.argtypes = [vp, dp]
    for fn in ("rbc_set_state_device", "rbc_get_state_device"):
        getattr(L, fn).argtypes = [vp] * 6
    for fn in ("rbc_nonlinear_device", "rbc_inverse_B_device", "rbc_implicit_device"):
        getattr(L, fn).argtypes = [vp] * 11
    L.rbc_stream.argtypes = [vp]
    L.rbc_stream.restype = vp
    L.rbc_init_state.argtypes = [vp, ctypes.c_int, ctypes.c_ulonglong, ctypes.c_double]
    L.rbc_init_pattern.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_ulonglong]
    L.rbc_characteristic_velocity.argtypes = [vp, ctypes.c_int, dp]
    L.rbc_set_penalization.argtypes = [vp, ctypes.c_double, ctypes.c_int, dp]
    L.rbc_freeze.argtypes = [vp, ctypes.c_int, ctypes.c_double]
    L.rbc_apply_freeze.argtypes = [vp]
    L.rbc_temp_mask.argtypes = [vp, ctypes.c_int, ctypes.c_double]
    L.rbc_apply_temp_mask.argtypes = [vp]
    L.rbc_mask_theta.argtypes = [vp] + [ctypes.c_double] * 4
    L.rbc_temperature_grid.argtypes = [vp, ctypes.c_int, dp, ctypes.POINTER(ctypes.c_float)]
    L.rbc_stream_packets.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                     ctypes.c_double, vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
    L.rbc_run_main.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p)]
    L.rbc_extract_vel_main.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p)]
    L.rbc_force_mean.argtypes = [vp, ctypes.c_double]
    L.rbc_profile_steps.argtypes = [vp, ctypes.c_int, dp]
    L.rbc_stage_name.argtypes = [ctypes.c_int]
    L.rbc_stage_name.restype = ctypes.c_char_p
    L.rbc_synchronize.argtypes = [vp]
    cpp = ctypes.POINTER(ctypes.c_char_p)
    fp = ctypes.POINTER(ctypes.c_float)
    L.rbc_read_state_files.argtypes = [vp, cpp, ctypes.c_int]
    L.rbc_write_state_files.argtypes = [vp, cpp, ctypes.c_int, ctypes.c_int]
    L.rbc_get_history.argtypes = [vp] + [dp] * 5
    L.rbc_set_history.argtypes = [vp] + [dp] * 5
    L.rbc_velocity_grid.argtypes = [vp, ctypes.c_int, dp, fp, fp, fp]
    L.rbc_velocity_at.argtypes = [vp, ctypes.c_int, fp, fp, fp, fp, fp, fp]
    L.rbc_velocity_at_device.argtypes = [vp, ctypes.c_int] + [vp] * 6
    L.rbc_velocity_at_device_noalloc.argtypes = [vp, ctypes.c_int] + [vp] * 6
    L.rbc_tracer_create.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                    ctypes.c_ulonglong, ctypes.POINTER(vp)]
    L.rbc_tracer_destroy.argtypes = [vp]
    L.rbc_tracer_destroy.restype = None
    L.rbc_tracer_last_error.argtypes = [vp]
    L.rbc_tracer_last_error.restype = ctypes.c_char_p
    L.rbc_tracer_init.argtypes = [vp, ctypes.c_int] + [ctypes.c_double] * 4
    L.rbc_tracer_set_positions.argtypes = [vp, fp, fp, fp]
    L.rbc_tracer_step.argtypes = [vp, ctypes.c_double]
    L.rbc_tracer_get_positions.argtypes = [vp, ctypes.c_int, fp, fp, fp, ip]
    L.rbc_tracer_current.argtypes = [vp, fp, fp, fp]
    L.rbc_tracer_info.argtypes = [vp, ip, ip]
    L.rbc_tracer_moments.argtypes = [vp, dp]
    L.rbc_energy_spectrum.argtypes = [vp, ctypes.c_int, dp, ip, dp]
    L.rbc_energy_spectrum_ref.argtypes = [vp, ctypes.c_double, ctypes.c_int, dp, ip]
    L.rbc_ensemble_create.argtypes = [ctypes.c_int, ctypes.POINTER(rbc_params), ctypes.POINTER(vp)]
    L.rbc_ensemble_destroy.argtypes = [vp]
    L.rbc_ensemble_destroy.restype = None
    L.rbc_ensemble_last_error.argtypes = [vp]
    L.rbc_ensemble_last_error.restype = ctypes.c_char_p
    L.rbc_ensemble_size.argtypes = [vp]
    L.rbc_ensemble_member.argtypes = [vp, ctypes.c_int]
    L.rbc_ensemble_member.restype = vp
    L.rbc_ensemble_tune.argtypes = [vp, dp]
    L.rbc_ensemble_step.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.rbc_ensemble_batching.argtypes = [vp, ctypes.c_int]
    L.rbc_ensemble_batched_members.argtypes = [vp]
    i_ = ctypes.c_int
    L.rbc_closure_create.argtypes = [vp, i_, i_, ctypes.POINTER(vp)]
    L.rbc_closure_destroy.argtypes = [vp]
    L.rbc_closure_destroy.restype = None
    L.rbc_closure_last_error.argtypes = [vp]
    L.rbc_closure_last_error.restype = ctypes.c_char_p
    L.rbc_closure_info.argtypes = [vp, ip, ip, ip, ip]
    L.rbc_closure_restrict_state.argtypes = [vp, vp]
    L.rbc_closure_sample.argtypes = [vp, vp, ip]
    L.rbc_closure_points.argtypes = [vp, vp, fp, fp]
    L.rbc_closure_dataset.argtypes = [vp, i_, i_, fp, fp]
    L.rbc_closure_load_dataset.argtypes = [vp, i_, fp, fp]
    L.rbc_closure_clear_dataset.argtypes = [vp]
    L.rbc_closure_fit_normalization.argtypes = [vp, i_]
    L.rbc_closure_get_normalization.argtypes = [vp, fp, fp, fp]
    L.rbc_closure_set_normalization.argtypes = [vp, fp, fp, fp]
    L.rbc_closure_get_weights.argtypes = [vp, fp]
    L.rbc_closure_set_weights.argtypes = [vp, fp]
    L.rbc_closure_gradient.argtypes = [vp, i_, fp, fp, fp, dp]
    L.rbc_closure_train.argtypes = [vp, i_, i_, ctypes.c_double, ctypes.c_ulonglong, i_, i_, dp]
    L.rbc_closure_predict.argtypes = [vp, i_, fp, fp]
    L.rbc_closure_eval.argtypes = [vp, i_, i_, dp, dp]
    L.rbc_closure_enable.argtypes = [vp, i_, ctypes.c_double]
    L.rbc_closure_interval.argtypes = [vp, i_]
    L.rbc_closure_damping.argtypes = [vp, ctypes.c_double, ctypes.c_double]
    L.rbc_closure_step.argtypes = [vp, i_, i_]
    L.rbc_closure_status.argtypes = [vp, ip, ip, dp]
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _paths(paths):
    if len(paths) != 5:
        raise ValueError("five paths are needed: theta, f, g, F, G")
    return (ctypes.c_char_p * 5)(*[os.fsencode(p) for p in paths])


class Stepper:
    """One RBC context on one GPU (≙ the operator set main.cu:431-454 builds)."""

    def __init__(self, M, N, L, Lx, Ly, Ra, Pr, dt, precision="f64", nq=0, device=0, _handle=None):
        self.lib = load_library()
        self.M, self.N, self.L = M, N, L
        self.mox = M // 2 + 1
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        self.params = rbc_params(M, N, L, Lx, Ly, Ra, Pr, dt, prec, nq, device)
        self._owned = _handle is None
        self._dependents = []  # tracers and closures, closed before this context
        if _handle is None:
            h = ctypes.c_void_p()
            rc = self.lib.rbc_create(ctypes.byref(self.params), ctypes.byref(h))
            if rc != 0:
                raise RbcError(f"rbc_create failed ({rc}): {self.lib.rbc_last_error(None).decode()}")
            self.h = h
        else:
            self.h = ctypes.c_void_p(_handle)
        nq_, kx, ky, ng = (ctypes.c_int() for _ in range(4))
        self.lib.rbc_get_info(self.h, ctypes.byref(nq_), ctypes.byref(kx), ctypes.byref(ky), ctypes.byref(ng))
        self.nq, self.kx_kept, self.ky_kept, self.n_unique_k = nq_.value, kx.value, ky.value, ng.value

    def _check(self, rc, what):
        if rc != 0:
            _raise(rc, f"{what} failed ({rc}): {self.lib.rbc_last_error(self.h).decode()}")

    def close(self):
        for ref in getattr(self, "_dependents", []):
            dep = ref()
            if dep is not None:
                dep.close()
        self._dependents = []
        if getattr(self, "h", None):
            if self._owned:
                self.lib.rbc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def zeros(self):
        return [np.zeros((self.L, self.N, self.mox), np.complex128) for _ in range(3)] + \
               [np.zeros(self.L, np.complex128) for _ in range(2)]

    def set_state(self, theta, f, g, F, G):
        arrs = [np.ascontiguousarray(a, np.complex128) for a in (theta, f, g, F, G)]
        self._check(self.lib.rbc_set_state(self.h, *[_dp(a.view(np.float64)) for a in arrs]), "rbc_set_state")

    def get_state(self):
        out = self.zeros()
        self._check(self.lib.rbc_get_state(self.h, *[_dp(a.view(np.float64)) for a in out]), "rbc_get_state")
        return out

    def step_time(self, nsteps=1):
        self._check(self.lib.rbc_step(self.h, int(nsteps)), "rbc_step")

    def synchronize(self):
        self._check(self.lib.rbc_synchronize(self.h), "rbc_synchronize")

    def reset_history(self):
        self._check(self.lib.rbc_reset_history(self.h), "rbc_reset_history")

    def _op(self, fn, name, ins):
        ins = [np.ascontiguousarray(a, np.complex128) for a in ins]
        out = self.zeros()
        self._check(fn(self.h, *[_dp(a.view(np.float64)) for a in ins], *[_dp(a.view(np.float64)) for a in out]),
                    name)
        return out

    def calculate_nonlinear(self, theta, f, g, F, G):
        return self._op(self.lib.rbc_nonlinear, "rbc_nonlinear", (theta, f, g, F, G))

    def calculate_inverse(self, theta, f, g, F, G):
        return self._op(self.lib.rbc_inverse_B, "rbc_inverse_B", (theta, f, g, F, G))

    def calculate_implicit(self, theta, f, g, F, G):
        return self._op(self.lib.rbc_implicit, "rbc_implicit", (theta, f, g, F, G))

    # ---- device-resident exchange (device pointers in the reference layout) ----------------
    # Each argument is an integer device address (e.g. torch.Tensor.data_ptr()) or None; the
    # element type is complex128 for precision "f64", complex64 otherwise.  The work is queued
    # on the context's stream: call synchronize() before reading the outputs.
    @staticmethod
    def _ptrs(ptrs):
        return [ctypes.c_void_p(int(p)) if p else None for p in ptrs]

    def set_state_device(self, theta, f, g, F, G):
        self._check(self.lib.rbc_set_state_device(self.h, *self._ptrs((theta, f, g, F, G))), "rbc_set_state_device")

    def get_state_device(self, theta, f, g, F, G):
        self._check(self.lib.rbc_get_state_device(self.h, *self._ptrs((theta, f, g, F, G))), "rbc_get_state_device")

    def calculate_nonlinear_device(self, ins, outs):
        self._check(self.lib.rbc_nonlinear_device(self.h, *self._ptrs(ins), *self._ptrs(outs)), "rbc_nonlinear_device")

    def calculate_inverse_device(self, ins, outs):
        self._check(self.lib.rbc_inverse_B_device(self.h, *self._ptrs(ins), *self._ptrs(outs)), "rbc_inverse_B_device")

    def calculate_implicit_device(self, ins, outs):
        self._check(self.lib.rbc_implicit_device(self.h, *self._ptrs(ins), *self._ptrs(outs)), "rbc_implicit_device")

    INIT_KINDS = {"zero": 0, "random": 1, "test": 2}

    def init_state(self, kind="random", seed=0, amplitude=1e-6):
        """Initial state on the device (rbc_init_state): "zero", "random" (init_random's
        distribution, project RNG) or "test" (init_test); make_real applied, history cleared."""
        k = self.INIT_KINDS[kind] if isinstance(kind, str) else int(kind)
        self._check(self.lib.rbc_init_state(self.h, k, int(seed), float(amplitude)), "rbc_init_state")

    PATTERNS = {"rectangle": 1, "dislocation": 2, "isr": 3, "sv": 4, "cr": 5, "giantspiral": 6}

    def init_pattern(self, kind, init_mode=None, seed=0):
        """Pattern start of -init_rectangle / -init_dislocation / -init_isr / -init_sv / -init_cr /
        -init_giantspiral (rbc_init_pattern): random theta at 1e-6 with vertical mode 0 replaced by
        the pattern's unnormalised R2C; init_mode defaults to Lx / 2 as in the reference."""
        k = self.PATTERNS[kind] if isinstance(kind, str) else int(kind)
        mode = self.params.Lx / 2.0 if init_mode is None else float(init_mode)
        self._check(self.lib.rbc_init_pattern(self.h, k, mode, int(seed)), "rbc_init_pattern")

    MASKS = {"rectangle": 0, "circle": 1, "diskarray": 2, "everywhere": 3, "custom": 4}
    PLANES = {"circle": 0, "rectangle": 1}

    def set_penalization(self, eta, shape="rectangle", mask=None):
        """Brinkman penalization of a PENALIZED build (eta = 0: off); mask (N, M) for "custom"."""
        sh = self.MASKS[shape] if isinstance(shape, str) else int(shape)
        m = None
        if mask is not None:
            m = np.ascontiguousarray(mask, np.float64)
            assert m.shape == (self.N, self.M)
        self._check(self.lib.rbc_set_penalization(self.h, float(eta), sh, _dp(m) if m is not None else None),
                    "rbc_set_penalization")

    def freeze(self, shape="circle", param=0.8):
        self._check(self.lib.rbc_freeze(self.h, self.PLANES[shape], float(param)), "rbc_freeze")

    def apply_freeze(self):
        self._check(self.lib.rbc_apply_freeze(self.h), "rbc_apply_freeze")

    def temp_mask(self, shape="rectangle", param=0.1):
        self._check(self.lib.rbc_temp_mask(self.h, self.PLANES[shape], float(param)), "rbc_temp_mask")

    def apply_temp_mask(self):
        self._check(self.lib.rbc_apply_temp_mask(self.h), "rbc_apply_temp_mask")

    def mask_theta(self, cx, cy, radius, factor):
        self._check(self.lib.rbc_mask_theta(self.h, float(cx), float(cy), float(radius), float(factor)),
                    "rbc_mask_theta")

    def force_mean(self, mean_F):
        """F_0 = mean_F (force_mean, util/util.cu:74-81), queued on the context's stream."""
        self._check(self.lib.rbc_force_mean(self.h, float(mean_F)), "rbc_force_mean")

    @property
    def stream(self):
        """The context's hipStream_t as an integer."""
        return int(self.lib.rbc_stream(self.h) or 0)

    def nusselt(self):
        nu = np.zeros(3)
        self._check(self.lib.rbc_nusselt(self.h, _dp(nu)), "rbc_nusselt")
        return nu

    def profile_steps(self, nsteps):
        ms = np.zeros(NSTAGES)
        self._check(self.lib.rbc_profile_steps(self.h, int(nsteps), _dp(ms)), "rbc_profile_steps")
        names = [self.lib.rbc_stage_name(i).decode() for i in range(NSTAGES)]
        return dict(zip(names, ms.tolist()))

    # ---- state files and AB2 history -------------------------------------------------------
    def read_state_files(self, paths, which=0):
        """Reads theta, f, g, F, G files in the reference format (which=1: the AB2 history)."""
        self._check(self.lib.rbc_read_state_files(self.h, _paths(paths), int(which)), "rbc_read_state_files")

    def write_state_files(self, paths, which=0, elem_bytes=0):
        self._check(self.lib.rbc_write_state_files(self.h, _paths(paths), int(which), int(elem_bytes)),
                    "rbc_write_state_files")

    def get_history(self):
        out = self.zeros()
        self._check(self.lib.rbc_get_history(self.h, *[_dp(a.view(np.float64)) for a in out]), "rbc_get_history")
        return out

    def set_history(self, theta, f, g, F, G):
        arrs = [np.ascontiguousarray(a, np.complex128) for a in (theta, f, g, F, G)]
        self._check(self.lib.rbc_set_history(self.h, *[_dp(a.view(np.float64)) for a in arrs]), "rbc_set_history")

    # ---- velocity and spectra ----------------------------------------------------------------
    def velocity_grid(self, z):
        """(u_x, u_y, u_z), each float32 (nz, N, M), of the current state at heights z."""
        z = np.ascontiguousarray(z, np.float64)
        out = [np.zeros((len(z), self.N, self.M), np.float32) for _ in range(3)]
        self._check(self.lib.rbc_velocity_grid(self.h, len(z), _dp(z), *[_fp(a) for a in out]), "rbc_velocity_grid")
        return out

    def velocity_at(self, x, y, z):
        x, y, z = (np.ascontiguousarray(a, np.float32) for a in (x, y, z))
        out = [np.zeros(len(x), np.float32) for _ in range(3)]
        self._check(self.lib.rbc_velocity_at(self.h, len(x), _fp(x), _fp(y), _fp(z), *[_fp(a) for a in out]),
                    "rbc_velocity_at")
        return out

    def velocity_at_device(self, n, x, y, z, u, v, w):
        """Velocity at n points on device float32 arrays (integer addresses); queued on the
        context's stream (synchronize() before reading u, v, w)."""
        self._check(self.lib.rbc_velocity_at_device(self.h, int(n), *self._ptrs((x, y, z, u, v, w))),
                    "rbc_velocity_at_device")

    def characteristic_velocity(self, nz=10):
        """<|u|> over the grid at nz heights linspace(-1/2, 1/2, nz) (main_extract_vel.cu)."""
        U = ctypes.c_double()
        self._check(self.lib.rbc_characteristic_velocity(self.h, int(nz), ctypes.byref(U)),
                    "rbc_characteristic_velocity")
        return U.value

    def velocity_at_device_noalloc(self, n, x, y, z, u, v, w):
        """velocity_at_device with the partial sums in the library's per-GPU static scratch (no
        allocation; n <= 1,398,101; overlapping calls on one GPU race)."""
        self._check(self.lib.rbc_velocity_at_device_noalloc(self.h, int(n), *self._ptrs((x, y, z, u, v, w))),
                    "rbc_velocity_at_device_noalloc")

    def temperature_grid(self, z):
        """theta(x, y, z) at heights z, float32 (nz, N, M) (calculate_temperature_operator)."""
        z = np.ascontiguousarray(z, np.float64)
        out = np.zeros((len(z), self.N, self.M), np.float32)
        self._check(self.lib.rbc_temperature_grid(self.h, len(z), _dp(z), _fp(out)), "rbc_temperature_grid")
        return out

    IMG = {"temperature": 1, "particles": 2, "temp_particles": 3}

    def stream_packets(self, img="temperature", res=0, Nt=0, eta=0.0, t_final=0.0, tracer=None):
        """The visual-stream packets of one frame (server_writer::Prepare) as bytes."""
        k = self.IMG[img] if isinstance(img, str) else int(img)
        th = tracer.h if tracer is not None else None
        n = ctypes.c_size_t()
        self._check(self.lib.rbc_stream_packets(self.h, th, k, int(res), int(Nt), float(eta), float(t_final), None, 0,
                                                ctypes.byref(n)), "rbc_stream_packets")
        buf = ctypes.create_string_buffer(n.value)
        self._check(self.lib.rbc_stream_packets(self.h, th, k, int(res), int(Nt), float(eta), float(t_final), buf,
                                                n.value, ctypes.byref(n)), "rbc_stream_packets")
        return buf.raw[:n.value]

    def energy_spectrum(self):
        """(E_h per |k| shell, (E_h total, E_z total)), FP64."""
        n = ctypes.c_int()
        self._check(self.lib.rbc_energy_spectrum(self.h, 0, None, ctypes.byref(n), None), "rbc_energy_spectrum")
        E = np.zeros(n.value)
        tot = np.zeros(2)
        self._check(self.lib.rbc_energy_spectrum(self.h, n.value, _dp(E), ctypes.byref(n), _dp(tot)),
                    "rbc_energy_spectrum")
        return E, tot

    def energy_spectrum_ref(self, z):
        """The reference's binning of |c_k|^2 at height z: array (3, nbins) for u_x, u_y, u_z."""
        n = ctypes.c_int()
        self._check(self.lib.rbc_energy_spectrum_ref(self.h, float(z), 0, None, ctypes.byref(n)),
                    "rbc_energy_spectrum_ref")
        E = np.zeros(3 * n.value)
        self._check(self.lib.rbc_energy_spectrum_ref(self.h, float(z), n.value, _dp(E), ctypes.byref(n)),
                    "rbc_energy_spectrum_ref")
        return E.reshape(3, n.value)


class Tracer:
    """Passive tracers advected by a Stepper's velocity (≙ particle_tracer,
    particle/particle_tracer.h:21-89).  mode "rk4_trilinear" is the reference's scheme,
    "spectral_rk2" evaluates the velocity exactly at each particle."""

    def __init__(self, stepper, num_particles, buffer_size=256, mode="rk4_trilinear", diff_coeff=0.0, seed=0):
        self.lib = stepper.lib
        self.stepper = stepper
        self.np = int(num_particles)
        self.buffer_size = int(buffer_size)
        m = TRACER_MODES[mode] if isinstance(mode, str) else int(mode)
        h = ctypes.c_void_p()
        rc = self.lib.rbc_tracer_create(stepper.h, self.np, self.buffer_size, m, float(diff_coeff), int(seed),
                                        ctypes.byref(h))
        if rc != 0:
            raise RbcError(f"rbc_tracer_create failed ({rc}): {self.lib.rbc_last_error(stepper.h).decode()}")
        self.h = h
        stepper._dependents.append(weakref.ref(self))

    def _check(self, rc, what):
        if rc != 0:
            raise RbcError(f"{what} failed ({rc}): {self.lib.rbc_tracer_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.rbc_tracer_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init_positions(self, kind="random", coord=(0.0, 0.0, 0.0), sigma=0.0):
        t = PART_INIT[kind] if isinstance(kind, str) else int(kind)
        self._check(self.lib.rbc_tracer_init(self.h, t, float(coord[0]), float(coord[1]), float(coord[2]),
                                             float(sigma)), "rbc_tracer_init")

    def set_positions(self, x, y, z):
        x, y, z = (np.ascontiguousarray(a, np.float32) for a in (x, y, z))
        self._check(self.lib.rbc_tracer_set_positions(self.h, _fp(x), _fp(y), _fp(z)), "rbc_tracer_set_positions")

    def step_particles(self, dt):
        self._check(self.lib.rbc_tracer_step(self.h, float(dt)), "rbc_tracer_step")

    def positions(self):
        out = [np.zeros(self.np, np.float32) for _ in range(3)]
        self._check(self.lib.rbc_tracer_current(self.h, *[_fp(a) for a in out]), "rbc_tracer_current")
        return out

    def get_particle_positions(self, clear=True):
        """(x, y, z) arrays shaped (num_timesteps, num_particles)."""
        cap = (self.buffer_size - 1) * self.np
        out = [np.zeros(cap, np.float32) for _ in range(3)]
        nt = ctypes.c_int()
        self._check(self.lib.rbc_tracer_get_positions(self.h, int(bool(clear)), *[_fp(a) for a in out],
                                                      ctypes.byref(nt)), "rbc_tracer_get_positions")
        return [a[:nt.value * self.np].reshape(nt.value, self.np) for a in out]

    def info(self):
        it, pt = ctypes.c_int(), ctypes.c_int()
        self._check(self.lib.rbc_tracer_info(self.h, ctypes.byref(it), ctypes.byref(pt)), "rbc_tracer_info")
        return it.value, pt.value

    def moments(self):
        out = np.zeros(4)
        self._check(self.lib.rbc_tracer_moments(self.h, _dp(out)), "rbc_tracer_moments")
        return out


class Ensemble:
    """Independent runs of uneven size on one device, stepped together through a captured HIP
    graph (rbc_ensemble_*).  members[i] is a Stepper view of member i (owned by the ensemble)."""

    def __init__(self, configs, device=0):
        """configs: list of dicts with M, N, L, Lx, Ly, Ra, Pr, dt and optional precision, nq."""
        self.lib = load_library()
        n = len(configs)
        arr = (rbc_params * n)()
        for i, c in enumerate(configs):
            prec = c.get("precision", "f32")
            prec = PRECISIONS[prec] if isinstance(prec, str) else int(prec)
            arr[i] = rbc_params(c["M"], c["N"], c["L"], c["Lx"], c["Ly"], c["Ra"], c["Pr"], c["dt"], prec,
                                c.get("nq", 0), device)
        h = ctypes.c_void_p()
        rc = self.lib.rbc_ensemble_create(n, arr, ctypes.byref(h))
        if rc != 0:
            raise RbcError(f"rbc_ensemble_create failed ({rc}): {self.lib.rbc_last_error(None).decode()}")
        self.h = h
        self.members = []
        for i, c in enumerate(configs):
            mh = self.lib.rbc_ensemble_member(self.h, i)
            self.members.append(Stepper(c["M"], c["N"], c["L"], c["Lx"], c["Ly"], c["Ra"], c["Pr"], c["dt"],
                                        precision=c.get("precision", "f32"), device=device, _handle=mh))

    def _check(self, rc, what):
        if rc != 0:
            raise RbcError(f"{what} failed ({rc}): {self.lib.rbc_ensemble_last_error(self.h).decode()}")

    def tune(self):
        rep = np.zeros(3 * len(self.members))
        self._check(self.lib.rbc_ensemble_tune(self.h, _dp(rep)), "rbc_ensemble_tune")
        return rep.reshape(-1, 3)

    def step(self, nsteps, steps_per_launch=10):
        self._check(self.lib.rbc_ensemble_step(self.h, int(nsteps), int(steps_per_launch)), "rbc_ensemble_step")

    def set_batching(self, on):
        """Size classes stepped as batches inside the graph (True, the default) or every member as
        a graph branch of its own (False)."""
        self._check(self.lib.rbc_ensemble_batching(self.h, 1 if on else 0), "rbc_ensemble_batching")

    @property
    def batched_members(self):
        """Members the current graph steps in batches (0 before the first graph launch)."""
        return int(self.lib.rbc_ensemble_batched_members(self.h))

    def close(self):
        if getattr(self, "h", None):
            for m in self.members:
                m.close()  # closes its tracers / closures; the context itself belongs to the ensemble
            self.lib.rbc_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Closure:
    """Learned closure attached to a coarse Stepper (rbc_closure_*): dataset generation from a
    finer Stepper, GPU training, and corrected coarse steps with fallback to the plain method."""

    def __init__(self, coarse, nz=0, capacity_points=0):
        self.lib = coarse.lib
        self.coarse = coarse
        h = ctypes.c_void_p()
        rc = self.lib.rbc_closure_create(coarse.h, int(nz), int(capacity_points), ctypes.byref(h))
        if rc != 0:
            raise RbcError(f"rbc_closure_create failed ({rc}): {self.lib.rbc_last_error(coarse.h).decode()}")
        self.h = h
        coarse._dependents.append(weakref.ref(self))
        a, b, c, d = (ctypes.c_int() for _ in range(4))
        self.lib.rbc_closure_info(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), ctypes.byref(d))
        self.npts, self.nweights, self.nz = a.value, b.value, d.value

    def _check(self, rc, what):
        if rc != 0:
            raise RbcError(f"{what} failed ({rc}): {self.lib.rbc_closure_last_error(self.h).decode()}")

    @property
    def count(self):
        c = ctypes.c_int()
        self.lib.rbc_closure_info(self.h, None, None, ctypes.byref(c), None)
        return c.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.rbc_closure_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def restrict_state(self, fine):
        self._check(self.lib.rbc_closure_restrict_state(self.h, fine.h), "rbc_closure_restrict_state")

    def sample(self, fine):
        n = ctypes.c_int()
        self._check(self.lib.rbc_closure_sample(self.h, fine.h, ctypes.byref(n)), "rbc_closure_sample")
        return n.value

    def points(self, fine):
        X = np.zeros((self.npts, CLOSURE_NIN), np.float32)
        Y = np.zeros((self.npts, CLOSURE_NOUT), np.float32)
        self._check(self.lib.rbc_closure_points(self.h, fine.h, _fp(X), _fp(Y)), "rbc_closure_points")
        return X, Y

    def dataset(self, first=0, n=None):
        n = self.count - first if n is None else n
        X = np.zeros((n, CLOSURE_NIN), np.float32)
        Y = np.zeros((n, CLOSURE_NOUT), np.float32)
        self._check(self.lib.rbc_closure_dataset(self.h, int(first), int(n), _fp(X), _fp(Y)), "rbc_closure_dataset")
        return X, Y

    def load_dataset(self, X, Y):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        self._check(self.lib.rbc_closure_load_dataset(self.h, len(X), _fp(X), _fp(Y)), "rbc_closure_load_dataset")

    def clear_dataset(self):
        self._check(self.lib.rbc_closure_clear_dataset(self.h), "rbc_closure_clear_dataset")

    def fit_normalization(self, n=0):
        self._check(self.lib.rbc_closure_fit_normalization(self.h, int(n)), "rbc_closure_fit_normalization")

    def normalization(self):
        m = np.zeros(CLOSURE_NIN, np.float32)
        i = np.zeros(CLOSURE_NIN, np.float32)
        s = np.zeros(CLOSURE_NOUT, np.float32)
        self._check(self.lib.rbc_closure_get_normalization(self.h, _fp(m), _fp(i), _fp(s)),
                    "rbc_closure_get_normalization")
        return m, i, s

    def set_normalization(self, mean, inv, scale):
        m, i, s = (np.ascontiguousarray(a, np.float32) for a in (mean, inv, scale))
        self._check(self.lib.rbc_closure_set_normalization(self.h, _fp(m), _fp(i), _fp(s)),
                    "rbc_closure_set_normalization")

    def weights(self):
        w = np.zeros(self.nweights, np.float32)
        self._check(self.lib.rbc_closure_get_weights(self.h, _fp(w)), "rbc_closure_get_weights")
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        assert w.size == self.nweights
        self._check(self.lib.rbc_closure_set_weights(self.h, _fp(w)), "rbc_closure_set_weights")

    def gradient(self, X, Y):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        g = np.zeros(self.nweights, np.float32)
        loss = ctypes.c_double()
        self._check(self.lib.rbc_closure_gradient(self.h, len(X), _fp(X), _fp(Y), _fp(g), ctypes.byref(loss)),
                    "rbc_closure_gradient")
        return g, loss.value

    def train(self, steps, batch=65536, lr=1e-3, seed=0, first=0, n=0):
        loss = np.zeros(max(1, steps))
        self._check(self.lib.rbc_closure_train(self.h, int(steps), int(batch), float(lr), int(seed), int(first),
                                               int(n), _dp(loss)), "rbc_closure_train")
        return loss[:steps]

    def predict(self, X):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.zeros((len(X), CLOSURE_NOUT), np.float32)
        self._check(self.lib.rbc_closure_predict(self.h, len(X), _fp(X), _fp(Y)), "rbc_closure_predict")
        return Y

    def evaluate(self, first, n):
        a, b = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.rbc_closure_eval(self.h, int(first), int(n), ctypes.byref(a), ctypes.byref(b)),
                    "rbc_closure_eval")
        return a.value, b.value

    def enable(self, on=True, max_ratio=1.0):
        self._check(self.lib.rbc_closure_enable(self.h, int(bool(on)), float(max_ratio)), "rbc_closure_enable")

    def set_interval(self, every):
        """Evaluate the network on every `every`-th corrected step; the steps between reuse the
        last evaluated correction."""
        self._check(self.lib.rbc_closure_interval(self.h, int(every)), "rbc_closure_interval")

    def set_damping(self, gain=1.0, kfrac=1.0):
        """Add gain times the correction on the modes with |k| <= kfrac * k_max, none above."""
        self._check(self.lib.rbc_closure_damping(self.h, float(gain), float(kfrac)), "rbc_closure_damping")

    def step(self, nsteps, check_every=100):
        self._check(self.lib.rbc_closure_step(self.h, int(nsteps), int(check_every)), "rbc_closure_step")

    def status(self):
        a, f, r = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        self._check(self.lib.rbc_closure_status(self.h, ctypes.byref(a), ctypes.byref(f), ctypes.byref(r)),
                    "rbc_closure_status")
        return bool(a.value), bool(f.value), r.value
