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

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsdc_rbc.so")

F64, F32, BF16X3 = 0, 1, 2
PRECISIONS = {"f64": F64, "f32": F32, "bf16x3": BF16X3}
NSTAGES = 8


class RbcError(RuntimeError):
    pass


class rbc_params(ctypes.Structure):
    _fields_ = [("M", ctypes.c_int), ("N", ctypes.c_int), ("L", ctypes.c_int),
                ("Lx", ctypes.c_double), ("Ly", ctypes.c_double),
                ("Ra", ctypes.c_double), ("Pr", ctypes.c_double), ("dt", ctypes.c_double),
                ("precision", ctypes.c_int), ("nq", ctypes.c_int), ("device", ctypes.c_int)]


_lib = None

EXPORTS = ["rbc_create", "rbc_destroy", "rbc_last_error", "rbc_get_info", "rbc_set_state",
           "rbc_get_state", "rbc_step", "rbc_reset_history", "rbc_nonlinear", "rbc_inverse_B",
           "rbc_implicit", "rbc_nusselt", "rbc_profile_steps", "rbc_stage_name", "rbc_synchronize"]


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
    L.rbc_profile_steps.argtypes = [vp, ctypes.c_int, dp]
    L.rbc_stage_name.argtypes = [ctypes.c_int]
    L.rbc_stage_name.restype = ctypes.c_char_p
    L.rbc_synchronize.argtypes = [vp]
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Stepper:
    """One RBC context on one GPU (≙ the operator set main.cu:431-454 builds)."""

    def __init__(self, M, N, L, Lx, Ly, Ra, Pr, dt, precision="f64", nq=0, device=0):
        self.lib = load_library()
        self.M, self.N, self.L = M, N, L
        self.mox = M // 2 + 1
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        self.params = rbc_params(M, N, L, Lx, Ly, Ra, Pr, dt, prec, nq, device)
        h = ctypes.c_void_p()
        rc = self.lib.rbc_create(ctypes.byref(self.params), ctypes.byref(h))
        if rc != 0:
            raise RbcError(f"rbc_create failed ({rc}): {self.lib.rbc_last_error(None).decode()}")
        self.h = h
        nq_, kx, ky, ng = (ctypes.c_int() for _ in range(4))
        self.lib.rbc_get_info(self.h, ctypes.byref(nq_), ctypes.byref(kx), ctypes.byref(ky), ctypes.byref(ng))
        self.nq, self.kx_kept, self.ky_kept, self.n_unique_k = nq_.value, kx.value, ky.value, ng.value

    def _check(self, rc, what):
        if rc != 0:
            raise RbcError(f"{what} failed ({rc}): {self.lib.rbc_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
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

    def nusselt(self):
        nu = np.zeros(3)
        self._check(self.lib.rbc_nusselt(self.h, _dp(nu)), "rbc_nusselt")
        return nu

    def profile_steps(self, nsteps):
        ms = np.zeros(NSTAGES)
        self._check(self.lib.rbc_profile_steps(self.h, int(nsteps), _dp(ms)), "rbc_profile_steps")
        names = [self.lib.rbc_stage_name(i).decode() for i in range(NSTAGES)]
        return dict(zip(names, ms.tolist()))
