"""ctypes view of the FP64 CPU oracle (oracle/rbc_oracle.c).

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
import this module.  The product package (sdc_amd) never does.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None

TWO_NAMES = ["Iac", "Iab", "Ia1", "Ia2"]
THREE_NAMES = ["Iac1", "Ibb1", "Iaab", "Iaad", "Iabc", "Ia11", "Ia12", "I111", "Ibbb", "Ib11",
               "Iaa1", "Iab2"]


def build(force=False):
    """Compile liboracle.so with gcc (the Makefile next to this file)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        L.orc_lambda.restype = ctypes.c_double
        L.orc_lambda.argtypes = [ctypes.c_int]
        L.orc_coeff_tables.argtypes = [ctypes.c_int, dp, dp]
        L.orc_gauss_legendre.argtypes = [ctypes.c_int, dp, dp]
        L.orc_fft_r2c.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp]
        L.orc_fft_c2r.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp]
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 5 + [ctypes.c_int]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        for fn in ("orc_set_state", "orc_get_state"):
            getattr(L, fn).argtypes = [ctypes.c_void_p] + [dp] * 5
        for fn in ("orc_nonlinear", "orc_inverse_B", "orc_implicit"):
            getattr(L, fn).argtypes = [ctypes.c_void_p] + [dp] * 10
        L.orc_dealias.argtypes = [ctypes.c_int] * 3 + [dp]
        L.orc_make_real.argtypes = [ctypes.c_int] * 3 + [dp]
        L.orc_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_reset_history.argtypes = [ctypes.c_void_p]
        L.orc_nusselt.argtypes = [ctypes.c_void_p, dp]
        L.orc_implicit_factor.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, dp]
        L.orc_set_penalization.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def lam(k):
    return lib().orc_lambda(k)


def coeff_tables(L):
    """Returns (two, three) dicts: two[name][i, j], three[name][i, j, k]."""
    two = np.zeros(4 * L * L)
    three = np.zeros(12 * L ** 3)
    rc = lib().orc_coeff_tables(L, _p(two), _p(three))
    assert rc == 0
    t2 = {n: two[i * L * L:(i + 1) * L * L].reshape(L, L).T.copy() for i, n in enumerate(TWO_NAMES)}
    t3 = {n: three[i * L ** 3:(i + 1) * L ** 3].reshape(L, L, L).transpose(2, 1, 0).copy()
          for i, n in enumerate(THREE_NAMES)}
    return t2, t3


def gauss_legendre(n):
    z = np.zeros(n)
    w = np.zeros(n)
    lib().orc_gauss_legendre(n, _p(z), _p(w))
    return z, w


def fft_r2c(phys):
    N, M = phys.shape
    out = np.zeros((N, M // 2 + 1), np.complex128)
    ph = np.ascontiguousarray(phys, np.float64)
    assert lib().orc_fft_r2c(M, N, _p(ph), _p(out.view(np.float64))) == 0
    return out


def fft_c2r(spec, M):
    N = spec.shape[0]
    out = np.zeros((N, M))
    sp = np.ascontiguousarray(spec, np.complex128)
    assert lib().orc_fft_c2r(M, N, _p(sp.view(np.float64)), _p(out)) == 0
    return out


def dealias(field):
    """field: complex128 array (L, N, mox), modified in place."""
    L, N, mox = field.shape
    lib().orc_dealias(mox, N, L, _p(field.view(np.float64)))
    return field


def init_test_state(M, N, L, fp32=False):
    """init_test_device (util/init.cu:263-301) on (L, N, M/2+1) arrays: theta and f hold
    1e-6 sin(x) (real part) in row y = 0 for columns x != 0, g = 0, F_z = 1e-6 cos(z), G = 0,
    evaluated in the build's real type (float32 for an FP32 build)."""
    mox = M // 2 + 1
    rt = np.float32 if fp32 else np.float64
    amp = rt(1e-6)
    row = np.zeros(mox, np.complex128)
    x = np.arange(1, mox)
    row[1:] = (amp * np.sin(x.astype(rt))).astype(np.float64)
    th = np.zeros((L, N, mox), np.complex128)
    th[:, 0, :] = row[None, :]
    f = th.copy()
    g = np.zeros_like(th)
    F = (amp * np.cos(np.arange(L).astype(rt))).astype(np.float64).astype(np.complex128)
    G = np.zeros(L, np.complex128)
    return [th, f, g, F, G]


def make_real(field):
    L, N, mox = field.shape
    lib().orc_make_real(mox, N, L, _p(field.view(np.float64)))
    return field


class Stepper:
    """FP64 reference stepper on the host.  Fields are numpy complex128 arrays shaped
    (L, N, M/2+1) (the reference layout read slowest-first); F, G are (L,)."""

    def __init__(self, M, N, L, Lx, Ly, Ra, Pr, dt, nthreads=1):
        self.M, self.N, self.L = M, N, L
        self.mox = M // 2 + 1
        self.h = lib().orc_create(M, N, L, Lx, Ly, Ra, Pr, dt, nthreads)
        if not self.h:
            raise RuntimeError("orc_create failed")

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _shape(self):
        return (self.L, self.N, self.mox)

    def zeros(self):
        return [np.zeros(self._shape(), np.complex128) for _ in range(3)] + \
               [np.zeros(self.L, np.complex128) for _ in range(2)]

    def set_state(self, th, f, g, F, G):
        arrs = [np.ascontiguousarray(a, np.complex128) for a in (th, f, g, F, G)]
        lib().orc_set_state(self.h, *[_p(a.view(np.float64)) for a in arrs])

    def get_state(self):
        out = self.zeros()
        lib().orc_get_state(self.h, *[_p(a.view(np.float64)) for a in out])
        return out

    def _op(self, fn, ins):
        ins = [np.ascontiguousarray(a, np.complex128) for a in ins]
        out = self.zeros()
        rc = fn(self.h, *[_p(a.view(np.float64)) for a in ins], *[_p(a.view(np.float64)) for a in out])
        assert rc == 0
        return out

    def nonlinear(self, *ins):
        return self._op(lib().orc_nonlinear, ins)

    def inverse_B(self, *ins):
        return self._op(lib().orc_inverse_B, ins)

    def implicit(self, *ins):
        return self._op(lib().orc_implicit, ins)

    def step(self, n=1):
        assert lib().orc_step(self.h, n) == 0

    def reset_history(self):
        lib().orc_reset_history(self.h)

    def nusselt(self):
        nu = np.zeros(3)
        lib().orc_nusselt(self.h, _p(nu))
        return nu

    MASKS = {"rectangle": 0, "circle": 1, "diskarray": 2, "everywhere": 3}

    def set_penalization(self, eta, shape="rectangle"):
        """Penalized build (PENALIZED): mask terms added to the nonlinear operator."""
        sh = self.MASKS[shape] if isinstance(shape, str) else int(shape)
        assert lib().orc_set_penalization(self.h, float(eta), sh) == 0

    def implicit_factor(self, col, row):
        out = np.zeros(4 * self.L * self.L)
        assert lib().orc_implicit_factor(self.h, col, row, _p(out)) == 0
        return out.reshape(4, self.L, self.L)


def time_step_sampled(stepper, pair_stride):
    """Sampled CPU timing of one reference step (see orc_time_step_sampled)."""
    f = lib().orc_time_step_sampled
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    out = np.zeros(4)
    assert f(stepper.h, int(pair_stride), _p(out)) == 0
    return {"s_per_step": out[0], "phys": out[1], "pairs_sampled": out[2], "rest": out[3]}
