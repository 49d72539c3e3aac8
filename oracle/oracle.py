"""ctypes front end of the FP64 CPU oracle (oracle/dns_oracle.c).

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP product path:
only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may
import it.  The product package ``optimizec_amd`` never imports it.

Arrays are numpy complex128 / float64 in the reference layouts:
om2/fnm1 [iz][ix] (N x N/2), uc [z][kx] (M x M/2+1), ur [z][x] (M x M).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


def build() -> None:
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


class Scalars(ctypes.Structure):
    _fields_ = [
        ("N", ctypes.c_int), ("M", ctypes.c_int), ("H", ctypes.c_int), ("NK", ctypes.c_int),
        ("it", ctypes.c_int), ("ifn", ctypes.c_int),
        ("Re", ctypes.c_double), ("K0", ctypes.c_double), ("visc", ctypes.c_double),
        ("t", ctypes.c_double), ("dt", ctypes.c_double), ("cn", ctypes.c_double),
        ("cnm1", ctypes.c_double), ("cflnum", ctypes.c_double), ("cflm", ctypes.c_double),
    ]


_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)


def _declare(L):
    L.or_fft.argtypes = [ctypes.c_int, ctypes.c_int, _D, _D]
    L.or_dft_brute.argtypes = [ctypes.c_int, ctypes.c_int, _D, _D]
    L.or_c2r_2d.argtypes = [ctypes.c_int, _D, _D]
    L.or_r2c_2d.argtypes = [ctypes.c_int, _D, _D]
    L.or_create.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double]
    L.or_create.restype = _P
    L.or_destroy.argtypes = [_P]
    L.or_get_scalars.argtypes = [_P, ctypes.POINTER(Scalars)]
    L.or_set_scalars.argtypes = [_P, ctypes.POINTER(Scalars)]
    L.or_init_pao.argtypes = [_P, ctypes.c_char_p, ctypes.c_int]
    L.or_init_pao.restype = ctypes.c_int
    L.or_init_config.argtypes = [_P, ctypes.c_uint64]
    L.or_ic_uniform.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int]
    L.or_ic_uniform.restype = ctypes.c_double
    for name in ("or_step2a", "or_step2b", "or_step3", "or_next_dt", "or_calcom", "or_step_adv",
                 "or_velocity_from_om2", "or_om2phys"):
        getattr(L, name).argtypes = [_P]
    L.or_start.argtypes = [_P, ctypes.c_int]
    L.or_one_step.argtypes = [_P, ctypes.c_int]
    L.or_run.argtypes = [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _D, ctypes.c_int]
    L.or_run.restype = ctypes.c_int
    L.or_run_timed.argtypes = [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _D, ctypes.c_int, _D]
    L.or_run_timed.restype = ctypes.c_int
    L.or_energy_enstrophy.argtypes = [_P, _D, _D]
    L.or_cflm.argtypes = [_P]
    L.or_cflm.restype = ctypes.c_double
    L.or_display_sigma.argtypes = [_P, ctypes.c_int]
    L.or_display_sigma.restype = ctypes.c_double
    L.or_pal_over_ens_kmax2.argtypes = [_P]
    L.or_pal_over_ens_kmax2.restype = ctypes.c_double
    L.or_nonlinear_adv.argtypes = [_P, _D]
    L.or_nonlinear_uiuj.argtypes = [_P, _D]
    for name in ("or_get_om2", "or_set_om2", "or_get_fnm1", "or_set_fnm1"):
        getattr(L, name).argtypes = [_P, _D]
    for name in ("or_get_ur", "or_get_uc", "or_set_uc"):
        getattr(L, name).argtypes = [_P, ctypes.c_int, _D]
    L.or_set_threads.argtypes = [ctypes.c_int]


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_D)


def fft(x, sign=-1) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    if lib().or_fft(len(x), sign, _ptr(x), _ptr(out)) != 0:
        raise ValueError("or_fft failed")
    return out


def dft_brute(x, sign=-1) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    lib().or_dft_brute(len(x), sign, _ptr(x), _ptr(out))
    return out


def c2r_2d(uc) -> np.ndarray:
    uc = np.ascontiguousarray(uc, dtype=np.complex128)
    M = uc.shape[0]
    assert uc.shape == (M, M // 2 + 1)
    out = np.empty((M, M), dtype=np.float64)
    lib().or_c2r_2d(M, _ptr(uc), _ptr(out))
    return out


def r2c_2d(ur) -> np.ndarray:
    ur = np.ascontiguousarray(ur, dtype=np.float64)
    M = ur.shape[0]
    out = np.empty((M, M // 2 + 1), dtype=np.complex128)
    lib().or_r2c_2d(M, _ptr(ur), _ptr(out))
    return out


def set_threads(n: int) -> None:
    lib().or_set_threads(int(n))


class Oracle:
    """One solver state (om2, fnm1, UC, UR and the scalars of DnsDeviceState)."""

    def __init__(self, N: int, Re: float, K0: float, cfl: float):
        self._h = lib().or_create(N, Re, K0, cfl)
        if not self._h:
            raise ValueError(f"or_create failed for N={N}")
        self.N, self.M, self.H, self.NK = N, 3 * N // 2, N // 2, 3 * N // 4 + 1

    def close(self):
        if self._h:
            lib().or_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    # -- scalars
    def scalars(self) -> Scalars:
        s = Scalars()
        lib().or_get_scalars(self._h, ctypes.byref(s))
        return s

    def set_scalars(self, s: Scalars) -> None:
        lib().or_set_scalars(self._h, ctypes.byref(s))

    # -- initial conditions
    def init_pao(self) -> tuple[int, str]:
        buf = ctypes.create_string_buffer(8192)
        seed = lib().or_init_pao(self._h, buf, len(buf))
        return seed, buf.value.decode()

    def init_config(self, seed: int) -> None:
        lib().or_init_config(self._h, ctypes.c_uint64(seed))

    # -- steps
    def start(self, form: int = 1) -> None:
        lib().or_start(self._h, form)

    def one_step(self, form: int = 1) -> None:
        lib().or_one_step(self._h, form)

    def run(self, steps: int, update: int, form: int = 1, maxrows: int = 100000) -> np.ndarray:
        buf = np.zeros((maxrows, 7), dtype=np.float64)
        n = lib().or_run(self._h, steps, update, form, _ptr(buf), maxrows)
        return buf[:n].copy()

    def run_timed(self, steps: int, update: int, form: int = 1, maxrows: int = 100000):
        """run(), plus the wall seconds of each step's loop body (diagnostics excluded)."""
        buf = np.zeros((maxrows, 7), dtype=np.float64)
        n_steps = max(0, steps - self.scalars().it)
        secs = np.zeros(max(1, n_steps), dtype=np.float64)
        n = lib().or_run_timed(self._h, steps, update, form, _ptr(buf), maxrows, _ptr(secs))
        return buf[:n].copy(), secs[:n_steps].copy()

    def step2a(self): lib().or_step2a(self._h)
    def step2b(self): lib().or_step2b(self._h)
    def step3(self): lib().or_step3(self._h)
    def next_dt(self): lib().or_next_dt(self._h)
    def calcom(self): lib().or_calcom(self._h)
    def step_adv(self): lib().or_step_adv(self._h)
    def velocity_from_om2(self): lib().or_velocity_from_om2(self._h)
    def om2phys(self): lib().or_om2phys(self._h)

    # -- diagnostics
    def energy_enstrophy(self) -> tuple[float, float]:
        e, z = ctypes.c_double(), ctypes.c_double()
        lib().or_energy_enstrophy(self._h, ctypes.byref(e), ctypes.byref(z))
        return e.value, z.value

    def cflm(self) -> float:
        return lib().or_cflm(self._h)

    def display_sigma(self, comp: int = 2) -> float:
        return lib().or_display_sigma(self._h, comp)

    def pal_over_ens_kmax2(self) -> float:
        return lib().or_pal_over_ens_kmax2(self._h)

    def nonlinear(self, form: int = 1) -> np.ndarray:
        out = np.empty((self.N, self.H), dtype=np.complex128)
        (lib().or_nonlinear_adv if form == 1 else lib().or_nonlinear_uiuj)(self._h, _ptr(out))
        return out

    # -- arrays
    def om2(self) -> np.ndarray:
        out = np.empty((self.N, self.H), dtype=np.complex128)
        lib().or_get_om2(self._h, _ptr(out))
        return out

    def set_om2(self, a) -> None:
        a = np.ascontiguousarray(a, dtype=np.complex128).reshape(self.N, self.H)
        lib().or_set_om2(self._h, _ptr(a))

    def fnm1(self) -> np.ndarray:
        out = np.empty((self.N, self.H), dtype=np.complex128)
        lib().or_get_fnm1(self._h, _ptr(out))
        return out

    def set_fnm1(self, a) -> None:
        a = np.ascontiguousarray(a, dtype=np.complex128).reshape(self.N, self.H)
        lib().or_set_fnm1(self._h, _ptr(a))

    def ur(self, comp: int) -> np.ndarray:
        out = np.empty((self.M, self.M), dtype=np.float64)
        lib().or_get_ur(self._h, comp, _ptr(out))
        return out

    def uc(self, comp: int) -> np.ndarray:
        out = np.empty((self.M, self.NK), dtype=np.complex128)
        lib().or_get_uc(self._h, comp, _ptr(out))
        return out

    def set_uc(self, comp: int, a) -> None:
        a = np.ascontiguousarray(a, dtype=np.complex128).reshape(self.M, self.NK)
        lib().or_set_uc(self._h, comp, _ptr(a))


def ic_uniform(seed: int, kx: int, kz: int) -> float:
    return lib().or_ic_uniform(ctypes.c_uint64(seed), kx, kz)
