"""optimizec_amd -- MI355X-native 2D pseudo-spectral vorticity DNS (the mem_cuda time step).

Python mirror of the C ABI in ``include/dns_hip.h`` (ctypes over the in-tree
``libdns_hip.so``).  The solver state lives on the GPU; this module only moves host
arrays across the boundary.  There is no CPU fallback: if the library or a HIP
device is missing, ``Solver`` raises.

Names follow the reference driver (cuda/cuda_dns.h, cuda/mem_cuda.cu): ``om2`` and
``fnm1`` are the spectral vorticity and the previous nonlinear term, [iz][ix] with
N rows of N/2 complex values; physical planes are M x M with M = 3N/2.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

__all__ = ["Solver", "Scalars", "DnsError", "library_path", "load_library", "ic_pao_host",
           "ic_k4exp_host", "supported_n", "FIELD_U", "FIELD_W", "FIELD_OMEGA", "FIELD_PSI",
           "FIELD_KINETIC", "IC_PAO_KM3", "IC_K4EXP", "PREC_FP64", "PREC_FP32", "PREC_BF16X3", "PREC_FP16",
           "PREC_FP16_STORE",
           "PATH_AUTO", "PATH_FUSED", "PATH_MULTIPASS", "PATH_FUSED_UIUJ", "PATH_MULTIPASS_UIUJ"]

_PKG = os.path.dirname(os.path.abspath(__file__))
IC_PAO_KM3, IC_K4EXP = 0, 1
PREC_FP64, PREC_FP32, PREC_BF16X3, PREC_FP16, PREC_FP16_STORE = 0, 1, 2, 3, 4
PATH_AUTO, PATH_FUSED, PATH_MULTIPASS, PATH_FUSED_UIUJ, PATH_MULTIPASS_UIUJ = 0, 1, 2, 3, 4
FIELD_U, FIELD_W, FIELD_OMEGA, FIELD_PSI, FIELD_KINETIC = 0, 1, 2, 3, 4
_STATUS = {0: "DNS_OK", 1: "DNS_E_ARG", 2: "DNS_E_UNSUPPORTED", 3: "DNS_E_HIP", 4: "DNS_E_NOMEM",
           5: "DNS_E_STATE", 6: "DNS_E_IO", 7: "DNS_E_FORMAT"}


class DnsError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{_STATUS.get(status, status)}: {msg}")
        self.status = status


class Config(ctypes.Structure):
    _fields_ = [("N", ctypes.c_int), ("Re", ctypes.c_double), ("K0", ctypes.c_double),
                ("cflnum", ctypes.c_double), ("precision", ctypes.c_int), ("device", ctypes.c_int),
                ("path", ctypes.c_int)]


class Scalars(ctypes.Structure):
    _fields_ = [
        ("N", ctypes.c_int), ("M", ctypes.c_int), ("H", ctypes.c_int), ("NK", ctypes.c_int),
        ("it", ctypes.c_int), ("ifn", ctypes.c_int),
        ("Re", ctypes.c_double), ("K0", ctypes.c_double), ("visc", ctypes.c_double),
        ("t", ctypes.c_double), ("dt", ctypes.c_double), ("cn", ctypes.c_double),
        ("cnm1", ctypes.c_double), ("cflnum", ctypes.c_double), ("cflm", ctypes.c_double),
    ]


class EnsMember(ctypes.Structure):
    _fields_ = [("N", ctypes.c_int), ("Re", ctypes.c_double), ("K0", ctypes.c_double),
                ("cflnum", ctypes.c_double), ("seed", ctypes.c_uint64)]


class EnsConfig(ctypes.Structure):
    _fields_ = [("nmembers", ctypes.c_int), ("members", ctypes.POINTER(EnsMember)),
                ("pair_every", ctypes.c_int), ("filter_ratio", ctypes.c_int), ("max_pairs", ctypes.c_int),
                ("use_graph", ctypes.c_int), ("device", ctypes.c_int), ("autotune", ctypes.c_int)]


_lib = None
_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_F = ctypes.POINTER(ctypes.c_float)

# every symbol include/dns_hip.h and include/dns_ensemble.h declare, with its ctypes signature
SIGNATURES = {
    "dns_create": (ctypes.c_int, [ctypes.POINTER(Config), ctypes.POINTER(_P)]),
    "dns_destroy": (None, [_P]),
    "dns_last_error": (ctypes.c_char_p, [_P]),
    "dns_supported_n": (ctypes.c_int, [ctypes.c_int]),
    "dns_supported_path": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
    "dns_init_ic": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_uint64, ctypes.c_char_p, ctypes.c_int]),
    "dns_start": (ctypes.c_int, [_P]),
    "dns_step": (ctypes.c_int, [_P, ctypes.c_int]),
    "dns_next_dt": (ctypes.c_int, [_P]),
    "dns_energy_enstrophy": (ctypes.c_int, [_P, _D, _D]),
    "dns_display_sigma": (ctypes.c_int, [_P, _D]),
    "dns_pal_over_ens_kmax2": (ctypes.c_int, [_P, _D]),
    "dns_fetch_field": (ctypes.c_int, [_P, ctypes.c_int, _D]),
    "dns_get_scalars": (ctypes.c_int, [_P, ctypes.POINTER(Scalars)]),
    "dns_set_scalars": (ctypes.c_int, [_P, ctypes.POINTER(Scalars)]),
    "dns_get_om2": (ctypes.c_int, [_P, _D]),
    "dns_set_om2": (ctypes.c_int, [_P, _D]),
    "dns_get_fnm1": (ctypes.c_int, [_P, _D]),
    "dns_set_fnm1": (ctypes.c_int, [_P, _D]),
    "dns_save_restart": (ctypes.c_int, [_P, ctypes.c_char_p]),
    "dns_load_restart": (ctypes.c_int, [_P, ctypes.c_char_p]),
    "dns_ic_pao_host": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_double, _D,
                                       ctypes.c_char_p, ctypes.c_int, _D]),
    "dns_ic_k4exp_host": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_uint64, _D]),
    "dns_set_kernel_timing": (ctypes.c_int, [_P, ctypes.c_int]),
    "dns_kernel_timing": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_char_p),
                                         ctypes.POINTER(ctypes.c_int64), _D, _D,
                                         ctypes.POINTER(ctypes.c_int)]),
    "dns_reset_kernel_timing": (ctypes.c_int, [_P]),
    "dns_synchronize": (ctypes.c_int, [_P]),
    "dns_device_bytes": (ctypes.c_size_t, [_P]),
    "dns_step_noalloc": (ctypes.c_int, [ctypes.POINTER(Scalars), _P, _P, ctypes.c_int, _P]),
    # batched ensemble (include/dns_ensemble.h)
    "dns_ens_create": (ctypes.c_int, [ctypes.POINTER(EnsConfig), ctypes.POINTER(_P)]),
    "dns_ens_destroy": (None, [_P]),
    "dns_ens_last_error": (ctypes.c_char_p, [_P]),
    "dns_ens_init": (ctypes.c_int, [_P]),
    "dns_ens_set_om2": (ctypes.c_int, [_P, ctypes.c_int, _D]),
    "dns_ens_advance": (ctypes.c_int, [_P, ctypes.c_int]),
    "dns_ens_synchronize": (ctypes.c_int, [_P]),
    "dns_ens_get_scalars": (ctypes.c_int, [_P, ctypes.c_int, ctypes.POINTER(Scalars)]),
    "dns_ens_set_scalars": (ctypes.c_int, [_P, ctypes.c_int, ctypes.POINTER(Scalars)]),
    "dns_ens_get_om2": (ctypes.c_int, [_P, ctypes.c_int, _D]),
    "dns_ens_get_fnm1": (ctypes.c_int, [_P, ctypes.c_int, _D]),
    "dns_ens_pair_info": (ctypes.c_int, [_P, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_int)]),
    "dns_ens_get_pairs": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_float)]),
    "dns_ens_device_bytes": (ctypes.c_size_t, [_P]),
    "dns_ens_tuning": (ctypes.c_int, [_P, ctypes.c_int, ctypes.POINTER(ctypes.c_int), _D,
                                      ctypes.POINTER(ctypes.c_int)]),
    # closure network (include/dns_closure.h)
    "dns_cnn_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(_P)]),
    "dns_cnn_destroy": (None, [_P]),
    "dns_cnn_last_error": (ctypes.c_char_p, [_P]),
    "dns_cnn_param_count": (ctypes.c_int, []),
    "dns_cnn_get_params": (ctypes.c_int, [_P, _F]),
    "dns_cnn_set_params": (ctypes.c_int, [_P, _F]),
    "dns_cnn_forward": (ctypes.c_int, [_P, _F, ctypes.c_int, ctypes.c_int, _F]),
    "dns_cnn_loss_grad": (ctypes.c_int, [_P, _F, _F, ctypes.c_int, ctypes.c_int, _D, _F]),
    "dns_cnn_adam_step": (ctypes.c_int, [_P, _F, _F, ctypes.c_int, ctypes.c_int, ctypes.c_float, _D]),
    "dns_cnn_train_on_ensemble": (ctypes.c_int, [_P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                                 ctypes.c_uint64, _D]),
    "dns_set_closure": (ctypes.c_int, [_P, _P]),
    "dns_closure_forcing": (ctypes.c_int, [_P, _D]),
    "dns_set_closure_band": (ctypes.c_int, [_P, ctypes.c_double, ctypes.c_double]),
    "dns_set_closure_limiter": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_double]),
}


def library_path() -> str:
    return os.path.join(_PKG, "libdns_hip.so")


def load_library():
    """Loads the in-tree libdns_hip.so (built by optimizec_amd/build.py); raises if absent."""
    global _lib
    if _lib is None:
        path = library_path()
        if not os.path.exists(path):
            raise DnsError(3, f"{path} is not built (run python -m optimizec_amd.build)")
        lib = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_D)


def supported_n(N: int) -> bool:
    return bool(load_library().dns_supported_n(N))


def supported_path(N: int, path: int) -> bool:
    return bool(load_library().dns_supported_path(N, path))


def ic_pao_host(N: int, Re: float, K0: float):
    """PAO-k^-3 initial om2 computed on the host: (om2, seed, printed block, visc)."""
    om = np.zeros((N, N // 2), dtype=np.complex128)
    buf = ctypes.create_string_buffer(8192)
    visc = ctypes.c_double()
    seed = load_library().dns_ic_pao_host(N, Re, K0, _ptr(om), buf, len(buf), ctypes.byref(visc))
    if seed < 0:
        raise DnsError(1, "dns_ic_pao_host rejected its arguments")
    return om, seed, buf.value.decode(), visc.value


def ic_k4exp_host(N: int, K0: float, seed: int) -> np.ndarray:
    om = np.zeros((N, N // 2), dtype=np.complex128)
    st = load_library().dns_ic_k4exp_host(N, K0, ctypes.c_uint64(seed), _ptr(om))
    if st:
        raise DnsError(st, "dns_ic_k4exp_host rejected its arguments")
    return om


class Solver:
    """One device state: the C handle of dns_create (dnsCudaCreate, cuda_dns.h:108)."""

    def __init__(self, N: int, Re: float, K0: float, cflnum: float, device: int = 0,
                 precision: int = PREC_FP64, path: int = PATH_AUTO):
        """precision: PREC_FP64, or PREC_FP32 (FP32 transforms, FP64 state/update/dt), or
        PREC_BF16X3 / PREC_FP16 (the same with the transforms' radix stages as DFT-matrix
        products on the matrix cores: three BF16 products, or one scaled FP16 product), or
        PREC_FP16_STORE (PREC_FP32 with the intermediates between the step's two kernels
        stored as power-of-two-scaled FP16 pairs; N=4096, fused path).
        path: PATH_AUTO (fused kernels where built, else multi-pass), PATH_FUSED or
        PATH_MULTIPASS (batched four-step line FFTs through HBM), or PATH_FUSED_UIUJ /
        PATH_MULTIPASS_UIUJ (those with the nonlinear term in the reference's u_i u_j form; FP64;
        the fused one at N=4096)."""
        L = load_library()
        cfg = Config(N, float(Re), float(K0), float(cflnum), int(precision), int(device), int(path))
        h = _P()
        st = L.dns_create(ctypes.byref(cfg), ctypes.byref(h))
        if st:
            raise DnsError(st, L.dns_last_error(None).decode())
        self._h = h
        self.N, self.M, self.H = N, 3 * N // 2, N // 2

    def _check(self, st: int) -> None:
        if st:
            raise DnsError(st, load_library().dns_last_error(self._h).decode())

    def close(self) -> None:
        if getattr(self, "_h", None):
            load_library().dns_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- initial condition and stepping (dnsCudaInit, next_dt_gpu, loop body)
    def init_ic(self, kind: int = IC_K4EXP, seed: int = 0) -> str:
        buf = ctypes.create_string_buffer(8192)
        self._check(load_library().dns_init_ic(self._h, kind, ctypes.c_uint64(seed), buf, len(buf)))
        return buf.value.decode()

    def start(self) -> None:
        self._check(load_library().dns_start(self._h))

    def step(self, n: int = 1) -> None:
        self._check(load_library().dns_step(self._h, n))

    def next_dt(self) -> None:
        self._check(load_library().dns_next_dt(self._h))

    def synchronize(self) -> None:
        self._check(load_library().dns_synchronize(self._h))

    def run(self, steps: int, update: int):
        """mem_cuda's loop (mem_cuda.cu:1848-1921) from the current it to `steps`; returns
        rows (it, t, dt, cn, cflm, E, Z) for the diagnostic steps."""
        rows = []
        s = self.scalars()
        for it in range(s.it + 1, steps + 1):
            self.step(1)
            if it == 1 or it % update == 0 or it == steps:
                self.next_dt()
                sc = self.scalars()
                E, Z = self.energy_enstrophy()
                rows.append((it, sc.t, sc.dt, sc.cn, sc.cflm, E, Z))
        return np.array(rows, dtype=np.float64).reshape(-1, 7)

    # -- diagnostics
    def energy_enstrophy(self):
        e, z = ctypes.c_double(), ctypes.c_double()
        self._check(load_library().dns_energy_enstrophy(self._h, ctypes.byref(e), ctypes.byref(z)))
        return e.value, z.value

    def pal_over_ens_kmax2(self) -> float:
        """Palinstrophy / (enstrophy * max K^2), the ADAPT_VISC input (mem_cuda.cu:872-935)."""
        v = ctypes.c_double()
        self._check(load_library().dns_pal_over_ens_kmax2(self._h, ctypes.byref(v)))
        return v.value

    def display_sigma(self) -> float:
        s = ctypes.c_double()
        self._check(load_library().dns_display_sigma(self._h, ctypes.byref(s)))
        return s.value

    def field(self, which: int) -> np.ndarray:
        out = np.empty((self.M, self.M), dtype=np.float64)
        self._check(load_library().dns_fetch_field(self._h, which, _ptr(out)))
        return out

    # -- state access
    def scalars(self) -> Scalars:
        s = Scalars()
        self._check(load_library().dns_get_scalars(self._h, ctypes.byref(s)))
        return s

    def set_scalars(self, s: Scalars) -> None:
        self._check(load_library().dns_set_scalars(self._h, ctypes.byref(s)))

    def om2(self) -> np.ndarray:
        out = np.empty((self.N, self.H), dtype=np.complex128)
        self._check(load_library().dns_get_om2(self._h, _ptr(out)))
        return out

    def set_om2(self, a) -> None:
        a = np.ascontiguousarray(a, dtype=np.complex128).reshape(self.N, self.H)
        self._check(load_library().dns_set_om2(self._h, _ptr(a)))

    def fnm1(self) -> np.ndarray:
        out = np.empty((self.N, self.H), dtype=np.complex128)
        self._check(load_library().dns_get_fnm1(self._h, _ptr(out)))
        return out

    def set_fnm1(self, a) -> None:
        a = np.ascontiguousarray(a, dtype=np.complex128).reshape(self.N, self.H)
        self._check(load_library().dns_set_fnm1(self._h, _ptr(a)))

    def save_restart(self, path: str) -> None:
        self._check(load_library().dns_save_restart(self._h, path.encode()))

    def load_restart(self, path: str) -> None:
        self._check(load_library().dns_load_restart(self._h, path.encode()))

    # -- measurement
    def set_kernel_timing(self, on: bool, every: int = 1) -> None:
        """Times the 1st, (every+1)-th, ... launch of each hot kernel with HIP events."""
        self._check(load_library().dns_set_kernel_timing(self._h, max(1, int(every)) if on else 0))

    def reset_kernel_timing(self) -> None:
        self._check(load_library().dns_reset_kernel_timing(self._h))

    def kernel_timing(self) -> dict:
        cap = 32
        n = ctypes.c_int(cap)
        names = (ctypes.c_char_p * cap)()
        launches = (ctypes.c_int64 * cap)()
        ms = (ctypes.c_double * cap)()
        bytes_ = (ctypes.c_double * cap)()
        self._check(load_library().dns_kernel_timing(self._h, names, launches, ms, bytes_, ctypes.byref(n)))
        return {names[i].decode(): {"launches": launches[i], "total_ms": ms[i], "bytes_per_launch": bytes_[i]}
                for i in range(min(cap, n.value))}

    def device_bytes(self) -> int:
        return int(load_library().dns_device_bytes(self._h))

    # -- closure
    def set_closure(self, net) -> None:
        """Attach a ClosureNet (its forcing is added in every step) or None to detach."""
        self._closure = net
        self._check(load_library().dns_set_closure(self._h, net._h if net is not None else None))

    def set_closure_band(self, k_lo: float, k_hi: float) -> None:
        self._check(load_library().dns_set_closure_band(self._h, float(k_lo), float(k_hi)))

    def set_closure_limiter(self, enabled: bool, nu_dt_max: float = 0.5) -> None:
        self._check(load_library().dns_set_closure_limiter(self._h, 1 if enabled else 0, float(nu_dt_max)))

    def closure_forcing(self) -> np.ndarray:
        out = np.empty((self.N, self.H), dtype=np.complex128)
        self._check(load_library().dns_closure_forcing(self._h, _ptr(out)))
        return out


class Ensemble:
    """Many independent runs advanced together (include/dns_ensemble.h): members given as
    (N, Re, K0, cflnum, seed) tuples; training pairs every ``pair_every`` steps on the
    N / ``filter_ratio`` grid, channels (filtered omega, u, w, subgrid term).

    ``autotune``: True (or 1) times the full step of each launch configuration whose results
    are bitwise the default's and keeps the fastest of each size; 2 also times the radix-16
    plans (results then agree with single runs to roundoff); ``configuration=v`` runs
    configuration v (0..8) on every size without timing."""

    def __init__(self, members, pair_every: int = 20, filter_ratio: int = 8, max_pairs: int = 100,
                 use_graph: int = 1, device: int = 0, autotune: bool = False,
                 configuration: int | None = None):
        L = load_library()
        self.members = [tuple(m) for m in members]
        arr = (EnsMember * len(self.members))()
        for i, (N, Re, K0, cfl, seed) in enumerate(self.members):
            arr[i] = EnsMember(int(N), float(Re), float(K0), float(cfl), int(seed))
        self._arr = arr
        cfg = EnsConfig(len(self.members), arr, int(pair_every), int(filter_ratio), int(max_pairs),
                        int(use_graph), int(device),
                        3 + int(configuration) if configuration is not None
                        else (0 if not autotune else (2 if int(autotune) == 2 else 1)))
        h = _P()
        st = L.dns_ens_create(ctypes.byref(cfg), ctypes.byref(h))
        if st:
            raise DnsError(st, L.dns_ens_last_error(None).decode())
        self._h = h
        self.pair_every = pair_every

    def _check(self, st: int) -> None:
        if st:
            raise DnsError(st, load_library().dns_ens_last_error(self._h).decode())

    def close(self) -> None:
        if getattr(self, "_h", None):
            load_library().dns_ens_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def init(self) -> None:
        self._check(load_library().dns_ens_init(self._h))

    def advance(self, nsteps: int) -> None:
        self._check(load_library().dns_ens_advance(self._h, nsteps))

    def synchronize(self) -> None:
        self._check(load_library().dns_ens_synchronize(self._h))

    def scalars(self, m: int) -> Scalars:
        s = Scalars()
        self._check(load_library().dns_ens_get_scalars(self._h, m, ctypes.byref(s)))
        return s

    def set_scalars(self, m: int, s: Scalars) -> None:
        self._check(load_library().dns_ens_set_scalars(self._h, m, ctypes.byref(s)))

    def om2(self, m: int) -> np.ndarray:
        N = self.members[m][0]
        out = np.empty((N, N // 2), dtype=np.complex128)
        self._check(load_library().dns_ens_get_om2(self._h, m, _ptr(out)))
        return out

    def set_om2(self, m: int, a) -> None:
        N = self.members[m][0]
        a = np.ascontiguousarray(a, dtype=np.complex128).reshape(N, N // 2)
        self._check(load_library().dns_ens_set_om2(self._h, m, _ptr(a)))

    def fnm1(self, m: int) -> np.ndarray:
        N = self.members[m][0]
        out = np.empty((N, N // 2), dtype=np.complex128)
        self._check(load_library().dns_ens_get_fnm1(self._h, m, _ptr(out)))
        return out

    def pair_info(self, m: int):
        c, n = ctypes.c_int(), ctypes.c_int()
        self._check(load_library().dns_ens_pair_info(self._h, m, ctypes.byref(c), ctypes.byref(n)))
        return c.value, n.value

    def pairs(self, m: int, first: int = 0, count: int = None) -> np.ndarray:
        """Training pairs of member m: float32 array [count][4][n][n]."""
        have, n = self.pair_info(m)
        if count is None:
            count = have - first
        out = np.empty((count, 4, n, n), dtype=np.float32)
        self._check(load_library().dns_ens_get_pairs(self._h, m, first, count,
                                                    out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def device_bytes(self) -> int:
        return int(load_library().dns_ens_device_bytes(self._h))

    def tuning(self, m: int):
        """(chosen launch configuration, [ms of each configuration]) for member m's size."""
        v, n = ctypes.c_int(), ctypes.c_int(16)
        ms = (ctypes.c_double * 16)()
        self._check(load_library().dns_ens_tuning(self._h, m, ctypes.byref(v), ms, ctypes.byref(n)))
        return v.value, [ms[i] for i in range(min(16, n.value))]


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(_F)


class ClosureNet:
    """The subgrid-closure network (include/dns_closure.h): trained on the GPU, applied
    in coarse runs through Solver.set_closure."""

    def __init__(self, seed: int = 0, device: int = 0):
        L = load_library()
        h = _P()
        st = L.dns_cnn_create(int(device), ctypes.c_uint64(seed), ctypes.byref(h))
        if st:
            raise DnsError(st, L.dns_cnn_last_error(None).decode())
        self._h = h
        self.nparams = int(L.dns_cnn_param_count())

    def _check(self, st: int) -> None:
        if st:
            raise DnsError(st, load_library().dns_cnn_last_error(self._h).decode())

    def close(self) -> None:
        if getattr(self, "_h", None):
            load_library().dns_cnn_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def params(self) -> np.ndarray:
        out = np.empty(self.nparams, dtype=np.float32)
        self._check(load_library().dns_cnn_get_params(self._h, _fptr(out)))
        return out

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparams
        self._check(load_library().dns_cnn_set_params(self._h, _fptr(p)))

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        B, c, n, _ = x.shape
        y = np.empty((B, n, n), dtype=np.float32)
        self._check(load_library().dns_cnn_forward(self._h, _fptr(x), B, n, _fptr(y)))
        return y

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        B, _, n, _ = x.shape
        g = np.empty(self.nparams, dtype=np.float32)
        loss = ctypes.c_double()
        self._check(load_library().dns_cnn_loss_grad(self._h, _fptr(x), _fptr(t), B, n, ctypes.byref(loss), _fptr(g)))
        return loss.value, g

    def adam_step(self, x, t, lr: float = 1e-3) -> float:
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        B, _, n, _ = x.shape
        loss = ctypes.c_double()
        self._check(load_library().dns_cnn_adam_step(self._h, _fptr(x), _fptr(t), B, n, ctypes.c_float(lr),
                                                     ctypes.byref(loss)))
        return loss.value

    def train_on_ensemble(self, ens, iterations: int, batch: int = 64, lr: float = 1e-3, seed: int = 0):
        hist = np.zeros(iterations, dtype=np.float64)
        self._check(load_library().dns_cnn_train_on_ensemble(self._h, ens._h, iterations, batch, ctypes.c_float(lr),
                                                             ctypes.c_uint64(seed), _ptr(hist)))
        return hist
