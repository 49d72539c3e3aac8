"""CPU-side checks of the product library: it loads, it exports exactly what
include/dns_hip.h declares, its host-side initial conditions agree with the oracle,
and device entry points fail loudly (not silently) where no GPU is present."""
import ctypes
import os
import re

import numpy as np
import pytest

import optimizec_amd as oa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    names = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            with open(os.path.join(ROOT, "include", h), encoding="utf-8") as f:
                names |= set(re.findall(r"\b(dns_[a-z0-9_]+)\s*\(", f.read()))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = oa.load_library()
    names = _declared_symbols()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    # the Python mirror binds every declared entry point
    assert sorted(oa.SIGNATURES) == names


def test_supported_sizes():
    for N in (16, 32, 64, 128, 256, 512, 768, 1024, 2048, 3072, 4096, 6144, 8192, 9216, 10240, 12288, 16384,
              20480, 24576, 29400):
        assert oa.supported_n(N), N
    for N in (10, 30, 100, 6000, 29404, 30000):
        assert not oa.supported_n(N), N
    # fused kernels up to N = 4096; the multi-pass path for the large grids and for
    # the sizes the parity tests force through it
    for N in (4096, 3072, 16):
        assert oa.supported_path(N, oa.PATH_FUSED), N
    for N in (6144, 8192, 9216, 16384, 20480, 29400):
        assert not oa.supported_path(N, oa.PATH_FUSED), N
    for N in (64, 128, 256, 512, 1024, 4096, 8192, 16384, 29400):
        assert oa.supported_path(N, oa.PATH_MULTIPASS), N
    assert not oa.supported_path(2048, oa.PATH_MULTIPASS)
    assert not oa.supported_path(512, 7)


def test_pao_host_ic_matches_oracle(oracle_mod):
    om, seed, log, visc = oa.ic_pao_host(512, 1e4, 10.0)
    o = oracle_mod.Oracle(512, 1e4, 10.0, 0.25)
    oseed, olog = o.init_pao()
    assert seed == oseed == 149
    assert log == olog
    assert np.array_equal(om, o.om2())
    assert visc == o.scalars().visc


def test_k4exp_host_ic_matches_oracle(oracle_mod):
    for N, K0, seed in ((64, 4.0, 0), (256, 8.0, 17), (512, 4.0, 3)):
        om = oa.ic_k4exp_host(N, K0, seed)
        o = oracle_mod.Oracle(N, 1e4, K0, 0.1)
        o.init_config(seed)
        assert np.array_equal(om, o.om2()), (N, K0, seed)


def test_create_fails_loudly_without_device():
    import torch  # noqa: F401  (only to ask whether a device is present)
    if torch.cuda.is_available():
        pytest.skip("a GPU is present; covered by the gpu tests")
    with pytest.raises(oa.DnsError) as ei:
        oa.Solver(64, 1e3, 4.0, 0.1)
    assert ei.value.status == 3  # DNS_E_HIP
    assert "device" in str(ei.value).lower()


def test_bad_arguments_are_rejected():
    lib = oa.load_library()
    cfg = oa.Config(30, 1e3, 4.0, 0.1, 0, 0, 0)
    h = ctypes.c_void_p()
    assert lib.dns_create(ctypes.byref(cfg), ctypes.byref(h)) == 1  # DNS_E_ARG: N % 4
    assert b"multiple of 4" in lib.dns_last_error(None)
    cfg = oa.Config(6000, 1e3, 4.0, 0.1, 0, 0, 0)
    assert lib.dns_create(ctypes.byref(cfg), ctypes.byref(h)) == 2  # DNS_E_UNSUPPORTED
    cfg = oa.Config(8192, 1e3, 4.0, 0.1, 0, 0, oa.PATH_FUSED)
    assert lib.dns_create(ctypes.byref(cfg), ctypes.byref(h)) == 2  # no fused kernels for N=8192
    cfg = oa.Config(512, 1e3, 4.0, 0.1, 0, 0, 9)
    assert lib.dns_create(ctypes.byref(cfg), ctypes.byref(h)) == 1  # unknown path
    # FP16 storage of the intermediates: N=4096 on the fused path only
    for N, path in ((1024, oa.PATH_AUTO), (4096, oa.PATH_MULTIPASS), (29400, oa.PATH_AUTO)):
        cfg = oa.Config(N, 1e3, 4.0, 0.1, oa.PREC_FP16_STORE, 0, path)
        assert lib.dns_create(ctypes.byref(cfg), ctypes.byref(h)) == 2, (N, path)
        assert b"FP16 storage" in lib.dns_last_error(None)
    cfg = oa.Config(512, 1e3, 4.0, 0.1, oa.PREC_FP16_STORE + 1, 0, 0)
    assert lib.dns_create(ctypes.byref(cfg), ctypes.byref(h)) == 2  # precision not built
    # the u_i u_j form of the fused kernels: FP64 at N=4096 only
    for N, prec in ((2048, oa.PREC_FP64), (29400, oa.PREC_FP64), (4096, oa.PREC_FP32), (4096, oa.PREC_FP16_STORE)):
        cfg = oa.Config(N, 1e3, 4.0, 0.1, prec, 0, oa.PATH_FUSED_UIUJ)
        assert lib.dns_create(ctypes.byref(cfg), ctypes.byref(h)) == 2, (N, prec)
    assert lib.dns_supported_path(4096, oa.PATH_FUSED_UIUJ) == 1
    assert lib.dns_supported_path(2048, oa.PATH_FUSED_UIUJ) == 0
    cfg = oa.Config(29400, 1e3, 4.0, 0.1, oa.PREC_FP32, 0, oa.PATH_MULTIPASS_UIUJ)
    assert lib.dns_create(ctypes.byref(cfg), ctypes.byref(h)) == 2  # the multi-pass u_i u_j form is FP64
    assert lib.dns_supported_path(29400, oa.PATH_MULTIPASS_UIUJ) == 1
    cfg = oa.Config(4096, 1e3, 4.0, 0.1, 0, 0, oa.PATH_MULTIPASS_UIUJ + 1)
    assert lib.dns_create(ctypes.byref(cfg), ctypes.byref(h)) == 1  # unknown path
    with pytest.raises(oa.DnsError):
        oa.ic_k4exp_host(30, 4.0, 0)


def test_ensemble_rejects_bad_configs_before_touching_a_device():
    with pytest.raises(oa.DnsError) as ei:
        oa.Ensemble([(100, 1e3, 4.0, 0.1, 0)])  # no fused kernels for N=100
    assert ei.value.status == 2
    with pytest.raises(oa.DnsError) as ei:
        oa.Ensemble([(1024, 1e3, 4.0, 0.1, 0)], filter_ratio=8)  # coarse grid 128 > 64
    assert ei.value.status == 2
    with pytest.raises(oa.DnsError) as ei:
        oa.Ensemble([(256, -1.0, 4.0, 0.1, 0)])
    assert ei.value.status == 1


def test_ensemble_create_fails_loudly_without_device():
    import torch  # noqa: F401  (only to ask whether a device is present)
    if torch.cuda.is_available():
        pytest.skip("a GPU is present; covered by the gpu tests")
    with pytest.raises(oa.DnsError) as ei:
        oa.Ensemble([(256, 1e3, 4.0, 0.1, 0)])
    assert ei.value.status == 3


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs one GPU; --gpus above 1 exits with status 2 and says why, before any
    device work."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "4"],
                       capture_output=True, text=True, timeout=120, cwd=root)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and r.stdout == ""
