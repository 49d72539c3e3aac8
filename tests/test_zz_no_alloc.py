"""dns_step_noalloc (the allocation-free step) against dns_step on a handle, through the
C ABI: two calls in a row on caller-owned device arrays, the second with a smaller grid,
each bitwise equal to the handle's result from the same state and scalars."""
import ctypes

import numpy as np
import pytest

import optimizec_amd as oa

pytestmark = pytest.mark.gpu


def _handle_state(N, seed, warm):
    s = oa.Solver(N, 1e4, 4.0, 0.1)
    s.init_ic(oa.IC_K4EXP, seed)
    s.start()
    s.step(warm)
    return s


def _noalloc(lib, torch, sc, om, fn, nsteps, stream):
    dom = torch.from_numpy(np.ascontiguousarray(om).view(np.float64)).to("cuda:0")
    dfn = torch.from_numpy(np.ascontiguousarray(fn).view(np.float64)).to("cuda:0")
    torch.cuda.synchronize()
    st = lib.dns_step_noalloc(ctypes.byref(sc), ctypes.c_void_p(dom.data_ptr()),
                              ctypes.c_void_p(dfn.data_ptr()), nsteps, ctypes.c_void_p(stream))
    assert st == 0, lib.dns_last_error(None)
    torch.cuda.synchronize()
    om_out = dom.cpu().numpy().view(np.complex128).reshape(om.shape)
    fn_out = dfn.cpu().numpy().view(np.complex128).reshape(fn.shape)
    return om_out, fn_out


def test_noalloc_step_matches_dns_step_twice():
    torch = pytest.importorskip("torch")
    lib = oa.load_library()
    stream = torch.cuda.current_stream().cuda_stream
    for N, seed, nsteps in ((512, 3, 12), (128, 4, 7)):  # second call: a smaller grid
        print(f"noalloc: N={N} seed={seed} steps={nsteps}", flush=True)
        s = _handle_state(N, seed, 5)
        sc = s.scalars()
        om, fn = s.om2(), s.fnm1()
        om_n, fn_n = _noalloc(lib, torch, sc, om, fn, nsteps, stream)
        s.step(nsteps)
        ref = s.scalars()
        assert np.array_equal(om_n, s.om2())
        assert np.array_equal(fn_n, s.fnm1())
        assert (sc.it, sc.t, sc.cnm1) == (ref.it, ref.t, ref.cnm1)
        assert not np.array_equal(om_n, om)  # it stepped
        s.close()


def test_noalloc_rejects_sizes_beyond_the_scratch():
    torch = pytest.importorskip("torch")
    lib = oa.load_library()
    sc = oa.Scalars()
    sc.N, sc.dt, sc.visc = 1024, 1e-3, 1e-4
    buf = torch.zeros(2 * 1024 * 512, dtype=torch.float64, device="cuda:0")
    st = lib.dns_step_noalloc(ctypes.byref(sc), ctypes.c_void_p(buf.data_ptr()),
                              ctypes.c_void_p(buf.data_ptr()), 1, None)
    assert st == 1  # DNS_E_ARG
    assert b"exceeds the scratch limit" in lib.dns_last_error(None)
    sc.N = 96  # not a built grid
    assert lib.dns_step_noalloc(ctypes.byref(sc), ctypes.c_void_p(buf.data_ptr()),
                                ctypes.c_void_p(buf.data_ptr()), 1, None) == 2
