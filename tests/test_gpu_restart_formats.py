"""restart.bin files of the reference's FP32 build (RestartHeaderV1 / RestartHeader with
real = float, mem_cuda.cu:1019-1054, 1236-1394) loaded by the FP64 solver: the file is
written here from the format (header fields, payload order) with float values, and the
loaded om2, fnm1 and scalars must equal those values upcast to FP64.  Files the loader must
refuse (bad magic, unknown version, wrong header size, dimensions or counts, truncated
payload) fail with an error and leave the state untouched."""
import os
import struct

import numpy as np
import pytest

import optimizec_amd as oa

pytestmark = pytest.mark.gpu

N = 64
H, M, NK = N // 2, 3 * N // 2, 3 * N // 4 + 1


def _fields(seed):
    rng = np.random.default_rng(seed)
    om = (rng.standard_normal((N, H)) + 1j * rng.standard_normal((N, H))).astype(np.complex64)
    fn = (rng.standard_normal((N, H)) + 1j * rng.standard_normal((N, H))).astype(np.complex64)
    om[0, 0] = 0
    return om, fn


SC = dict(Re=2500.0, K0=6.0, visc=4e-4, t=0.125, dt=1.5e-3, cn=0.97, cnm1=0.98, cflnum=0.2, cflm=12.5)


def _write(path, version, om, fn, *, magic=b"CTRB_RESTART", header_bytes=None, n=N, counts=None, cut=0):
    """An FP32 restart file: 72 bytes of magic/u32/i32 fields, 9 float scalars, padding to
    8 bytes, u64 counts (uc, om2, fnm1[, ur]), then the payload (v1: UC, om2, fnm1;
    v2: UR, om2, fnm1)."""
    hb = (136 if version == 1 else 144) if header_bytes is None else header_bytes
    ints = struct.pack("<16s4I10i", magic.ljust(16, b"\0"), version, hb, 4, 8,
                       n, n, n, n, 3 * n // 4 + 1, 3 * n // 2, 3 * n // 2, 3 * n // 4 + 1, 17, 1)
    sc = struct.pack("<9f", *(SC[k] for k in ("Re", "K0", "visc", "t", "dt", "cn", "cnm1", "cflnum", "cflm")))
    pad = b"\0" * ((-(72 + 36)) % 8)
    uc_n, om_n, ur_n = 3 * M * NK, N * H, 3 * M * M
    if counts is None:
        counts = (uc_n, om_n, om_n) if version == 1 else (0, om_n, om_n, ur_n)
    cnt = struct.pack(f"<{len(counts)}Q", *counts)
    head = ints + sc + pad + cnt
    assert len(head) == (136 if version == 1 else 144)
    rng = np.random.default_rng(1)
    first = (rng.standard_normal(2 * uc_n) if version == 1 else rng.standard_normal(ur_n)).astype(np.float32)
    payload = first.tobytes() + om.tobytes() + fn.tobytes()
    with open(path, "wb") as f:
        f.write(head[:hb] if header_bytes is not None else head)
        f.write(payload[: len(payload) - cut])


@pytest.mark.parametrize("version", [1, 2])
def test_fp32_restart_loads_upcast(tmp_path, version):
    print(f"FP32 restart v{version} -> FP64 solver", flush=True)
    om, fn = _fields(version)
    p = str(tmp_path / "restart.bin")
    _write(p, version, om, fn)
    s = oa.Solver(N, 1e4, 4.0, 0.1)
    s.load_restart(p)
    assert np.array_equal(s.om2(), om.astype(np.complex128))
    assert np.array_equal(s.fnm1(), fn.astype(np.complex128))
    sc = s.scalars()
    for k, v in SC.items():
        assert getattr(sc, k) == float(np.float32(v)), k
    assert sc.it == 17 and sc.ifn == 1
    s.step(2)  # the loaded state steps
    assert np.all(np.isfinite(s.om2()))
    s.close()


def test_bad_restart_files_are_refused(tmp_path):
    print("malformed restart files", flush=True)
    om, fn = _fields(3)
    s = oa.Solver(N, 1e4, 4.0, 0.1)
    s.init_ic(oa.IC_K4EXP, 0)
    before = s.om2()
    cases = {
        "magic": dict(magic=b"NOT_A_RESTART"),
        "size": dict(header_bytes=140),
        "dims": dict(n=2 * N),
        "counts": dict(counts=(0, N * H + 1, N * H, 3 * M * M)),
        "truncated": dict(cut=100),
    }
    for name, kw in cases.items():
        p = str(tmp_path / f"{name}.bin")
        _write(p, 2, om, fn, **kw)
        with pytest.raises(oa.DnsError):
            s.load_restart(p)
    p = str(tmp_path / "v3.bin")
    _write(p, 2, om, fn)
    with open(p, "r+b") as f:
        f.seek(16)
        f.write(struct.pack("<I", 3))
    with pytest.raises(oa.DnsError):
        s.load_restart(p)
    assert np.array_equal(s.om2(), before)
    s.close()
