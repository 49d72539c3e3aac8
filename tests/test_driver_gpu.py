"""The mem_hip driver on the real solver: its stdout against the reference's own log
(sim_output_512_1001_gpu.txt, FP32 on an RTX 3090) and a SIGRTMIN snapshot +
RESTART round trip through the FP64 restart.bin."""
import json
import os
import signal
import struct
import subprocess
import time

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "optimizec_amd", "mem_hip")
GOLD = os.path.join(ROOT, "tests", "golden", "sim_output_kats.json")


def test_driver_matches_reference_log(tmp_path):
    with open(GOLD, encoding="utf-8") as f:
        kat = json.load(f)["sim_output_512_1001_gpu"]
    r = subprocess.run([DRIVER, "512", "1e4", "10", "3", "0.25", "100", "0"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = r.stdout.splitlines()
    for text in kat["pao_text"].values():
        assert text in out, text
    assert f" PAO seed    ={kat['pao_seed']:12d}" in out
    init = [ln for ln in out if ln.startswith(" [NEXTDT INIT]")][0]
    cflm = float(init.split("CFLM=")[1].split()[0])
    assert abs(cflm / kat["nextdt_init"]["CFLM"] - 1) < 3e-7  # FP32 reference
    it1 = [ln for ln in out if ln.startswith(" ITERATION      1 ")][0]
    fields = dict(tok.split("=") for tok in it1.split()[2:] if "=" in tok)
    assert fields["T"] == f"{kat['it1']['T']:.10f}"
    assert abs(float(fields["CN"]) - kat["it1"]["CN"]) < 1e-7
    assert abs(float(fields["CFLM"]) / kat["it1"]["CFLM"] - 1) < 3e-7
    assert int(fields["SIG"]) == kat["it1"]["SIG"]


def _iterations(stdout):
    rows = {}
    for ln in stdout.splitlines():
        if ln.startswith(" ITERATION"):
            f = dict(tok.split("=") for tok in ln.split()[2:] if "=" in tok)
            rows[int(ln.split()[1])] = f
    return rows


def test_driver_multipass_lines_match_fused(tmp_path):
    """The driver on the multi-pass path (DNS_PATH=multipass, N=1024): the step after each
    diagnostic step reuses the velocity rows NEXTDT leaves in place (display sigma runs
    before NEXTDT), and the ITERATION lines equal the fused path's to the printed digits."""
    args = ["1024", "1e4", "4", "60", "0.1", "20", "0"]
    env = dict(os.environ, DNS_IC="k4exp", DNS_SEED="1")
    runs = {}
    for path in ("fused", "multipass"):
        r = subprocess.run([DRIVER, *args], cwd=tmp_path, env=dict(env, DNS_PATH=path), capture_output=True,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        runs[path] = _iterations(r.stdout)
    assert sorted(runs["fused"]) == [1, 20, 40, 60]
    for it, f in runs["fused"].items():
        m = runs["multipass"][it]
        for key in ("T", "DT", "CN", "CFLM", "SIG"):
            assert m[key] == f[key], (it, key, m[key], f[key])


@pytest.mark.parametrize("prec,banner,tol", [("fp32", "precision: FP32 transforms, FP64 state", 1e-6),
                                             ("bf16x3", "precision: FP32 transforms with 3xBF16 matrix-core", 1e-6),
                                             ("fp16", "precision: FP32 transforms with FP16 matrix-core", 1e-4)])
def test_driver_mixed_precision_knob(tmp_path, prec, banner, tol):
    """DNS_PREC=fp32 / bf16x3 / fp16 runs the mixed-precision step from the command line: the
    ITERATION lines follow the FP64 run's (T, DT, CN, CFLM to 1e-6 relative, 1e-4 with the
    FP16 products; NEXTDT runs in FP64) and the run announces its precision."""
    args = ["512", "1e4", "4", "200", "0.1", "50", "0"]
    env = dict(os.environ, DNS_IC="k4exp", DNS_SEED="2")
    r64 = subprocess.run([DRIVER, *args], cwd=tmp_path, env=env, capture_output=True, text=True, timeout=120)
    r32 = subprocess.run([DRIVER, *args], cwd=tmp_path, env=dict(env, DNS_PREC=prec), capture_output=True,
                         text=True, timeout=120)
    assert r64.returncode == 0 and r32.returncode == 0, r64.stderr + r32.stderr
    assert banner in r32.stdout and "precision:" not in r64.stdout
    a, b = _iterations(r64.stdout), _iterations(r32.stdout)
    assert sorted(a) == sorted(b) == [1, 50, 100, 150, 200]
    for it in a:
        for key in ("T", "DT", "CN", "CFLM"):
            x, y = float(a[it][key]), float(b[it][key])
            assert abs(y - x) <= tol * abs(x), (it, key, x, y)


def test_driver_restart_roundtrip(tmp_path):
    """SIGRTMIN writes output_..._restart/restart.bin (v2, 8-byte reals); RESTART continues
    to the same state as an uninterrupted run (bitwise: final T line)."""
    env = dict(os.environ, DNS_IC="k4exp", DNS_SEED="3")
    args = ["256", "1e4", "4", "0", "0.1", "5", "0"]
    p = subprocess.Popen([DRIVER, *args], cwd=tmp_path, env=env, stdout=subprocess.PIPE, text=True)
    t_end = time.time() + 60
    while time.time() < t_end:
        ln = p.stdout.readline()
        if ln.startswith(" ITERATION") and int(ln.split()[1]) >= 20:
            break
    p.send_signal(signal.SIGRTMIN)
    t_end = time.time() + 60
    while time.time() < t_end:
        ln = p.stdout.readline()
        if ln.startswith("[CSV] Wrote") and "meta.csv" in ln:
            break
    p.send_signal(signal.SIGINT)
    p.communicate(timeout=60)
    assert p.returncode == 128 + signal.SIGINT
    folder = [d for d in os.listdir(tmp_path) if d.endswith("_restart")][0]
    path = tmp_path / folder / "restart.bin"
    with open(path, "rb") as f:
        head = f.read(176)
    magic = head[:16].rstrip(b"\0")
    version, header_bytes, real_bytes, cplx_bytes = struct.unpack("<4I", head[16:32])
    assert magic == b"CTRB_RESTART" and (version, header_bytes, real_bytes, cplx_bytes) == (2, 176, 8, 16)
    it_saved = struct.unpack("<10i", head[32:72])[8]
    N, M = 256, 384
    assert os.path.getsize(path) == 176 + 8 * 3 * M * M + 2 * 16 * N * (N // 2)
    stop = it_saved + 13
    r1 = subprocess.run([DRIVER, "RESTART", folder], cwd=tmp_path, env=dict(env, RESTART_STEPS=str(stop)),
                        capture_output=True, text=True, timeout=120)
    assert r1.returncode == 0, r1.stdout + r1.stderr
    r2 = subprocess.run([DRIVER, "256", "1e4", "4", str(stop), "0.1", "5", "0"], cwd=tmp_path, env=env,
                        capture_output=True, text=True, timeout=120)
    assert r2.returncode == 0
    last1 = [ln for ln in r1.stdout.splitlines() if ln.startswith(" ITERATION")][-1]
    last2 = [ln for ln in r2.stdout.splitlines() if ln.startswith(" ITERATION")][-1]
    strip = lambda ln: ln.split(" FPS=")[0]  # noqa: E731
    assert strip(last1) == strip(last2)
    final1 = [ln for ln in r1.stdout.splitlines() if ln.startswith(" Final T=")][0]
    final2 = [ln for ln in r2.stdout.splitlines() if ln.startswith(" Final T=")][0]
    assert final1 == final2


def test_driver_output_files_match_oracle(oracle_mod, tmp_path):
    """SIGTERM dump of a real run (N=64, k4exp seed 2): the five PGMs and
    energy_spectrum.csv against the oracle's restatement of the reference writers
    (oracle/outputs.py) applied to the oracle's fields at the same iteration."""
    import numpy as np

    import outputs  # oracle/ (test infrastructure)

    N, M = 64, 96
    print("mem_hip output files vs oracle writers, N=64", flush=True)
    env = dict(os.environ, DNS_IC="k4exp", DNS_SEED="2")
    p = subprocess.Popen([DRIVER, str(N), "1e4", "4", "0", "0.1", "5", "0"], cwd=tmp_path, env=env,
                         stdout=subprocess.PIPE, text=True)
    t_end = time.time() + 60
    while time.time() < t_end:
        ln = p.stdout.readline()
        if ln.startswith(" ITERATION") and int(ln.split()[1]) >= 10:
            break
    p.send_signal(signal.SIGTERM)
    p.communicate(timeout=60)
    assert p.returncode == 128 + signal.SIGTERM
    folder = [d for d in os.listdir(tmp_path) if d.startswith("output_64_")][0]
    it = int(folder.rsplit("_", 1)[1])
    o = oracle_mod.Oracle(N, 1e4, 4.0, 0.1)
    o.init_config(2)
    o.start(form=1)
    o.run(it, 5, form=1)
    o.velocity_from_om2()
    u, w = o.ur(0).copy(), o.ur(1).copy()
    o.om2phys()
    om_phys = o.ur(2).copy()
    om = o.om2()
    kx = np.arange(N // 2)[None, :]
    kz = np.where(np.arange(N) < N // 2, np.arange(N), np.arange(N) - N)[:, None]
    k2 = (kx ** 2 + kz ** 2).astype(float)
    g = np.zeros((M, 3 * N // 4 + 1), dtype=np.complex128)
    psi = np.where(k2 > 0, om / np.where(k2 > 0, k2, 1.0), 0.0)
    g[: N // 2, : N // 2] = psi[: N // 2]
    g[N:, : N // 2] = psi[N // 2:]
    fields = {"u_velocity.pgm": (u, False), "v_velocity.pgm": (w, False),
              "kinetic.pgm": (np.hypot(u, w), False), "omega.pgm": (om_phys, True),
              "stream.pgm": (oracle_mod.c2r_2d(g), False)}
    for name, (f, clip) in fields.items():
        with open(tmp_path / folder / name, "rb") as fh:
            data = fh.read()
        head = f"P5\n{M} {M}\n255\n".encode()
        assert data.startswith(head)
        got = np.frombuffer(data[len(head):], dtype=np.uint8).reshape(M, M).astype(int)
        want = outputs.pgm_pixels(f, clipped=clip).astype(int)
        diff = np.abs(got - want)
        assert diff.max() <= 1 and np.mean(diff > 0) < 1e-3, (name, diff.max(), np.mean(diff > 0))
    with open(tmp_path / folder / "energy_spectrum.csv") as fh:
        fh.readline()
        got = [tuple(float(x) for x in ln.split(",")) for ln in fh if ln.strip()]
    want = outputs.energy_spectrum(u, w)
    total = sum(e for _, e, _ in want)
    # bins below 1e-10 of the total hold the R2C's roundoff (oracle) or the far spectral
    # tail (both); compare the rest, to 1e-9 where the shell carries 1e-6 of the energy
    got = [r for r in got if r[1] > 1e-10 * total]
    want = [r for r in want if r[1] > 1e-10 * total]
    assert len(got) == len(want) and len(want) > 20
    for (rg, eg, cg), (rw, ew, cw) in zip(got, want):
        assert rg == rw and int(cg) == cw
        assert abs(eg / ew - 1) < (1e-9 if ew > 1e-6 * total else 1e-3), (rg, eg, ew)
    o.close()


def test_driver_mov_frames_match_oracle(oracle_mod, tmp_path):
    """MOV frames of a real run (N=64, k4exp seed 4, UPDATE=5): the frame of step 10
    against the oracle's frame renderer on the oracle's omega at step 10."""
    import numpy as np

    import outputs  # oracle/ (test infrastructure)

    N, M = 64, 96
    print("mem_hip MOV frames vs oracle, N=64", flush=True)
    env = dict(os.environ, DNS_IC="k4exp", DNS_SEED="4", MOV_DIR=str(tmp_path / "mov"))
    r = subprocess.run([DRIVER, str(N), "1e4", "4", "10", "0.1", "5", "0", "MOV"], cwd=tmp_path, env=env,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    folder = tmp_path / "mov" / "hip_64_4_1E04_0.100_U5"
    assert sorted(p.name for p in folder.iterdir()) == ["omega_inferno_00001.png", "omega_inferno_00002.png"]
    o = oracle_mod.Oracle(N, 1e4, 4.0, 0.1)
    o.init_config(4)
    o.start(form=1)
    o.run(10, 5, form=1)
    o.om2phys()
    want, sigma = outputs.movie_frame(o.ur(2), outputs.movie_scale_f(N))
    w, h, rgb = outputs.read_png_rgb((folder / "omega_inferno_00002.png").read_bytes())
    assert (w, h) == (M, M)
    same = np.mean(np.all(rgb == want, axis=2))
    print(f"sigma {sigma:.2f}, identical pixels {same:.5f}")
    assert same > 0.999
    assert f"(Inferno PNG, {M}x{M}, F=1, comp=2, sigma={int(sigma)})" in r.stdout
    o.close()
