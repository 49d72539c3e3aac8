"""CPU tests of the mem_hip driver's boundary behaviour (mem_cuda.cu:1654-2082):
command line, exit codes, signals, output folder names and contents, CSV headers,
RESTART.  The driver is run against tests/mock_backend/mock_dns.c (a toy host
implementation of the same C ABI), copied next to the binary so its $ORIGIN rpath
picks the mock up.  The real solver's parts of these paths run in tests/test_driver_gpu.py."""
import os
import shutil
import signal
import subprocess
import time

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "optimizec_amd", "mem_hip")


@pytest.fixture()
def rundir(tmp_path):
    if not os.path.exists(DRIVER):
        pytest.skip("mem_hip not built")
    shutil.copy(DRIVER, tmp_path / "mem_hip")
    subprocess.run(["gcc", "-O1", "-shared", "-fPIC", "-o", str(tmp_path / "libdns_hip.so"),
                    os.path.join(ROOT, "tests", "mock_backend", "mock_dns.c"), "-lm"], check=True)
    return tmp_path


def run(rundir, args, env=None, timeout=60):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([str(rundir / "mem_hip"), *args], cwd=rundir, env=e, capture_output=True,
                          text=True, timeout=timeout)


def start(rundir, args, env=None):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.Popen([str(rundir / "mem_hip"), *args], cwd=rundir, env=e, stdout=subprocess.PIPE,
                            stderr=subprocess.PIPE, text=True)


def wait_for_iteration(p, it):
    """read stdout until the ITERATION line of `it` appears; returns the lines read"""
    lines = []
    t_end = time.time() + 30
    while time.time() < t_end:
        ln = p.stdout.readline()
        if not ln:
            break
        lines.append(ln)
        if ln.startswith(" ITERATION") and int(ln.split()[1]) >= it:
            return lines
    raise AssertionError("iteration line not seen:\n" + "".join(lines))


def test_usage_error(rundir):
    r = run(rundir, ["RESTART"])
    assert r.returncode == 2
    assert "Usage" in r.stderr


def test_restart_folder_without_meta(rundir):
    os.makedirs(rundir / "nofolder", exist_ok=True)
    r = run(rundir, ["RESTART", "nofolder"])
    assert r.returncode == 1


def test_normal_completion_writes_nothing(rundir):
    r = run(rundir, ["16", "1000", "4", "5", "0.1", "2", "0"])
    assert r.returncode == 0, r.stderr
    its = [int(ln.split()[1]) for ln in r.stdout.splitlines() if ln.startswith(" ITERATION")]
    assert its == [1, 2, 4, 5]  # it == 1, it % UPDATE == 0, it == STEPS
    assert "[OUTPUT] Normal completion; no output files written." in r.stdout
    assert " [NEXTDT INIT] CFLM=" in r.stdout and " Final T=" in r.stdout
    assert not [d for d in os.listdir(rundir) if d.startswith("output_")]


def test_defaults_for_nonpositive_update_and_steps(rundir):
    r = run(rundir, ["16", "1000", "4", "3", "0.1", "0", "0"])
    assert "UPDATE=100" in r.stdout


def _check_folder(folder, N):
    M = 3 * N // 2
    for name in ("u_velocity.pgm", "v_velocity.pgm", "kinetic.pgm", "omega.pgm", "stream.pgm",
                 "energy_spectrum.csv", "metrics.csv", "meta.csv"):
        assert os.path.exists(folder / name), name
    with open(folder / "omega.pgm", "rb") as f:
        data = f.read()
    header = f"P5\n{M} {M}\n255\n".encode()
    assert data.startswith(header) and len(data) == len(header) + M * M
    with open(folder / "meta.csv") as f:
        head, row = f.read().splitlines()[:2]
    assert head == "timestamp,N,K0,Re,CFL,visc,T,IT,Update,sigma,10K_pal_over_Zkmax2,minutes,FPS,backend"
    cols = row.split(",")
    assert cols[1] == str(N) and cols[-1] == "HIP"
    with open(folder / "metrics.csv") as f:
        assert f.readline().strip() == "N,K0,Re,CFL,VISC,DT,STEPS,PALIN,SIG,TIME,MINUTES,FPS"
    with open(folder / "energy_spectrum.csv") as f:
        assert f.readline().strip() == "normalized_radius,shell_sum_energy,count"
    return cols


def test_sigterm_dumps_and_exits_143(rundir):
    p = start(rundir, ["16", "1000", "4", "0", "0.1", "5", "0"], {"MOCK_STEP_MS": "5"})
    wait_for_iteration(p, 10)
    p.send_signal(signal.SIGTERM)
    out, _ = p.communicate(timeout=30)
    assert p.returncode == 128 + signal.SIGTERM
    assert "[SIGNAL] Received SIGTERM (15); dumping current state before exit." in out
    folders = [d for d in os.listdir(rundir) if d.startswith("output_16_4_1E03_0.10_")]
    assert len(folders) == 1
    cols = _check_folder(rundir / folders[0], 16)
    assert folders[0] == f"output_16_4_1E03_0.10_{cols[7]}"


def test_sigint_exits_without_saving(rundir):
    p = start(rundir, ["16", "1000", "4", "0", "0.1", "5", "0"], {"MOCK_STEP_MS": "5"})
    wait_for_iteration(p, 5)
    p.send_signal(signal.SIGINT)
    out, _ = p.communicate(timeout=30)
    assert p.returncode == 128 + signal.SIGINT
    assert "exiting without saving" in out
    assert not [d for d in os.listdir(rundir) if d.startswith("output_")]


def test_usr1_snapshot_then_continue(rundir):
    p = start(rundir, ["16", "1000", "4", "0", "0.1", "5", "0"], {"MOCK_STEP_MS": "5"})
    wait_for_iteration(p, 5)
    p.send_signal(signal.SIGUSR1)
    wait_for_iteration(p, 40)  # still running after the snapshot
    p.send_signal(signal.SIGINT)
    out, _ = p.communicate(timeout=30)
    folders = [d for d in os.listdir(rundir) if d.startswith("output_")]
    assert len(folders) == 1 and not folders[0].endswith("_restart")
    _check_folder(rundir / folders[0], 16)
    assert not os.path.exists(rundir / folders[0] / "restart.bin")


def test_rtmin_restart_folder_and_resume(rundir):
    p = start(rundir, ["16", "1000", "4", "0", "0.1", "5", "0"], {"MOCK_STEP_MS": "5"})
    wait_for_iteration(p, 5)
    p.send_signal(signal.SIGRTMIN)
    wait_for_iteration(p, 40)
    p.send_signal(signal.SIGINT)
    p.communicate(timeout=30)
    folders = [d for d in os.listdir(rundir) if d.endswith("_restart")]
    assert len(folders) == 1
    folder = folders[0]
    assert os.path.exists(rundir / folder / "restart.bin")
    cols = _check_folder(rundir / folder, 16)
    saved_it = int(cols[7])
    r = run(rundir, ["RESTART", folder], {"RESTART_STEPS": str(saved_it + 7)})
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"[LOAD] Loaded restart binary {folder}/restart.bin at iteration {saved_it}" in r.stdout
    its = [int(ln.split()[1]) for ln in r.stdout.splitlines() if ln.startswith(" ITERATION")]
    assert its[-1] == saved_it + 7 and its[0] > saved_it


def test_usr2_pause_and_cont(rundir):
    p = start(rundir, ["16", "1000", "4", "0", "0.1", "5", "0"], {"MOCK_STEP_MS": "5"})
    wait_for_iteration(p, 5)
    p.send_signal(signal.SIGUSR2)
    lines = []
    while True:
        ln = p.stdout.readline()
        lines.append(ln)
        if "pausing at iteration" in ln:
            break
    time.sleep(0.6)
    p.send_signal(signal.SIGCONT)
    while True:
        ln = p.stdout.readline()
        if "resuming after" in ln:
            paused = float(ln.split("after")[1].split()[0])
            break
    assert paused >= 0.5
    p.send_signal(signal.SIGINT)
    p.communicate(timeout=30)
    assert p.returncode == 128 + signal.SIGINT


def test_trailing_restart_argument(rundir):
    p = start(rundir, ["16", "1000", "4", "0", "0.1", "5", "0"], {"MOCK_STEP_MS": "5"})
    wait_for_iteration(p, 10)
    p.send_signal(signal.SIGRTMIN)
    wait_for_iteration(p, 40)
    p.send_signal(signal.SIGINT)
    p.communicate(timeout=30)
    folder = [d for d in os.listdir(rundir) if d.endswith("_restart")][0]
    r = run(rundir, ["16", "1000", "4", "60", "0.1", "5", "0", f"{folder}/restart.bin"])
    assert r.returncode == 0
    assert "[LOAD] Loaded restart binary" in r.stdout
    # a restart file of another N is rejected with exit code 1
    r = run(rundir, ["32", "1000", "4", "60", "0.1", "5", "0", f"{folder}/restart.bin"])
    assert r.returncode == 1


def test_adapt_visc_controller_and_palin_columns(rundir):
    """ADAPT_VISC=1: on every diagnostic step the driver feeds the backend's palinstrophy
    ratio (0.003 in the mock: 10000 x ratio = 30, above the target 20) to the controller
    (mem_cuda.cu:492-528), whose correction is clamped to +1%, so VISC grows by exactly
    exp(0.01) per diagnostic step; PALIN = 30 in the ITERATION lines and metrics.csv, and
    meta.csv's 10K_pal_over_Zkmax2 column carries it."""
    import math

    p = start(rundir, ["16", "1000", "4", "0", "0.1", "5", "1"], {"MOCK_STEP_MS": "5"})
    lines = wait_for_iteration(p, 20)
    p.send_signal(signal.SIGTERM)
    out, _ = p.communicate(timeout=30)
    assert p.returncode == 128 + signal.SIGTERM
    import re

    its = [dict(re.findall(r"(\w+)=\s*(\S+)", ln)) for ln in lines + out.splitlines(True)
           if ln.startswith(" ITERATION")]
    visc = [float(kv["VISC"]) for kv in its]
    palin = [int(kv["PALIN"]) for kv in its]
    assert all(pv == 30 for pv in palin)
    assert visc[0] == pytest.approx(1e-3 * math.exp(0.01), rel=1e-5)  # the first diagnostic step
    for a, b in zip(visc, visc[1:]):
        assert b / a == pytest.approx(math.exp(0.01), rel=1e-4)
    folder = [d for d in os.listdir(rundir) if d.startswith("output_16_4_1E03_0.10_")][0]
    with open(rundir / folder / "metrics.csv") as f:
        rows = [ln.strip().split(",") for ln in f][1:]
    assert rows and all(r[7] == "30" for r in rows)
    with open(rundir / folder / "meta.csv") as f:
        f.readline()
        meta = f.readline().strip().split(",")
    assert float(meta[10]) == pytest.approx(30.0)


def test_mov_frames(rundir):
    """MOV: a PNG of omega on every diagnostic step that is a multiple of UPDATE, numbered on
    from the last frame already in $MOV_DIR/hip_<N>_<K0>_<Re>_<CFL>_U<UPDATE>/
    (mem_cuda.cu:266-332, 1834-1910); the images decode (CRCs, zlib) and match the oracle's
    restatement of the frame renderer on the backend's field."""
    import numpy as np

    import outputs  # oracle/ (test infrastructure)

    movdir = rundir / "mov"
    folder = movdir / "hip_16_4_1E03_0.100_U5"
    folder.mkdir(parents=True)
    (folder / "omega_inferno_00007.png").write_bytes(b"")
    r = run(rundir, ["16", "1000", "4", "12", "0.1", "5", "0", "MOV"], {"MOV_DIR": str(movdir)})
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"[MOV] Saving omega_inferno_%05d.png every 5 steps to {folder} (F=1, next=00008)" in r.stdout
    names = sorted(p.name for p in folder.iterdir())
    assert names == ["omega_inferno_00007.png", "omega_inferno_00008.png", "omega_inferno_00009.png"]
    M = 24
    x = np.arange(M)
    field = np.sin(0.05 * x[None, :] * 3) * np.cos(0.03 * x[:, None])  # the mock's omega, up to a positive factor
    want, _ = outputs.movie_frame(field, 1)
    for n in ("omega_inferno_00008.png", "omega_inferno_00009.png"):
        w, h, rgb = outputs.read_png_rgb((folder / n).read_bytes())
        assert (w, h) == (M, M)
        assert np.mean(np.all(rgb == want, axis=2)) > 0.97
