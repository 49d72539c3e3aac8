"""Register DFTs of the LDS FFT engine (optimizec_amd/csrc/fft_lds.hpp), checked on the
host: each radix, including the prime-factor composites 10, 14, 15, 21, 35 that the
multi-pass path's 210-point lines use and the 4 x 4 radix 16, against a direct DFT in long
double (forward and inverse, max abs error over unit-size inputs), and the radix plans:
210 = 15 x 14, 2048 = 8 x 8 x 8 x 4 (default) or 16 x 16 x 8 (radices up to 16),
1536 = 8 x 8 x 8 x 3 and 128 = 16 x 8 (radices up to 16); and the in-place
(decimation in frequency / in time) stages of fft_lines_ip with their position map ip_pos,
run butterfly by butterfly on the host, against a direct DFT."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_register_dfts_match_direct_dft(tmp_path):
    exe = str(tmp_path / "dft_check")
    subprocess.run([HIPCC, "-std=c++17", "-O1", f"-I{ROOT}/optimizec_amd/csrc",
                    os.path.join(ROOT, "tests", "native", "dft_check.cpp"), "-o", exe],
                   check=True, capture_output=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    radices = {}
    inplace = {}
    for line in out:
        f = line.split()
        if f and f[0] == "ip":
            inplace[(int(f[1]), int(f[2]))] = [float(x) for x in f[3:]]
        elif f and f[0].isdigit():
            radices[int(f[0])] = (float(f[1]), float(f[2]))
        elif f and f[0] == "plan210":
            assert [int(x) for x in f[1:]] == [2, 15, 14]
        elif f and f[0] == "plan2048":
            assert [int(x) for x in f[1:]] == [4, 8, 8, 8, 4]
        elif f and f[0] == "plan2048r16":
            assert [int(x) for x in f[1:]] == [3, 16, 16, 8]
        elif f and f[0] == "plan1536r16":
            assert [int(x) for x in f[1:]] == [4, 8, 8, 8, 3]
        elif f and f[0] == "plan128r16":
            assert [int(x) for x in f[1:]] == [2, 16, 8]
    assert sorted(radices) == [2, 3, 4, 5, 7, 8, 10, 14, 15, 16, 21, 35]
    for r, (fwd, inv) in radices.items():
        assert fwd < 1e-14 * r and inv < 1e-14 * r, (r, fwd, inv)
    # in-place plans: DIF (natural -> ip_pos order) and DIT (ip_pos order -> natural), both
    # signs; error per element of a unit-size input over sqrt(L)
    assert sorted(inplace) == [(96, 8), (128, 16), (210, 8), (768, 8), (1024, 8), (1536, 8), (2048, 8), (2048, 16)]
    for key, errs in inplace.items():
        assert len(errs) == 4 and max(errs) < 2e-14, (key, errs)
