"""Writes the committed golden fixtures under tests/golden/.

* fft_vectors.npz -- seeded random inputs and numpy (pocketfft, FP64) outputs for the
  1D complex DFT at the transform sizes the solver uses (and small 7-smooth sizes), and
  for the 2D R2C/C2R pair on small M x M grids.  Conventions match the solver:
  forward unnormalized e^{-i k x}, inverse unnormalized e^{+i k x}; the 2D inverse is
  M^2 * numpy.fft.irfft2 (imaginary parts of the kx=0 and kx=M/2 bins ignored).
* sim_output_kats.json -- the known-answer values printed by the reference runs
  (sim_output_*_gpu.txt, RTX 3090, FP32): the PAO initial-condition block, the final
  PAO LCG seed, the [NEXTDT INIT] line and the it=1 ITERATION line.  Only the numbers
  are kept (data), not the log files.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"

FFT_SIZES = [1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 14, 15, 16, 21, 24, 25, 35, 48, 49, 64, 96, 105, 225, 1225,
             128, 192, 210, 256, 384, 441, 512, 768, 1536, 2048, 2205, 3072]
GRID_SIZES = [12, 24, 48, 96]


def make_fft() -> None:
    rng = np.random.default_rng(20240611)
    arrays = {}
    for n in FFT_SIZES:
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        arrays[f"x_{n}"] = x
        arrays[f"fwd_{n}"] = np.fft.fft(x)
        arrays[f"inv_{n}"] = np.fft.ifft(x) * n
    for M in GRID_SIZES:
        ur = rng.standard_normal((M, M))
        arrays[f"ur_{M}"] = ur
        arrays[f"r2c_{M}"] = np.fft.fft(np.fft.rfft(ur, axis=1), axis=0)
        uc = rng.standard_normal((M, M // 2 + 1)) + 1j * rng.standard_normal((M, M // 2 + 1))
        arrays[f"uc_{M}"] = uc
        arrays[f"c2r_{M}"] = np.fft.irfft2(uc, s=(M, M)) * (M * M)
    np.savez_compressed(os.path.join(HERE, "fft_vectors.npz"), **arrays)


_PAO_KEYS = {
    "Reynolds n.": "Re", "Energy": "Energy", "WiWi": "WiWi", "Epsilon": "Epsilon", "a11": "a11",
    "e11": "e11", "Time scale": "Time_scale", "Kolmogorov": "Kolmogorov", "Viscosity": "Viscosity",
    "dx/Kol.": "dx_Kol", "2Pi/Nlamda": "2Pi_Nlamda", "2Pi/Lux": "2Pi_Lux", "2Pi/Luz": "2Pi_Luz",
    "2Pi/Lwx": "2Pi_Lwx", "2Pi/Lwz": "2Pi_Lwz", "Deps.": "Deps", "Ceps2": "Ceps2",
}


def parse_log(path: str) -> dict:
    out: dict = {"pao_text": {}, "pao": {}}
    with open(path, encoding="utf-8") as f:
        lines = f.read().splitlines()
    for ln in lines:
        m = re.match(r"^ N=(\d+), Re=(\d+), K0=(\d+), STEPS=(\d+), CFL=([\d.]+), UPDATE=(\d+)", ln)
        if m:
            out["args"] = {"N": int(m.group(1)), "Re": float(m.group(2)), "K0": float(m.group(3)),
                           "STEPS": int(m.group(4)), "CFL": float(m.group(5)),
                           "UPDATE": int(m.group(6))}
        m = re.match(r"^ (.+?)\s*=\s*(-?[\d.]+)\s*$", ln)
        if m and m.group(1).strip() in _PAO_KEYS:
            key = _PAO_KEYS[m.group(1).strip()]
            out["pao"][key] = float(m.group(2))
            out["pao_text"][key] = ln
        m = re.match(r"^ PAO seed\s*=\s*(\d+)", ln)
        if m:
            out["pao_seed"] = int(m.group(1))
        m = re.match(r"^ \[NEXTDT INIT\] CFLM=\s*([\d.]+)\s+DT=\s*([\d.]+)", ln)
        if m:
            out["nextdt_init"] = {"CFLM": float(m.group(1)), "DT": float(m.group(2))}
        m = re.match(r"^ ITERATION\s+1 T=([\d.]+) DT=([\d.]+) CN=([\d.]+) CFLM=([\d.]+) .* SIG=(\d+)", ln)
        if m:
            out["it1"] = {"T": float(m.group(1)), "DT": float(m.group(2)), "CN": float(m.group(3)),
                          "CFLM": float(m.group(4)), "SIG": int(m.group(5))}
    return out


def make_kats() -> None:
    kats = {}
    for name in ["sim_output_512_1001_gpu.txt", "sim_output_4096_501_gpu.txt",
                 "sim_output_16384_21_gpu.txt", "sim_output_16384_101_gpu.txt"]:
        d = parse_log(os.path.join(REF, name))
        d["source"] = name
        kats[name.replace(".txt", "")] = d
    with open(os.path.join(HERE, "sim_output_kats.json"), "w", encoding="utf-8") as f:
        json.dump(kats, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    make_fft()
    if os.path.isdir(REF):
        make_kats()
