"""Builds the gfx950 solver library ``optimizec_amd/libdns_hip.so`` and the host driver
``optimizec_amd/mem_hip`` in-tree.

* ``csrc/dns_kernels.hip`` (fused kernels), ``csrc/dns_big.hip`` (multi-pass path) --
  device code, hipcc --offload-arch=gfx950 (packed FP32 ops disabled for every HIP
  translation unit); the two compile in parallel.
* ``csrc/dns_api.cpp``, ``csrc/dns_ensemble.cpp``, ``csrc/ic_host.cpp`` -- host code, g++ with -ffp-contract=off so
  the FP32 parts of the reference initial condition round exactly like the reference.
* ``driver/mem_hip.cpp`` -- the command-line driver, linked against the library.

Usage: ``python -m optimizec_amd.build`` (or ``python optimizec_amd/build.py``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
ARCH = os.environ.get("DNS_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function"]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra",
              "-Wno-unused-parameter", "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include"]

LIB = os.path.join(PKG, "libdns_hip.so")
DRIVER = os.path.join(PKG, "mem_hip")


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    procs = []
    for src in ("dns_kernels.hip", "dns_big.hip", "dns_ens_kernels.hip", "cnn_kernels.hip", "closure_kernels.hip",
                "dns_static.hip"):
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        cmd = [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        print("+", " ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd)))
        objs.append(obj)
    for cmd, p in procs:
        if p.wait() != 0:
            raise subprocess.CalledProcessError(p.returncode, cmd)
    for src in ("dns_api.cpp", "dns_ensemble.cpp", "dns_closure.cpp", "ic_host.cpp"):
        obj = os.path.join(BUILD, src.replace(".cpp", ".o"))
        _run(["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, LIB)
    drv_src = os.path.join(PKG, "driver", "mem_hip.cpp")
    if os.path.exists(drv_src):
        _run(["g++", *HOST_FLAGS, "-O2", drv_src, "-o", DRIVER, f"-L{PKG}", "-ldns_hip",
              "-Wl,-rpath,$ORIGIN", "-lpthread"])
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
