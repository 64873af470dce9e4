"""Builds liblts.so (HIP for gfx950) in-tree with hipcc, and the CLI drop-in `lts_render`.

Floating-point flags are part of the product's definition: every fp32 operation is rounded on
its own (-ffp-contract=off), divide and sqrt are IEEE, denormals are kept, and packed-fp32
instructions are disabled (-packed-fp32-ops).
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "liblts.so")
CLI = os.path.join(PKG, "lts_render")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("LTS_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-Wno-unused-result",
          "-I", os.path.join(ROOT, "include"), "-I", CSRC]
DEVICE = [f"--offload-arch={ARCH}", "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

SOURCES = ["bvh.hip", "render.hip", "bdpt.hip", "wavefront.hip", "nrc.hip", "mesh.hip", "lts_capi.cpp", "scene_host.cpp",
           "image_io.cpp", "jpeg.cpp"]


def _obj(src: str) -> str:
    return os.path.join(BUILD, os.path.splitext(src)[0] + ".o")


def _compile(src: str) -> tuple[str, int, str]:
    path = os.path.join(CSRC, src)
    cmd = [HIPCC] + COMMON
    if src.endswith(".hip"):
        cmd += ["-x", "hip"] + DEVICE
    if src == "nrc.hip":
        # MFMA results in VGPRs: the network's activations feed the next layer's operands directly
        cmd += ["-mllvm", "-amdgpu-mfma-vgpr-form"]
    else:
        cmd += ["-D__HIP_PLATFORM_AMD__"]
    cmd += ["-c", path, "-o", _obj(src)]
    p = subprocess.run(cmd, capture_output=True, text=True)
    return src, p.returncode, p.stdout + p.stderr


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        results = list(ex.map(_compile, SOURCES))
    errors = [(s, out) for s, rc, out in results if rc != 0]
    for s, rc, out in results:
        if verbose and out.strip():
            print(f"--- {s}\n{out}")
    if errors:
        msg = "\n".join(f"--- {s}\n{out}" for s, out in errors)
        raise RuntimeError(f"hipcc failed:\n{msg}")
    tmp = LIB + ".tmp"
    link = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", tmp] + [_obj(s) for s in SOURCES] + ["-lz"]
    p = subprocess.run(link, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed:\n" + p.stdout + p.stderr)
    os.replace(tmp, LIB)
    cli_src = os.path.join(CSRC, "lts_render_main.cpp")
    if os.path.exists(cli_src):
        p = subprocess.run([HIPCC] + COMMON + ["-D__HIP_PLATFORM_AMD__", cli_src, "-o", CLI + ".tmp", "-L", PKG, "-llts",
                            "-Wl,-rpath,$ORIGIN"], capture_output=True, text=True)
        if p.returncode != 0:
            raise RuntimeError("CLI build failed:\n" + p.stdout + p.stderr)
        os.replace(CLI + ".tmp", CLI)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
