"""Writes tests/golden/jpeg_stb_sha256.json: the SHA-256 of the RGB bytes that the reference's
image decoder, stb_image 2.27 (extern/stb_image.h, compiled into oracle/_ref/libstb_image.so by
oracle/Makefile), returns for the reference's res/blackDragon.jpg and for every synthetic JPEG of
tests/jpeg_synth.py, with 3 components requested as stbi_loadf does (core/ImageI.cuh:14).

Run from the repository root after `make -C oracle`: python tests/golden/make_jpeg_golden.py
"""
import ctypes
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))

import jpeg_synth  # noqa: E402


def stb_rgb(lib, data: bytes):
    x, y, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    p = lib.stbi_load_from_memory(data, len(data), ctypes.byref(x), ctypes.byref(y), ctypes.byref(c), 3)
    if not p:
        return None
    a = np.ctypeslib.as_array(p, shape=(y.value, x.value, 3)).copy()
    lib.stbi_image_free(p)
    return a


def load_stb():
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libstb_image.so"))
    lib.stbi_load_from_memory.restype = ctypes.POINTER(ctypes.c_ubyte)
    return lib


def digest(rgb: np.ndarray) -> str:
    return f"{rgb.shape[1]}x{rgb.shape[0]}:" + hashlib.sha256(rgb.tobytes()).hexdigest()[:32]


def main():
    lib = load_stb()
    out = {}
    with open(os.path.join(HERE, "res", "blackDragon.jpg"), "rb") as f:
        out["res/blackDragon.jpg"] = digest(stb_rgb(lib, f.read()))
    for name, kw in jpeg_synth.cases():
        rgb = stb_rgb(lib, jpeg_synth.encode(**kw))
        out[name] = digest(rgb)
    with open(os.path.join(HERE, "jpeg_stb_sha256.json"), "w") as f:
        json.dump(out, f, indent=0, sort_keys=False)
        f.write("\n")
    print(f"{len(out)} digests")


if __name__ == "__main__":
    main()
