"""The texture loader's JPEG decode (lts_load_image; the reference decodes with stb_image 2.27,
extern/stb_image.h, through stbi_loadf at core/ImageI.cuh:14) on the host.

Pinned three ways:
- against the reference's own decoder, compiled from extern/stb_image.h into
  oracle/_ref/libstb_image.so (run when that library exists);
- against tests/golden/jpeg_stb_sha256.json, the digests of that decoder's bytes
  (tests/golden/make_jpeg_golden.py), which hold without the reference tree;
- against closed forms: a file whose every block carries only a DC term decodes to the flat
  value the inverse DCT gives a constant block, and JFIF's conversion of grey (Cb = Cr = 128).

Inputs: the reference's res/blackDragon.jpg (512 x 512, 4:2:0, the texture of blackDragon.subject)
and synthetic files from tests/jpeg_synth.py covering every sampling layout, scan layout, restart
interval and quantisation precision of a baseline file.
"""
import ctypes
import hashlib
import json
import os

import numpy as np
import pytest

import jpeg_synth
from lighttransportsimulation_amd import LtsError, load_image

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
STB = os.path.join(ROOT, "oracle", "_ref", "libstb_image.so")
with open(os.path.join(HERE, "golden", "jpeg_stb_sha256.json")) as _f:
    GOLDEN = json.load(_f)

LUT = np.array([np.float32(np.float64(np.float32(c / np.float32(255))) ** np.float64(np.float32(2.2)))
                for c in range(256)], np.float32)


def _bytes_of(linear: np.ndarray) -> np.ndarray:
    """RGB bytes, top row first, of the loader's linear floats (bottom row first); the table is
    strictly increasing, so the bytes are recovered exactly."""
    idx = np.searchsorted(LUT, linear)
    assert np.array_equal(LUT[idx], linear), "decoded value outside the 8-bit table"
    return idx.astype(np.uint8)[::-1]


def _digest(rgb: np.ndarray) -> str:
    return f"{rgb.shape[1]}x{rgb.shape[0]}:" + hashlib.sha256(np.ascontiguousarray(rgb).tobytes()).hexdigest()[:32]


def _decode(tmp_path, data: bytes) -> np.ndarray:
    p = tmp_path / "img.jpg"
    p.write_bytes(data)
    return _bytes_of(load_image(str(p)))


@pytest.fixture(scope="module")
def stb():
    if not os.path.exists(STB):
        pytest.skip("oracle/_ref/libstb_image.so not built (needs the reference tree)")
    lib = ctypes.CDLL(STB)
    lib.stbi_load_from_memory.restype = ctypes.POINTER(ctypes.c_ubyte)

    def decode(data: bytes):
        x, y, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        p = lib.stbi_load_from_memory(data, len(data), ctypes.byref(x), ctypes.byref(y), ctypes.byref(c), 3)
        if not p:
            return None
        a = np.ctypeslib.as_array(p, shape=(y.value, x.value, 3)).copy()
        lib.stbi_image_free(p)
        return a
    return decode


def test_reference_texture_matches_the_golden_digest(tmp_path):
    with open(os.path.join(HERE, "golden", "res", "blackDragon.jpg"), "rb") as f:
        got = _decode(tmp_path, f.read())
    assert got.shape == (512, 512, 3)
    assert _digest(got) == GOLDEN["res/blackDragon.jpg"]


def test_reference_texture_matches_stb_image(tmp_path, stb):
    with open(os.path.join(HERE, "golden", "res", "blackDragon.jpg"), "rb") as f:
        data = f.read()
    assert np.array_equal(_decode(tmp_path, data), stb(data))


@pytest.mark.parametrize("layout", list(jpeg_synth.LAYOUTS))
def test_synthetic_files_match_the_golden_digests(tmp_path, layout):
    n = 0
    for name, kw in jpeg_synth.cases():
        if not name.startswith(layout + "-"):
            continue
        got = _decode(tmp_path, jpeg_synth.encode(**kw))
        assert _digest(got) == GOLDEN[name], name
        n += 1
    assert n > 0


@pytest.mark.parametrize("layout", list(jpeg_synth.LAYOUTS))
def test_synthetic_files_match_stb_image(tmp_path, stb, layout):
    for name, kw in jpeg_synth.cases():
        if not name.startswith(layout + "-"):
            continue
        data = jpeg_synth.encode(**kw)
        ref = stb(data)
        assert ref is not None, name
        assert np.array_equal(_decode(tmp_path, data), ref), name


@pytest.mark.parametrize("app,ids", [("adobe0", None), ("adobe1", None), (None, [82, 71, 66]),
                                     ("jfif", [82, 71, 66]), ("adobe0", [82, 71, 66]), (None, None)])
def test_colour_transform_flags_match_stb_image(tmp_path, stb, app, ids):
    """Adobe transform 0 without JFIF, or component ids 'R','G','B', mean RGB samples; otherwise
    YCbCr (stb_image.h:3835)."""
    data = jpeg_synth.encode(20, 13, [(2, 2), (1, 1), (1, 1)], seed=5, app=app, ids=ids)
    assert np.array_equal(_decode(tmp_path, data), stb(data))


def test_dc_only_values_follow_the_fixed_point_formula(tmp_path):
    """A block holding only a DC term is flat at 128 + DC / 8 (the inverse DCT of a constant block
    under JPEG's normalisation); in the decoder's fixed point that is 128 + ((d << 14) + 2^16) >> 17
    for the dequantised DC d, i.e. d / 8 rounded half up, then clamped to a byte."""
    for d_quant, q in [(0, 1), (1, 1), (3, 1), (4, 1), (-4, 1), (-5, 1), (100, 7), (-128, 8), (127, 8)]:
        data = _single_block_jpeg(d_quant, q)
        got = _decode(tmp_path, data)
        d = d_quant * q
        want = min(max(128 + (((d << 14) + (1 << 16)) >> 17), 0), 255)
        assert got.shape == (8, 8, 3)
        assert np.all(got == want), (d_quant, q, got[0, 0], want)


def _single_block_jpeg(dc: int, q: int) -> bytes:
    """An 8x8 grey JPEG of one block whose DC is `dc` (quantised) with DC quantiser q, written by
    hand: one DC code per category (all of length 4), an AC table holding only EOB."""
    import struct
    seg = jpeg_synth._segment
    out = bytearray(b"\xff\xd8")
    out += seg(0xDB, bytes([0]) + bytes([q] + [1] * 63))
    # DC table: 12 symbols of length 4
    out += seg(0xC4, bytes([0x00]) + bytes([0, 0, 0, 12] + [0] * 12) + bytes(range(12)))
    # AC table: EOB only, length 1
    out += seg(0xC4, bytes([0x10]) + bytes([1] + [0] * 15) + bytes([0]))
    out += seg(0xC0, bytes([8]) + struct.pack(">HH", 8, 8) + bytes([1, 1, 0x11, 0]))
    out += seg(0xDA, bytes([1, 1, 0x00, 0, 63, 0]))
    bw = jpeg_synth.BitWriter()
    s = jpeg_synth._category(dc)
    bw.put(s, 4)
    if s:
        bw.put(jpeg_synth._bits_of(dc, s), s)
    bw.put(0, 1)  # EOB
    bw.flush()
    out += bytes(bw.out) + b"\xff\xd9"
    return bytes(out)


@pytest.mark.parametrize("dc_y", [-1024, -40, 0, 40, 1016])
def test_grey_chroma_converts_to_equal_channels(tmp_path, dc_y):
    """JFIF conversion of Cb = Cr = 128 (chroma DC 0, no AC) gives R = G = B = Y."""
    got = _decode(tmp_path, _single_block_three_component(dc_y))
    y = min(max(128 + (((dc_y << 14) + (1 << 16)) >> 17), 0), 255)
    assert np.all(got == y)


def _single_block_three_component(dc_y: int) -> bytes:
    import struct
    seg = jpeg_synth._segment
    out = bytearray(b"\xff\xd8")
    out += seg(0xE0, b"JFIF\x00\x01\x01\x00\x00\x01\x00\x01\x00\x00")
    out += seg(0xDB, bytes([0]) + bytes([1] * 64))
    out += seg(0xC4, bytes([0x00]) + bytes([0, 0, 0, 12] + [0] * 12) + bytes(range(12)))
    out += seg(0xC4, bytes([0x10]) + bytes([1] + [0] * 15) + bytes([0]))
    out += seg(0xC0, bytes([8]) + struct.pack(">HH", 8, 8) + bytes([3, 1, 0x11, 0, 2, 0x11, 0, 3, 0x11, 0]))
    out += seg(0xDA, bytes([3, 1, 0x00, 2, 0x00, 3, 0x00, 0, 63, 0]))
    bw = jpeg_synth.BitWriter()
    for dc in (dc_y, 0, 0):
        s = jpeg_synth._category(dc)
        bw.put(s, 4)
        if s:
            bw.put(jpeg_synth._bits_of(dc, s), s)
        bw.put(0, 1)
    bw.flush()
    out += bytes(bw.out) + b"\xff\xd9"
    return bytes(out)


@pytest.mark.parametrize("label,kw,msg", [
    ("progressive", dict(dct_process=0xC2), "progressive"),
    ("12-bit", dict(precision=12), "8-bit"),
    ("lossless", dict(dct_process=0xC3), "lossless"),
])
def test_unsupported_processes_are_errors(tmp_path, label, kw, msg):
    data = jpeg_synth.encode(20, 13, [(2, 2), (1, 1), (1, 1)], seed=6, **kw)
    p = tmp_path / "img.jpg"
    p.write_bytes(data)
    with pytest.raises(LtsError, match=msg):
        load_image(str(p))


def test_cmyk_is_an_error(tmp_path):
    data = jpeg_synth.encode(16, 16, [(1, 1)] * 4, seed=2, app="adobe0")
    p = tmp_path / "img.jpg"
    p.write_bytes(data)
    with pytest.raises(LtsError, match="CMYK"):
        load_image(str(p))


def test_extended_sequential_frames_decode(tmp_path, stb):
    data = jpeg_synth.encode(20, 13, [(2, 2), (1, 1), (1, 1)], seed=6, dct_process=0xC1)
    assert np.array_equal(_decode(tmp_path, data), stb(data))


@pytest.mark.parametrize("cut", [0.1, 0.5, 0.9, -5, -2])
def test_truncated_files_are_errors(tmp_path, cut):
    data = jpeg_synth.encode(40, 33, [(2, 2), (1, 1), (1, 1)], seed=7, restart=2)
    n = int(len(data) * cut) if cut > 0 else len(data) + cut
    p = tmp_path / "img.jpg"
    p.write_bytes(data[:n])
    with pytest.raises(LtsError):
        load_image(str(p))


def test_missing_restart_marker_is_an_error(tmp_path):
    data = jpeg_synth.encode(40, 33, [(2, 2), (1, 1), (1, 1)], seed=7, restart=2)
    i = data.find(b"\xff\xd0")
    p = tmp_path / "img.jpg"
    p.write_bytes(data[:i] + b"\xff\xd9" + data[i + 2:])
    with pytest.raises(LtsError, match="restart"):
        load_image(str(p))


def test_undefined_huffman_table_is_an_error(tmp_path):
    data = jpeg_synth.encode(16, 16, [(1, 1)] * 3, seed=4)
    # drop the chroma AC table (the fourth DHT segment)
    at = -1
    for _ in range(4):
        at = data.find(b"\xff\xc4", at + 1)
    seg_len = int.from_bytes(data[at + 2:at + 4], "big")
    p = tmp_path / "img.jpg"
    p.write_bytes(data[:at] + data[at + 2 + seg_len:])
    with pytest.raises(LtsError, match="Huffman"):
        load_image(str(p))
