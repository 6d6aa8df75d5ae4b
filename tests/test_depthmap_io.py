"""The depthmap files of opensfm/dataset.py:1080-1149 (np.savez in an LZ4
frame) written and read without the lz4 module: the LZ4 block decoder on
hand-assembled blocks from the specification (literal-only, overlapping
match, extended lengths, a match into the previous block of a linked
frame), the well-known empty frame of the lz4 tool, round trips through
the compressor (incompressible, repetitive, depth-like data, several
blocks), checksum failures, and the raw / clean depthmap files."""
import struct

import numpy as np
import pytest

from opensfm_amd import depthmap_io as io


def _frame(flg, bd, blocks, content_checksum=None):
    desc = bytes([flg, bd])
    hc = (io._xxh32(desc) >> 8) & 0xFF
    out = struct.pack("<I", io.MAGIC) + desc + bytes([hc])
    for raw, data in blocks:
        out += struct.pack("<I", len(data) | (0x80000000 if raw else 0)) + data
    out += struct.pack("<I", 0)
    if content_checksum is not None:
        out += struct.pack("<I", content_checksum)
    return out


def test_empty_frame_of_the_lz4_tool():
    # `lz4` on empty input: FLG 0x64 (v01, independent, content checksum),
    # BD 0x40, HC 0xA7, end mark, xxh32("") = 0x02CC5D05
    empty = bytes.fromhex("04224d186440a700000000055dcc02")
    assert io.lz4_frame_decompress(empty) == b""


def test_block_decoder_on_specification_sequences():
    # "abcd", then a match of 8 at offset 4, then 5 literals
    blk = bytes([0x44]) + b"abcd" + bytes([4, 0]) + bytes([0x50]) + b"efghi"
    assert io.lz4_frame_decompress(_frame(0x60, 0x40, [(False, blk)])) == b"abcdabcdabcdefghi"
    # overlapping match: one literal repeated (offset 1), match length 10
    blk = bytes([0x16]) + b"a" + bytes([1, 0]) + bytes([0x10]) + b"b"
    assert io.lz4_frame_decompress(_frame(0x60, 0x40, [(False, blk)])) == b"a" * 11 + b"b"
    # extended lengths: 20 literals (15 + 5), match 4 + 15 + 255 + 1 = 275
    lit = bytes(range(20))
    blk = bytes([0xFF, 5]) + lit + bytes([20, 0, 255, 1]) + bytes([0x00])
    want = lit + (lit * 14)[:275]
    assert io.lz4_frame_decompress(_frame(0x60, 0x40, [(False, blk)])) == want
    # linked blocks: the second block's match reaches into the first
    b1 = bytes([0xC0]) + b"hello world!"
    b2 = bytes([0x08, 12, 0, 0x10]) + b"."
    assert (io.lz4_frame_decompress(_frame(0x40, 0x40, [(False, b1), (False, b2)]))
            == b"hello world!hello world!.")
    # a stored (uncompressed) block and concatenated frames
    f = _frame(0x60, 0x40, [(True, b"xyz")])
    assert io.lz4_frame_decompress(f + f) == b"xyzxyz"


@pytest.mark.parametrize("kind", ["random", "zeros", "depth", "multi", "empty", "tiny"])
def test_round_trip(kind):
    rng = np.random.default_rng(0)
    if kind == "random":
        data = rng.integers(0, 256, 100000, dtype=np.uint8).tobytes()
    elif kind == "zeros":
        data = bytes(300000)
    elif kind == "depth":
        d = (5 + np.sin(np.arange(480 * 640) / 500.0)).astype(np.float32)
        d[rng.random(d.size) < 0.3] = 0
        data = d.tobytes()
    elif kind == "multi":
        data = (b"0123456789abcdef" * 20000) + rng.integers(0, 4, 70000, dtype=np.uint8).tobytes()
    elif kind == "empty":
        data = b""
    else:
        data = b"abc"
    for code in (4, 7):
        f = io.lz4_frame_compress(data, block_code=code)
        assert io.lz4_frame_decompress(f) == data
    if kind in ("zeros", "multi"):
        assert len(io.lz4_frame_compress(data)) < len(data) // 2
    if kind == "depth":                          # 30 % zeros: runs of 4-byte zeros match
        assert len(io.lz4_frame_compress(data)) < 0.9 * len(data)


def test_corrupt_frames_fail():
    f = bytearray(io.lz4_frame_compress(b"hello hello hello hello hello"))
    g = bytearray(f)
    g[6] ^= 1                                   # header (content size) byte
    with pytest.raises(ValueError):
        io.lz4_frame_decompress(bytes(g))
    g = bytearray(f)
    g[-1] ^= 1                                  # content checksum
    with pytest.raises(ValueError):
        io.lz4_frame_decompress(bytes(g))
    with pytest.raises(ValueError):
        io.lz4_frame_decompress(b"\x00\x01\x02\x03rest")


def test_raw_and_clean_depthmap_files(tmp_path):
    rng = np.random.default_rng(1)
    h, w = 60, 80
    depth = rng.uniform(1, 9, (h, w)).astype(np.float32)
    plane = rng.normal(0, 1, (h, w, 3)).astype(np.float32)
    score = rng.uniform(0, 1, (h, w)).astype(np.float32)
    nghbr = np.zeros((0,), np.int32)
    conf = rng.uniform(0, 1, (h, w)).astype(np.float32)
    p = str(tmp_path / "a.raw.npz")
    io.save_raw_depthmap(p, depth, plane, score, nghbr, ["s1", "s2"], conf)
    d, pl, sc, ng, ngs, cf = io.load_raw_depthmap(p)
    assert np.array_equal(d, depth) and np.array_equal(pl, plane) and np.array_equal(sc, score)
    assert ng.shape == (0,) and list(ngs) == ["s1", "s2"] and np.array_equal(cf, conf)
    p = str(tmp_path / "a.clean.npz")
    io.save_clean_depthmap(p, depth, plane, score)
    d, pl, sc, cf = io.load_clean_depthmap(p)
    assert np.array_equal(d, depth) and cf is None
    with open(p, "rb") as f:
        assert f.read(4) == bytes.fromhex("04224d18")
