"""The depthmap files of the reference pipeline, without the ``lz4`` module.

``opensfm/dataset.py:1080-1149`` stores raw and clean depthmaps as
``np.savez`` archives wrapped in an LZ4 frame (``lz4.frame``).  This module
writes and reads the same files: the LZ4 frame layer (magic 0x184D2204,
frame descriptor with its xxh32 header checksum, blocks with the
uncompressed flag, end mark, optional checksums, linked or independent
blocks) here, the LZ4 block codec in the engine library
(``pm_lz4_block_compress`` / ``pm_lz4_block_decompress``).  Written frames
use independent 4 MB blocks, the content size and a content checksum; any
valid LZ4 frame is read, including ``lz4.frame.compress``'s default (64 KB
linked blocks with the content size).
"""
from __future__ import annotations

import ctypes as C
import io
import struct
from typing import List, Optional

import numpy as np

from . import _capi

MAGIC = 0x184D2204
_BLOCK_MAX = {4: 64 << 10, 5: 256 << 10, 6: 1 << 20, 7: 4 << 20}


def _xxh32(data: bytes) -> int:
    import xxhash
    return xxhash.xxh32_intdigest(data, seed=0)


def lz4_frame_compress(data: bytes, block_code: int = 7) -> bytes:
    """One LZ4 frame: independent blocks of _BLOCK_MAX[block_code] bytes,
    each stored compressed when that is smaller, content size and
    checksum."""
    lib = _capi.load()
    data = bytes(data)
    flg = (1 << 6) | (1 << 5) | (1 << 3) | (1 << 2)   # version 01, B.Indep, C.Size, C.Checksum
    bd = block_code << 4
    desc = bytes([flg, bd]) + struct.pack("<Q", len(data))
    out = bytearray(struct.pack("<I", MAGIC) + desc + bytes([(_xxh32(desc) >> 8) & 0xFF]))
    bmax = _BLOCK_MAX[block_code]
    src = np.frombuffer(data, np.uint8)
    cap = bmax + bmax // 255 + 16
    dst = np.empty(cap, np.uint8)
    for start in range(0, len(data), bmax):
        chunk = src[start:start + bmax]
        n = C.c_int64(0)
        _capi.check(lib.pm_lz4_block_compress(
            _capi.ptr(np.ascontiguousarray(chunk), C.c_uint8), len(chunk),
            _capi.ptr(dst, C.c_uint8), cap, C.byref(n)))
        if n.value < len(chunk):
            out += struct.pack("<I", n.value) + dst[:n.value].tobytes()
        else:
            out += struct.pack("<I", len(chunk) | 0x80000000) + chunk.tobytes()
    out += struct.pack("<I", 0)
    out += struct.pack("<I", _xxh32(data))
    return bytes(out)


def lz4_frame_decompress(frame: bytes) -> bytes:
    """Decodes one or more concatenated LZ4 frames."""
    lib = _capi.load()
    frame = bytes(frame)
    pos = 0
    result = bytearray()
    while pos < len(frame):
        (magic,) = struct.unpack_from("<I", frame, pos)
        if 0x184D2A50 <= magic <= 0x184D2A5F:          # skippable frame
            (size,) = struct.unpack_from("<I", frame, pos + 4)
            pos += 8 + size
            continue
        if magic != MAGIC:
            raise ValueError("not an LZ4 frame")
        flg, bd = frame[pos + 4], frame[pos + 5]
        if flg >> 6 != 1:
            raise ValueError("unsupported LZ4 frame version")
        block_checksum = bool(flg & 0x10)
        has_size = bool(flg & 0x08)
        content_checksum = bool(flg & 0x04)
        has_dict = bool(flg & 0x01)
        q = pos + 6
        size = None
        if has_size:
            (size,) = struct.unpack_from("<Q", frame, q)
            q += 8
        if has_dict:
            raise ValueError("LZ4 frames with a dictionary are not supported")
        if (_xxh32(frame[pos + 4:q]) >> 8) & 0xFF != frame[q]:
            raise ValueError("LZ4 frame header checksum mismatch")
        q += 1
        bmax = _BLOCK_MAX.get((bd >> 4) & 7)
        if bmax is None:
            raise ValueError("bad LZ4 block size code")
        cap = size if size is not None else bmax
        out = np.empty(max(cap, 1), np.uint8)
        op = 0
        while True:
            (bsize,) = struct.unpack_from("<I", frame, q)
            q += 4
            if bsize == 0:
                break
            raw = bool(bsize & 0x80000000)
            bsize &= 0x7FFFFFFF
            block = frame[q:q + bsize]
            if len(block) != bsize:
                raise ValueError("truncated LZ4 block")
            q += bsize
            if block_checksum:
                if struct.unpack_from("<I", frame, q)[0] != _xxh32(block):
                    raise ValueError("LZ4 block checksum mismatch")
                q += 4
            if op + (bsize if raw else bmax) > len(out):
                grown = np.empty(max(2 * len(out), op + bmax), np.uint8)
                grown[:op] = out[:op]
                out = grown
            if raw:
                out[op:op + bsize] = np.frombuffer(block, np.uint8)
                op += bsize
            else:
                src = np.frombuffer(block, np.uint8)
                npos = C.c_int64(0)
                _capi.check(lib.pm_lz4_block_decompress(
                    _capi.ptr(np.ascontiguousarray(src), C.c_uint8), bsize,
                    _capi.ptr(out, C.c_uint8), op, len(out), C.byref(npos)))
                op = npos.value
        data = out[:op].tobytes()
        if size is not None and size != len(data):
            raise ValueError("LZ4 content size mismatch")
        if content_checksum:
            if struct.unpack_from("<I", frame, q)[0] != _xxh32(data):
                raise ValueError("LZ4 content checksum mismatch")
            q += 4
        result += data
        pos = q
    return bytes(result)


def _write(path: str, arrays: dict) -> None:
    buf = io.BytesIO()
    np.savez(buf, **arrays)
    with open(path, "wb") as f:
        f.write(lz4_frame_compress(buf.getvalue()))


def _read(path: str):
    with open(path, "rb") as f:
        raw = lz4_frame_decompress(f.read())
    return np.load(io.BytesIO(raw))


def save_raw_depthmap(path: str, depth, plane, score, nghbr, nghbrs: List[str],
                      confidence: Optional[np.ndarray] = None) -> None:
    """DataSet._write_raw_depthmap (dataset.py:1080-1099)."""
    arrays = dict(depth=depth, plane=plane, score=score, nghbr=nghbr, nghbrs=nghbrs)
    if confidence is not None:
        arrays["confidence"] = confidence
    _write(path, arrays)


def load_raw_depthmap(path: str):
    """DataSet.load_raw_depthmap (dataset.py:1101-1114): (depth, plane,
    score, nghbr, nghbrs, confidence or None)."""
    o = _read(path)
    out = (o["depth"], o["plane"], o["score"], o["nghbr"], o["nghbrs"],
           o["confidence"] if "confidence" in o else None)
    o.close()
    return out


def save_clean_depthmap(path: str, depth, plane, score,
                        confidence: Optional[np.ndarray] = None) -> None:
    """DataSet._write_clean_depthmap (dataset.py:1125-1137)."""
    arrays = dict(depth=depth, plane=plane, score=score)
    if confidence is not None:
        arrays["confidence"] = confidence
    _write(path, arrays)


def load_clean_depthmap(path: str):
    """DataSet.load_clean_depthmap (dataset.py:1139-1149): (depth, plane,
    score, confidence or None)."""
    o = _read(path)
    out = (o["depth"], o["plane"], o["score"],
           o["confidence"] if "confidence" in o else None)
    o.close()
    return out
