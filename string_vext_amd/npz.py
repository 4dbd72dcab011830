"""Compressed NPZ output written with parallel deflate.

``write_npz(handle, fields)`` writes the same archive format as
``numpy.savez_compressed`` -- a zip of ``<name>.npy`` members (NPY format
header + C-order data), each compressed with deflate -- which ``numpy.load``
reads unchanged.  The deflate strategy defaults to run-length matches only
(zlib's Z_RLE): the float32 energies of a field have near-random low
mantissa bits, so LZ77's string matching finds almost nothing in them (on a
60^3 x 64 field the default strategy makes the archive 0.8% smaller at a
quarter of the speed), while runs such as zero padding still compress.

Each member's deflate stream is produced in independent chunks on a thread
pool: chunk j is compressed with the last 32 KiB of chunk j - 1 as its preset
dictionary and ends with a sync flush (the last with the final block), so the
concatenated chunks form one valid deflate stream whose back-references never
reach past the 32 KiB window; the member's CRC-32 is combined from the
chunks' own.  zlib releases the GIL, so the chunks compress and checksum
concurrently on up to 16 threads (the 55 MB float32
orientation field of a 60^3 x 64 grid is 53 chunks).

The archive layout follows the zip application note (local file headers,
central directory, end record), with zip64 extra fields for members or
offsets beyond 4 GiB.
"""
from __future__ import annotations

import io
import os
import queue
import struct
import threading
import time
import zlib
from concurrent.futures import ThreadPoolExecutor

import numpy as np

CHUNK = 1 << 20          # bytes of input per deflate chunk
WINDOW = 1 << 15         # deflate window: a chunk's preset dictionary
LEVEL = 6                # zlib's default level, as numpy.savez_compressed uses
STRATEGY = zlib.Z_RLE    # matches of distance 1 only (runs); Z_DEFAULT_STRATEGY is numpy's
_MARK32 = 0xFFFFFFFF   # a 32-bit field whose value is in the zip64 extra field
ZIP64_LIMIT = 0xFFFFFFFF   # sizes and offsets from here on take zip64 records
ZIP64_COUNT = 0xFFFF


def _npy_bytes(array) -> memoryview:
    """The .npy encoding of `array` (header + C-order data) as one buffer."""
    a = np.asarray(array)
    if a.dtype.hasobject:
        raise ValueError("object arrays cannot be written without pickling")
    if not a.flags.c_contiguous:
        a = a.copy(order="C")
    header = io.BytesIO()
    np.lib.format.write_array_header_1_0(header, np.lib.format.header_data_from_array_1_0(a))
    data = a.reshape(-1).view(np.uint8)
    out = np.empty(header.tell() + data.size, dtype=np.uint8)
    out[:header.tell()] = np.frombuffer(header.getvalue(), dtype=np.uint8)
    out[header.tell():] = data
    return memoryview(out)


# CRC-32 of a concatenation from the CRCs of its parts.  The CRC is affine
# over GF(2) in the data, so crc(A + B) = Z^len(B) crc(A) xor crc(B), where Z
# is the 32 x 32 bit matrix that appends one zero byte to a finished CRC (the
# identity behind zlib's crc32_combine).  Z's columns are the CRCs of one zero
# byte continued from each single-bit start value, less that from zero.
_BITS = np.arange(32, dtype=np.uint64)


def _zero_byte_operator() -> np.ndarray:
    base = zlib.crc32(b"\0", 0)
    cols = np.array([zlib.crc32(b"\0", 1 << i) ^ base for i in range(32)], dtype=np.uint64)
    return ((cols[None, :] >> _BITS[:, None]) & 1).astype(np.int64)


_ZERO_BYTE = _zero_byte_operator()
_ZERO_RUNS: dict = {}


def _zeros_operator(n: int) -> np.ndarray:
    """Z^n by binary powering (cached per length)."""
    if n not in _ZERO_RUNS:
        out, m, k = np.eye(32, dtype=np.int64), _ZERO_BYTE, n
        while k:
            if k & 1:
                out = (out @ m) & 1
            m = (m @ m) & 1
            k >>= 1
        _ZERO_RUNS[n] = out
    return _ZERO_RUNS[n]


def crc32_combine(crc_a: int, crc_b: int, len_b: int) -> int:
    """CRC-32 of A + B from crc(A), crc(B) and the length of B."""
    bits = ((np.uint64(crc_a) >> _BITS) & np.uint64(1)).astype(np.int64)
    shifted = (_zeros_operator(len_b) @ bits) & 1
    return int((shifted.astype(np.uint64) << _BITS).sum()) ^ crc_b


def _deflate_chunk(buf: memoryview, lo: int, hi: int, last: bool, level: int,
                   strategy: int):
    kwargs = {"zdict": bytes(buf[max(0, lo - WINDOW):lo])} if lo > 0 else {}
    comp = zlib.compressobj(level, zlib.DEFLATED, -15, 9, strategy, **kwargs)
    body = comp.compress(buf[lo:hi])
    return body + comp.flush(zlib.Z_FINISH if last else zlib.Z_SYNC_FLUSH), zlib.crc32(buf[lo:hi])


def _deflate_chunks(buf: memoryview, pool: ThreadPoolExecutor, level: int, strategy: int):
    """The chunks of `buf`'s raw deflate stream in order, as (length of input,
    compressed bytes, CRC-32 of the input), each as soon as it and the ones
    before it are done."""
    n = len(buf)
    bounds = list(range(0, n, CHUNK)) or [0]
    jobs = [pool.submit(_deflate_chunk, buf, lo, min(lo + CHUNK, n), lo + CHUNK >= n, level,
                        strategy) for lo in bounds]
    for lo, job in zip(bounds, jobs):
        body, c = job.result()
        yield min(lo + CHUNK, n) - lo, body, c


def deflate_crc_parallel(buf: memoryview, pool: ThreadPoolExecutor, level: int = LEVEL,
                         strategy: int = STRATEGY):
    """One raw deflate stream of `buf` and its CRC-32, both formed chunk by
    chunk on `pool`."""
    parts, crc = [], 0
    for j, (length, body, c) in enumerate(_deflate_chunks(buf, pool, level, strategy)):
        parts.append(body)
        crc = crc32_combine(crc, c, length) if j else c
    return b"".join(parts), crc


def deflate_parallel(buf: memoryview, pool: ThreadPoolExecutor, level: int = LEVEL,
                     strategy: int = STRATEGY) -> bytes:
    """One raw deflate stream of `buf`, compressed chunk by chunk on `pool`."""
    return deflate_crc_parallel(buf, pool, level, strategy)[0]


class _Digester:
    """Feeds written blocks, in order, to a hashlib object on a thread of its
    own (hashlib releases the GIL on large blocks), so that hashing overlaps
    the compression of the blocks after them."""

    def __init__(self, digest):
        self.digest = digest
        self.error = None
        self.blocks = queue.Queue(maxsize=256)
        self.thread = threading.Thread(target=self._run, daemon=True)
        self.thread.start()

    def _run(self):
        # keeps draining after a failed update, so that feed never blocks for good
        while True:
            block = self.blocks.get()
            if block is None:
                return
            if self.error is None:
                try:
                    self.digest.update(block)
                except Exception as exc:   # reported by close()
                    self.error = exc

    def feed(self, block: bytes):
        self.blocks.put(block)

    def close(self):
        self.blocks.put(None)
        self.thread.join()
        if self.error is not None:
            raise RuntimeError("hashing the written archive failed") from self.error


def _dos_time(t: float):
    y, mo, d, h, mi, s = time.localtime(t)[:6]
    y = max(1980, min(2107, y))
    return (h << 11) | (mi << 5) | (s // 2), ((y - 1980) << 9) | (mo << 4) | d


class _Member:
    def __init__(self, name: bytes, crc: int, csize: int, usize: int, offset: int):
        self.name, self.crc, self.csize, self.usize, self.offset = name, crc, csize, usize, offset


def write_npz(handle, fields: dict, level: int = LEVEL, threads: int | None = None,
              strategy: int = STRATEGY, digest=None) -> None:
    """Write `fields` (name -> array) to the binary file object `handle` as a
    compressed NPZ (members ``<name>.npy``, deflate level `level` with zlib
    strategy `strategy`).

    Each member is written as its chunks finish: its local header carries
    zero sizes and CRC with general-purpose flag bit 3, and a data descriptor
    after the data holds them (the central directory has them too, as zip
    readers use).  With `digest` (a hashlib object), every byte written is
    also fed to it, in order, so that the archive's hash is ready when the
    archive is."""
    workers = threads or min(16, os.cpu_count() or 1)
    dtime, ddate = _dos_time(time.time())
    flags = 0x0808   # bit 3: sizes and CRC in a data descriptor; bit 11: UTF-8 names
    members = []
    pos = 0
    digester = _Digester(digest) if digest is not None else None

    def emit(data: bytes) -> None:
        nonlocal pos
        handle.write(data)
        if digester is not None:
            digester.feed(data)
        pos += len(data)

    try:
        with ThreadPoolExecutor(max_workers=workers) as pool:
            for key, value in fields.items():
                name = (str(key) + ".npy").encode("utf-8")
                raw = _npy_bytes(value)
                usize = len(raw)
                # zip64 sizes when the member or its deflate stream (which
                # can exceed the input by a few bytes per block) may pass 4 GiB
                big = usize + usize // 64 + 4096 >= ZIP64_LIMIT
                extra = struct.pack("<HHQQ", 0x0001, 16, 0, 0) if big else b""
                offset = pos
                emit(struct.pack("<IHHHHHIIIHH", 0x04034B50, 45 if big else 20, flags, 8, dtime,
                                 ddate, 0, _MARK32 if big else 0, _MARK32 if big else 0,
                                 len(name), len(extra)) + name + extra)
                csize = crc = 0
                for j, (length, body, c) in enumerate(_deflate_chunks(raw, pool, level, strategy)):
                    emit(body)
                    csize += len(body)
                    crc = crc32_combine(crc, c, length) if j else c
                crc &= 0xFFFFFFFF
                if not big and csize >= ZIP64_LIMIT:
                    raise RuntimeError("deflate stream outgrew its member's 32-bit size fields")
                emit(struct.pack("<IIQQ", 0x08074B50, crc, csize, usize) if big
                     else struct.pack("<IIII", 0x08074B50, crc, csize, usize))
                members.append(_Member(name, crc, csize, usize, offset))
        cd_start = pos
        for m in members:
            sizes64 = m.usize >= ZIP64_LIMIT or m.csize >= ZIP64_LIMIT
            offset64 = m.offset >= ZIP64_LIMIT
            values = ([m.usize, m.csize] if sizes64 else []) + ([m.offset] if offset64 else [])
            extra = (struct.pack(f"<HH{len(values)}Q", 0x0001, 8 * len(values), *values)
                     if values else b"")
            version = 45 if values else 20
            # made by: Unix host (3), so that the external attributes are permissions
            emit(struct.pack("<IHHHHHHIIIHHHHHII", 0x02014B50, (3 << 8) | version, version, flags,
                             8, dtime, ddate, m.crc, _MARK32 if sizes64 else m.csize,
                             _MARK32 if sizes64 else m.usize, len(m.name), len(extra), 0, 0, 0,
                             0o644 << 16, _MARK32 if offset64 else m.offset) + m.name + extra)
        cd_size = pos - cd_start
        n = len(members)
        zip64 = n >= ZIP64_COUNT or cd_start >= ZIP64_LIMIT or cd_size >= ZIP64_LIMIT
        if zip64:
            eocd64 = pos
            emit(struct.pack("<IQHHIIQQQQ", 0x06064B50, 44, (3 << 8) | 45, 45, 0, 0, n, n, cd_size,
                             cd_start))
            emit(struct.pack("<IIQI", 0x07064B50, 0, eocd64, 1))
        emit(struct.pack("<IHHHHIIH", 0x06054B50, 0, 0, 0xFFFF if zip64 else n,
                         0xFFFF if zip64 else n, _MARK32 if zip64 else cd_size,
                         _MARK32 if zip64 else cd_start, 0))
    finally:
        if digester is not None:
            digester.close()
