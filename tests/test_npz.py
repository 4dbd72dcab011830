"""The parallel-deflate NPZ writer (string_vext_amd.npz) against numpy's reader
and the zip format's own checks (CPU only)."""
import io
import zipfile

import numpy as np
import pytest

from string_vext_amd import npz


def _fields(rng):
    big = (rng.normal(size=(5, 70, 70, 70)) * 100).astype(np.float32)   # 6.9 MB: 7 chunks
    return {
        "orientation_K": big,
        "marginal_K": big[0].copy(),
        "site_K": np.broadcast_to(big[1], (2, 70, 70, 70)),                # not contiguous
        "fortran": np.asfortranarray(rng.normal(size=(5, 7))),
        "strided": rng.normal(size=(9, 9))[::2, ::3],
        "scalar": np.array(3.5),
        "empty": np.zeros((0, 3)),
        "ints": np.arange(10, dtype=np.int64),
        "cell_matrix": np.eye(3),
        "one_chunk_exactly": np.zeros(npz.CHUNK - 128, dtype=np.uint8),  # with its header: 1 MiB
    }


def _check(buf, fields):
    buf.seek(0)
    with zipfile.ZipFile(buf) as z:
        assert z.testzip() is None                 # every member's CRC-32 over its data
        assert [i.filename for i in z.infolist()] == [k + ".npy" for k in fields]
        assert all(i.compress_type == zipfile.ZIP_DEFLATED for i in z.infolist())
    buf.seek(0)
    with np.load(buf) as d:
        for k, v in fields.items():
            assert d[k].dtype == v.dtype and d[k].shape == v.shape, k
            np.testing.assert_array_equal(d[k], v, err_msg=k)


def test_round_trip_through_numpy_and_zipfile():
    fields = _fields(np.random.default_rng(1))
    buf = io.BytesIO()
    npz.write_npz(buf, fields, threads=4)
    _check(buf, fields)
    # the same archive format and compression as numpy.savez_compressed
    ref = io.BytesIO()
    np.savez_compressed(ref, **fields)
    assert abs(buf.tell() - ref.tell()) < 0.01 * ref.tell()


def test_zip64_records_when_forced(monkeypatch):
    """Members, offsets and the directory beyond the zip64 limit (the limit is
    lowered so that small arrays take the zip64 records of a > 4 GiB field)."""
    monkeypatch.setattr(npz, "ZIP64_LIMIT", 1000)
    monkeypatch.setattr(npz, "ZIP64_COUNT", 3)
    fields = _fields(np.random.default_rng(2))
    buf = io.BytesIO()
    npz.write_npz(buf, fields, threads=3)
    _check(buf, fields)


@pytest.mark.parametrize("strategy", ["rle", "default"])
def test_chunks_form_one_deflate_stream(strategy):
    """Chunks joined by sync flushes decompress as one stream; with LZ77 string
    matches (numpy's strategy) the back-references across a chunk boundary go
    through the preset dictionary."""
    import zlib
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(3)
    data = memoryview(np.repeat(rng.integers(0, 7, size=300_000, dtype=np.uint8), 11))
    st = {"rle": zlib.Z_RLE, "default": zlib.Z_DEFAULT_STRATEGY}[strategy]
    with ThreadPoolExecutor(4) as pool:
        stream = npz.deflate_parallel(data, pool, strategy=st)
    assert zlib.decompress(stream, -15) == bytes(data)


def test_strategies_round_trip_and_runs_compress():
    """Both strategies give archives numpy reads back; the default (runs only)
    still compresses a long run to a small fraction of its size."""
    import zlib
    fields = {"zeros": np.zeros(1 << 20, dtype=np.float32),
              "noise": np.random.default_rng(4).normal(size=50_000).astype(np.float32)}
    sizes = {}
    for st in (zlib.Z_RLE, zlib.Z_DEFAULT_STRATEGY, zlib.Z_HUFFMAN_ONLY):
        buf = io.BytesIO()
        npz.write_npz(buf, fields, threads=2, strategy=st)
        _check(buf, fields)
        sizes[st] = buf.tell()
    assert sizes[zlib.Z_RLE] < 0.05 * (4 << 20)
    assert sizes[zlib.Z_RLE] < sizes[zlib.Z_HUFFMAN_ONLY]


def test_corrupted_member_fails_its_crc():
    fields = {"a": np.arange(100_000, dtype=np.float64)}
    buf = io.BytesIO()
    npz.write_npz(buf, fields)
    raw = bytearray(buf.getvalue())
    at = raw.index(b"a.npy") + 5 + 4000        # inside the compressed data
    raw[at] ^= 0x01
    with zipfile.ZipFile(io.BytesIO(bytes(raw))) as z:
        with pytest.raises(Exception):
            z.read("a.npy")


def test_object_arrays_refused():
    with pytest.raises(ValueError):
        npz.write_npz(io.BytesIO(), {"o": np.array([object()])})


def test_crc32_combine_equals_whole_buffer_crc():
    """The chunk CRCs combined over GF(2) give zlib's CRC-32 of the whole
    buffer, for empty, short, chunk-sized and ragged parts."""
    import zlib
    rng = np.random.default_rng(5)
    for la, lb in ((0, 0), (0, 7), (5, 0), (1, 1), (1000, 1 << 20), ((1 << 20) + 3, 12345)):
        a = rng.integers(0, 256, la, dtype=np.uint8).tobytes()
        b = rng.integers(0, 256, lb, dtype=np.uint8).tobytes()
        assert npz.crc32_combine(zlib.crc32(a), zlib.crc32(b), lb) == zlib.crc32(a + b)
    # a multi-chunk member: the stream's CRC is that of the whole buffer
    from concurrent.futures import ThreadPoolExecutor
    data = memoryview(rng.integers(0, 256, 3 * npz.CHUNK + 777, dtype=np.uint8))
    with ThreadPoolExecutor(3) as pool:
        stream, crc = npz.deflate_crc_parallel(data, pool)
    assert crc == zlib.crc32(data) and zlib.decompress(stream, -15) == bytes(data)


def test_digest_is_the_hash_of_the_written_bytes(monkeypatch):
    """write_npz(..., digest=h) feeds h every byte written, in order: its
    digest equals the hash of the archive (plain and with zip64 records)."""
    import hashlib
    fields = _fields(np.random.default_rng(6))
    for limit in (npz.ZIP64_LIMIT, 1000):
        monkeypatch.setattr(npz, "ZIP64_LIMIT", limit)
        buf = io.BytesIO()
        h = hashlib.sha256()
        npz.write_npz(buf, fields, threads=4, digest=h)
        assert h.hexdigest() == hashlib.sha256(buf.getvalue()).hexdigest()
        _check(buf, fields)


def test_failing_digest_is_reported_not_hung():
    """A digest whose update raises ends the write with an error (the hashing
    thread keeps draining, so the writer never waits on it for good)."""
    class Broken:
        def update(self, block):
            raise ValueError("broken digest")
    fields = {"a": np.arange(300_000, dtype=np.float64)}
    with pytest.raises(RuntimeError, match="hashing"):
        npz.write_npz(io.BytesIO(), fields, threads=2, digest=Broken())
