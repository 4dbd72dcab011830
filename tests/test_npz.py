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


def test_chunks_form_one_deflate_stream():
    import zlib
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(3)
    data = memoryview(np.repeat(rng.integers(0, 7, size=300_000, dtype=np.uint8), 11))
    with ThreadPoolExecutor(4) as pool:
        stream = npz.deflate_parallel(data, pool)
    assert zlib.decompress(stream, -15) == bytes(data)


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
