"""Raw state-file container (no GPU): the reference's layout of an 8-byte header holding
ZSTD_compressBound(payload size) followed by one zstd frame (RawBinary.cpp:42-72), checked against
zstd's published bound formula and against an independent zstd implementation (pyarrow)."""
import struct

import numpy as np
import pytest

from subrosadg_amd import _lib as L

pa = pytest.importorskip("pyarrow")


def _bound(n: int) -> int:
    """ZSTD_COMPRESSBOUND of the zstd 1.4 public header."""
    return n + (n >> 8) + (((128 << 10) - n) >> 11 if n < (128 << 10) else 0)


@pytest.mark.parametrize("n", [0, 8, 4096, 200_000])
def test_roundtrip_and_header(tmp_path, n):
    rng = np.random.default_rng(n)
    payload = rng.standard_normal(n // 8).tobytes() if n else b""
    path = str(tmp_path / "x.zst")
    L.raw_file_write(path, payload)
    raw = open(path, "rb").read()
    (header,) = struct.unpack("<Q", raw[:8])
    assert header == _bound(len(payload))
    assert raw[8:12] == bytes.fromhex("28b52ffd")  # zstd frame magic
    assert L.raw_file_read(path) == payload
    # an independent decoder reads the frame
    assert pa.decompress(raw[8:], decompressed_size=len(payload), codec="zstd").to_pybytes() == payload


def test_reads_a_frame_from_an_independent_encoder(tmp_path):
    payload = np.linspace(-1.0, 1.0, 12345).tobytes()
    frame = pa.Codec("zstd", compression_level=1).compress(payload, asbytes=True)
    path = tmp_path / "y.zst"
    path.write_bytes(struct.pack("<Q", _bound(len(payload))) + frame)
    assert L.raw_file_read(str(path)) == payload


def test_missing_file_is_an_error(tmp_path):
    with pytest.raises(L.SDGError):
        L.raw_file_read(str(tmp_path / "absent.zst"))
