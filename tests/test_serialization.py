"""Host side of the stream and cache format (roughpy_amd.serialization):
field-exact round trips, the documented byte layout on a hand-assembled
example, and the errors for foreign, truncated or newer data."""
import pickle
import struct

import numpy as np
import pytest

from roughpy_amd import serialization as ser
from roughpy_amd import streams as rs


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_lie_increment_roundtrip_fields(rng, dtype):
    data = rng.normal(size=(37, 3)).astype(dtype)
    idx = np.sort(rng.uniform(0, 2, size=37))
    s = rs.LieIncrementStream.from_increments(data, indices=idx, depth=4, resolution=5)
    for s2 in (pickle.loads(pickle.dumps(s)), ser.loads(ser.dumps(s))):
        assert (s2.width, s2.depth, s2.resolution, s2.support) == (3, 4, 5, s.support)
        assert s2.dtype == s.dtype
        assert np.array_equal(s2._params, s._params)
        assert np.array_equal(s2._rows, s._rows)


def test_empty_stream_roundtrip():
    s = rs.LieIncrementStream.from_increments(np.zeros((0, 2)), width=2, depth=3)
    s2 = ser.loads(ser.dumps(s))
    assert s2._rows.shape == (0, 2) and s2.support == s.support


def test_byte_layout_known_answer():
    """The layout the module documents, assembled by hand for the reference
    test's two-row stream (tests/streams/test_stream.py:11-22)."""
    data = np.array([[0.4, 0.5, 0.6], [0.7, 0.8, 0.9]])
    s = rs.LieIncrementStream.from_increments(data, width=3, depth=4)
    body = struct.pack("<iiii", 3, 4, 0, 8) + struct.pack("<dd", 0.0, 2.0) + struct.pack("<q", 2)
    body += struct.pack("<2d", 0.0, 1.0) + struct.pack("<6d", *data.ravel())
    expected = b"RPYAMD\x00\x01" + struct.pack("<IIQ", 2, 1, len(body)) + body
    assert ser.dumps(s) == expected


def test_format_errors():
    s = rs.LieIncrementStream.from_increments(np.ones((2, 2)), depth=2)
    good = ser.dumps(s)
    with pytest.raises(ValueError, match="magic"):
        ser.loads(b"NOTRPY\x00\x01" + good[8:])
    with pytest.raises(ValueError, match="version"):
        ser.loads(good[:8] + struct.pack("<I", 99) + good[12:])
    with pytest.raises(ValueError, match="length"):
        ser.loads(good[:-1])
    with pytest.raises(ValueError, match="kind"):
        ser.loads(good[:12] + struct.pack("<I", 42) + good[16:])
    with pytest.raises(ValueError):
        ser.loads(good[:10])
    with pytest.raises(TypeError):
        ser.dumps(object())
