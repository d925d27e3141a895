"""The RMCLS reader and writer against the reference's own fixture files
(Test/MatrixData1.txt, Test/MatrixData2.txt; CPU only)."""
from pathlib import Path

import numpy as np
import pytest

from mrmath_amd import persistence as ps

GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.mark.parametrize("name,first", [("MatrixData1.txt", 1.0), ("MatrixData2.txt", 2.0)])
def test_reads_the_reference_fixtures(name, first):
    m = ps.read_matrix(GOLDEN / name)
    assert (m.Width, m.Height) == (5, 10)
    assert m.sub == (0, 0, 5, 10)
    assert m.name == ""
    # the fixtures hold consecutive integers, row major
    assert np.array_equal(m.data.ravel(), first + np.arange(50.0))
    assert np.array_equal(m.SubMatrix(), m.data)


@pytest.mark.parametrize("name", ["MatrixData1.txt", "MatrixData2.txt"])
def test_writer_reproduces_the_fixture_bytes(name):
    raw = (GOLDEN / name).read_bytes()
    assert ps.dumps(ps.loads(raw)) == raw


def test_round_trip_with_name_and_window():
    m = ps.PersistedMatrix(np.arange(12.0).reshape(3, 4) / 7, "abc", (1, 0, 2, 3))
    r = ps.loads(ps.dumps(m))
    assert r.name == "abc" and r.sub == (1, 0, 2, 3)
    assert np.array_equal(r.data, m.data)
    assert np.array_equal(r.SubMatrix(), m.data[:, 1:3])


def test_malformed_streams_raise():
    raw = (GOLDEN / "MatrixData1.txt").read_bytes()
    with pytest.raises(ValueError):
        ps.loads(b"XXXXX" + raw[5:])
    with pytest.raises(ValueError):
        ps.loads(raw[:-7])  # no end section
    assert raw[49:57] == b"\x11\x00\x00\x00\x05\x06\x00H"  # the Height section
    bad = bytearray(raw)
    bad[53] = 0x06  # Height declared as a begin-list section
    with pytest.raises(ValueError, match="not part of a matrix"):
        ps.loads(bytes(bad))
    bad = bytearray(raw)
    bad[52] = 0x7F  # a section longer than the stream
    with pytest.raises(ValueError, match="bad section length"):
        ps.loads(bytes(bad))
