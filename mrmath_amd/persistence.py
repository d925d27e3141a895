"""Reader and writer of mrMath's binary persistence format ("RMCLS") for
TDoubleMatrix, the format of the reference's test fixtures
(Test/MatrixData1.txt, Test/MatrixData2.txt, loaded by
TestTDoubleMatrix.SetUp, Test/TestMatrixClass.pas:146-150).

Format (src/BinaryReaderWriter.pas:73-88, 135-230): the 5 bytes "RMCLS", a
32-bit version (1), then sections until an end-of-stream section.  A section
is a packed header {int32 Len; uint8 type; uint16 NameLen}, NameLen bytes of
UTF-8 name, and Len - 7 - NameLen bytes of data (all little endian).  A
matrix is written (src/Matrix.pas:873-885, src/DblMatrix.pas:3354-3367) as
the object descriptor "obj" = "MTX", the string property Name, the integer
properties Width and Height, the double array "data" (the full matrix, row
major), the integer properties SubWidth, SubHeight, OffsetX, OffsetY; the end-of-stream section
closes the top-level object (object-end sections close nested objects only).

Only the sections a matrix uses are decoded; any other section type raises
ValueError.  Nothing in a file is executed.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

HEADER = b"RMCLS"
VERSION = 1

# TBinarySectionType (src/BinaryReaderWriter.pas:77-79), one byte each
(bsBinary, bsString, bsDouble, bsDoubleArr, bsIntArr, bsInteger, bsBeginList, bsListIntArr, bsListDoubleArr,
 bsEndList, bsObjectDescription, bsObjectBegin, bsObjectEnd, bsEndaStream, bsComplex, bsComplexArr) = range(16)

_SECT = struct.Struct("<iBH")  # TBinarySectionRec, packed: Len, SectType, NameLen


@dataclass
class PersistedMatrix:
    """A TDoubleMatrix as stored: the full matrix plus its sub-matrix window."""
    data: np.ndarray                  # Height x Width, float64
    name: str = ""
    sub: tuple = field(default=(0, 0, 0, 0))   # (OffsetX, OffsetY, SubWidth, SubHeight)

    @property
    def Width(self) -> int:
        return self.data.shape[1]

    @property
    def Height(self) -> int:
        return self.data.shape[0]

    def SubMatrix(self) -> np.ndarray:
        """The window the matrix object operates on (its SubWidth x SubHeight
        block at OffsetX, OffsetY)."""
        ox, oy, sw, sh = self.sub
        return self.data[oy:oy + sh, ox:ox + sw]


def loads(buf: bytes) -> PersistedMatrix:
    """Decodes one TDoubleMatrix (TBinaryReaderWriter.LoadFromStream)."""
    if buf[:5] != HEADER:
        raise ValueError("Wrong file format")
    pos = 5
    (version,) = struct.unpack_from("<I", buf, pos)
    pos += 4
    if version != VERSION:
        raise ValueError(f"unsupported RMCLS version {version}")
    cls = None
    ints: dict = {}
    name = ""
    values = None
    while True:
        if pos + _SECT.size > len(buf):
            raise ValueError("truncated stream: no end section")
        ln, typ, nlen = _SECT.unpack_from(buf, pos)
        head = _SECT.size + nlen
        if ln < head or pos + ln > len(buf):
            raise ValueError(f"bad section length {ln} at byte {pos}")
        sname = buf[pos + _SECT.size:pos + head].decode("utf-8")
        data = buf[pos + head:pos + ln]
        pos += ln
        if typ == bsEndaStream:
            break
        if typ == bsObjectDescription:
            if cls is not None:
                raise ValueError("nested objects are not part of a matrix")
            cls = data.decode("utf-8")
        elif typ == bsString:
            if sname.lower() == "name":
                name = data.decode("utf-8")
        elif typ == bsInteger:
            if len(data) != 4:
                raise ValueError("Error wrong initialized section")
            ints[sname.lower()] = struct.unpack("<i", data)[0]
        elif typ == bsDoubleArr:
            if len(data) % 8:
                raise ValueError("Error wrong initialized section")
            if sname.lower() == "data":
                values = np.frombuffer(data, dtype="<f8").astype(np.float64)
        else:
            raise ValueError(f"section type {typ} is not part of a matrix")
    if cls != "MTX":
        raise ValueError(f"not a TDoubleMatrix (class {cls!r})")
    w, h = ints.get("width", 0), ints.get("height", 0)
    if values is None:
        values = np.zeros(0)
    if values.size != w * h:
        raise ValueError("The loaded array does not fit the expected size.")
    m = PersistedMatrix(values.reshape(h, w).copy() if w * h else np.zeros((h, w)), name)
    m.sub = (ints.get("offsetx", 0), ints.get("offsety", 0), ints.get("subwidth", w), ints.get("subheight", h))
    return m


def read_matrix(path) -> PersistedMatrix:
    """ReadObjFromFile for a file holding a TDoubleMatrix."""
    return loads(Path(path).read_bytes())


def _section(typ: int, name: str, data: bytes = b"") -> bytes:
    nb = name.encode("utf-8")
    return _SECT.pack(_SECT.size + len(nb) + len(data), typ, len(nb)) + nb + data


def dumps(m: PersistedMatrix) -> bytes:
    """Encodes a TDoubleMatrix as the reference's writer does."""
    h, w = m.data.shape
    ox, oy, sw, sh = m.sub
    out = [HEADER, struct.pack("<I", VERSION),
           _section(bsObjectDescription, "obj", b"MTX"),
           _section(bsString, "Name", m.name.encode("utf-8")),
           _section(bsInteger, "Width", struct.pack("<i", w)),
           _section(bsInteger, "Height", struct.pack("<i", h)),
           _section(bsDoubleArr, "data", np.ascontiguousarray(m.data, dtype="<f8").tobytes()),
           _section(bsInteger, "SubWidth", struct.pack("<i", sw)),
           _section(bsInteger, "SubHeight", struct.pack("<i", sh)),
           _section(bsInteger, "OffsetX", struct.pack("<i", ox)),
           _section(bsInteger, "OffsetY", struct.pack("<i", oy)),
           _section(bsEndaStream, "")]
    return b"".join(out)


def write_matrix(path, m: PersistedMatrix) -> None:
    Path(path).write_bytes(dumps(m))
