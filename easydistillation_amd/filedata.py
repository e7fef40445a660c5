"""On-disk formats feeding elemental generation.

* ILDG gauge configurations in LIME containers (the reader of
  lattice/filedata/ildg.py:76-104, used through ``GaugeFieldIldg``,
  lattice/preset.py:151-160).  A LIME record is a 144-byte header -- magic
  0x456789ab, version 1, MB/ME flags, big-endian 64-bit data length, 128-byte
  type string -- followed by the data padded to 8 bytes.  The
  ``ildg-format`` record holds XML with ``precision``, ``lx .. lt``; the
  ``ildg-binary-data`` record holds the links big-endian in
  [t][z][y][x][mu][a][b] order (complex64 or complex128 by precision).
* QDP LazyDiskMapObj files (lattice/filedata/timeslice.py:102-129, used
  through ``EigenvectorTimeSlice``, lattice/preset.py:55-64): a length-prefixed
  magic string ``XXXXQDPLazyDiskMapObjFileXXXX``, a version, user XML
  (``lattSize``, ``decay_dir``), the position of a record map, and the map
  (number of records, then per record an int tuple key and a position).
  Each record holds one array of the element shape (big-endian).

* Headerless binary arrays (lattice/filedata/binary.py:15-60, used through
  ``ElementalBinary``, lattice/preset.py:129-137): the array in C order in
  the file's dtype (``"<c16"`` for elementals), read back in that dtype.

Readers memory-map the file and return native complex arrays for the
requested key, as the reference's ``FileData.__getitem__``.  ``write_ildg``
and ``write_qdp_lazy`` produce the same layouts (for tests and for moving
synthetic data through the same path).
"""

from __future__ import annotations

import mmap
import re
import struct
import xml.etree.ElementTree as ET
from typing import Dict, Sequence, Tuple

import numpy as np

LIME_MAGIC = 0x456789AB
QDP_MAGIC = "XXXXQDPLazyDiskMapObjFileXXXX"


def _prod(a) -> int:
    p = 1
    for v in a:
        p *= int(v)
    return p


def _elem_bytes(dtype: str) -> int:
    m = re.match(r"^[<>=]?[iufc](?P<bytes>\d+)$", dtype)
    if not m:
        raise ValueError(f"unsupported dtype {dtype!r}")
    return int(m.group("bytes"))


def _mapped(path: str, offset: int, shape: Sequence[int], dtype: str, key, out_dtype) -> np.ndarray:
    start = offset - offset % mmap.ALLOCATIONGRANULARITY
    with open(path, "rb") as f:
        with mmap.mmap(f.fileno(), offset - start + _prod(shape) * _elem_bytes(dtype), access=mmap.ACCESS_READ,
                       offset=start) as mm:
            arr = np.ndarray(tuple(shape), dtype=np.dtype(dtype), buffer=mm, offset=offset - start)
            return np.ascontiguousarray(arr[key]).astype(out_dtype)


# ----------------------------------------------------------------------------
# LIME / ILDG
# ----------------------------------------------------------------------------
def read_lime_records(path: str) -> Dict[str, Tuple[int, int]]:
    """{record type: (data offset, padded data length)} of a LIME file."""
    recs: Dict[str, Tuple[int, int]] = {}
    with open(path, "rb") as f:
        while True:
            head = f.read(8)
            if len(head) < 8 or head == b"\x0a":
                break
            magic, version = struct.unpack(">IH", head[:6])
            if magic != LIME_MAGIC or version != 1:
                raise ValueError(f"{path}: not a LIME record at offset {f.tell() - 8}")
            length = struct.unpack(">Q", f.read(8))[0]
            rtype = f.read(128).strip(b"\x00").decode("utf-8")
            padded = (length + 7) // 8 * 8
            recs[rtype] = (f.tell(), padded)
            f.seek(padded, 1)
    return recs


class IldgFileData:
    """Links of one ILDG file; ``[key]`` returns complex128 (ildg.py:20-73)."""

    def __init__(self, file: str, shape: Sequence[int], dtype: str, offset: Tuple[int, int], xml_tree: ET.ElementTree):
        self.file = file
        self.shape = list(shape)
        self.dtype = dtype
        self.offset = offset[0]
        root = xml_tree.getroot()
        m = re.match(r"\{.*\}", root.tag)
        tag = m.group(0) if m else ""
        self.latt_size = [int(root.find(f"{tag}{k}").text) for k in ("lx", "ly", "lz", "lt")]
        self.bytes = _elem_bytes(dtype)
        precision = int(root.find(f"{tag}precision").text)
        if self.bytes != precision // 8 * 2:
            raise ValueError(f"{file}: ILDG precision {precision} does not match dtype {dtype}")
        if _prod(self.shape) * self.bytes > offset[1]:
            raise ValueError(f"{file}: binary record holds {offset[1]} bytes, shape {self.shape} needs more")
        self.time_in_sec = 0.0
        self.size_in_byte = 0

    def __getitem__(self, key):
        ret = _mapped(self.file, self.offset, self.shape, self.dtype, key, np.complex128)
        self.size_in_byte += ret.nbytes
        return ret


class IldgFile:
    def __init__(self) -> None:
        self.file = None
        self.data = None

    def get_file_data(self, name: str, shape: Sequence[int], dtype: str) -> IldgFileData:
        if self.file != name:
            recs = read_lime_records(name)
            if "ildg-format" not in recs or "ildg-binary-data" not in recs:
                raise ValueError(f"{name}: no ildg-format / ildg-binary-data records")
            off, length = recs["ildg-format"]
            with open(name, "rb") as f:
                f.seek(off)
                text = f.read(length).strip(b"\x00").decode("utf-8")
            self.data = IldgFileData(name, shape, dtype, recs["ildg-binary-data"], ET.ElementTree(ET.fromstring(text)))
            self.file = name
        return self.data


def _lime_record(rtype: str, data: bytes, mb: bool, me: bool) -> bytes:
    flags = (0x8000 if mb else 0) | (0x4000 if me else 0)
    head = struct.pack(">IHHQ", LIME_MAGIC, 1, flags, len(data)) + rtype.encode("utf-8").ljust(128, b"\x00")
    return head + data + b"\x00" * ((-len(data)) % 8)


def write_ildg(path: str, U: np.ndarray, precision: int = 64) -> None:
    """Write links U[Lt, Lz, Ly, Lx, 4, 3, 3] as an ILDG LIME file."""
    U = np.asarray(U)
    Lt, Lz, Ly, Lx = U.shape[:4]
    if U.shape[4:] != (4, 3, 3):
        raise ValueError("U must be [Lt, Lz, Ly, Lx, 4, 3, 3]")
    dt = ">c16" if precision == 64 else ">c8"
    xml = (
        '<?xml version="1.0" encoding="UTF-8"?>'
        '<ildgFormat xmlns="http://www.lqcd.org/ildg">'
        f"<version>1.0</version><field>su3gauge</field><precision>{precision}</precision>"
        f"<lx>{Lx}</lx><ly>{Ly}</ly><lz>{Lz}</lz><lt>{Lt}</lt></ildgFormat>"
    ).encode("utf-8")
    with open(path, "wb") as f:
        f.write(_lime_record("ildg-format", xml, True, False))
        f.write(_lime_record("ildg-binary-data", U.astype(dt).tobytes(), False, False))
        f.write(_lime_record("ildg-data-lfn", b"synthetic", False, True))


# ----------------------------------------------------------------------------
# QDP LazyDiskMapObj
# ----------------------------------------------------------------------------
def _read_str(f) -> str:
    n = struct.unpack(">i", f.read(4))[0]
    return f.read(n).decode("utf-8")


def _read_tuple(f) -> Tuple[int, ...]:
    n = struct.unpack(">i", f.read(4))[0] // 4
    return struct.unpack(">" + "i" * n, f.read(4 * n))


def _read_pos(f) -> int:
    return struct.unpack(">qq", f.read(16))[1]


class QDPLazyDiskMapObjFileData:
    """Records keyed by the leading ``extra`` indices (timeslice.py:36-100).

    ``[key]`` with ``key[:extra]`` a stored record key returns the record
    (or a part of it, by the remaining indices) as complex64; other keys raise
    IndexError.
    """

    def __init__(self, file: str, shape: Sequence[int], dtype: str, extra: int, offsets, xml_tree: ET.ElementTree):
        self.file = file
        self.extra = int(extra)
        self.extra_shape = list(shape[: self.extra])
        self.shape = list(shape[self.extra :])
        self.dtype = dtype
        self.offsets = offsets
        root = xml_tree.getroot()
        self.latt_size = [int(v) for v in root.find("lattSize").text.split()]
        if int(root.find("decay_dir").text) != 3:
            raise ValueError(f"{file}: decay_dir must be 3")
        self.time_in_sec = 0.0
        self.size_in_byte = 0

    def __getitem__(self, key):
        if not isinstance(key, tuple):
            key = (key,)
        rec = tuple(int(k) for k in key[: self.extra]) if all(isinstance(k, (int, np.integer)) for k in key[: self.extra]) else None
        if rec is None or rec not in self.offsets:
            raise IndexError(f"index {key} is out of bounds for axes")
        ret = _mapped(self.file, self.offsets[rec], self.shape, self.dtype, key[self.extra :], np.complex64)
        self.size_in_byte += ret.nbytes
        return ret


class QDPLazyDiskMapObjFile:
    def __init__(self) -> None:
        self.file = None
        self.version = 0
        self.data = None

    def get_file_data(self, name: str, shape: Sequence[int], dtype: str, extra: int) -> QDPLazyDiskMapObjFileData:
        if self.file != name:
            with open(name, "rb") as f:
                if _read_str(f) != QDP_MAGIC:
                    raise ValueError(f"{name}: not a QDP LazyDiskMapObj file")
                self.version = struct.unpack(">i", f.read(4))[0]
                xml_tree = ET.ElementTree(ET.fromstring(_read_str(f)))
                f.seek(_read_pos(f))
                n = struct.unpack(">I", f.read(4))[0]
                offsets = {}
                for _ in range(n):
                    k = _read_tuple(f)
                    offsets[k] = _read_pos(f)
            self.data = QDPLazyDiskMapObjFileData(name, shape, dtype, extra, offsets, xml_tree)
            self.file = name
        return self.data


def write_qdp_lazy(path: str, records: Dict[Tuple[int, ...], np.ndarray], latt_size: Sequence[int], dtype: str = ">c8",
                   version: int = 1) -> None:
    """Write ``{key tuple: array}`` as a QDP LazyDiskMapObj file (decay_dir 3)."""

    def s(b: bytes) -> bytes:
        return struct.pack(">i", len(b)) + b

    xml = (f"<MODMetaData><lattSize>{' '.join(str(int(v)) for v in latt_size)}</lattSize>"
           "<decay_dir>3</decay_dir></MODMetaData>").encode("utf-8")
    head = s(QDP_MAGIC.encode("utf-8")) + struct.pack(">i", version) + s(xml)
    pos = len(head) + 16
    body = b""
    where = {}
    for k, arr in records.items():
        where[tuple(int(v) for v in k)] = pos + len(body)
        body += np.asarray(arr).astype(dtype).tobytes()
    table = struct.pack(">I", len(where))
    for k, p in where.items():
        table += struct.pack(">i", 4 * len(k)) + struct.pack(">" + "i" * len(k), *k) + struct.pack(">qq", 0, p)
    with open(path, "wb") as f:
        f.write(head)
        f.write(struct.pack(">qq", 0, pos + len(body)))
        f.write(body)
        f.write(table)


# ----------------------------------------------------------------------------
# Headerless binary arrays
# ----------------------------------------------------------------------------
class BinaryFileData:
    """One headerless array file: ``shape`` in C order, elements of ``dtype``.

    ``data[key]`` maps the file and returns a copy of ``array[key]`` in the
    file's dtype (lattice/filedata/binary.py:29-60); a file shorter than the
    shape raises ``ValueError``.
    """

    def __init__(self, file: str, shape: Sequence[int], dtype: str):
        self.file = file
        self.shape = [int(v) for v in shape]
        self.dtype = dtype
        self.bytes = _elem_bytes(dtype)
        self.stride = [_prod(self.shape[i:]) for i in range(1, len(self.shape))] + [1]
        self.time_in_sec = 0.0
        self.size_in_byte = 0

    def get_count(self, key: Tuple[int, ...]) -> int:
        return self.stride[len(key) - 1]

    def get_offset(self, key: Tuple[int, ...]) -> int:
        return sum(int(a) * b for a, b in zip(key, self.stride)) * self.bytes

    def __getitem__(self, key):
        import os
        import time

        t0 = time.time()
        need = _prod(self.shape) * self.bytes
        have = os.path.getsize(self.file)
        if have < need:
            raise ValueError(f"{self.file}: {have} bytes, the shape {self.shape} of {self.dtype} needs {need}")
        out = _mapped(self.file, 0, self.shape, self.dtype, key, np.dtype(self.dtype))
        self.time_in_sec += time.time() - t0
        self.size_in_byte += out.nbytes
        return out


class BinaryFile:
    """Keeps the last opened file (lattice/filedata/binary.py:63-72)."""

    def __init__(self):
        self.file = None
        self.data = None

    def get_file_data(self, name: str, shape: Sequence[int], dtype: str) -> BinaryFileData:
        if self.file != name:
            self.data = BinaryFileData(name, shape, dtype)
            self.file = name
        return self.data


def write_binary(path: str, array: np.ndarray, dtype: str = "<c16") -> None:
    """Write ``array`` in C order as ``dtype``, no header."""
    with open(path, "wb") as f:
        f.write(np.ascontiguousarray(np.asarray(array).astype(dtype)).tobytes())
