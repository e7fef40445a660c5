"""On-disk formats feeding and consuming elemental generation.

Three container formats, each described here from its published layout and
read through a memory map so that only the bytes asked for leave the disk:

* **LIME / ILDG** gauge configurations (read by the reference in
  lattice/filedata/ildg.py:76-104 for ``GaugeFieldIldg``,
  lattice/preset.py:151-160).  A LIME file is a sequence of records.  Each
  record is a 144-byte big-endian header --

      offset  size  field
      0       4     magic 0x456789ab
      4       2     LIME version (1)
      6       2     flags: bit 15 message begin, bit 14 message end
      8       8     payload length in bytes
      16      128   record type, NUL padded ("ildg-format", ...)

  -- followed by the payload, zero padded to a multiple of 8 bytes.  The
  ``ildg-format`` payload is XML (namespace http://www.lqcd.org/ildg) with
  ``precision`` (32 or 64) and the extents ``lx ly lz lt``; the
  ``ildg-binary-data`` payload holds the links as big-endian complex numbers
  of that precision, site-major [t][z][y][x], then direction mu, row a,
  column b.
* **QDP LazyDiskMapObj** files (lattice/filedata/timeslice.py:102-129, used by
  ``EigenvectorTimeSlice`` and ``GaugeFieldTimeSlice``,
  lattice/preset.py:44-64).  Big-endian throughout:

      string   magic "XXXXQDPLazyDiskMapObjFileXXXX"  (int32 byte count + bytes)
      int32    format version
      string   user XML (<lattSize>Lx Ly Lz Lt</lattSize>, <decay_dir>3</decay_dir>)
      pos      file offset of the record map
      ...      record payloads
      map:     uint32 record count, then per record
               key  (int32 byte count, then that many bytes of int32 indices)
               pos  file offset of the payload

  where ``pos`` is a 16-byte field whose last 8 bytes are the big-endian
  int64 offset.  Every payload is one array of the element shape.  As the
  reference reader does, payloads are returned as complex64 whatever their
  stored precision.
* **Headerless binary arrays** (lattice/filedata/binary.py:15-72, used by
  ``ElementalBinary`` and ``GaugeFieldBinary``, lattice/preset.py:129-170): the
  array in C order in the file's dtype (``"<c16"`` for elementals), returned
  in that dtype.

Indexing a file-data object (``data[key]``) returns a fresh array, the
``FileData.__getitem__`` contract of lattice/filedata/abstract.py:12-20.
``write_ildg``, ``write_qdp_lazy`` and ``write_binary`` produce the same
layouts (tests, and moving synthetic data through the same path).
"""

from __future__ import annotations

import os
import struct
import time
import xml.etree.ElementTree as ET
from typing import Dict, List, NamedTuple, Sequence, Tuple

import numpy as np

LIME_MAGIC = 0x456789AB
LIME_HEADER = struct.Struct(">IHHQ128s")  # magic, version, flags, payload length, type
QDP_MAGIC = b"XXXXQDPLazyDiskMapObjFileXXXX"
ILDG_NS = "http://www.lqcd.org/ildg"


class _Counters:
    """I/O accounting kept by every file-data object (abstract.py:12-16)."""

    def _account(self, started: float, arr: np.ndarray) -> np.ndarray:
        self.time_in_sec += time.perf_counter() - started
        self.size_in_byte += arr.nbytes
        return arr


def _window(path: str, start: int, shape: Sequence[int], dtype) -> np.memmap:
    """Read-only map of one array stored at byte ``start`` of ``path``."""
    dt = np.dtype(dtype)
    need = start + int(np.prod(shape, dtype=np.int64)) * dt.itemsize
    size = os.path.getsize(path)
    if size < need:
        raise ValueError(f"{path}: {size} bytes, the {list(shape)} {dt.str} array at byte {start} ends at {need}")
    return np.memmap(path, dtype=dt, mode="r", offset=start, shape=tuple(int(s) for s in shape))


# ----------------------------------------------------------------------------
# LIME / ILDG
# ----------------------------------------------------------------------------
class LimeRecord(NamedTuple):
    type: str
    start: int  # byte offset of the payload
    length: int  # payload bytes (without padding)
    begin: bool  # message-begin flag
    end: bool  # message-end flag


def lime_records(path: str) -> List[LimeRecord]:
    """Every record of a LIME file, in file order."""
    out: List[LimeRecord] = []
    size = os.path.getsize(path)
    with open(path, "rb") as f:
        at = 0
        while at + LIME_HEADER.size <= size:
            f.seek(at)
            magic, version, flags, length, rtype = LIME_HEADER.unpack(f.read(LIME_HEADER.size))
            if magic != LIME_MAGIC:
                raise ValueError(f"{path}: byte {at} does not start a LIME record (magic {magic:#x})")
            if version != 1:
                raise ValueError(f"{path}: LIME version {version} at byte {at} (1 expected)")
            start = at + LIME_HEADER.size
            if start + length > size:
                raise ValueError(f"{path}: record {rtype.rstrip(bytes(1))!r} runs past the end of the file")
            out.append(LimeRecord(rtype.rstrip(bytes(1)).decode("ascii"), start, length, bool(flags & 0x8000),
                                  bool(flags & 0x4000)))
            at = start + (length + 7) // 8 * 8
    return out


def read_lime_records(path: str) -> Dict[str, Tuple[int, int]]:
    """{record type: (payload offset, payload length)}; the first record of a type wins."""
    table: Dict[str, Tuple[int, int]] = {}
    for r in lime_records(path):
        table.setdefault(r.type, (r.start, r.length))
    return table


def _ildg_field(root: ET.Element, name: str) -> str:
    node = root.find(f"{{{ILDG_NS}}}{name}")
    if node is None:
        node = root.find(name)  # files written without the namespace
    if node is None or node.text is None:
        raise ValueError(f"ILDG header has no <{name}>")
    return node.text.strip()


class IldgFileData(_Counters):
    """The ``ildg-binary-data`` payload of one ILDG file as an array of
    ``shape`` (e.g. [Lt, Lz, Ly, Lx, 4, 3, 3]); ``[key]`` returns complex128
    (the reference's cast, ildg.py:70)."""

    def __init__(self, file: str, shape: Sequence[int], dtype: str):
        self.file = file
        self.shape = [int(s) for s in shape]
        self.dtype = dtype
        self.time_in_sec = 0.0
        self.size_in_byte = 0
        table = read_lime_records(file)
        for need in ("ildg-format", "ildg-binary-data"):
            if need not in table:
                raise ValueError(f"{file}: no {need} record")
        fmt_at, fmt_len = table["ildg-format"]
        with open(file, "rb") as f:
            f.seek(fmt_at)
            root = ET.fromstring(f.read(fmt_len).rstrip(bytes(1)))
        self.latt_size = [int(_ildg_field(root, k)) for k in ("lx", "ly", "lz", "lt")]
        self.precision = int(_ildg_field(root, "precision"))
        elem = np.dtype(dtype)
        if elem.kind != "c" or elem.itemsize != self.precision // 4:
            raise ValueError(f"{file}: ILDG precision {self.precision} does not match dtype {dtype}")
        self.offset, length = table["ildg-binary-data"]
        need = int(np.prod(self.shape, dtype=np.int64)) * elem.itemsize
        if need > length:
            raise ValueError(f"{file}: binary payload holds {length} bytes, shape {self.shape} needs {need}")

    def __getitem__(self, key):
        t0 = time.perf_counter()
        view = _window(self.file, self.offset, self.shape, self.dtype)
        return self._account(t0, np.array(view[key], dtype=np.complex128))


class IldgFile:
    """Opens ILDG files, keeping the most recent one parsed."""

    def __init__(self) -> None:
        self.file = None
        self.data = None

    def get_file_data(self, name: str, shape: Sequence[int], dtype: str) -> IldgFileData:
        if self.data is None or self.data.file != name:
            self.data = IldgFileData(name, shape, dtype)
            self.file = name
        return self.data


def _lime_record(rtype: str, payload: bytes, begin: bool, end: bool) -> bytes:
    flags = (0x8000 if begin else 0) | (0x4000 if end else 0)
    head = LIME_HEADER.pack(LIME_MAGIC, 1, flags, len(payload), rtype.encode("ascii"))
    return head + payload + bytes(-len(payload) % 8)


def write_ildg(path: str, U: np.ndarray, precision: int = 64) -> None:
    """Write links U[Lt, Lz, Ly, Lx, 4, 3, 3] as an ILDG LIME file."""
    U = np.asarray(U)
    if U.ndim != 7 or U.shape[4:] != (4, 3, 3):
        raise ValueError("U must be [Lt, Lz, Ly, Lx, 4, 3, 3]")
    if precision not in (32, 64):
        raise ValueError("precision is 32 or 64")
    Lt, Lz, Ly, Lx = U.shape[:4]
    fields = dict(version="1.0", field="su3gauge", precision=precision, lx=Lx, ly=Ly, lz=Lz, lt=Lt)
    xml = ('<?xml version="1.0" encoding="UTF-8"?>' f'<ildgFormat xmlns="{ILDG_NS}">'
           + "".join(f"<{k}>{v}</{k}>" for k, v in fields.items()) + "</ildgFormat>")
    payload = U.astype(">c16" if precision == 64 else ">c8").tobytes()
    with open(path, "wb") as f:
        f.write(_lime_record("ildg-format", xml.encode("utf-8"), True, False))
        f.write(_lime_record("ildg-binary-data", payload, False, False))
        f.write(_lime_record("ildg-data-lfn", b"synthetic", False, True))


# ----------------------------------------------------------------------------
# QDP LazyDiskMapObj
# ----------------------------------------------------------------------------
class _BigEndian:
    """Sequential big-endian reads of the LazyDiskMapObj header and map."""

    def __init__(self, f):
        self.f = f

    def take(self, n: int) -> bytes:
        b = self.f.read(n)
        if len(b) != n:
            raise ValueError("LazyDiskMapObj file ends inside its header or record map")
        return b

    def int32(self) -> int:
        return struct.unpack(">i", self.take(4))[0]

    def uint32(self) -> int:
        return struct.unpack(">I", self.take(4))[0]

    def blob(self) -> bytes:
        return self.take(self.int32())

    def offset(self) -> int:
        return struct.unpack(">q", self.take(16)[8:])[0]

    def key(self) -> Tuple[int, ...]:
        raw = self.blob()
        return tuple(struct.unpack(f">{len(raw) // 4}i", raw))


class QDPLazyDiskMapObjFileData(_Counters):
    """Records of one LazyDiskMapObj file, addressed by their int keys.

    ``shape`` is the full logical shape: the first ``extra`` axes are the
    record key (e.g. (t, e) for eigenvectors, (t, mu) for links), the rest the
    shape of one record.  ``data[k0, .., k_extra-1, *rest]`` reads record
    (k0, ..) and applies ``rest`` to it; ``data[:]`` (or ``data[()]``) reads
    every record into one array of the full shape.  Values come back as
    complex64 (timeslice.py:96).  A key with no record raises IndexError.
    """

    def __init__(self, file: str, shape: Sequence[int], dtype: str, extra: int, offsets: Dict[Tuple[int, ...], int],
                 xml_root: ET.Element, version: int = 0):
        self.file = file
        self.extra = int(extra)
        self.extra_shape = [int(s) for s in shape[: self.extra]]
        self.shape = [int(s) for s in shape[self.extra :]]
        self.dtype = dtype
        self.offsets = offsets
        self.version = version
        self.time_in_sec = 0.0
        self.size_in_byte = 0
        latt = xml_root.find("lattSize")
        decay = xml_root.find("decay_dir")
        if latt is None or decay is None:
            raise ValueError(f"{file}: user XML lacks lattSize / decay_dir")
        self.latt_size = [int(v) for v in latt.text.split()]
        if int(decay.text) != 3:
            raise ValueError(f"{file}: decay_dir {decay.text.strip()} (only 3, the time direction, is supported)")

    def record(self, rec: Tuple[int, ...]) -> np.memmap:
        """Read-only map of record ``rec`` in its stored dtype."""
        if rec not in self.offsets:
            raise IndexError(f"{self.file}: no record with key {rec}")
        return _window(self.file, self.offsets[rec], self.shape, self.dtype)

    def __getitem__(self, key):
        t0 = time.perf_counter()
        if key == slice(None) or key == () or key is Ellipsis:
            out = np.empty(self.extra_shape + self.shape, dtype=np.complex64)
            for rec in np.ndindex(*self.extra_shape):
                out[rec] = self.record(tuple(int(i) for i in rec))
            return self._account(t0, out)
        if not isinstance(key, tuple):
            key = (key,)
        head = key[: self.extra]
        if len(head) < self.extra or not all(isinstance(k, (int, np.integer)) for k in head):
            raise IndexError(f"{self.file}: the first {self.extra} indices must be integers (record key), got {key}")
        rec = tuple(int(k) for k in head)
        return self._account(t0, np.array(self.record(rec)[key[self.extra :]], dtype=np.complex64))


class QDPLazyDiskMapObjFile:
    """Opens LazyDiskMapObj files, keeping the most recent one parsed."""

    def __init__(self) -> None:
        self.file = None
        self.version = 0
        self.data = None

    def get_file_data(self, name: str, shape: Sequence[int], dtype: str, extra: int) -> QDPLazyDiskMapObjFileData:
        if self.data is None or self.data.file != name:
            with open(name, "rb") as f:
                r = _BigEndian(f)
                if r.blob() != QDP_MAGIC:
                    raise ValueError(f"{name}: not a QDP LazyDiskMapObj file")
                version = r.int32()
                root = ET.fromstring(r.blob())
                f.seek(r.offset())
                offsets = {}
                for _ in range(r.uint32()):
                    k = r.key()
                    offsets[k] = r.offset()
            self.data = QDPLazyDiskMapObjFileData(name, shape, dtype, extra, offsets, root, version)
            self.file = name
            self.version = version
        return self.data


def write_qdp_lazy(path: str, records: Dict[Tuple[int, ...], np.ndarray], latt_size: Sequence[int], dtype: str = ">c8",
                   version: int = 1) -> None:
    """Write ``{key tuple: array}`` as a QDP LazyDiskMapObj file (decay_dir 3)."""

    def blob(b: bytes) -> bytes:
        return struct.pack(">i", len(b)) + b

    def pos(p: int) -> bytes:
        return bytes(8) + struct.pack(">q", p)

    xml = (f"<MODMetaData><lattSize>{' '.join(str(int(v)) for v in latt_size)}</lattSize>"
           "<decay_dir>3</decay_dir></MODMetaData>").encode("utf-8")
    head = blob(QDP_MAGIC) + struct.pack(">i", version) + blob(xml)
    first = len(head) + 16
    payloads, table = [], [struct.pack(">I", len(records))]
    at = first
    for k, arr in records.items():
        data = np.asarray(arr).astype(dtype).tobytes()
        ints = tuple(int(v) for v in k)
        table.append(blob(struct.pack(f">{len(ints)}i", *ints)) + pos(at))
        payloads.append(data)
        at += len(data)
    with open(path, "wb") as f:
        f.write(head + pos(at))
        for p in payloads:
            f.write(p)
        f.write(b"".join(table))


# ----------------------------------------------------------------------------
# Headerless binary arrays
# ----------------------------------------------------------------------------
class BinaryFileData(_Counters):
    """One headerless array file of ``shape`` in C order and ``dtype``;
    ``[key]`` returns that part of the array in the file's dtype.  A file too
    short for the shape raises ValueError."""

    def __init__(self, file: str, shape: Sequence[int], dtype: str):
        self.file = file
        self.shape = [int(s) for s in shape]
        self.dtype = dtype
        self.time_in_sec = 0.0
        self.size_in_byte = 0

    def __getitem__(self, key):
        t0 = time.perf_counter()
        view = _window(self.file, 0, self.shape, self.dtype)
        return self._account(t0, np.array(view[key], dtype=view.dtype))


class BinaryFile:
    """Opens headerless array files, keeping the most recent one."""

    def __init__(self):
        self.file = None
        self.data = None

    def get_file_data(self, name: str, shape: Sequence[int], dtype: str) -> BinaryFileData:
        if self.data is None or self.data.file != name:
            self.data = BinaryFileData(name, shape, dtype)
            self.file = name
        return self.data


def write_binary(path: str, array: np.ndarray, dtype: str = "<c16") -> None:
    """Write ``array`` in C order as ``dtype``, no header."""
    with open(path, "wb") as f:
        f.write(np.ascontiguousarray(np.asarray(array).astype(dtype)).tobytes())
