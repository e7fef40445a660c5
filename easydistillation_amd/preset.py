"""Input/output containers with the reference's ``load(key)`` contract.

``ElementalGenerator.load(key)`` needs (lattice/generator/elemental.py:102-105,
lattice/filedata/abstract.py:12-20):

* ``gauge_field.load(key)[:]`` -> array [Lt, Lz, Ly, Lx, Nd, Nc, Nc];
* ``eigenvector.Ne`` and ``eigenvector.load(key)[t, e]`` -> [Lz, Ly, Lx, Nc].

The classes here provide that contract for numpy ``.npy`` files (read through
a memory map, so only the slices asked for are read), for arrays already in
memory, for gauge configurations as ILDG files (``GaugeFieldIldg``,
lattice/preset.py:151-160), QDP LazyDiskMapObj files with one record per
(t, mu) (``GaugeFieldTimeSlice``, lattice/preset.py:44-53) or headerless
binary arrays (``GaugeFieldBinary``, lattice/preset.py:162-170), and for QDP
LazyDiskMapObj eigenvector files (``EigenvectorTimeSlice``,
lattice/preset.py:55-64).  A gauge container whose layout is not
[Lt, Lz, Ly, Lx, Nd, Nc, Nc] provides ``spatial_links(key, latt_size)``,
which ``ElementalGenerator.load`` uses instead of ``load(key)[:]``.  ``ElementalNpy`` and
``ElementalBinary`` read elementals in the reference's file layout
[D, P, Lt, Ne, Ne] (lattice/preset.py:173-181, 129-137: .npy, and headerless
little-endian complex128); ``save_elemental`` writes either.
"""

from __future__ import annotations

from typing import List, Sequence

import numpy as np

from .constant import Nc, Nd
from .filedata import (BinaryFile, BinaryFileData, IldgFile, IldgFileData, QDPLazyDiskMapObjFile,
                       QDPLazyDiskMapObjFileData, write_binary)


class FileMetaData:
    def __init__(self, shape: List[int], dtype: str = "<c16", extra=None):
        self.shape = shape
        self.dtype = dtype
        self.extra = extra


class ArrayFileData:
    """``FileData`` over a numpy array or read-only memory map."""

    def __init__(self, array: np.ndarray, file: str | None = None):
        self._a = array
        self.file = file
        self.shape = list(array.shape)
        self.dtype = array.dtype.str
        self.time_in_sec = 0.0
        self.size_in_byte = 0

    def __getitem__(self, key):
        out = np.array(self._a[key], copy=True)
        self.size_in_byte += out.nbytes
        return out


class _NpyFile:
    def __init__(self, prefix: str, suffix: str, shape: Sequence[int] | None):
        self.prefix = prefix
        self.suffix = suffix
        self.shape = None if shape is None else list(shape)
        self._cache_name = None
        self._cache = None

    def _open(self, key: str) -> ArrayFileData:
        name = f"{self.prefix}{key}{self.suffix}"
        if name != self._cache_name:
            arr = np.load(name, mmap_mode="r", allow_pickle=False)
            if self.shape is not None and list(arr.shape) != self.shape:
                raise ValueError(f"{name}: shape {list(arr.shape)} does not match {self.shape}")
            self._cache_name, self._cache = name, ArrayFileData(arr, name)
        return self._cache


class GaugeField:
    pass


class Eigenvector:
    def __init__(self, Ne: int):
        self.Ne = Ne


class GaugeFieldNpy(_NpyFile, GaugeField):
    """Gauge configuration stored as .npy with shape [Lt, Lz, Ly, Lx, Nd, Nc, Nc]."""

    def __init__(self, prefix: str, suffix: str = ".npy", shape: Sequence[int] | None = None):
        _NpyFile.__init__(self, prefix, suffix, shape)

    def load(self, key: str) -> ArrayFileData:
        return self._open(key)


class GaugeFieldHostmem(GaugeField):
    """Gauge configuration held in memory, [Lt, Lz, Ly, Lx, Nd, Nc, Nc]."""

    def __init__(self, array: np.ndarray):
        self.data = ArrayFileData(np.asarray(array))

    def load(self, key: str | None = None) -> ArrayFileData:
        return self.data


class EigenvectorNpy(_NpyFile, Eigenvector):
    """Eigenvectors stored as .npy [Lt, totNe, Lz, Ly, Lx, Nc] (lattice/preset.py:66-74)."""

    def __init__(self, prefix: str, suffix: str = ".npy", shape: Sequence[int] | None = None, totNe: int = 70):
        _NpyFile.__init__(self, prefix, suffix, shape)
        Eigenvector.__init__(self, totNe)

    def load(self, key: str) -> ArrayFileData:
        return self._open(key)


class EigenvectorHostmem(Eigenvector):
    """Eigenvectors held in memory, [Lt, Ne, Lz, Ly, Lx, Nc] (lattice/preset.py:77-89)."""

    def __init__(self, host_ndarray: np.ndarray, totNe: int | None = None):
        arr = np.asarray(host_ndarray)
        Eigenvector.__init__(self, arr.shape[1] if totNe is None else totNe)
        self.data = ArrayFileData(arr)

    def load(self, key: str | None = None) -> ArrayFileData:
        return self.data


class GaugeFieldIldg(IldgFile, GaugeField):
    """ILDG (LIME) gauge configuration files ``{prefix}{key}{suffix}`` (lattice/preset.py:151-160).

    ``shape`` is the binary record's shape, [Lt, Lz, Ly, Lx, Nd, Nc, Nc] for
    elemental generation; the record is big-endian complex128 (``dtype``
    ">c16", ILDG precision 64) or complex64 (">c8", precision 32).
    """

    def __init__(self, prefix: str, suffix: str | None = ".lime", shape: Sequence[int] = (128, 16**3, 4, 3, 3),
                 dtype: str = ">c16"):
        IldgFile.__init__(self)
        self.prefix = prefix
        self.suffix = ".lime" if suffix is None else suffix
        self.shape = list(shape)
        self.dtype = dtype

    def load(self, key: str) -> IldgFileData:
        return self.get_file_data(f"{self.prefix}{key}{self.suffix}", self.shape, self.dtype)


class GaugeFieldTimeSlice(QDPLazyDiskMapObjFile, GaugeField):
    """QDP LazyDiskMapObj gauge files with one record per (t, mu) (lattice/preset.py:44-53).

    ``shape`` = [Lt, Nd, *per-record shape], the record being the links of
    direction mu on timeslice t, [Lz*Ly*Lx, Nc, Nc] (row a, column b, x
    fastest); records are stored big-endian complex128 and read back as
    complex64, as the reference's LazyDiskMapObj reader returns them.
    """

    def __init__(self, prefix: str, suffix: str | None, shape: Sequence[int] = (128, 4, 16**3, 3, 3)):
        QDPLazyDiskMapObjFile.__init__(self)
        self.prefix = prefix
        self.suffix = ".stout.n20.f0.12.mod" if suffix is None else suffix
        self.shape = list(shape)

    def load(self, key: str) -> QDPLazyDiskMapObjFileData:
        return self.get_file_data(f"{self.prefix}{key}{self.suffix}", self.shape, ">c16", 2)

    def spatial_links(self, key: str, latt_size: Sequence[int]) -> np.ndarray:
        """U[3, Lt, Lz, Ly, Lx, 3, 3] from the records (t, mu), mu = x, y, z."""
        Lx, Ly, Lz, Lt = (int(v) for v in latt_size)
        data = self.load(key)
        if data.extra_shape[0] != Lt or int(np.prod(data.shape)) != Lz * Ly * Lx * Nc * Nc:
            raise ValueError(f"records {data.extra_shape} x {data.shape} do not hold a {latt_size} lattice")
        U = np.empty((Nd - 1, Lt, Lz, Ly, Lx, Nc, Nc), dtype=np.complex64)
        for t in range(Lt):
            for mu in range(Nd - 1):
                U[mu, t] = data[t, mu].reshape(Lz, Ly, Lx, Nc, Nc)
        return U


class GaugeFieldBinary(BinaryFile, GaugeField):
    """Gauge configuration as a headerless array ``{prefix}{key}{suffix}``
    (lattice/preset.py:162-170) of ``shape`` (e.g. [Lt, Lz*Ly*Lx, Nd, Nc, Nc])
    in ``dtype``.  A real dtype holds interleaved (re, im) pairs, so its last
    axis is 2 (e.g. shape [Lt, V3, Nd, Nc, Nc, 2], dtype "<f8")."""

    def __init__(self, prefix: str, suffix: str | None, shape: Sequence[int] = (128, 16**3, 4, 3, 3),
                 dtype: str = "<f8"):
        BinaryFile.__init__(self)
        self.prefix = prefix
        self.suffix = ".dat" if suffix is None else suffix
        self.shape = list(shape)
        self.dtype = dtype

    def load(self, key: str) -> BinaryFileData:
        return self.get_file_data(f"{self.prefix}{key}{self.suffix}", self.shape, self.dtype)


class EigenvectorTimeSlice(QDPLazyDiskMapObjFile, Eigenvector):
    """QDP LazyDiskMapObj eigenvector files, one record per (t, e) (lattice/preset.py:55-64).

    ``shape`` = [Lt, totNe, *per-record shape], e.g. [Lt, Ne, Lz, Ly, Lx, Nc];
    records are big-endian complex64.  ``load(key)[t, e]`` reads one
    eigenvector.
    """

    def __init__(self, prefix: str, suffix: str | None, shape: Sequence[int] = (128, 70, 16**3, 3), totNe: int = 70):
        QDPLazyDiskMapObjFile.__init__(self)
        Eigenvector.__init__(self, totNe)
        self.prefix = prefix
        self.suffix = ".stout.n20.f0.12.laplace_eigs.3d.mod" if suffix is None else suffix
        self.shape = list(shape)

    def load(self, key: str) -> QDPLazyDiskMapObjFileData:
        return self.get_file_data(f"{self.prefix}{key}{self.suffix}", self.shape, ">c8", 2)


class ElementalNpy(_NpyFile):
    """Elementals in the reference's file layout [D, P, Lt, Ne, Ne] (lattice/preset.py:173-181)."""

    def __init__(self, prefix: str, suffix: str, shape: Sequence[int] | None = None, totNe: int = 70):
        _NpyFile.__init__(self, prefix, suffix, shape)
        self.Ne = totNe

    def load(self, key: str) -> ArrayFileData:
        return self._open(key)


class ElementalBinary(BinaryFile):
    """Elementals as a headerless little-endian complex128 array [D, P, Lt, Ne, Ne]
    in ``{prefix}{key}{suffix}`` (lattice/preset.py:129-137)."""

    def __init__(self, prefix: str, suffix: str | None, shape: Sequence[int] = (40, 27, 128, 70, 70),
                 totNe: int = 70):
        BinaryFile.__init__(self)
        self.prefix = prefix
        self.suffix = ".stout.n20.f0.12.nev70.meson" if suffix is None else suffix
        self.shape = list(shape)
        self.Ne = totNe

    def load(self, key: str) -> BinaryFileData:
        return self.get_file_data(f"{self.prefix}{key}{self.suffix}", self.shape, "<c16")


def save_elemental(path: str, data_t_major: np.ndarray, fmt: str = "npy") -> None:
    """Write [Lt, D, P, Ne, Ne] as the file layout [D, P, Lt, Ne, Ne] (tests/test_elemental.py:48):
    ``fmt`` "npy" (``ElementalNpy``) or "binary" (``ElementalBinary``, complex128, no header)."""
    data = np.asarray(data_t_major).transpose(1, 2, 0, 3, 4)
    if fmt == "npy":
        np.save(path, np.ascontiguousarray(data))
    elif fmt == "binary":
        write_binary(str(path), data, "<c16")
    else:
        raise ValueError(f"fmt must be 'npy' or 'binary', not {fmt!r}")
