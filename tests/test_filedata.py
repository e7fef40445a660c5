"""ILDG/LIME and QDP LazyDiskMapObj readers (lattice/filedata/ildg.py:76-104,
lattice/filedata/timeslice.py:102-129) against files assembled byte by byte
from the published layouts in this test (not by the package's writers), and
the writers against the readers; the headerless elemental binary layout
(lattice/preset.py:129-137) byte by byte; the gauge containers
GaugeFieldTimeSlice (records per (t, mu), lattice/preset.py:44-53) and
GaugeFieldBinary (lattice/preset.py:162-170) on spec-built files.
"""

import struct

import numpy as np
import pytest

from easydistillation_amd import EigenvectorTimeSlice, GaugeFieldBinary, GaugeFieldIldg, GaugeFieldTimeSlice
from easydistillation_amd.filedata import lime_records, read_lime_records, write_ildg, write_qdp_lazy


def lime(rtype, data, flags):
    # LIME record: magic 0x456789ab, version 1, MB/ME flags, 64-bit length, 128-byte type, data padded to 8
    return (
        bytes.fromhex("456789ab") + b"\x00\x01" + struct.pack(">H", flags) + struct.pack(">Q", len(data))
        + rtype.encode().ljust(128, b"\x00") + data + b"\x00" * ((-len(data)) % 8)
    )


def make_lime(path, arr, precision, ns=True):
    Lt, Lz, Ly, Lx = arr.shape[:4]
    xmlns = ' xmlns="http://www.lqcd.org/ildg"' if ns else ""
    xml = (f"<?xml version='1.0'?><ildgFormat{xmlns}><version>1.0</version><field>su3gauge</field>"
           f"<precision>{precision}</precision><lx>{Lx}</lx><ly>{Ly}</ly><lz>{Lz}</lz><lt>{Lt}</lt></ildgFormat>")
    dt = ">c16" if precision == 64 else ">c8"
    with open(path, "wb") as f:
        f.write(lime("ildg-format", xml.encode(), 0x8000))
        f.write(lime("odd-length-record", b"abc", 0))  # exercises the 8-byte padding
        f.write(lime("ildg-binary-data", arr.astype(dt).tobytes(), 0))
        f.write(lime("ildg-data-lfn", b"lfn://x", 0x4000))


@pytest.mark.parametrize("precision,ns", [(64, True), (32, False)])
def test_ildg_reader_on_spec_built_file(tmp_path, precision, ns):
    rng = np.random.default_rng(1)
    shape = (3, 2, 4, 5, 4, 3, 3)
    arr = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    make_lime(tmp_path / "cfg.lime", arr, precision, ns)
    recs = read_lime_records(str(tmp_path / "cfg.lime"))
    assert set(recs) == {"ildg-format", "odd-length-record", "ildg-binary-data", "ildg-data-lfn"}
    g = GaugeFieldIldg(str(tmp_path) + "/", ".lime", list(shape), ">c16" if precision == 64 else ">c8")
    data = g.load("cfg")
    assert data.latt_size == [5, 4, 2, 3]
    got = data[:]
    want = arr if precision == 64 else arr.astype(np.complex64).astype(np.complex128)
    assert got.dtype == np.complex128 and np.array_equal(got, want)
    assert np.array_equal(data[1, 0], want[1, 0])


def test_ildg_precision_mismatch_raises(tmp_path):
    arr = np.zeros((1, 2, 2, 2, 4, 3, 3), complex)
    make_lime(tmp_path / "c.lime", arr, 32)
    with pytest.raises(ValueError):
        GaugeFieldIldg(str(tmp_path) + "/", ".lime", list(arr.shape), ">c16").load("c")


def test_ildg_writer_round_trip(tmp_path):
    rng = np.random.default_rng(2)
    shape = (2, 3, 2, 4, 4, 3, 3)
    arr = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    write_ildg(str(tmp_path / "w.lime"), arr)
    assert np.array_equal(GaugeFieldIldg(str(tmp_path) + "/", ".lime", list(shape)).load("w")[:], arr)


def make_qdp(path, records, latt, dt=">c8"):
    # length-prefixed magic, version, user XML, position of the map, records, map
    def s(b):
        return struct.pack(">i", len(b)) + b

    xml = f"<MODMetaData><lattSize>{' '.join(map(str, latt))}</lattSize><decay_dir>3</decay_dir></MODMetaData>"
    head = s(b"XXXXQDPLazyDiskMapObjFileXXXX") + struct.pack(">i", 1) + s(xml.encode())
    blobs, offs, pos = [], {}, len(head) + 16
    for k, a in records.items():
        offs[k] = pos
        b = a.astype(dt).tobytes()
        blobs.append(b)
        pos += len(b)
    table = struct.pack(">I", len(offs))
    for k, o in offs.items():
        table += struct.pack(">i", 4 * len(k)) + struct.pack(">" + "i" * len(k), *k) + struct.pack(">qq", 7, o)
    with open(path, "wb") as f:
        f.write(head + struct.pack(">qq", 7, pos) + b"".join(blobs) + table)


def test_qdp_reader_on_spec_built_file(tmp_path):
    rng = np.random.default_rng(3)
    Lt, Ne, rec = 3, 4, (2, 3, 5, 3)
    recs = {(t, e): (rng.normal(size=rec) + 1j * rng.normal(size=rec)) for t in range(Lt) for e in range(Ne)}
    make_qdp(tmp_path / "e.mod", recs, [5, 3, 2, Lt])
    ev = EigenvectorTimeSlice(str(tmp_path) + "/", ".mod", [Lt, Ne, *rec], Ne)
    data = ev.load("e")
    assert data.latt_size == [5, 3, 2, Lt]
    for k, a in recs.items():
        got = data[k]
        assert got.dtype == np.complex64 and np.array_equal(got, a.astype(np.complex64))
    assert np.array_equal(data[1, 2, 0], recs[(1, 2)][0].astype(np.complex64))
    with pytest.raises(IndexError):
        data[Lt, 0]
    with pytest.raises(IndexError):
        data[0, slice(None)]


def test_qdp_writer_round_trip(tmp_path):
    rng = np.random.default_rng(4)
    recs = {(t, e): rng.normal(size=(4, 3)) + 0j for t in range(2) for e in range(3)}
    write_qdp_lazy(str(tmp_path / "w.mod"), recs, [2, 2, 1, 2])
    data = EigenvectorTimeSlice(str(tmp_path) + "/", ".mod", [2, 3, 4, 3], 3).load("w")
    for k, a in recs.items():
        assert np.array_equal(data[k], a.astype(np.complex64))


def test_elemental_binary_layout_and_reader(tmp_path):
    # headerless little-endian complex128 in [D, P, Lt, Ne, Ne] C order
    # (lattice/preset.py:129-137, lattice/filedata/binary.py:15-60)
    from easydistillation_amd import ElementalBinary, save_elemental

    rng = np.random.default_rng(7)
    Lt, D, P, Ne = 3, 2, 4, 5
    E = rng.normal(size=(Lt, D, P, Ne, Ne)) + 1j * rng.normal(size=(Lt, D, P, Ne, Ne))
    path = tmp_path / "cfg.meson"
    save_elemental(path, E, fmt="binary")
    raw = path.read_bytes()
    assert len(raw) == Lt * D * P * Ne * Ne * 16
    for t, d, p, e, f in [(0, 0, 0, 0, 0), (2, 1, 3, 4, 4), (1, 0, 2, 3, 1), (2, 1, 0, 0, 3)]:
        off = ((((d * P + p) * Lt + t) * Ne + e) * Ne + f) * 16
        assert struct.unpack("<dd", raw[off:off + 16]) == (E[t, d, p, e, f].real, E[t, d, p, e, f].imag)
    eb = ElementalBinary(f"{tmp_path}/", ".meson", [D, P, Lt, Ne, Ne], Ne)
    data = eb.load("cfg")
    assert eb.Ne == Ne and eb.load("cfg") is data
    assert np.array_equal(data[:], E.transpose(1, 2, 0, 3, 4))
    assert np.array_equal(data[1, 2], E[:, 1, 2])
    assert data[1, 2].dtype == np.dtype("<c16")
    with pytest.raises(ValueError):
        ElementalBinary(f"{tmp_path}/", ".meson", [D, P, Lt + 1, Ne, Ne], Ne).load("cfg")[0]
    with pytest.raises(ValueError):
        save_elemental(path, E, fmt="hdf5")


def test_lime_record_flags_and_errors(tmp_path):
    arr = np.zeros((1, 2, 2, 2, 4, 3, 3), complex)
    make_lime(tmp_path / "c.lime", arr, 64)
    recs = lime_records(str(tmp_path / "c.lime"))
    assert [r.type for r in recs] == ["ildg-format", "odd-length-record", "ildg-binary-data", "ildg-data-lfn"]
    assert [(r.begin, r.end) for r in recs] == [(True, False), (False, False), (False, False), (False, True)]
    assert recs[1].length == 3 and recs[2].start == recs[1].start + 8 + 144
    raw = (tmp_path / "c.lime").read_bytes()
    (tmp_path / "bad.lime").write_bytes(b"\x00" * 4 + raw[4:])
    with pytest.raises(ValueError):
        lime_records(str(tmp_path / "bad.lime"))
    # a record whose payload runs past the end of the file
    (tmp_path / "cut.lime").write_bytes(raw[: recs[2].start + 100])
    with pytest.raises(ValueError):
        lime_records(str(tmp_path / "cut.lime"))


def test_qdp_full_read_and_bad_magic(tmp_path):
    rng = np.random.default_rng(5)
    recs = {(t, e): rng.normal(size=(2, 3)) + 1j * rng.normal(size=(2, 3)) for t in range(2) for e in range(3)}
    make_qdp(tmp_path / "a.mod", recs, [1, 1, 2, 2])
    data = EigenvectorTimeSlice(str(tmp_path) + "/", ".mod", [2, 3, 2, 3], 3).load("a")
    full = data[:]
    assert full.shape == (2, 3, 2, 3) and full.dtype == np.complex64
    for k, a in recs.items():
        assert np.array_equal(full[k], a.astype(np.complex64))
    raw = (tmp_path / "a.mod").read_bytes()
    (tmp_path / "b.mod").write_bytes(raw[:4] + b"Y" + raw[5:])
    with pytest.raises(ValueError):
        EigenvectorTimeSlice(str(tmp_path) + "/", ".mod", [2, 3, 2, 3], 3).load("b")


def test_gauge_timeslice_records_per_t_mu(tmp_path):
    # one big-endian complex128 record [V3, 3, 3] per (t, mu), read back as complex64
    rng = np.random.default_rng(6)
    Lx, Ly, Lz, Lt = 3, 2, 4, 2
    V3 = Lx * Ly * Lz
    links = rng.normal(size=(Lt, 4, V3, 3, 3)) + 1j * rng.normal(size=(Lt, 4, V3, 3, 3))
    recs = {(t, mu): links[t, mu] for t in range(Lt) for mu in range(4)}
    make_qdp(tmp_path / "g.mod", recs, [Lx, Ly, Lz, Lt], dt=">c16")
    g = GaugeFieldTimeSlice(str(tmp_path) + "/", ".mod", [Lt, 4, V3, 3, 3])
    data = g.load("g")
    assert g.load("g") is data and data.latt_size == [Lx, Ly, Lz, Lt]
    got = data[1, 2]
    assert got.dtype == np.complex64 and np.array_equal(got, links[1, 2].astype(np.complex64))
    assert np.array_equal(data[0, 3, 5, 1], links[0, 3, 5, 1].astype(np.complex64))
    U = g.spatial_links("g", [Lx, Ly, Lz, Lt])
    want = links[:, :3].reshape(Lt, 3, Lz, Ly, Lx, 3, 3).transpose(1, 0, 2, 3, 4, 5, 6).astype(np.complex64)
    assert U.shape == (3, Lt, Lz, Ly, Lx, 3, 3) and np.array_equal(U, want)
    with pytest.raises(ValueError):
        g.spatial_links("g", [Lx, Ly, Lz, Lt + 1])
    with pytest.raises(IndexError):
        data[Lt, 0]


def test_gauge_binary_complex_and_interleaved(tmp_path):
    rng = np.random.default_rng(8)
    Lt, V3 = 2, 6
    links = rng.normal(size=(Lt, V3, 4, 3, 3)) + 1j * rng.normal(size=(Lt, V3, 4, 3, 3))
    # complex128 little-endian, no header
    (tmp_path / "c.dat").write_bytes(links.astype("<c16").tobytes())
    g = GaugeFieldBinary(str(tmp_path) + "/", ".dat", [Lt, V3, 4, 3, 3], "<c16")
    data = g.load("c")
    assert data[:].dtype == np.dtype("<c16") and np.array_equal(data[:], links)
    assert np.array_equal(data[1, 4, 2], links[1, 4, 2])
    # real float64 with (re, im) interleaved on a trailing axis of 2 -- the default dtype
    inter = np.stack([links.real, links.imag], axis=-1)
    raw = struct.pack(f"<{inter.size}d", *inter.ravel())
    (tmp_path / "f.dat").write_bytes(raw)
    gf = GaugeFieldBinary(str(tmp_path) + "/", None, [Lt, V3, 4, 3, 3, 2])
    assert gf.dtype == "<f8" and gf.suffix == ".dat"
    arr = gf.load("f")[:]
    assert arr.dtype == np.dtype("<f8")
    assert np.array_equal(np.ascontiguousarray(arr).view(np.complex128)[..., 0], links)
    with pytest.raises(ValueError):
        GaugeFieldBinary(str(tmp_path) + "/", ".dat", [Lt + 1, V3, 4, 3, 3], "<c16").load("c")[0]
