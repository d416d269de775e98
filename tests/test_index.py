"""The v4 index reader (saber_hip_index_open, host code): residues and
subject offsets read back from a file laid out as IndexBuilder writes it
(tests/index_v4.py), the CRC-32 of crc32.rs (its known vector, and zlib on
random data), and the rejections of SeedIndex::open that the reference's own
test checks (index.rs:2160-2200: truncation, a format-3 file with rebuild
advice, a flipped metadata byte) plus the other checks of open()."""
import struct

import numpy as np
import pytest

import saber_amd as S
from saber_amd import _ffi
from tests import index_v4
from tests.gen import Lcg, rand_seq


def subjects(seed=1, n=40):
    r = Lcg(seed)
    return [rand_seq(r, 1 + r.next(300)) for _ in range(n)] + [np.zeros(0, np.uint8)]


def write(tmp_path, data, name="db.sdx"):
    p = tmp_path / name
    p.write_bytes(data)
    return p


def open_error(path, verify=False):
    with pytest.raises(_ffi.HipError) as e:
        S.IndexFile(path, verify=verify)
    return str(e.value)


def test_crc32_known_vector_and_zlib():
    # crc32.rs:55-61: crc32("123456789") == 0xCBF43926
    assert S.crc32(b"123456789") == 0xCBF43926
    import zlib
    rng = np.random.default_rng(5)
    for n in (0, 1, 7, 4096, 100003):
        b = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        assert S.crc32(b) == zlib.crc32(b)


def test_round_trip(tmp_path):
    subs = subjects()
    data, meta = index_v4.build(subs)
    p = write(tmp_path, data)
    with S.IndexFile(p, verify=True) as idx:
        assert idx.num_subjects == len(subs) and idx.protein and not idx.translated
        assert idx.total_residues == sum(len(s) for s in subs)
        for i, s in enumerate(subs):
            a, b = int(idx.offsets[i]), int(idx.offsets[i + 1])
            assert np.array_equal(idx.residues[a:b], s)


def test_empty_index(tmp_path):
    data, _ = index_v4.build([])
    with S.IndexFile(write(tmp_path, data), verify=True) as idx:
        assert idx.num_subjects == 0 and idx.total_residues == 0 and list(idx.offsets) == [0]


def test_rejections_of_the_reference_test(tmp_path):
    data, _ = index_v4.build(subjects())
    # truncation is detected (index.rs:2180-2185)
    e = open_error(write(tmp_path, data[:-10], "t.sdx"))
    assert "truncated" in e
    # old format versions are rejected with migration advice (:2186-2191)
    v3 = bytearray(data)
    v3[8:12] = struct.pack("<I", 3)
    e = open_error(write(tmp_path, bytes(v3), "v3.sdx"))
    assert "rebuild" in e.lower() and "v3" in e
    # a flipped byte in the metadata is detected (:2192-2197)
    bad = bytearray(data)
    bad[-5] ^= 0x55
    assert "metadata checksum" in open_error(write(tmp_path, bytes(bad), "m.sdx"))


def test_other_checks_of_open(tmp_path):
    subs = subjects()
    data, meta = index_v4.build(subs)
    assert "not a SABER index" in open_error(write(tmp_path, b"SABERIDY" + data[8:], "magic.sdx"))
    b = bytearray(data)
    b[12:16] = struct.pack("<I", 0x04030201)
    assert "byte order" in open_error(write(tmp_path, index_v4.reheader(b), "endian.sdx"))
    b = bytearray(data)
    b[100] ^= 1
    assert "header checksum" in open_error(write(tmp_path, bytes(b), "hdr.sdx"))
    b = bytearray(data) + b"\0" * 8
    assert "header says" in open_error(write(tmp_path, bytes(b), "long.sdx"))
    # a corrupt section is found only when sections are verified
    enc = next(s for s in meta["sections"] if s["name"] == "enc")
    b = bytearray(data)
    b[enc["offset"] + 3] ^= 0x01
    p = write(tmp_path, bytes(b), "sec.sdx")
    S.IndexFile(p).close()
    assert "section 'enc' checksum mismatch" in open_error(p, verify=True)


def test_offsets_against_the_residues(tmp_path):
    """validate_structure (index.rs:824-836): the offset section's length,
    first and last offsets, and monotonicity; a metadata section out of bounds."""
    import json
    import zlib
    subs = subjects(3, 10)

    def with_meta(edit):
        data, meta = index_v4.build(subs)
        meta_off = struct.unpack_from("<Q", data, 24)[0]
        edit(meta)
        js = json.dumps(meta, separators=(",", ":")).encode()
        b = bytearray(data[:meta_off]) + js
        struct.pack_into("<QQQI", b, 16, len(b), meta_off, len(js), zlib.crc32(js))
        return index_v4.reheader(b)

    def more_subjects(m):
        m["num_subjects"] += 1
    assert "enc_off section has the wrong length" in open_error(write(tmp_path, with_meta(more_subjects), "a.sdx"))

    def fewer_residues(m):
        m["total_residues"] -= 1
    assert "do not match the residue section" in open_error(write(tmp_path, with_meta(fewer_residues), "b.sdx"))

    def misalign(m):
        m["sections"][0]["offset"] += 1
    assert "out of bounds or misaligned" in open_error(write(tmp_path, with_meta(misalign), "c.sdx"))

    def drop_ids(m):
        m["sections"] = [s for s in m["sections"] if s["name"] != "ids"]
    assert "missing section 'ids'" in open_error(write(tmp_path, with_meta(drop_ids), "d.sdx"))

    # offsets not monotonic: swap two interior offsets in the file
    data, meta = index_v4.build(subs)
    eo = next(s for s in meta["sections"] if s["name"] == "enc_off")
    b = bytearray(data)
    o1, o2 = eo["offset"] + 8 * 3, eo["offset"] + 8 * 4
    b[o1:o1 + 8], b[o2:o2 + 8] = b[o2:o2 + 8], b[o1:o1 + 8]
    assert "not monotonic" in open_error(write(tmp_path, bytes(b), "e.sdx"))


def test_library_exports():
    lib = _ffi.lib()
    for name in ("saber_hip_index_open", "saber_hip_index_close", "saber_hip_upload_index", "saber_hip_crc32"):
        assert hasattr(lib, name)
