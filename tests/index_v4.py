"""A writer of SABER index files, format v4, for tests: the file layout of
IndexBuilder::finish_to_file (src/search/index.rs:1679-1713) -- a 256-byte
header (magic, version, byte-order mark, file size, metadata offset, length
and CRC-32, header CRC-32 of bytes [0, 252) at 252), sections 8-byte
aligned in the order they are pushed (Sink::push, :1063-1078), then the
metadata JSON with IndexMeta's field names (:269-320; SectionMeta :245-250,
PartitionMeta :252-267) -- for an index without seed postings (postings =
false: residues, identifiers and offsets only, searched with the scan
schedule). CRC-32 is zlib's, the same IEEE CRC as src/search/crc32.rs.
"""
import json
import struct
import zlib

import numpy as np

MAGIC = b"SABERIDX"
ENDIAN_MARK = 0x01020304
HEADER_BYTES = 256


def build(subjects, ids=None):
    """The bytes of a v4 index of subjects (arrays of alignment codes)."""
    ids = ids or [f"s{i}" for i in range(len(subjects))]
    enc = np.concatenate([np.asarray(s, np.uint8) for s in subjects]) if subjects else np.zeros(0, np.uint8)
    off = np.zeros(len(subjects) + 1, np.uint64)
    np.cumsum([len(s) for s in subjects], out=off[1:])
    idb = b"".join(i.encode() for i in ids)
    idoff = np.zeros(len(ids) + 1, np.uint64)
    np.cumsum([len(i.encode()) for i in ids], out=idoff[1:])
    body = bytearray(HEADER_BYTES)
    sections = []

    def push(name, data):
        pad = (-len(body)) % 8
        body.extend(b"\0" * pad)
        sections.append({"name": name, "offset": len(body), "len": len(data), "crc32": zlib.crc32(data)})
        body.extend(data)

    push("enc", enc.tobytes())
    push("enc_off", off.tobytes())
    push("ids", idb)
    push("id_off", idoff.tobytes())
    n = len(subjects)
    meta = {
        "format_version": 4, "saber_version": "1.4.0", "kind": "Protein", "alphabet_name": "murphy10",
        "alphabet_size": 10, "alphabet_map": [0] * 24, "shapes": ["1111"], "minimizer_window": 1,
        "masking": "none", "masking_matrix": "blosum62", "num_subjects": n, "total_residues": int(off[-1]),
        "max_subject_len": max((len(s) for s in subjects), default=0),
        "partitions": [{"first_subject": 0, "num_subjects": n, "residues": int(off[-1]),
                        "max_len": max((len(s) for s in subjects), default=0), "compact": True,
                        "postings": [0], "sparse": []}],
        "sections": sections, "source": "test", "total_postings": [0], "translated": False,
        "genetic_code": "", "nt_sequences": 0, "nt_residues": 0, "postings": False, "scan_totals": [],
        "scan_freq_min": 128,
    }
    js = json.dumps(meta, separators=(",", ":")).encode()
    body.extend(b"\0" * ((-len(body)) % 8))
    meta_off = len(body)
    body.extend(js)
    hdr = bytearray(HEADER_BYTES)
    hdr[0:8] = MAGIC
    struct.pack_into("<IIQQQI", hdr, 8, 4, ENDIAN_MARK, len(body), meta_off, len(js), zlib.crc32(js))
    struct.pack_into("<I", hdr, 252, zlib.crc32(bytes(hdr[0:252])))
    body[0:HEADER_BYTES] = hdr
    return bytes(body), meta


def reheader(data):
    """data with its header CRC recomputed (after editing header fields)."""
    b = bytearray(data)
    struct.pack_into("<I", b, 252, zlib.crc32(bytes(b[0:252])))
    return bytes(b)
