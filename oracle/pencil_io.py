"""CPU restatement of the reference's single-dataset file format
(TEST INFRASTRUCTURE ONLY: only tests/ import this module, as the checker of
the device writer and reader).

Follows crates/pencil-io/src/mpi_io.rs: Header::for_write / encode_into
(:1492-1607), parse_prefix (:1610-1631), decode (:1633-1690), V1Fields
(:1749-1825), V1Arrays (:1856-1890), matches_view (:1692-1748), the read
flow's size checks (:636-760) and the physical decode (format.rs:209-330);
element codes format.rs:50-126.
"""
from __future__ import annotations

import struct
from typing import List, Sequence

import numpy as np

MAGIC = b"PENCILIO"
VERSION = 1
COMMIT = 0x434F4D4D49545445
INCOMPLETE = 0x494E434F4D504C45
PREFIX = 96
MAX_HEADER = 1024 * 1024
MAX_RANK = 1024

CODES = {np.dtype(np.int8): 1, np.dtype(np.uint8): 2, np.dtype(np.int16): 3, np.dtype(np.uint16): 4,
         np.dtype(np.int32): 5, np.dtype(np.uint32): 6, np.dtype(np.int64): 7, np.dtype(np.uint64): 8,
         np.dtype(np.float32): 9, np.dtype(np.float64): 10, np.dtype(np.complex64): 11,
         np.dtype(np.complex128): 12}
WIDTHS = {1: 1, 2: 1, 3: 2, 4: 2, 5: 4, 6: 4, 7: 8, 8: 8, 9: 4, 10: 8, 11: 8, 12: 16}


class IoError(Exception):
    """IoError variant with its string field (reason / field / what)."""

    def __init__(self, variant: str, detail: str = ""):
        self.variant, self.detail = variant, detail
        super().__init__(f"{variant}: {detail}")


def u64(b: bytes, off: int) -> int:
    return struct.unpack_from("<Q", b, off)[0]


def canonical_payload(memory: np.ndarray, global_shape: Sequence[int], extra: Sequence[int],
                      perm: Sequence[int]) -> bytes:
    """pack_view (format.rs:156-207): memory-order values [extra][memory axes]
    to canonical [extra][logical axes] little-endian bytes."""
    n = len(global_shape)
    mem_shape = list(extra) + [global_shape[perm[q]] for q in range(n)]
    a = memory.reshape(mem_shape)
    e = len(extra)
    # memory position q holds logical axis perm[q]; logical axis a sits at q with perm[q] == a
    order = list(range(e)) + [e + list(perm).index(ax) for ax in range(n)]
    return np.ascontiguousarray(a.transpose(order)).astype(memory.dtype.newbyteorder("<")).tobytes()


def encode(memory: np.ndarray, global_shape, extra, perm, grid: Sequence[int]) -> bytes:
    """Header::for_write + encode_into, then the payload, with the commit
    marker set (the state of a completed write)."""
    n, dt = len(global_shape), memory.dtype
    count = n + len(extra) + len(grid) + n
    header_len = PREFIX + 8 * count
    payload = canonical_payload(memory, global_shape, extra, perm)
    h = bytearray(header_len)
    h[0:8] = MAGIC
    for off, v in ((8, VERSION), (16, header_len), (24, COMMIT), (32, n), (40, CODES[dt]), (48, dt.itemsize),
                   (56, len(extra)), (64, header_len), (72, len(payload)), (80, len(grid)), (88, 0)):
        struct.pack_into("<Q", h, off, v)
    off = PREFIX
    for vals in (global_shape, extra, grid, perm):
        for v in vals:
            struct.pack_into("<Q", h, off, int(v))
            off += 8
    return bytes(h) + payload


def _valid_dimensions(v: List[int]) -> bool:
    if not v:
        return False
    p = 1
    for x in v:
        p *= x
        if p >= 1 << 64:
            return False
    return True


def decode_header(b: bytes):
    """read flow + parse_prefix + decode: returns (n, code, width, global,
    extra, grid, perm, payload_offset, payload_bytes) or raises IoError in
    the reference's order of checks."""
    if len(b) < PREFIX:
        raise IoError("InvalidFile", "short header")
    if b[0:8] != MAGIC:
        raise IoError("InvalidFile", "magic")
    if u64(b, 8) != VERSION:
        raise IoError("InvalidFile", "version")
    marker = u64(b, 24)
    if marker == INCOMPLETE:
        raise IoError("IncompleteFile")
    if marker != COMMIT:
        raise IoError("InvalidFile", "commit marker")
    header_len, n, code, width = u64(b, 16), u64(b, 32), u64(b, 40), u64(b, 48)
    extra_rank, poff, pbytes, grid_rank = u64(b, 56), u64(b, 64), u64(b, 72), u64(b, 80)
    if n > MAX_RANK or extra_rank > MAX_RANK or n + extra_rank > MAX_RANK or grid_rank > MAX_RANK or grid_rank > n:
        raise IoError("InvalidFile", "rank bound")
    if header_len != PREFIX + 8 * (2 * n + extra_rank + grid_rank) or header_len > MAX_HEADER:
        raise IoError("InvalidFile", "header length")
    if poff != header_len:
        raise IoError("InvalidFile", "payload offset")
    if WIDTHS.get(code) != width:
        raise IoError("InvalidFile", "type descriptor")
    if header_len > len(b):
        raise IoError("InvalidFile", "truncated header")
    if u64(b, 88) != 0:
        raise IoError("InvalidFile", "reserved field")
    vals = [u64(b, PREFIX + 8 * i) for i in range(2 * n + extra_rank + grid_rank)]
    glob, extra = vals[:n], vals[n:n + extra_rank]
    grid, perm = vals[n + extra_rank:n + extra_rank + grid_rank], vals[n + extra_rank + grid_rank:]
    grid_ok = _valid_dimensions(grid) and all(g > 0 for g in grid) and len(grid) <= len(glob)
    if not _valid_dimensions(glob) or not grid_ok or sorted(perm) != list(range(n)):
        raise IoError("InvalidFile", "header metadata")
    elems = 1
    for v in glob + extra:
        elems *= v
    if elems >= 1 << 64:
        raise IoError("InvalidFile", "payload element count")
    if elems * width != pbytes:
        raise IoError("InvalidFile", "payload bytes")
    return n, code, width, glob, extra, grid, perm, poff, pbytes


def read(b: bytes, dtype, global_shape, extra, perm) -> np.ndarray:
    """read_mpi at one rank: header checks, matches_view, exact file size,
    then the physical decode into the destination's memory order."""
    n, code, width, glob, ex, _grid, _wperm, poff, pbytes = decode_header(b)
    dt = np.dtype(dtype)
    if n != len(global_shape) or code != CODES[dt] or width != dt.itemsize:
        raise IoError("MetadataMismatch", "type or rank")
    if list(ex) != list(extra):
        raise IoError("MetadataMismatch", "extra shape")
    if list(glob) != list(global_shape):
        raise IoError("MetadataMismatch", "global shape")
    expected = int(np.prod(global_shape, dtype=np.int64)) * int(np.prod(extra, dtype=np.int64)) * dt.itemsize
    if expected != pbytes:
        raise IoError("MetadataMismatch", "payload size")
    if poff + pbytes != len(b):
        raise IoError("InvalidFile", "trailing or truncated payload")
    canon = np.frombuffer(b[poff:], dtype=dt.newbyteorder("<")).astype(dt)
    e = len(extra)
    a = canon.reshape(list(extra) + list(global_shape))
    order = list(range(e)) + [e + perm[q] for q in range(len(perm))]
    return np.ascontiguousarray(a.transpose(order)).reshape(-1)


# ---------------------------------------------------------------------------
# Named (v2) append-only container (crates/pencil-io/src/named_mpi.rs)
# ---------------------------------------------------------------------------
NAMED_MAGIC = b"PIONAM02"
RECORD_MAGIC = b"PIOREC02"
CONTAINER = 32
RECORD_HEADER = 72


def container_header() -> bytes:
    """container_header (named_mpi.rs:982-988)."""
    return NAMED_MAGIC + struct.pack("<QQQ", 2, CONTAINER, 0)


def valid_name(name: bytes) -> bool:
    """valid_name_bytes (lib.rs:343-348)."""
    if not name or len(name) > 1024 or b"\0" in name:
        return False
    try:
        name.decode("utf-8", errors="strict")
    except UnicodeDecodeError:
        return False
    return True


def encode_record(name: str, memory: np.ndarray, global_shape, extra, perm) -> bytes:
    """record (named_mpi.rs:920-980) + payload + commit word."""
    nb = name.encode("utf-8")
    payload = canonical_payload(memory, global_shape, extra, perm)
    meta = struct.pack("<QQ", len(global_shape), len(extra)) + b"".join(struct.pack("<Q", int(v))
                                                                        for v in list(global_shape) + list(extra))
    total = RECORD_HEADER + len(nb) + len(meta) + len(payload) + 8
    dt = memory.dtype
    head = RECORD_MAGIC + struct.pack("<QQQQQQQQ", 2, len(nb), len(meta), len(payload), total, CODES[dt],
                                      dt.itemsize, 0)
    return head + nb + meta + payload + struct.pack("<Q", COMMIT)


def scan(b: bytes):
    """scan_root (named_mpi.rs:333-505): the committed records
    [(name, code, width, global, extra, payload_offset, payload_len)] and the
    end of the committed prefix; raises for a bad container header."""
    size = len(b)
    if size < CONTAINER or b[0:8] != NAMED_MAGIC or u64(b, 8) != 2 or u64(b, 16) != CONTAINER or u64(b, 24) != 0:
        raise IoError("InvalidFile", "named container header")
    p, out, names = CONTAINER, [], set()
    while p < size:
        if size - p < RECORD_HEADER:
            break
        if b[p:p + 8] != RECORD_MAGIC or u64(b, p + 8) != 2 or u64(b, p + 64) != 0:
            break
        nl, ml, payload, total = u64(b, p + 16), u64(b, p + 24), u64(b, p + 32), u64(b, p + 40)
        if not (0 < nl <= 1024) or ml > MAX_HEADER:
            break
        end = p + total
        if end > size or end >= 1 << 63:
            break
        if RECORD_HEADER + nl + ml + payload + 8 != total:
            break
        name = b[p + RECORD_HEADER:p + RECORD_HEADER + nl]
        if not valid_name(name) or name in names:
            break
        meta = p + RECORD_HEADER + nl
        nr, er = u64(b, meta), u64(b, meta + 8)
        if not (0 < nr <= MAX_RANK) or er > MAX_RANK or nr + er > MAX_RANK or 16 + 8 * (nr + er) != ml:
            break
        dims = [u64(b, meta + 16 + 8 * i) for i in range(nr + er)]
        elems = 1
        for d in dims:
            elems *= d
        code, width = u64(b, p + 48), u64(b, p + 56)
        if WIDTHS.get(code) != width or elems * width != payload:
            break
        if u64(b, end - 8) != COMMIT:
            break
        names.add(name)
        out.append((name.decode(), code, width, dims[:nr], dims[nr:], meta + ml, payload))
        p = end
    return out, p


def read_named(b: bytes, name: str, dtype, global_shape, extra, perm) -> np.ndarray:
    """read_mpi_named at one rank (named_mpi.rs:1430-1470)."""
    if not valid_name(name.encode("utf-8")):
        raise IoError("InvalidName")
    recs, _ = scan(b)
    hit = [r for r in recs if r[0] == name]
    if not hit:
        raise IoError("NotFound")
    _, code, width, glob, ex, off, plen = hit[0]
    dt = np.dtype(dtype)
    if code != CODES[dt] or width != dt.itemsize or list(glob) != list(global_shape) or list(ex) != list(extra):
        raise IoError("MetadataMismatch", "named record")
    canon = np.frombuffer(b[off:off + plen], dtype=dt.newbyteorder("<")).astype(dt)
    e = len(extra)
    a = canon.reshape(list(extra) + list(global_shape))
    order = list(range(e)) + [e + perm[q] for q in range(len(perm))]
    return np.ascontiguousarray(a.transpose(order)).reshape(-1)
