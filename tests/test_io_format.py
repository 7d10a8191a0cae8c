"""CPU tests of the on-file format: the oracle (oracle/pencil_io.py) pinned by
the reference's own file test (crates/pencil-io/tests/mpi_io.rs:55-360: the
[4, 5] i32 array with extras [2] under the [1, 0] permutation, its header
bytes, its canonical payload at byte 152, and every corrupted-file case with
the exact error it must give), and the C ABI's header validation, catalog,
exclusive creation and path checks, which all run before any device work.
"""
import ctypes
import errno
import os
import struct

import numpy as np
import pytest

import pencil_io as PIO

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def kat_memory():
    """e * 1000 + x * 100 + y at logical (e, x, y), stored [e][y][x]."""
    v = np.array([[[e * 1000 + x * 100 + y for y in range(5)] for x in range(4)] for e in range(2)], dtype=np.int32)
    return np.ascontiguousarray(v.transpose(0, 2, 1)).reshape(-1)


def kat_file():
    return PIO.encode(kat_memory(), (4, 5), (2,), (1, 0), (1, 1))


def patch(b: bytearray, off: int, v: int):
    struct.pack_into("<Q", b, off, v)


def test_kat_bytes():
    b = kat_file()
    assert b[0:8] == b"PENCILIO"
    assert b[24:32] == (0x434F4D4D49545445).to_bytes(8, "little")
    expected = b"".join(int(e * 1000 + x * 100 + y).to_bytes(4, "little", signed=True)
                        for e in range(2) for x in range(4) for y in range(5))
    assert b[152:] == expected
    assert np.array_equal(PIO.read(b, np.int32, (4, 5), (2,), (1, 0)), kat_memory())
    # read back into another layout: memory [e][x][y]
    canon = PIO.read(b, np.int32, (4, 5), (2,), (0, 1))
    assert np.array_equal(canon, np.frombuffer(expected, dtype="<i4"))


def _drain(b, lo, hi):
    del b[lo:hi]


# (mutation, variant, detail): tests/mpi_io.rs:176-360
CASES = [
    (lambda b: b.__setitem__(0, ord("X")), "InvalidFile", "magic"),
    (lambda b: patch(b, 8, 2), "InvalidFile", "version"),
    (lambda b: patch(b, 40, 99), "InvalidFile", "type descriptor"),
    (lambda b: (_drain(b, 112, 136), patch(b, 16, 128), patch(b, 40, 99), patch(b, 56, 0), patch(b, 64, 128),
                patch(b, 80, 0)), "InvalidFile", "type descriptor"),
    (lambda b: patch(b, 88, 1), "InvalidFile", "reserved field"),
    (lambda b: patch(b, 120, 0), "InvalidFile", "header metadata"),
    (lambda b: (patch(b, 80, 3), b.__setitem__(slice(136, 136), (1).to_bytes(8, "little")), patch(b, 16, 160),
                patch(b, 64, 160)), "InvalidFile", "rank bound"),
    (lambda b: (patch(b, 96, 3), patch(b, 72, 2 * 3 * 5 * 4), b.__delitem__(slice(152 + 120, None))),
     "MetadataMismatch", "global shape"),
    (lambda b: patch(b, 24, 7), "InvalidFile", "commit marker"),
    (lambda b: patch(b, 24, 0x494E434F4D504C45), "IncompleteFile", None),
    (lambda b: b.__delitem__(slice(8, None)), "InvalidFile", "short header"),
    (lambda b: b.pop(), "InvalidFile", "trailing or truncated payload"),
    (lambda b: b.append(0), "InvalidFile", "trailing or truncated payload"),
    (lambda b: patch(b, 40, 6), "MetadataMismatch", "type or rank"),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_oracle_rejects_corrupted_files(case):
    mutate, variant, detail = CASES[case]
    b = bytearray(kat_file())
    mutate(b)
    with pytest.raises(PIO.IoError) as e:
        PIO.read(bytes(b), np.int32, (4, 5), (2,), (1, 0))
    assert e.value.variant == variant
    if detail:
        assert e.value.detail == detail


# ------------------------------------------------ C ABI (no device work) --
@pytest.fixture(scope="module")
def pf():
    import subprocess

    import pencil_fft_hip as pf

    if not os.path.exists(os.path.join(ROOT, "pencil-rs_amd", "libpencilfft_hip.so")):
        subprocess.run(["make", "-C", os.path.join(ROOT, "pencil-rs_amd"), "-j8"], check=True)
    pf.load_library()
    return pf


def fake_array(pf, shape=(4, 5), perm=(1, 0), extra=(2,), code=5, decomposition=(0, 1)):
    a = pf._Array()
    a.data = 0x1000
    a.scalar_type = code
    a.extra_ndim = len(extra)
    for i, v in enumerate(extra):
        a.extra_shape[i] = v
    a.pencil = pf.PencilLayout(shape, perm, decomposition)._c()
    return a


def abi_read(pf, path, arr=None, raw=None):
    arr = fake_array(pf) if arr is None else arr
    ws = ctypes.c_void_p(0x2000)
    if raw is None:
        st = pf._lib.pfh_io_read(os.fsencode(str(path)), ctypes.byref(arr), ws, 1 << 20, None)
    else:
        st = pf._lib.pfh_io_read_raw(os.fsencode(str(path)), ctypes.byref(arr), raw[0], raw[1], ws, 1 << 20, None)
    info = pf._ErrorInfo()
    pf._lib.pfh_last_error(ctypes.byref(info))
    return pf.STATUS_NAMES[st], info


@pytest.mark.parametrize("case", range(len(CASES)))
def test_abi_rejects_corrupted_files_before_any_device_work(pf, tmp_path, case):
    mutate, variant, detail = CASES[case]
    b = bytearray(kat_file())
    mutate(b)
    path = tmp_path / "case.pio"
    path.write_bytes(bytes(b))
    name, info = abi_read(pf, path)
    assert name == variant
    if detail:
        assert info.message.decode() == detail


def test_abi_catalog_and_metadata_mismatches(pf, tmp_path):
    path = tmp_path / "array.pio"
    path.write_bytes(kat_file())
    info = pf.read_mpi_catalog(path)
    assert info.global_shape == (4, 5) and info.extra_shape == (2,) and info.writer_grid == (1, 1)
    assert info.writer_permutation == (1, 0) and info.payload_bytes == 160
    import torch

    assert info.scalar_type == torch.int32
    # a destination of another extra shape, global shape or type
    name, i = abi_read(pf, path, fake_array(pf, extra=(3,)))
    assert (name, i.message.decode()) == ("MetadataMismatch", "extra shape")
    name, i = abi_read(pf, path, fake_array(pf, shape=(5, 4)))
    assert (name, i.message.decode()) == ("MetadataMismatch", "global shape")
    name, i = abi_read(pf, path, fake_array(pf, code=9))
    assert (name, i.message.decode()) == ("MetadataMismatch", "type or rank")
    # missing file
    name, i = abi_read(pf, tmp_path / "missing.pio")
    assert name == "Native" and i.actual == errno.ENOENT


def test_abi_write_is_exclusive_and_paths_are_checked(pf, tmp_path):
    path = tmp_path / "array.pio"
    path.write_bytes(kat_file())
    before = path.read_bytes()
    arr = fake_array(pf)
    st = pf._lib.pfh_io_write(os.fsencode(str(path)), ctypes.byref(arr), ctypes.c_void_p(0x2000), 1 << 20, None)
    info = pf._ErrorInfo()
    pf._lib.pfh_last_error(ctypes.byref(info))
    assert pf.STATUS_NAMES[st] == "Native" and info.message.decode() == "open" and info.actual == errno.EEXIST
    assert path.read_bytes() == before
    long_path = os.fsencode("p" * (100 * 1024))
    st = pf._lib.pfh_io_write(long_path, ctypes.byref(arr), ctypes.c_void_p(0x2000), 1 << 20, None)
    assert pf.STATUS_NAMES[st] == "SizeLimit"
    st = pf._lib.pfh_io_read(long_path, ctypes.byref(arr), ctypes.c_void_p(0x2000), 1 << 20, None)
    assert pf.STATUS_NAMES[st] == "SizeLimit"
    st = pf._lib.pfh_io_read(b"", ctypes.byref(arr), ctypes.c_void_p(0x2000), 1 << 20, None)
    assert pf.STATUS_NAMES[st] == "InvalidPath"
    # too small a staging buffer is reported before the file is touched
    st = pf._lib.pfh_io_read(os.fsencode(str(path)), ctypes.byref(arr), ctypes.c_void_p(0x2000), 100, None)
    pf._lib.pfh_last_error(ctypes.byref(info))
    assert pf.STATUS_NAMES[st] == "WorkspaceTooSmall" and (info.required, info.actual) == (160, 100)


def test_abi_raw_read_checks(pf, tmp_path):
    path = tmp_path / "raw.bin"
    path.write_bytes(bytes(100))
    name, i = abi_read(pf, path, raw=(0, 0))  # 160 bytes wanted, 100 present
    assert (name, i.message.decode()) == ("InvalidFile", "truncated raw payload")
    name, _ = abi_read(pf, path, raw=(0, 7))
    assert name == "InvalidInput"
    name, i = abi_read(pf, path, raw=(2 ** 64 - 8, 1))
    assert (name, i.message.decode()) == ("SizeLimit", "raw payload end")


# ------------------------------------------------------- named container --
def named_a_memory():
    """e * 10000 + x * 100 + y at logical (e, x, y) of a [4, 5] array with
    extras [2], stored [e][y][x] (the [1, 0] permutation)."""
    v = np.array([[[e * 10000 + x * 100 + y for y in range(5)] for x in range(4)] for e in range(2)], dtype=np.int32)
    return np.ascontiguousarray(v.transpose(0, 2, 1)).reshape(-1)


def named_file():
    a = PIO.encode_record("A", named_a_memory(), (4, 5), (2,), (1, 0))
    b = PIO.encode_record("B", np.full(12, 3.25), (3, 4), (), (0, 1))
    return PIO.container_header() + a + b, len(PIO.container_header()) + len(a)


def test_named_oracle_matches_the_reference_raw_layout():
    """raw_oracle (crates/pencil-io/tests/named_mpi.rs:82-140), restated."""
    raw, bstart = named_file()
    word = lambda p: struct.unpack_from("<Q", raw, p)[0]
    assert raw[:8] == b"PIONAM02" and word(8) == 2 and word(16) == 32
    p = 32
    for key, code, width, dims, extra in (("A", 5, 4, [4, 5], [2]), ("B", 10, 8, [3, 4], [])):
        assert raw[p:p + 8] == b"PIOREC02" and word(p + 8) == 2
        nl, ml, payload, total = word(p + 16), word(p + 24), word(p + 32), word(p + 40)
        assert (word(p + 48), word(p + 56), word(p + 64)) == (code, width, 0)
        assert raw[p + 72:p + 72 + nl] == key.encode()
        meta = p + 72 + nl
        assert word(meta) == 2 and word(meta + 8) == len(extra) and ml == 16 + 8 * (2 + len(extra))
        for i, d in enumerate(dims + extra):
            assert word(meta + 16 + 8 * i) == d
        data = meta + ml
        assert payload == int(np.prod(dims)) * int(np.prod(extra, dtype=np.int64)) * width
        assert total == 72 + nl + ml + payload + 8
        assert raw[p + total - 8:p + total] == (0x434F4D4D49545445).to_bytes(8, "little")
        if key == "A":
            expected = [e * 10000 + x * 100 + y for e in range(2) for x in range(4) for y in range(5)]
            assert list(np.frombuffer(raw[data:data + payload], dtype="<i4")) == expected
        else:
            assert np.all(np.frombuffer(raw[data:data + payload], dtype="<f8") == 3.25)
        p += total
    assert p == len(raw)
    assert p - total == bstart


def bad_tails():
    """tests/named_mpi.rs:556-575: record B's header fields corrupted, or the
    file cut inside record B."""
    raw, bstart = named_file()
    out = []
    for off, value in ((8, 99), (16, 2 ** 64 - 1), (24, 2 ** 64 - 1), (32, 1), (40, 2 ** 64 - 1), (48, 99),
                       (56, 1), (64, 1), (73, 2 ** 64 - 1)):
        c = bytearray(raw)
        c[bstart + off:bstart + off + 8] = value.to_bytes(8, "little")
        out.append(bytes(c))
    for cut in (bstart + 1, bstart + 72, len(raw) - 8, len(raw) - 1):
        out.append(raw[:cut])
    return out


@pytest.mark.parametrize("case", range(13))
def test_named_bad_tails(pf, tmp_path, case):
    corrupt = bad_tails()[case]
    recs, tail = PIO.scan(corrupt)
    assert [r[0] for r in recs] == ["A"] and tail < len(corrupt)
    assert np.array_equal(PIO.read_named(corrupt, "A", np.int32, (4, 5), (2,), (1, 0)), named_a_memory())
    with pytest.raises(PIO.IoError) as e:
        PIO.read_named(corrupt, "B", np.float64, (3, 4), (), (0, 1))
    assert e.value.variant == "NotFound"
    path = tmp_path / "named.pio"
    path.write_bytes(corrupt)
    arr = fake_array(pf)
    ws = ctypes.c_void_p(0x2000)
    st = pf._lib.pfh_io_append_named(os.fsencode(str(path)), b"new", ctypes.byref(arr), ws, 1 << 20, None)
    assert pf.STATUS_NAMES[st] == "InvalidTail"
    assert path.read_bytes() == corrupt
    b_arr = fake_array(pf, shape=(3, 4), perm=(0, 1), extra=(), code=10)
    st = pf._lib.pfh_io_read_named(os.fsencode(str(path)), b"B", ctypes.byref(b_arr), ws, 1 << 20, None)
    assert pf.STATUS_NAMES[st] == "NotFound"


def test_named_abi_checks(pf, tmp_path):
    raw, _ = named_file()
    path = tmp_path / "named.pio"
    path.write_bytes(raw)
    arr = fake_array(pf)
    ws = ctypes.c_void_p(0x2000)
    f = os.fsencode(str(path))

    def call(fn, name, a=arr):
        st = fn(f, name, ctypes.byref(a), ws, 1 << 20, None)
        info = pf._ErrorInfo()
        pf._lib.pfh_last_error(ctypes.byref(info))
        return pf.STATUS_NAMES[st], info.message.decode()

    for bad in (b"", b"x" * 1025, b"\xff\xfe", b"\xed\xa0\x80", b"\xc0\xaf"):
        assert not PIO.valid_name(bad)
        assert call(pf._lib.pfh_io_append_named, bad)[0] == "InvalidName"
        assert call(pf._lib.pfh_io_read_named, bad)[0] == "InvalidName"
    assert PIO.valid_name("utf8/α".encode()) and PIO.valid_name(b"x" * 1024)
    assert call(pf._lib.pfh_io_append_named, b"A")[0] == "DuplicateName"
    assert call(pf._lib.pfh_io_read_named, b"absent")[0] == "NotFound"
    assert call(pf._lib.pfh_io_read_named, b"A", fake_array(pf, extra=(3,))) == ("MetadataMismatch", "named record")
    assert call(pf._lib.pfh_io_read_named, b"B")[0] == "MetadataMismatch"
    assert path.read_bytes() == raw
    # an exclusive write onto an existing container fails without touching it
    assert call(pf._lib.pfh_io_write_named, b"C")[0] == "Native"
    assert path.read_bytes() == raw
    bad = bytearray(raw)
    bad[24] = 1
    path.write_bytes(bytes(bad))
    assert call(pf._lib.pfh_io_read_named, b"A") == ("InvalidFile", "named container header")
    with pytest.raises(PIO.IoError):
        PIO.scan(bytes(bad))
