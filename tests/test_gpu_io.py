"""Device arrays through the on-file format, against the oracle's encoder and
decoder (oracle/pencil_io.py): the written file bit-exact (header and
canonical payload, NaN payloads included) for every element type and
several memory orders, reads into another memory order bit-exact, rejected
reads leave the destination unchanged, headerless raw input in both byte
orders, and a 512^3 f32 field written and read back.
"""
import os
import time

import numpy as np
import pytest
import torch

import pencil_fft_hip as pf
import pencil_io as PIO

pytestmark = pytest.mark.gpu

TYPES = [np.int8, np.uint8, np.int16, np.uint16, np.int32, np.uint32, np.int64, np.uint64, np.float32, np.float64,
         np.complex64, np.complex128]


def values(rng, dt, n):
    """Random bit patterns of the type (floats include NaN payloads and
    denormals), so the round trip is checked bit for bit."""
    dt = np.dtype(dt)
    return np.frombuffer(rng.bytes(n * dt.itemsize), dtype=dt).copy()


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def bits(t: torch.Tensor) -> np.ndarray:
    return t.cpu().contiguous().view(torch.uint8).numpy()


@pytest.mark.parametrize("dt", TYPES)
@pytest.mark.parametrize("layout", [((6, 5, 4), (0, 1, 2), (2,)), ((6, 5, 4), (2, 0, 1), (3, 1)),
                                    ((7, 3), (1, 0), ())])
def test_write_bit_exact_and_read_into_another_order(gpu, tmp_path, dt, layout):
    shape, perm, extra = layout
    rng = np.random.default_rng(PIO.CODES[np.dtype(dt)] * 10 + len(perm) + perm[0])
    n = int(np.prod(shape)) * int(np.prod(extra, dtype=np.int64))
    mem = values(rng, dt, n)
    lay = pf.PencilLayout(shape, perm, decomposition=(0, 1))
    src = pf.PencilArray(dev(mem), lay, extra)
    path = tmp_path / "a.pio"
    pf.write_mpi(path, src)
    want = PIO.encode(mem, shape, extra, perm, (1, 1))
    got = path.read_bytes()
    assert got == want
    # read into the reversed memory order
    rperm = tuple(reversed(perm))
    dst = pf.PencilArray(torch.zeros(n, dtype=src.data.dtype, device="cuda"), pf.PencilLayout(shape, rperm), extra)
    pf.read_mpi(path, dst)
    ref = PIO.read(got, dt, shape, extra, rperm)
    assert np.array_equal(bits(dst.data), ref.view(np.uint8))
    info = pf.read_mpi_catalog(path)
    assert info.global_shape == shape and info.extra_shape == tuple(extra) and info.writer_permutation == perm


def test_rejected_reads_leave_the_destination_unchanged(gpu, tmp_path):
    rng = np.random.default_rng(3)
    shape, perm, extra = (8, 6), (1, 0), (2,)
    mem = rng.standard_normal(96).astype(np.float64)
    path = tmp_path / "a.pio"
    pf.write_mpi(path, pf.PencilArray(dev(mem), pf.PencilLayout(shape, perm, (0,)), extra))
    good = path.read_bytes()
    before = torch.full((96,), -7.0, dtype=torch.float64, device="cuda")
    dst = pf.PencilArray(before.clone(), pf.PencilLayout(shape, perm), extra)
    for k, mutate in enumerate([lambda b: b.__setitem__(slice(24, 32), bytes(8)), lambda b: b.pop(),
                                lambda b: b.__setitem__(slice(0, 1), b"Q")]):
        b = bytearray(good)
        mutate(b)
        p = tmp_path / f"bad{k}.pio"
        p.write_bytes(bytes(b))
        with pytest.raises(pf.PfhError) as e:
            pf.read_mpi(p, dst)
        assert e.value.variant == "InvalidFile"
        torch.cuda.synchronize()
        assert torch.equal(dst.data, before)
    with pytest.raises(pf.PfhError) as e:
        pf.read_mpi(path, pf.PencilArray(dst.data, pf.PencilLayout(shape, perm), (1, 2)))
    assert e.value.variant == "MetadataMismatch" and e.value.detail == "extra shape"
    assert torch.equal(dst.data, before)
    with pytest.raises(pf.PfhError) as e:
        pf.write_mpi(path, dst)
    assert e.value.variant == "Native"
    assert path.read_bytes() == good
    pf.read_mpi(path, dst)
    assert np.array_equal(dst.data.cpu().numpy(), PIO.read(good, np.float64, shape, extra, perm))


@pytest.mark.parametrize("dt", [np.float32, np.complex128, np.int16])
@pytest.mark.parametrize("order", ["little", "big", "native"])
def test_raw_input(gpu, tmp_path, dt, order):
    rng = np.random.default_rng(11)
    shape, perm, extra = (5, 4, 3), (2, 1, 0), (2,)
    n = 120
    canon = values(rng, dt, n)
    # numpy's byteswap reverses each real component (both parts of a complex value)
    payload = canon.byteswap().tobytes() if order == "big" else canon.tobytes()
    path = tmp_path / "raw.bin"
    path.write_bytes(b"prefix-13byte" + payload + b"trailing records")
    dst = pf.PencilArray(torch.zeros(n, dtype=dev(canon).dtype, device="cuda"), pf.PencilLayout(shape, perm), extra)
    pf.read_mpi_raw(path, dst, offset=13, byte_order=order)
    want = np.ascontiguousarray(canon.reshape(2, 5, 4, 3).transpose(0, 3, 2, 1)).reshape(-1)
    assert np.array_equal(bits(dst.data), want.view(np.uint8))


@pytest.mark.slow
def test_full_size_field_round_trip(gpu, tmp_path):
    """A 512^3 f32 field in the R2C output's [2, 1, 0] memory order written
    and read back; prints the device-to-file and file-to-device rates."""
    shape = (512, 512, 512)
    g = torch.Generator(device="cuda")
    g.manual_seed(2)
    x = torch.randn(512 ** 3, generator=g, device="cuda", dtype=torch.float32)
    lay = pf.PencilLayout(shape, (2, 1, 0), (0, 1))
    ws = torch.empty(x.numel() * 4, dtype=torch.uint8, device="cuda")
    path = tmp_path / "field.pio"
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pf.write_mpi(path, pf.PencilArray(x, lay), workspace=ws)
    t1 = time.perf_counter()
    y = torch.empty_like(x)
    pf.read_mpi(path, pf.PencilArray(y, lay), workspace=ws)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    gb = x.numel() * 4 / 1e9
    print(f"512^3 f32: write {gb / (t1 - t0):.2f} GB/s, read {gb / (t2 - t1):.2f} GB/s", flush=True)
    assert torch.equal(x, y)
    assert os.path.getsize(path) == 96 + 8 * (3 + 0 + 2 + 3) + x.numel() * 4
    # the payload is canonical: element (i, j, k) at ((i * 512 + j) * 512 + k)
    with open(path, "rb") as f:
        f.seek(96 + 64 + 4 * ((3 * 512 + 5) * 512 + 7))
        v = np.frombuffer(f.read(4), dtype="<f4")[0]
    assert v == x.reshape(512, 512, 512)[7, 5, 3].item()
    os.remove(path)


def test_named_container_write_append_read(gpu, tmp_path):
    """write_mpi_named A, append_mpi_named B: the file equals the oracle's
    container bit for bit; both records read back into other memory orders;
    a duplicate name and a corrupted tail are refused without touching the
    file (tests/named_mpi.rs:320-590 at one rank)."""
    a_log = np.array([[[e * 10000 + x * 100 + y for y in range(5)] for x in range(4)] for e in range(2)],
                     dtype=np.int32)
    a_mem = np.ascontiguousarray(a_log.transpose(0, 2, 1)).reshape(-1)
    A = pf.PencilArray(dev(a_mem), pf.PencilLayout((4, 5), (1, 0), (0,)), (2,))
    B = pf.PencilArray(torch.full((12,), 3.25, dtype=torch.float64, device="cuda"), pf.PencilLayout((3, 4)))
    path = tmp_path / "named.pio"
    pf.write_mpi_named(path, "A", A)
    base = path.read_bytes()
    assert base == PIO.container_header() + PIO.encode_record("A", a_mem, (4, 5), (2,), (1, 0))
    pf.append_mpi_named(path, "B", B)
    raw = path.read_bytes()
    assert raw[:len(base)] == base
    assert raw == base + PIO.encode_record("B", np.full(12, 3.25), (3, 4), (), (0, 1))
    ar = pf.PencilArray(torch.full((40,), -77, dtype=torch.int32, device="cuda"), pf.PencilLayout((4, 5)), (2,))
    pf.read_mpi_named(path, "A", ar)
    assert np.array_equal(ar.data.cpu().numpy(), a_log.reshape(-1))
    br = pf.PencilArray(torch.full((12,), -2.0, dtype=torch.float64, device="cuda"), pf.PencilLayout((3, 4), (1, 0)))
    pf.read_mpi_named(path, "B", br)
    assert bool(torch.all(br.data == 3.25))
    with pytest.raises(pf.PfhError) as e:
        pf.append_mpi_named(path, "A", A)
    assert e.value.variant == "DuplicateName" and path.read_bytes() == raw
    pf.append_mpi_named(path, "utf8/α", A)
    u = pf.PencilArray(torch.full((40,), -5, dtype=torch.int32, device="cuda"), pf.PencilLayout((4, 5), (1, 0)), (2,))
    pf.read_mpi_named(path, "utf8/α", u)
    assert np.array_equal(u.data.cpu().numpy(), a_mem)
    # a record cut short: the committed prefix still reads, append is refused
    cut = path.read_bytes()[:-3]
    path.write_bytes(cut)
    pf.read_mpi_named(path, "B", br)
    before = u.data.clone()
    with pytest.raises(pf.PfhError) as e:
        pf.read_mpi_named(path, "utf8/α", u)
    assert e.value.variant == "NotFound" and torch.equal(u.data, before)
    with pytest.raises(pf.PfhError) as e:
        pf.append_mpi_named(path, "C", A)
    assert e.value.variant == "InvalidTail" and path.read_bytes() == cut
    # empty extras are valid records
    z = pf.PencilArray(torch.zeros(0, dtype=torch.int32, device="cuda"), pf.PencilLayout((2, 2)), (0,))
    zpath = tmp_path / "zero.pio"
    pf.write_mpi_named(zpath, "zero", z)
    pf.read_mpi_named(zpath, "zero", z)
