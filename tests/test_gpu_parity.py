"""GPU parity: the HIP compressor against the oracle, bit for bit.

Compress: every record payload (mask + survivors, after zerobits/shuffle)
must equal the oracle's payload for the same block.  Decompress: the decoded
field must equal the oracle's decoded field bit for bit.
"""
import numpy as np
import pytest

from cubismz_amd import czw, fields

pytestmark = pytest.mark.gpu

SHAPE = (64, 96, 128)  # 2 x 3 x 4 = 24 blocks, non-cubic


def _field(torch, seed, shape=SHAPE, kind="bubble", dtype=None):
    dtype = dtype or torch.float32
    if kind == "bubble":
        nb = max(4, int(200 * np.prod(shape) / 512 ** 3 * 8))
        return fields.bubble_field(shape, seed=seed, nbubbles=nb, dtype=dtype)
    return fields.bumps_field(shape, seed=seed, dtype=dtype)


def _payloads(records, offsets):
    return czw.records_to_payloads(records.cpu().numpy(), offsets.cpu().numpy())


@pytest.mark.parametrize("wtype", [1, 2, 3])
@pytest.mark.parametrize("zerobits,shuffle", [(0, False), (4, False), (0, True), (4, True)])
@pytest.mark.parametrize("eps", [1e-2, 1e-3, 1e-4])
def test_field_compress_bitexact(torch_gpu, oracle, wtype, zerobits, shuffle, eps):
    torch = torch_gpu
    print(f"[case] field {SHAPE} wtype={wtype} zb={zerobits} sh={shuffle} eps={eps}", flush=True)
    f = _field(torch, seed=11 + wtype)
    rec, off = czw.compress_field(f.cuda(), eps, wtype, zerobits, shuffle)
    torch.cuda.synchronize()
    got = _payloads(rec, off)
    slots, nbytes = oracle.compress_field(f.numpy(), eps, wtype, zerobits, shuffle)
    assert len(got) == len(nbytes)
    for b, p in enumerate(got):
        assert len(p) == nbytes[b], f"block {b}: size {len(p)} vs {nbytes[b]}"
        assert p == slots[b, :nbytes[b]].tobytes(), f"block {b}: payload differs"
    # records are contiguous and framed with their sizes
    o = off.cpu().numpy()
    assert o[0] == 0 and np.all(np.diff(o) == 4 + nbytes.astype(np.int64))


@pytest.mark.parametrize("wtype", [1, 2, 3])
@pytest.mark.parametrize("shuffle", [False, True])
def test_field_roundtrip_bitexact(torch_gpu, oracle, wtype, shuffle):
    torch = torch_gpu
    print(f"[case] roundtrip {SHAPE} wtype={wtype} sh={shuffle}", flush=True)
    f = _field(torch, seed=21 + wtype, kind="bumps")
    for eps, zb in ((1e-3, 4), (1e-4, 0)):
        rec, off = czw.compress_field(f.cuda(), eps, wtype, zb, shuffle)
        out = czw.decompress_field(rec, off, SHAPE, torch.float32, wtype, shuffle).cpu().numpy()
        slots, nbytes = oracle.compress_field(f.numpy(), eps, wtype, zb, shuffle)
        ref = oracle.decompress_field(slots, nbytes, SHAPE, wtype, "f32", shuffle)
        assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("wtype", [1, 3])
def test_blocks_layout(torch_gpu, oracle, wtype):
    torch = torch_gpu
    print(f"[case] blocks wtype={wtype}", flush=True)
    rng = np.random.default_rng(5)
    blocks = (rng.standard_normal((6, 32, 32, 32)) * np.logspace(-3, 0, 32)[None, :, None, None]).astype(np.float32)
    rec, off = czw.compress_blocks(torch.from_numpy(blocks).cuda(), 1e-2, wtype, 0, False)
    got = _payloads(rec, off)
    for b in range(6):
        assert got[b] == oracle.compress(blocks[b], 1e-2, wtype).tobytes()
    out = czw.decompress_blocks(rec, off, 6, torch.float32, wtype).cpu().numpy()
    for b in range(6):
        ref = oracle.decompress(got[b], wtype, "f32")
        assert np.array_equal(out[b].view(np.uint32), ref.view(np.uint32))


def test_eps_zero_keeps_every_nonzero(torch_gpu, oracle):
    """eps = 0: every non-zero coefficient survives and decode equals iwt(fwt(x))."""
    torch = torch_gpu
    rng = np.random.default_rng(9)
    blocks = rng.standard_normal((3, 32, 32, 32)).astype(np.float32)
    rec, off = czw.compress_blocks(torch.from_numpy(blocks).cuda(), 0.0, 3)
    got = _payloads(rec, off)
    out = czw.decompress_blocks(rec, off, 3, torch.float32, 3).cpu().numpy()
    for b in range(3):
        coef = oracle.fwt(blocks[b], 3)
        assert len(got[b]) == 4096 + 4 * int(np.count_nonzero(coef))
        assert np.array_equal(out[b], oracle.iwt(coef, 3))


def test_corrupt_record_sets_error(torch_gpu):
    torch = torch_gpu
    f = _field(torch, seed=3)
    rec, off = czw.compress_field(f.cuda(), 1e-3, 3)
    bad = rec.clone()
    o = int(off[1].item())
    bad[o:o + 4] = torch.tensor([1, 2, 0, 0], dtype=torch.uint8)  # nbytes not 4096 + 4k
    with pytest.raises(czw.CZWError):
        czw.decompress_field(bad, off, SHAPE, torch.float32, 3)
    bad = rec.clone()
    bad[o + 4 + 100] ^= 0xFF  # flip mask bits: popcount no longer matches
    with pytest.raises(czw.CZWError):
        czw.decompress_field(bad, off, SHAPE, torch.float32, 3)
    for b, shift in ((1, 1), (2, rec.numel())):  # a misaligned record, one past the end
        badoff = off.clone()
        badoff[b] += shift
        with pytest.raises(czw.CZWError):
            czw.decompress_field(rec, badoff, SHAPE, torch.float32, 3)
    czw.decompress_field(rec, off, SHAPE, torch.float32, 3)  # the good stream still decodes


@pytest.mark.parametrize("wtype", [1, 2, 3])
def test_against_compiled_reference(torch_gpu, oracle, wtype):
    """Where oracle/_ref was built, the GPU payloads equal the reference's own."""
    torch = torch_gpu
    if not oracle.ref_available("f32", 4, True):
        pytest.skip("oracle/_ref not built")
    f = _field(torch, seed=31)
    rec, off = czw.compress_field(f.cuda(), 1e-3, wtype, 4, True)
    got = _payloads(rec, off)
    fn = f.numpy()
    for b in (0, 7, 23):
        ix, iy, iz = b % 2, (b // 2) % 3, b // 6
        blk = np.ascontiguousarray(fn[32 * ix:32 * ix + 32, 32 * iy:32 * iy + 32, 32 * iz:32 * iz + 32].transpose(2, 1, 0))
        assert got[b] == oracle.ref_compress(blk, 1e-3, wtype, 4, True).tobytes()


@pytest.mark.parametrize("wtype", [1, 3])
@pytest.mark.parametrize("shuffle", [False, True])
def test_dense_and_mixed_blocks(torch_gpu, oracle, wtype, shuffle):
    """Blocks whose survivors overflow the on-chip stages take the dense code
    paths of both kernels; a field mixing dense noise and smooth regions
    exercises both paths in one launch."""
    torch = torch_gpu
    print(f"[case] dense wtype={wtype} sh={shuffle}", flush=True)
    rng = np.random.default_rng(17)
    f = fields.bumps_field(SHAPE, seed=5).numpy().copy()
    f[:32] += rng.standard_normal((32,) + SHAPE[1:]).astype(np.float32)  # blocks with ix = 0 become dense
    ft = torch.from_numpy(f)
    for eps in (1e-4, 0.0):
        rec, off = czw.compress_field(ft.cuda(), eps, wtype, 4, shuffle)
        got = _payloads(rec, off)
        slots, nbytes = oracle.compress_field(f, eps, wtype, 4, shuffle)
        assert max(nbytes) > 4096 + 4 * 16848  # at least one block past the stage capacity
        for b, p in enumerate(got):
            assert p == slots[b, :nbytes[b]].tobytes(), f"eps={eps} block {b}"
        out = czw.decompress_field(rec, off, SHAPE, torch.float32, wtype, shuffle).cpu().numpy()
        ref = oracle.decompress_field(slots, nbytes, SHAPE, wtype, "f32", shuffle)
        assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("wtype", [1, 2, 3])
def test_special_values_match_oracle(torch_gpu, oracle, prec, wtype):
    """Blocks of subnormals, signed zeros, overflowing predictors,
    infinities and NaNs beside ordinary ones: the same masks, survivors and
    decoded values as the oracle (NaN payloads aside; the oracle is checked
    against the compiled reference on the same blocks)."""
    from tests.test_oracle import nan_aware_equal, special_block

    torch = torch_gpu
    print(f"[case] special values {prec} wtype={wtype}", flush=True)
    dt = np.float32 if prec == "f32" else np.float64
    rng = np.random.default_rng(21)
    f = _field(torch, seed=4, kind="bumps", dtype=torch.float32 if prec == "f32" else torch.float64).numpy().copy()
    kinds = ["subnormal", "zeros", "huge", "inf", "nan"]
    for b, kind in enumerate(kinds * 2):  # blocks 0..9 special, the rest ordinary
        ix, iy, iz = b % 2, (b // 2) % 3, b // 6
        f[32 * ix:32 * ix + 32, 32 * iy:32 * iy + 32, 32 * iz:32 * iz + 32] = special_block(dt, rng, kind)
    for eps in (1e-3, 0.0):
        rec, off = czw.compress_field(torch.from_numpy(f).cuda(), eps, wtype, 0, False)
        got = _payloads(rec, off)
        slots, nbytes = oracle.compress_field(f, eps, wtype, 0, False)
        for b, p in enumerate(got):
            ref = slots[b, :nbytes[b]].tobytes()
            assert len(p) == len(ref) and p[:4096] == ref[:4096], f"eps={eps} block {b}"
            assert nan_aware_equal(np.frombuffer(p[4096:], np.uint8), np.frombuffer(ref[4096:], np.uint8), dt), \
                f"eps={eps} block {b}"
        out = czw.decompress_field(rec, off, SHAPE, torch.float32 if prec == "f32" else torch.float64, wtype,
                                   False).cpu().numpy()
        ref = oracle.decompress_field(slots, nbytes, SHAPE, wtype, prec, False)
        assert nan_aware_equal(out, ref, dt), f"eps={eps}"


@pytest.mark.parametrize("wtype", [1, 3])
def test_extreme_thresholds(torch_gpu, oracle, wtype):
    """eps above every coefficient keeps only the 4^3 base of each block; a
    negative eps keeps every coefficient, zeros included."""
    torch = torch_gpu
    print(f"[case] extreme eps wtype={wtype}", flush=True)
    f = _field(torch, seed=6)
    for eps, S in ((1e30, 64), (float("inf"), 64), (-1.0, 32768)):
        rec, off = czw.compress_field(f.cuda(), eps, wtype, 0, False)
        got = _payloads(rec, off)
        slots, nbytes = oracle.compress_field(f.numpy(), eps, wtype, 0, False)
        assert all(len(p) == 4096 + 4 * S for p in got), eps
        for b, p in enumerate(got):
            assert p == slots[b, :nbytes[b]].tobytes(), f"eps={eps} block {b}"
        out = czw.decompress_field(rec, off, SHAPE, torch.float32, wtype, False).cpu().numpy()
        ref = oracle.decompress_field(slots, nbytes, SHAPE, wtype, "f32", False)
        assert np.array_equal(out.view(np.uint32), ref.view(np.uint32)), eps


def test_buffer_validation(torch_gpu):
    """Caller buffers of the wrong size, dtype, shape or device are refused
    before any launch."""
    torch = torch_gpu
    print("[case] buffer validation", flush=True)
    f = _field(torch, seed=2).cuda()
    nb = 24
    with pytest.raises(czw.CZWError):  # records on the host
        czw.compress_field(f, 1e-3, 3, records=torch.empty(czw.records_bound(nb, 4), dtype=torch.uint8))
    with pytest.raises(czw.CZWError):  # offsets one short
        czw.compress_field(f, 1e-3, 3, offsets=torch.empty(nb, dtype=torch.int64, device="cuda"))
    with pytest.raises(czw.CZWError):  # workspace too small
        czw.compress_field(f, 1e-3, 3, workspace=torch.empty(16, dtype=torch.uint8, device="cuda"))
    rec, off = czw.compress_field(f, 1e-3, 3)
    with pytest.raises(czw.CZWError):  # output of another shape
        czw.decompress_field(rec, off, SHAPE, torch.float32, 3, out=torch.empty((64, 96, 96), device="cuda"))
    with pytest.raises(czw.CZWError):  # output of another dtype
        czw.decompress_field(rec, off, SHAPE, torch.float32, 3, out=torch.empty(SHAPE, dtype=torch.float64,
                                                                                  device="cuda"))
    with pytest.raises(czw.CZWError):  # offsets too short for the field
        czw.decompress_field(rec, off[:10], SHAPE, torch.float32, 3)
    with pytest.raises(czw.CZWError):  # size array too short for the slots stage
        czw.compress_field_slots(f, 1e-3, 3, 0, torch.empty(nb * czw.slot_bytes(4), dtype=torch.uint8, device="cuda"),
                                 torch.empty(nb, dtype=torch.int32, device="cuda"))
    out = czw.decompress_field(rec, off, SHAPE, torch.float32, 3)
    assert out.shape == SHAPE


@pytest.mark.parametrize("prec", [4, 8])
def test_two_stage_equals_one_call(torch_gpu, prec):
    """compress_field_slots + pack_records (the two stages timed apart) give
    the records of compress_field."""
    torch = torch_gpu
    print(f"[case] two-stage prec={prec}", flush=True)
    f = _field(torch, seed=8, dtype=torch.float32 if prec == 4 else torch.float64).cuda()
    nb = 24
    rec, off = czw.compress_field(f, 1e-4, 1, 0, True)
    slots = torch.empty(nb * czw.slot_bytes(prec), dtype=torch.uint8, device="cuda")
    nbytes = torch.empty(czw.size_words(prec) * nb, dtype=torch.int32, device="cuda")
    rec2 = torch.empty(czw.records_bound(nb, prec), dtype=torch.uint8, device="cuda")
    off2 = torch.empty(nb + 1, dtype=torch.int64, device="cuda")
    czw.compress_field_slots(f, 1e-4, 1, 0, slots, nbytes)
    czw.pack_records(slots, czw.slot_bytes(prec), nbytes, nb, True, rec2, off2)
    assert torch.equal(off, off2) and torch.equal(rec[: int(off[-1])], rec2[: int(off2[-1])])
