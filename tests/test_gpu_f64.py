"""GPU parity of the binary64 kernels (the reference's double build) against
the oracle, bit for bit: payloads (mask + survivors after zerobits/shuffle,
sizeof(Real) = 8) and decoded fields."""
import numpy as np
import pytest

from cubismz_amd import batch as B
from cubismz_amd import czw, fields

pytestmark = pytest.mark.gpu

SHAPE = (64, 96, 128)


def _field(torch, seed, shape=SHAPE, kind="bubble"):
    if kind == "bubble":
        return fields.bubble_field(shape, seed=seed, nbubbles=6, dtype=torch.float64)
    return fields.bumps_field(shape, seed=seed, dtype=torch.float64)


def _payloads(records, offsets):
    return czw.records_to_payloads(records.cpu().numpy(), offsets.cpu().numpy())


def _check_records(rec, off, slots, nbytes):
    got = _payloads(rec, off)
    assert len(got) == len(nbytes)
    for b, p in enumerate(got):
        assert len(p) == nbytes[b], f"block {b}: size {len(p)} vs {nbytes[b]}"
        assert p == slots[b, :nbytes[b]].tobytes(), f"block {b}: payload differs"


@pytest.mark.parametrize("wtype", [1, 2, 3])
@pytest.mark.parametrize("zerobits,shuffle", [(0, False), (4, False), (0, True), (16, True)])
@pytest.mark.parametrize("eps", [1e-3, 1e-6, 1e-8])  # 1e-8: the 1024^3 scale-up's densest setting
def test_field_compress_and_decode_bitexact_f64(torch_gpu, oracle, wtype, zerobits, shuffle, eps):
    torch = torch_gpu
    print(f"[case] f64 field {SHAPE} wtype={wtype} zb={zerobits} sh={shuffle} eps={eps}", flush=True)
    f = _field(torch, seed=11 + wtype)
    rec, off = czw.compress_field(f.cuda(), eps, wtype, zerobits, shuffle)
    torch.cuda.synchronize()
    slots, nbytes = oracle.compress_field(f.numpy(), eps, wtype, zerobits, shuffle)
    _check_records(rec, off, slots, nbytes)
    out = czw.decompress_field(rec, off, SHAPE, torch.float64, wtype, shuffle).cpu().numpy()
    ref = oracle.decompress_field(slots, nbytes, SHAPE, wtype, "f64", shuffle)
    assert np.array_equal(out.view(np.uint64), ref.view(np.uint64))


@pytest.mark.parametrize("wtype", [1, 3])
@pytest.mark.parametrize("shuffle", [False, True])
def test_blocks_and_dense_f64(torch_gpu, oracle, wtype, shuffle):
    """Block layout; noisy blocks keep (nearly) every coefficient, smooth ones few."""
    torch = torch_gpu
    print(f"[case] f64 blocks wtype={wtype} sh={shuffle}", flush=True)
    rng = np.random.default_rng(5)
    blocks = (rng.standard_normal((6, 32, 32, 32)) * np.logspace(-6, 0, 32)[None, :, None, None])
    blocks[3:] = np.linspace(0, 1, 32)[None, None, None, :] ** 2 + 0 * blocks[3:]
    for eps in (1e-4, 0.0):
        rec, off = czw.compress_blocks(torch.from_numpy(blocks).cuda(), eps, wtype, 0, shuffle)
        got = _payloads(rec, off)
        for b in range(6):
            assert got[b] == oracle.compress(blocks[b], eps, wtype, 0, shuffle).tobytes(), (eps, b)
        out = czw.decompress_blocks(rec, off, 6, torch.float64, wtype, shuffle).cpu().numpy()
        for b in range(6):
            ref = oracle.decompress(got[b], wtype, "f64", shuffle)
            assert np.array_equal(out[b].view(np.uint64), ref.view(np.uint64)), (eps, b)


def test_eps_zero_f64_is_fwt_iwt(torch_gpu, oracle):
    torch = torch_gpu
    rng = np.random.default_rng(9)
    blocks = rng.standard_normal((3, 32, 32, 32))
    rec, off = czw.compress_blocks(torch.from_numpy(blocks).cuda(), 0.0, 3)
    got = _payloads(rec, off)
    out = czw.decompress_blocks(rec, off, 3, torch.float64, 3).cpu().numpy()
    for b in range(3):
        coef = oracle.fwt(blocks[b], 3)
        assert len(got[b]) == 4096 + 8 * int(np.count_nonzero(coef))
        assert np.array_equal(out[b], oracle.iwt(coef, 3))


def test_corrupt_record_sets_error_f64(torch_gpu):
    torch = torch_gpu
    f = _field(torch, seed=3)
    rec, off = czw.compress_field(f.cuda(), 1e-4, 3)
    bad = rec.clone()
    o = int(off[1].item())
    bad[o:o + 4] = torch.tensor([4, 16, 0, 0], dtype=torch.uint8)  # 4096 + 4: not a whole double
    with pytest.raises(czw.CZWError):
        czw.decompress_field(bad, off, SHAPE, torch.float64, 3)
    bad = rec.clone()
    bad[o + 4 + 100] ^= 0xFF
    with pytest.raises(czw.CZWError):
        czw.decompress_field(bad, off, SHAPE, torch.float64, 3)


@pytest.mark.parametrize("wtype", [1, 2, 3])
def test_f64_against_compiled_reference(torch_gpu, oracle, wtype):
    torch = torch_gpu
    if not oracle.ref_available("f64", 4, True):
        pytest.skip("oracle/_ref not built")
    f = _field(torch, seed=31)
    rec, off = czw.compress_field(f.cuda(), 1e-5, wtype, 4, True)
    got = _payloads(rec, off)
    fn = f.numpy()
    for b in (0, 7, 23):
        ix, iy, iz = b % 2, (b // 2) % 3, b // 6
        blk = np.ascontiguousarray(fn[32 * ix:32 * ix + 32, 32 * iy:32 * iy + 32, 32 * iz:32 * iz + 32].transpose(2, 1, 0))
        assert got[b] == oracle.ref_compress(blk, 1e-5, wtype, 4, True).tobytes()


@pytest.mark.parametrize("queue", [0, 1])
def test_batch_f64_equals_per_field(torch_gpu, queue):
    torch = torch_gpu
    shapes = [(64, 32, 96), (32, 32, 32), (96, 64, 32)]
    eps = [1e-5, 1e-7, 1e-4]
    fs = [fields.bubble_field(s, seed=40 + k, nbubbles=4, dtype=torch.float64, device="cuda")
          for k, s in enumerate(shapes)]
    bt = B.Batch(fs, eps, wtype=1, shuffle=True)
    bt.set_launch(3 if queue else 0, queue)
    bt.compress()
    bt.decompress()
    torch.cuda.synchronize()
    for k, f in enumerate(fs):
        rec, off = czw.compress_field(f, eps[k], 1, 0, True)
        brec, boff = bt.field_records(k)
        assert torch.equal(boff, off)
        assert torch.equal(brec, rec[: int(off[-1])])
        dec = czw.decompress_field(rec, off, f.shape, torch.float64, 1, True)
        assert torch.equal(bt.outputs[k].view(torch.int64), dec.view(torch.int64))
