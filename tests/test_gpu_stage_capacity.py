"""GPU parity at the compressors' LDS stage capacities.

The binary32 whole-block compressor stages a block's level-32 survivors in
LDS when there are at most 16848 of them and writes them to the slot
directly otherwise; the half-block compressor (binary64, and the binary32
batch tile) stages the survivors of the x < 16 half (level-32 codes 1, 4
and 5) when there are at most 8656.  Blocks with exactly capacity - 1,
capacity and capacity + 1 such survivors are built by choosing eps from the
oracle's mask, and both sides of each switch must match the oracle bit for
bit.
"""
import numpy as np
import pytest

from cubismz_amd import batch as B
from cubismz_amd import czw

pytestmark = pytest.mark.gpu

STAGE32 = 16848  # whole-block compressor: level-32 survivors staged in LDS
STAGE0 = 8656    # half-block compressor: x < 16 half's level-32 survivors staged in LDS


def _words(payload):
    return np.frombuffer(payload[:4096], dtype=np.uint32)


def _popc(a):
    return int(np.unpackbits(np.ascontiguousarray(a, dtype=np.uint32).view(np.uint8)).sum())


def _level32(payload):
    """Level-32 survivors: the mask's popcount outside the [0,16)^3 corner
    (mask word x + 32 y, bit z; the corner is x, y, z < 16)."""
    w = _words(payload).reshape(32, 32)  # [y][x]
    return _popc(w) - _popc(w[:16, :16] & 0xFFFF)


def _x_low_half(payload):
    """Level-32 survivors with x < 16 (codes 1, 4, 5): z >= 16 for y < 16, all z for y >= 16."""
    w = _words(payload).reshape(32, 32)
    return _popc(w[:16, :16] >> 16) + _popc(w[16:, :16])


def _eps_for(oracle, block, count, target):
    lo, hi = -8.0, 2.0  # log10 eps: many survivors at lo, few at hi
    for _ in range(100):
        mid = 0.5 * (lo + hi)
        eps = float(np.float32(10.0 ** mid))
        n = count(oracle.compress(block, eps, 1))
        if n == target:
            return eps
        if n > target:
            lo = mid
        else:
            hi = mid
    raise AssertionError(f"no eps gives {target} survivors")


def _cases(oracle, dtype, count, cap, seed):
    rng = np.random.default_rng(seed)
    out = []
    for target in (cap - 1, cap, cap + 1):
        blk = rng.standard_normal((32, 32, 32)).astype(dtype)
        out.append((blk, _eps_for(oracle, blk, count, target)))
    return out


def _compress_and_decode(torch, oracle, blk, eps, prec):
    """One block through compress_blocks and the field decoder, against the oracle."""
    rec, off = czw.compress_blocks(torch.from_numpy(blk[None]).cuda(), eps, 1)
    (p,) = czw.records_to_payloads(rec.cpu().numpy(), off.cpu().numpy())
    ref = oracle.compress(blk, eps, 1)
    assert bytes(p) == ref.tobytes()
    dt = torch.float32 if prec == "f32" else torch.float64
    ui = np.uint32 if prec == "f32" else np.uint64
    out = czw.decompress_field(rec, off, (32, 32, 32), dt, 1, False).cpu().numpy()
    want = oracle.decompress(ref, 1, prec)
    assert np.array_equal(out.transpose(2, 1, 0).view(ui), want.view(ui))


def test_whole_block_compressor_at_stage_capacity(torch_gpu, oracle):
    torch = torch_gpu
    for blk, eps in _cases(oracle, np.float32, _level32, STAGE32, seed=1):
        print(f"[case] f32 block, eps {eps:.6g}", flush=True)
        _compress_and_decode(torch, oracle, blk, eps, "f32")


def test_half_block_compressor_at_stage_capacity_f64(torch_gpu, oracle):
    torch = torch_gpu
    for blk, eps in _cases(oracle, np.float64, _x_low_half, STAGE0, seed=2):
        print(f"[case] f64 block, eps {eps:.6g}", flush=True)
        _compress_and_decode(torch, oracle, blk, eps, "f64")


def test_half_block_tile_batch_at_stage_capacity_f32(torch_gpu, oracle):
    torch = torch_gpu
    cases = _cases(oracle, np.float32, _x_low_half, STAGE0, seed=3)
    # one 32^3 field per case, HDF5 order f[x][y][z] = block[z][y][x]
    fs = [torch.from_numpy(np.ascontiguousarray(blk.transpose(2, 1, 0))).cuda() for blk, _ in cases]
    bt = B.Batch(fs, [eps for _, eps in cases], wtype=1)
    bt.set_launch(stage="compress", tile=B.TILE_HALF)
    assert bt.launch_shape["compress"]["tile"] == "half"
    bt.compress(check=True)
    torch.cuda.synchronize()
    for k, (blk, eps) in enumerate(cases):
        print(f"[case] f32 batch field {k}, half tile, eps {eps:.6g}", flush=True)
        rec, off = bt.field_records(k)
        (p,) = czw.records_to_payloads(rec.cpu().numpy(), off.cpu().numpy())
        ref = oracle.compress(blk, eps, 1)
        assert bytes(p) == ref.tobytes()
