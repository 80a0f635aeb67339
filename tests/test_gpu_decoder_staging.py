"""GPU parity at the binary32 decoder's LDS staging capacity.

The decoder copies a block's level-32 survivors into LDS when there are at
most 1728 of them (1088 in a batch launch, whose field table takes the rest
of the LDS) and gathers them from the record otherwise.  Blocks with exactly
capacity - 1, capacity and capacity + 1 level-32 survivors are built by
choosing eps from the oracle's mask, and both sides of the switch must
decode bit for bit like the oracle.
"""
import numpy as np
import pytest

from cubismz_amd import batch as B
from cubismz_amd import czw

pytestmark = pytest.mark.gpu

CAP_FIELD, CAP_BATCH = 1728, 1088


def _level32_survivors(payload):
    """Level-32 survivors of a payload: the mask's popcount outside the
    [0,16)^3 corner (mask word x + 32 y, bit z; the corner words' low 16 bits)."""
    words = np.frombuffer(payload[:4096], dtype=np.uint32)
    total = int(np.unpackbits(words.view(np.uint8)).sum())
    corner = np.array([words[x + 32 * y] & 0xFFFF for y in range(16) for x in range(16)], dtype=np.uint32)
    return total - int(np.unpackbits(corner.view(np.uint8)).sum())


def _eps_for(oracle, block, target):
    """A float32 eps that leaves exactly `target` level-32 survivors in `block`
    ([z][y][x]), by bisection on log eps against the oracle's compressor."""
    lo, hi = -12.0, 4.0  # log10 eps: many survivors at lo, few at hi
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        eps = float(np.float32(10.0 ** mid))
        n = _level32_survivors(oracle.compress(block, eps, 1))
        if n == target:
            return eps
        if n > target:
            lo = mid
        else:
            hi = mid
    raise AssertionError(f"no eps gives {target} level-32 survivors")


def _blocks(oracle, targets, seed):
    rng = np.random.default_rng(seed)
    out = []
    for t in targets:
        blk = rng.standard_normal((32, 32, 32)).astype(np.float32)
        out.append((blk, _eps_for(oracle, blk, t)))
    return out


def _as_field(torch, blk):
    # one 32^3 block as an HDF5-ordered field: f[x][y][z] = block[z][y][x]
    return torch.from_numpy(np.ascontiguousarray(blk.transpose(2, 1, 0))).cuda()


@pytest.mark.parametrize("cap", [CAP_FIELD, CAP_BATCH])
def test_single_field_decoder_at_capacity(torch_gpu, oracle, cap):
    torch = torch_gpu
    for blk, eps in _blocks(oracle, (cap - 1, cap, cap + 1), seed=cap):
        print(f"[case] one block, eps {eps:.6g}", flush=True)
        f = _as_field(torch, blk)
        rec, off = czw.compress_field(f, eps, 1, 0, False)
        (p,) = czw.records_to_payloads(rec.cpu().numpy(), off.cpu().numpy())
        ref = oracle.compress(blk, eps, 1)
        assert bytes(p) == ref.tobytes()
        out = czw.decompress_field(rec, off, f.shape, torch.float32, 1, False).cpu().numpy()
        want = oracle.decompress(ref, 1, "f32")
        assert np.array_equal(out.transpose(2, 1, 0).view(np.uint32), want.view(np.uint32))


def test_batch_decoder_at_capacity(torch_gpu, oracle):
    torch = torch_gpu
    cases = _blocks(oracle, (CAP_BATCH - 1, CAP_BATCH, CAP_BATCH + 1, CAP_FIELD, CAP_FIELD + 1), seed=7)
    fs = [_as_field(torch, blk) for blk, _ in cases]
    bt = B.Batch(fs, [eps for _, eps in cases], wtype=1)
    bt.compress(check=True)
    bt.decompress()
    torch.cuda.synchronize()
    for k, (blk, eps) in enumerate(cases):
        print(f"[case] batch field {k}, eps {eps:.6g}", flush=True)
        rec, off = bt.field_records(k)
        (p,) = czw.records_to_payloads(rec.cpu().numpy(), off.cpu().numpy())
        ref = oracle.compress(blk, eps, 1)
        assert bytes(p) == ref.tobytes()
        want = oracle.decompress(ref, 1, "f32")
        assert np.array_equal(bt.outputs[k].cpu().numpy().transpose(2, 1, 0).view(np.uint32), want.view(np.uint32))
