"""czw_compress_field_lowmem: the binary32 compressor that frames its own
records (record offsets by device look-back) with a small scratch.  Its
records and offsets must equal czw_compress_field's byte for byte (that path
is pinned against the oracle in test_gpu_parity.py); one case is checked
against the oracle here as well.  The look-back's bounded wait is forced to
give up from the host (a re-read limit of 1, on a field whose every 16th
block is dense and slow), and must end with the error bit set and the kernel
finished, never a hang."""
import numpy as np
import pytest

from cubismz_amd import czw, fields

pytestmark = pytest.mark.gpu


def _smooth_plus_noise(torch, shape, seed, noise_frac):
    """A cheap device field: smooth waves, with uniform noise (dense blocks,
    most level-32 details kept) in the first noise_frac of the X extent."""
    g = torch.Generator(device="cuda").manual_seed(seed)
    nx, ny, nz = shape
    x = torch.arange(nx, device="cuda", dtype=torch.float32).view(-1, 1, 1)
    y = torch.arange(ny, device="cuda", dtype=torch.float32).view(1, -1, 1)
    z = torch.arange(nz, device="cuda", dtype=torch.float32).view(1, 1, -1)
    f = 1.0 + 0.3 * torch.sin(x * 0.05 + y * 0.031) * torch.cos(z * 0.043) + 0.01 * torch.sin(x * 0.7 + z * 0.9)
    n = int(nx * noise_frac) // 32 * 32
    if n:
        f[:n] += 0.5 * torch.rand((n, ny, nz), device="cuda", generator=g)
    return f.contiguous()


def _same(torch, a, b):
    ra, oa = a
    rb, ob = b
    assert torch.equal(oa, ob)
    n = int(oa[-1])
    assert torch.equal(ra[:n], rb[:n])


@pytest.mark.parametrize("shape,wtype,eps,zb,sh", [
    ((64, 64, 96), 1, 1e-3, 0, False),     # 24 blocks
    ((64, 64, 96), 3, 1e-4, 4, True),
    ((96, 32, 64), 2, 0.0, 0, False),      # eps 0: every coefficient survives (overflow path)
    ((512, 512, 512), 1, 1e-3, 0, False),  # 4096 blocks
])
def test_lowmem_equals_two_stage(torch_gpu, shape, wtype, eps, zb, sh):
    torch = torch_gpu
    print(f"[case] lowmem {shape} wtype={wtype} eps={eps} zb={zb} sh={sh}", flush=True)
    f = fields.bubble_field(shape, seed=3, nbubbles=8, dtype=torch.float32, device="cuda") if shape[0] < 512 else \
        _smooth_plus_noise(torch, shape, 1, 0.125)
    want = czw.compress_field(f, eps, wtype, zb, sh)
    got = czw.compress_field_lowmem(f, eps, wtype, zb, sh)
    _same(torch, got, want)


def test_lowmem_32768_blocks(torch_gpu):
    """1024^3 float32 (32768 blocks, more than the look-back reads at once
    and more blocks than workgroups by far), dense noise in the first 1/16."""
    torch = torch_gpu
    print("[case] lowmem 1024^3", flush=True)
    f = _smooth_plus_noise(torch, (1024, 1024, 1024), 2, 1.0 / 16)
    want = czw.compress_field(f, 1e-3, 1, 0, False)
    got = czw.compress_field_lowmem(f, 1e-3, 1, 0, False)
    _same(torch, got, want)
    del want, got, f
    torch.cuda.empty_cache()


def test_lowmem_against_oracle(torch_gpu, oracle):
    torch = torch_gpu
    fh = fields.bubble_field((64, 96, 64), seed=12, nbubbles=6, dtype=torch.float32)
    rec, off = czw.compress_field_lowmem(fh.cuda(), 1e-3, 3, 4, False)
    got = czw.records_to_payloads(rec.cpu().numpy(), off.cpu().numpy())
    slots, nb = oracle.compress_field(fh.numpy(), 1e-3, 3, 4)
    for b, p in enumerate(got):
        assert p == slots[b, :nb[b]].tobytes(), b


def test_lowmem_refuses_f64_and_small_scratch(torch_gpu):
    torch = torch_gpu
    f = torch.ones((32, 32, 32), dtype=torch.float64, device="cuda")
    with pytest.raises(czw.CZWError):
        czw.compress_field_lowmem(f, 1e-3)
    with pytest.raises(czw.CZWError):
        czw.compress_field_lowmem(f.float(), 1e-3, scratch=torch.empty(64, dtype=torch.uint8, device="cuda"))


def test_lookback_give_up(torch_gpu):
    """With one re-read allowed, blocks whose predecessors are still being
    compressed give up: the call ends (no hang), the error word says so, and
    the next call with the default limit is correct again.  A block is framed
    two of its workgroup's blocks after it was compressed, by when its
    predecessors normally have their sizes published; the dense blocks of
    the first x-slab (every 16th block of the queue, all survivors kept)
    take longer than that."""
    torch = torch_gpu
    f = _smooth_plus_noise(torch, (512, 512, 512), 4, 1.0 / 16)
    want = czw.compress_field(f, 1e-3, 1, 0, False)
    scratch = torch.empty(czw.lowmem_bytes(4096, 4), dtype=torch.uint8, device="cuda")
    prev = czw.set_lookback_limit(1)
    errs = 0
    try:
        for rep in range(4):
            print(f"[case] look-back limit 1, run {rep}", flush=True)
            rec, off = czw.compress_field_lowmem(f, 1e-3, 1, 0, False, scratch=scratch, check=False)
            torch.cuda.synchronize()
            e = czw.workspace_errors(scratch)
            if e:
                assert e == czw.ERR_LOOKBACK
                o = off.cpu().numpy().view(np.uint64)
                assert (o == np.uint64(czw.OFFSET_INVALID)).any()
                errs += 1
            else:  # every wait happened to end in time: the output is the normal one
                _same(torch, (rec, off), want)
    finally:
        czw.set_lookback_limit(prev)
    assert errs > 0
    _same(torch, czw.compress_field_lowmem(f, 1e-3, 1, 0, False, scratch=scratch), want)


PARK = 2880  # survivors of a block that can be parked in LDS between its compute and its record


def _survivors(payload):
    return int(np.unpackbits(np.frombuffer(payload[:4096], dtype=np.uint8)).sum())


@pytest.mark.parametrize("target", [PARK - 1, PARK, PARK + 1])
def test_parked_block_capacity(torch_gpu, oracle, target):
    """Blocks with exactly PARK - 1, PARK and PARK + 1 survivors (eps chosen
    from the oracle's mask), 640 copies of one block in a field so that every
    workgroup compresses several and parks all but its last (or, above the
    capacity, looks each up at once): records equal the two-stage path's,
    and the first one the oracle's payload."""
    torch = torch_gpu
    rng = np.random.default_rng(target)
    blk = rng.standard_normal((32, 32, 32)).astype(np.float32)
    lo, hi = -8.0, 2.0  # log10 eps: many survivors at lo, few at hi
    for _ in range(100):
        mid = 0.5 * (lo + hi)
        eps = float(np.float32(10.0 ** mid))
        n = _survivors(oracle.compress(blk, eps, 1))
        if n == target:
            break
        if n > target:
            lo = mid
        else:
            hi = mid
    else:
        raise AssertionError(f"no eps gives {target} survivors")
    print(f"[case] 640 copies of a block with {target} survivors, eps {eps:.6g}", flush=True)
    one = np.ascontiguousarray(blk.transpose(2, 1, 0))  # f[x][y][z] = blk[z][y][x]
    f = torch.from_numpy(np.tile(one, (8, 8, 10))).cuda()
    a = czw.compress_field_lowmem(f, eps, 1)
    b = czw.compress_field(f, eps, 1)
    torch.cuda.synchronize()
    _same(torch, a, b)
    rec, off = a
    first = czw.records_to_payloads(rec.cpu().numpy(), off.cpu().numpy())[0]
    assert bytes(first) == oracle.compress(blk, eps, 1).tobytes()
