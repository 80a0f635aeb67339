"""GPU tests of batched mode: one persistent launch per stage over the blocks
of many uneven fields, each with its own eps.  Every field's records and
reconstruction must equal, byte for byte, what the per-field entry points
give for it alone (those are pinned against the oracle elsewhere; one field
is re-checked against the oracle here)."""
import numpy as np
import pytest

from cubismz_amd import batch as B
from cubismz_amd import czw, fields

pytestmark = pytest.mark.gpu

SHAPES = [(64, 32, 96), (32, 32, 32), (96, 64, 32), (32, 96, 64), (64, 64, 64), (32, 32, 128), (128, 32, 32)]
EPS = [1e-3, 1e-4, 5e-3, 2e-4, 1e-2, 3e-3, 7e-4]


def _fields(torch, shapes=SHAPES, seed0=0):
    return [fields.bubble_field(s, seed=seed0 + k, nbubbles=5, dtype=torch.float32, device="cuda")
            for k, s in enumerate(shapes)]


def _check_against_per_field(torch, bt, fs, eps, wtype, shuffle, zerobits=0):
    torch.cuda.synchronize()
    for k, f in enumerate(fs):
        rec, off = czw.compress_field(f, eps[k], wtype, zerobits, shuffle)
        n = int(off[-1])
        brec, boff = bt.field_records(k)
        assert torch.equal(boff, off), k
        assert torch.equal(brec, rec[:n]), k
        dec = czw.decompress_field(rec, off, f.shape, torch.float32, wtype, shuffle)
        assert torch.equal(bt.outputs[k].view(torch.int32), dec.view(torch.int32)), k


@pytest.mark.parametrize("wtype,shuffle,zerobits", [(3, 0, 0), (1, 1, 0), (2, 0, 4)])
@pytest.mark.parametrize("shape", [(0, 0, 0), (0, 1, 0), (3, 1, 0), (5, 0, 0), (0, 0, 1), (0, 1, 1), (5, 1, 1)])
def test_batch_equals_per_field(torch_gpu, wtype, shuffle, zerobits, shape):
    """Both compress tiles (whole block, half block) in every launch shape."""
    torch = torch_gpu
    groups, queue, tile = shape
    print(f"[case] batch wtype={wtype} shuffle={shuffle} zb={zerobits} groups={groups} queue={queue} tile={tile}",
          flush=True)
    fs = _fields(torch)
    bt = B.Batch(fs, EPS, wtype=wtype, zerobits=zerobits, shuffle=shuffle)
    bt.set_launch(groups, queue, tile=tile)
    bt.compress()
    bt.decompress()
    _check_against_per_field(torch, bt, fs, EPS, wtype, shuffle, zerobits)


def test_batch_field_against_oracle(torch_gpu, oracle):
    torch = torch_gpu
    fs = _fields(torch, SHAPES[:3], seed0=11)
    bt = B.Batch(fs, EPS[:3], wtype=3)
    bt.set_launch(0, 1)
    bt.compress()
    bt.decompress()
    torch.cuda.synchronize()
    k = 2
    fh = fs[k].cpu().numpy()
    slots, nb = oracle.compress_field(fh, EPS[k], 3)
    rec, off = (t.cpu().numpy() for t in bt.field_records(k))
    for b in range(len(nb)):
        o = int(off[b])
        assert int(rec[o:o + 4].view(np.int32)[0]) == int(nb[b])
        assert np.array_equal(rec[o + 4:o + 4 + int(nb[b])], slots[b, :nb[b]])
    ref = oracle.decompress_field(slots, nb, fh.shape, 3, "f32")
    assert np.array_equal(bt.outputs[k].cpu().numpy().view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("tile", [B.TILE_BLOCK, B.TILE_HALF])
def test_batch_graph_replays(torch_gpu, tile):
    torch = torch_gpu
    print(f"[case] batch graph tile={tile}", flush=True)
    fs = _fields(torch, seed0=30)
    bt = B.Batch(fs, EPS, wtype=1)
    bt.set_launch(0, 1, tile=tile)
    bt.compress()
    bt.decompress()
    torch.cuda.synchronize()
    want_rec = bt.records[: int(bt.offsets[-1])].clone()
    want_out = [o.clone() for o in bt.outputs]
    g = bt.graph(("compress", "decompress"))
    try:
        for rep in range(3):
            bt.records.zero_()
            for o in bt.outputs:
                o.fill_(float("nan"))
            torch.cuda.synchronize()
            g.launch()
            torch.cuda.synchronize()
            assert int(bt.err.item()) == 0
            assert torch.equal(bt.records[: want_rec.numel()], want_rec), rep
            for k in range(len(fs)):
                assert torch.equal(bt.outputs[k].view(torch.int32), want_out[k].view(torch.int32)), (rep, k)
            # eager work between replays
            _ = sum(float(f.sum()) for f in fs)
    finally:
        g.close()


def test_batch_autotune_keeps_results(torch_gpu):
    torch = torch_gpu
    fs = _fields(torch, seed0=50)
    bt = B.Batch(fs, EPS, wtype=3)
    ms_c = bt.autotune("compress")
    shape_c = dict(bt.launch_shape["compress"])
    ms_d = bt.autotune("decompress")
    assert ms_c > 0 and ms_d > 0
    assert shape_c["groups"] > 0 and bt.launch_shape["decompress"]["groups"] > 0
    assert shape_c["tile"] in ("block", "half")
    # tuning one stage leaves the other stage's shape as it was
    assert bt.launch_shape["compress"] == shape_c
    bt.compress()
    bt.decompress()
    _check_against_per_field(torch, bt, fs, EPS, 3, 0)


def test_batch_corrupt_record_sets_error(torch_gpu):
    torch = torch_gpu
    fs = _fields(torch, SHAPES[:3], seed0=70)
    bt = B.Batch(fs, EPS[:3], wtype=3)
    bt.compress()
    torch.cuda.synchronize()
    b0 = int(bt.table[1]["block0"])
    o = int(bt.offsets[b0])
    bt.records[o:o + 4] = torch.tensor([255, 255, 255, 127], dtype=torch.uint8, device="cuda")
    with pytest.raises(czw.CZWError):
        bt.decompress()


@pytest.mark.parametrize("case", range(12))
def test_random_batches(torch_gpu, case):
    """Seeded random batches: 1..12 fields of random shapes and eps, one
    precision, random wavelet type, zerobits, shuffle and launch shape; each
    field equals its per-field compress and decode byte for byte."""
    torch = torch_gpu
    rng = np.random.default_rng(500 + case)
    nf = int(rng.integers(1, 13))
    dtype = torch.float64 if rng.random() < 0.3 else torch.float32
    shapes = [tuple(int(32 * rng.integers(1, 4)) for _ in range(3)) for _ in range(nf)]
    eps = [float(10 ** rng.uniform(-5, -2)) for _ in range(nf)]
    wtype, zerobits, shuffle = int(rng.integers(1, 4)), int(rng.choice([0, 4])), bool(rng.random() < 0.5)
    groups, queue = int(rng.choice([0, 1, 7])), int(rng.integers(0, 2))
    tile = int(rng.integers(0, 2))
    print(f"[case] random batch {case}: {nf} fields {dtype} wtype={wtype} zb={zerobits} sh={shuffle} "
          f"groups={groups} queue={queue} tile={tile}", flush=True)
    fs = [fields.bubble_field(s, seed=900 + 31 * case + k, nbubbles=int(rng.integers(1, 8)), dtype=dtype,
                              device="cuda") for k, s in enumerate(shapes)]
    bt = B.Batch(fs, eps, wtype=wtype, zerobits=zerobits, shuffle=shuffle)
    bt.set_launch(groups, queue, tile=tile)
    bt.compress()
    bt.decompress(check=True)
    torch.cuda.synchronize()
    iv = torch.int32 if dtype == torch.float32 else torch.int64
    for k, f in enumerate(fs):
        rec, off = czw.compress_field(f, eps[k], wtype, zerobits, shuffle)
        brec, boff = bt.field_records(k)
        assert torch.equal(boff, off) and torch.equal(brec, rec[: int(off[-1])]), k
        dec = czw.decompress_field(rec, off, f.shape, dtype, wtype, shuffle)
        assert torch.equal(bt.outputs[k].view(iv), dec.view(iv)), k


@pytest.mark.parametrize("tile", [B.TILE_BLOCK, B.TILE_HALF])
def test_batch_64_fields_of_the_benchmark_shapes(torch_gpu, oracle, tile):
    """A full table: 64 fields with edges from {64, ..., 384} and per-field
    eps log-uniform in [1e-4, 1e-2] (the batched benchmark's generator), every
    field equal to its per-field calls, one small field against the oracle."""
    torch = torch_gpu
    shapes, eps = fields.batched_shapes(64, seed=7)
    print(f"[case] 64-field batch, {sum(int(np.prod(s)) for s in shapes) // 32768} blocks, tile={tile}", flush=True)
    fs = [fields.bubble_field(s, seed=100 + k, nbubbles=fields.bubbles_for(s), dtype=torch.float32, device="cuda")
          for k, s in enumerate(shapes)]
    bt = B.Batch(fs, [float(e) for e in eps], wtype=1)
    bt.set_launch(0, 1, stage="compress", tile=tile)
    bt.compress()
    bt.decompress()
    _check_against_per_field(torch, bt, fs, [float(e) for e in eps], 1, False)
    k = int(np.argmin([np.prod(s) for s in shapes]))
    fh = fs[k].cpu().numpy()
    slots, nb = oracle.compress_field(fh, float(eps[k]), 1)
    rec, off = (t.cpu().numpy() for t in bt.field_records(k))
    for b in range(len(nb)):
        o = int(off[b])
        assert np.array_equal(rec[o + 4:o + 4 + int(nb[b])], slots[b, :nb[b]]), b
    ref = oracle.decompress_field(slots, nb, fh.shape, 1, "f32")
    assert np.array_equal(bt.outputs[k].cpu().numpy().view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("dtype_name,tile", [("float32", 0), ("float32", 1), ("float64", 0)])
@pytest.mark.parametrize("queue", [0, 1])
def test_batch_more_groups_than_resident(torch_gpu, dtype_name, tile, queue):
    """A launch shape far above the resident workgroups (4096) is clamped to
    them, for both stages and both binary32 tiles, in static and queue order."""
    torch = torch_gpu
    dtype = getattr(torch, dtype_name)
    shapes = SHAPES[:5]
    fs = [fields.bubble_field(s, seed=300 + k, nbubbles=5, dtype=dtype, device="cuda") for k, s in enumerate(shapes)]
    bt = B.Batch(fs, EPS[:5], wtype=3, shuffle=True)
    bt.set_launch(4096, queue, tile=tile)
    print(f"[case] batch {dtype_name} groups=4096 queue={queue} tile={tile}", flush=True)
    bt.compress()
    bt.decompress()
    torch.cuda.synchronize()
    iv = torch.int32 if dtype == torch.float32 else torch.int64
    for k, f in enumerate(fs):
        rec, off = czw.compress_field(f, EPS[k], 3, 0, True)
        brec, boff = bt.field_records(k)
        assert torch.equal(boff, off) and torch.equal(brec, rec[: int(off[-1])]), k
        dec = czw.decompress_field(rec, off, f.shape, dtype, 3, True)
        assert torch.equal(bt.outputs[k].view(iv), dec.view(iv)), k


def test_batch_f64_autotuned_both_stages(torch_gpu):
    torch = torch_gpu
    shapes = [(64, 32, 96), (32, 32, 32), (96, 64, 32)]
    eps = [1e-5, 1e-7, 1e-4]
    fs = [fields.bubble_field(s, seed=60 + k, nbubbles=4, dtype=torch.float64, device="cuda")
          for k, s in enumerate(shapes)]
    bt = B.Batch(fs, eps, wtype=2)
    bt.autotune("compress")
    shape_c = dict(bt.launch_shape["compress"])
    bt.autotune("decompress")
    assert bt.launch_shape["compress"] == shape_c
    bt.compress()
    bt.decompress()
    torch.cuda.synchronize()
    for k, f in enumerate(fs):
        rec, off = czw.compress_field(f, eps[k], 2, 0, False)
        brec, boff = bt.field_records(k)
        assert torch.equal(boff, off) and torch.equal(brec, rec[: int(off[-1])]), k
        dec = czw.decompress_field(rec, off, f.shape, torch.float64, 2, False)
        assert torch.equal(bt.outputs[k].view(torch.int64), dec.view(torch.int64)), k
