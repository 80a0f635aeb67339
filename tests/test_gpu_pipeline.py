"""GPU tests of the block-range compressor and of the device-field .cz writer
that overlaps the host zlib stage with the device (czh_compress_device_field):
its file must be byte-identical to writing the whole field's records at once."""
import numpy as np
import pytest

from cubismz_amd import czfile, czw, fields

pytestmark = pytest.mark.gpu

SHAPE = (96, 64, 128)  # 3 x 2 x 4 = 24 blocks


@pytest.mark.parametrize("dt", ["f32", "f64"])
def test_range_equals_whole_field(torch_gpu, dt):
    torch = torch_gpu
    dtype = torch.float32 if dt == "f32" else torch.float64
    f = fields.bubble_field(SHAPE, seed=4, nbubbles=5, dtype=dtype, device="cuda")
    rec, off = czw.compress_field(f, 1e-4, 1, 4, True)
    o = off.cpu().numpy()
    r = rec[: int(o[-1])].cpu().numpy()
    for b0, n in ((0, 24), (5, 7), (13, 11), (23, 1)):
        rr, ro = czw.compress_field_range(f, 1e-4, 1, 4, True, b0, n)
        ro = ro.cpu().numpy()
        assert np.array_equal(ro, o[b0:b0 + n + 1] - o[b0]), (b0, n)
        assert np.array_equal(rr[: int(ro[-1])].cpu().numpy(), r[o[b0]:o[b0 + n]]), (b0, n)
    with pytest.raises(czw.CZWError):
        czw.compress_field_range(f, 1e-4, 1, 4, True, 20, 5)


@pytest.mark.parametrize("dt,slab", [("f32", 15), ("f32", 0), ("f64", 7), ("f64", 14)])
@pytest.mark.parametrize("overlap", [True, False])
def test_device_writer_equals_whole_field_write(torch_gpu, tmp_path, dt, slab, overlap):
    torch = torch_gpu
    dtype = torch.float32 if dt == "f32" else torch.float64
    prec = 4 if dt == "f32" else 8
    print(f"[case] device writer {dt} slab={slab} overlap={overlap}", flush=True)
    f = fields.bubble_field(SHAPE, seed=8, nbubbles=5, dtype=dtype, device="cuda")
    rec, off = czw.compress_field(f, 1e-4, 3, 0, False)
    o = off.cpu().numpy()
    ref = tmp_path / "ref.cz"
    czfile.write(str(ref), rec[: int(o[-1])].cpu().numpy(), o, (3, 2, 4), prec, 3, 1e-4)
    got = tmp_path / "got.cz"
    nbytes, times = czfile.compress_device_field(str(got), f, 1e-4, 3, slab_blocks=slab, overlap=overlap)
    assert nbytes == got.stat().st_size
    assert got.read_bytes() == ref.read_bytes()
    assert times["wall"] > 0 and times["device"] > 0


@pytest.mark.parametrize("dt", ["f32", "f64"])
@pytest.mark.parametrize("encoder", ["lz4", "none"])
def test_device_writer_other_encoders(torch_gpu, tmp_path, dt, encoder):
    """The overlapped writer with the LZ4 and the uncompressed second stage
    writes the bytes of the whole-field writer with that stage."""
    torch = torch_gpu
    dtype = torch.float32 if dt == "f32" else torch.float64
    prec = 4 if dt == "f32" else 8
    print(f"[case] device writer {dt} encoder={encoder}", flush=True)
    f = fields.bubble_field(SHAPE, seed=9, nbubbles=5, dtype=dtype, device="cuda")
    rec, off = czw.compress_field(f, 1e-4, 1, 0, False)
    o = off.cpu().numpy()
    r = rec[: int(o[-1])].cpu().numpy()
    ref = tmp_path / "ref.cz"
    czfile.write(str(ref), r, o, (3, 2, 4), prec, 1, 1e-4, encoder=encoder)
    got = tmp_path / "got.cz"
    czfile.compress_device_field(str(got), f, 1e-4, 1, encoder=encoder, slab_blocks=2 * czfile.entries_per_chunk(prec))
    assert got.read_bytes() == ref.read_bytes()
    with czfile.CzFile(str(got)) as cf:
        assert cf.encoder == encoder
        r2, o2 = cf.records()
    assert np.array_equal(r2, r) and np.array_equal(o2, o)


def test_device_writer_size_only_and_bad_path(torch_gpu, tmp_path):
    """path None reports the size the file would have; a path that cannot be
    opened is an error before any device work, and leaves no file."""
    torch = torch_gpu
    print("[case] device writer size-only / bad path", flush=True)
    f = fields.bubble_field(SHAPE, seed=10, nbubbles=5, dtype=torch.float32, device="cuda")
    got = tmp_path / "got.cz"
    nbytes, _ = czfile.compress_device_field(str(got), f, 1e-4, 1)
    size_only, _ = czfile.compress_device_field(None, f, 1e-4, 1)
    assert size_only == nbytes == got.stat().st_size
    bad = tmp_path / "missing_dir" / "x.cz"
    with pytest.raises(czfile.CzFileError):
        czfile.compress_device_field(str(bad), f, 1e-4, 1)
    assert not bad.exists()
