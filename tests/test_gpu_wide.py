"""Fields whose 32 x-planes exceed the 2 GiB range of the buffer-addressed
loads (32*ny*nz*sizeof(Real) >= 2^31) take the 64-bit-address variant of the
field kernels; its records and decode are checked against the oracle on a
slab of blocks."""
import numpy as np
import pytest

from cubismz_amd import czw

pytestmark = pytest.mark.gpu


def _smooth_field(torch, shape, dtype):
    nx, ny, nz = shape
    x = torch.arange(nx, device="cuda", dtype=torch.float64).view(nx, 1, 1) / nx
    y = torch.arange(ny, device="cuda", dtype=torch.float64).view(1, ny, 1) / 97.0
    z = torch.arange(nz, device="cuda", dtype=torch.float64).view(1, 1, nz) / 131.0
    f = 1.0 + 0.3 * torch.tanh((torch.sin(y) + torch.cos(z) - 0.5 + x) * 6.0) + 0.01 * torch.sin(7 * y * z)
    return f.to(dtype)


@pytest.mark.parametrize("dt,shape", [("f32", (32, 4096, 4096)), ("f64", (32, 2048, 4096))])
def test_wide_field_matches_oracle(torch_gpu, oracle, dt, shape):
    torch = torch_gpu
    dtype = torch.float32 if dt == "f32" else torch.float64
    prec = 4 if dt == "f32" else 8
    assert 32 * shape[1] * shape[2] * prec >= 1 << 31  # the wide variant is the one under test
    print(f"[case] wide field {shape} {dt}", flush=True)
    f = _smooth_field(torch, shape, dtype)
    rec, off = czw.compress_field(f, 1e-4, 1, 0, True)
    out = czw.decompress_field(rec, off, shape, dtype, 1, True)
    torch.cuda.synchronize()
    nb = 256  # blocks [0, 256): iz in {0, 1}, every iy
    fh = f.cpu().numpy()
    slots, nbytes = oracle.compress_field(fh, 1e-4, 1, 0, True, 0, nb)
    o = off[: nb + 1].cpu().numpy()
    recs = rec[: int(o[-1])].cpu().numpy()
    for b in range(nb):
        ob = int(o[b])
        assert int(recs[ob:ob + 4].view(np.int32)[0]) == int(nbytes[b]), b
        assert np.array_equal(recs[ob + 4:ob + 4 + int(nbytes[b])], slots[b, :nbytes[b]]), b
    ref = oracle.decompress_field(slots, nbytes, shape, 1, dt, True, 0, nb)
    got = out[:, :, :64].cpu().numpy()
    iview = np.uint32 if dt == "f32" else np.uint64
    assert np.array_equal(got.view(iview), np.ascontiguousarray(ref[:, :, :64]).view(iview))
