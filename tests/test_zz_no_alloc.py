"""czw_compress_field_noalloc (the allocation-free variant of
czw_compress_field_lowmem, scratch in a fixed device array of the library)
against czw_compress_field_lowmem through the C ABI: two calls in a row, the
second with a smaller field, records and offsets equal byte for byte."""
import pytest

from cubismz_amd import czw, fields

pytestmark = pytest.mark.gpu


def _check(torch, f, eps, wtype, zb, sh):
    nb = f.numel() // 32768
    rec = torch.empty(czw.records_bound(nb, 4), dtype=torch.uint8, device="cuda")
    off = torch.empty(nb + 1, dtype=torch.int64, device="cuda")
    czw.compress_field_noalloc(f, eps, wtype, zb, sh, rec, off)
    want_rec, want_off = czw.compress_field_lowmem(f, eps, wtype, zb, sh)
    assert torch.equal(off, want_off)
    n = int(off[-1])
    assert torch.equal(rec[:n], want_rec[:n])


def test_noalloc_equals_lowmem_twice(torch_gpu):
    torch = torch_gpu
    big = fields.bubble_field((256, 256, 192), seed=21, nbubbles=20, dtype=torch.float32, device="cuda")
    small = fields.bubble_field((64, 32, 96), seed=22, nbubbles=4, dtype=torch.float32, device="cuda")
    print("[case] noalloc 256x256x192 then 64x32x96", flush=True)
    _check(torch, big, 1e-3, 1, 0, False)
    _check(torch, small, 1e-4, 3, 4, True)


def test_noalloc_refuses_too_many_blocks(torch_gpu):
    """33792 blocks (a 1056x1024x1024 field): refused with CZW_ETOOBIG before
    any pointer is used, so placeholders stand in for the buffers."""
    import ctypes as C

    L = czw.load_library()
    dummy = C.c_void_p(1 << 20)  # never dereferenced: the size check comes first
    rc = L.czw_compress_field_noalloc(dummy, 1056, 1024, 1024, 4, 1e-3, 1, 0, 0, dummy, dummy, None)
    assert rc == -5  # CZW_ETOOBIG
