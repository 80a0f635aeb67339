"""The per-block C++ adapter (include/czw_compressor.hpp), driven the way the
reference's writer and reader drive WaveletCompressorGeneric, against the
oracle: payload bytes and decoded block bit for bit."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native", "adapter_check")


@pytest.mark.parametrize("dt", ["f32", "f64"])
@pytest.mark.parametrize("wtype,zb,sh", [(1, 0, 0), (3, 4, 1), (2, 0, 1)])
def test_adapter_matches_oracle(torch_gpu, oracle, tmp_path, dt, wtype, zb, sh):
    if not os.path.exists(EXE):
        pytest.skip("tests/native/adapter_check not built")
    rng = np.random.default_rng(wtype + 10 * zb)
    z, y, x = np.meshgrid(*(np.linspace(0, 1, 32),) * 3, indexing="ij")
    blk = (1 + np.tanh((x - 0.4 * y + 0.2 * z - 0.3) * 12) + 1e-3 * rng.standard_normal(x.shape))
    blk = blk.astype(np.float32 if dt == "f32" else np.float64)
    src = tmp_path / "blk.raw"
    blk.tofile(src)
    pay, dec = tmp_path / "pay.bin", tmp_path / "dec.raw"
    eps = 1e-3
    print(f"[case] adapter {dt} wtype={wtype} zb={zb} sh={sh}", flush=True)
    r = subprocess.run([EXE, dt, str(src), repr(eps), str(wtype), str(zb), str(sh), str(pay), str(dec)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    want = oracle.compress(blk, eps, wtype, zb, bool(sh))
    assert pay.read_bytes() == want.tobytes()
    got = np.fromfile(dec, blk.dtype).reshape(32, 32, 32)
    ref = oracle.decompress(want.tobytes(), wtype, dt, bool(sh))
    iv = np.uint32 if dt == "f32" else np.uint64
    assert np.array_equal(got.view(iv), ref.view(iv))


@pytest.mark.parametrize("dt", ["f32", "f64"])
def test_adapter_per_block_objects_share_buffers(torch_gpu, oracle, tmp_path, dt):
    """64 blocks, each with its own compressor and decompressor object, as the
    reference's writer and reader construct them: identical results, and the
    device buffers are allocated once for the thread."""
    if not os.path.exists(EXE):
        pytest.skip("tests/native/adapter_check not built")
    rng = np.random.default_rng(5)
    blk = (1 + 0.01 * rng.standard_normal((32, 32, 32))).astype(np.float32 if dt == "f32" else np.float64)
    src = tmp_path / "blk.raw"
    blk.tofile(src)
    pay, dec = tmp_path / "pay.bin", tmp_path / "dec.raw"
    print(f"[case] adapter per-block objects {dt}", flush=True)
    r = subprocess.run([EXE, dt, str(src), "1e-3", "3", "0", "0", str(pay), str(dec), "64"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.split() == ["allocations", "1"]
    assert pay.read_bytes() == oracle.compress(blk, 1e-3, 3, 0, False).tobytes()
