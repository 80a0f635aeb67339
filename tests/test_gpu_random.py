"""Seeded random cases against the oracle: field shape (each edge a
multiple of 32, non-cubic), precision, wavelet type, eps, zerobits and
shuffle drawn per case; records and decode bit-exact."""
import numpy as np
import pytest

from cubismz_amd import czw, fields

pytestmark = pytest.mark.gpu

CASES = 48


@pytest.mark.parametrize("case", range(CASES))
def test_random_case_matches_oracle(torch_gpu, oracle, case):
    torch = torch_gpu
    rng = np.random.default_rng(1000 + case)
    shape = tuple(int(32 * rng.integers(1, 5)) for _ in range(3))
    prec = "f64" if rng.random() < 0.3 else "f32"
    wtype = int(rng.integers(1, 4))
    eps = float(10 ** rng.uniform(-5, -1.5))
    zerobits = int(rng.choice([0, 4, 8]))
    shuffle = bool(rng.random() < 0.5)
    kind = "bubble" if rng.random() < 0.5 else "bumps"
    print(f"[case] random {case}: {shape} {prec} wtype={wtype} eps={eps:.3g} zb={zerobits} sh={shuffle} {kind}",
          flush=True)
    dtype = torch.float32 if prec == "f32" else torch.float64
    if kind == "bubble":
        f = fields.bubble_field(shape, seed=case, nbubbles=int(rng.integers(2, 12)), dtype=dtype)
    else:
        f = fields.bumps_field(shape, seed=case, dtype=dtype)
    rec, off = czw.compress_field(f.cuda(), eps, wtype, zerobits, shuffle)
    got = czw.records_to_payloads(rec.cpu().numpy(), off.cpu().numpy())
    slots, nbytes = oracle.compress_field(f.numpy(), eps, wtype, zerobits, shuffle)
    for b, p in enumerate(got):
        assert p == slots[b, :nbytes[b]].tobytes(), f"block {b}"
    out = czw.decompress_field(rec, off, shape, dtype, wtype, shuffle).cpu().numpy()
    ref = oracle.decompress_field(slots, nbytes, shape, wtype, prec, shuffle)
    it = np.uint32 if prec == "f32" else np.uint64
    assert np.array_equal(out.view(it), ref.view(it))
