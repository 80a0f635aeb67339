"""The oracle (C restatement) pinned against the reference's own outputs and
against what the mathematics fixes.  CPU only.

Golden vectors (tests/golden, written by make_golden.py from the reference
compressor compiled from its sources) pin the exact binary32/binary64
arithmetic; the property tests below pin what a plausible mistake in the
restatement would break (a wrong stencil weight, a swapped pass order, a
wrong stream or mask index, a broken shuffle).
"""
import hashlib
import json
import lzma
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def demo():
    out = {}
    for prec, dt in (("f32", "<f4"), ("f64", "<f8")):
        raw = lzma.decompress(open(os.path.join(GOLD, f"demo_{prec}.raw.xz"), "rb").read())
        out[prec] = np.frombuffer(raw, dt).reshape(128, 128, 128)
    return out


@pytest.fixture(scope="module")
def blocks(demo):
    from tests.golden.make_golden import block_inputs

    return block_inputs(demo["f32"], demo["f64"])


def test_demo_fixture_matches_reference_data(demo):
    # demo.h5 values lie in [0.00677, 0.8593] (the range read from Tests/Data/data.tar.xz)
    assert demo["f32"].min() == pytest.approx(0.0067658233, abs=1e-9)
    assert demo["f32"].max() == pytest.approx(0.8592596, abs=1e-7)
    assert np.allclose(demo["f64"], demo["f32"], atol=1e-7)


def test_rows_match_reference(oracle):
    g = np.load(os.path.join(GOLD, "rows.npz"))
    for prec in ("f32", "f64"):
        for n in (8, 16, 32):
            x = g[f"in_{prec}_{n}"]
            for w in (1, 2, 3):
                for fwd, key in ((True, "fwd"), (False, "inv")):
                    want = g[f"{key}_{prec}_{n}_w{w}"]
                    got = np.array([oracle.row(r, w, fwd) for r in x])
                    assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), (prec, n, w, key)


def test_blocks_match_reference(oracle, blocks):
    table = json.load(open(os.path.join(GOLD, "blocks.json")))
    seen_fwt = {}
    for c in table["cases"]:
        blk = blocks[c["prec"]][c["block"]]
        key = (c["prec"], c["block"], c["wtype"])
        if key not in seen_fwt:
            seen_fwt[key] = _sha(oracle.fwt(blk, c["wtype"]))
        assert seen_fwt[key] == c["fwt_sha256"], key
        pay = oracle.compress(blk, c["eps"], c["wtype"], c["zerobits"], bool(c["shuffle"]))
        assert pay.size == c["nbytes"] and _sha(pay) == c["payload_sha256"], c
        dec = oracle.decompress(pay, c["wtype"], c["prec"], bool(c["shuffle"]))
        assert _sha(dec) == c["decoded_sha256"], c


def test_full_payloads_match_reference(oracle, blocks):
    g = np.load(os.path.join(GOLD, "payloads.npz"))
    for key in g.files:
        prec, w, zb, sh = key.split("_")
        pay = oracle.compress(blocks[prec][1], 1e-3, int(w[1:]), int(zb[2:]), sh == "sh1")
        assert np.array_equal(pay, g[key]), key


# --- properties fixed by the mathematics -------------------------------------

@pytest.mark.parametrize("n", [8, 16, 32])
@pytest.mark.parametrize("w", [1, 2])
def test_interpolating_details_vanish_on_cubics(oracle, n, w):
    """4-point interpolating predictors reproduce cubics (all four boundary
    variants), so the details of integer cubic samples are exactly zero."""
    i = np.arange(n, dtype=np.float64)
    for coefs in ((1, 0, 0, 0), (0, 1, 0, 0), (0, 0, 1, 0), (3, -2, 5, 1)):
        x = coefs[0] * i ** 3 + coefs[1] * i ** 2 + coefs[2] * i + coefs[3]
        y = oracle.row(x, w, True)
        assert np.all(y[n // 2:] == 0.0), (coefs, y[n // 2:])
        if w == 1:
            assert np.array_equal(y[:n // 2], x[0::2])


@pytest.mark.parametrize("n", [8, 16, 32])
def test_average_interpolating_details_vanish_on_quadratics(oracle, n):
    """3rd-order average-interpolating wavelets reproduce quadratics in cell
    averages: for cell averages of 3x^2 (integers 3i^2+3i+1) and of lower
    degree, the details are exactly zero and the scalings are the coarse
    cell averages."""
    i = np.arange(n, dtype=np.float64)
    for x in (3 * i ** 2 + 3 * i + 1, 2 * i + 1, np.full(n, 7.0)):
        y = oracle.row(x, 3, True)
        assert np.all(y[n // 2:] == 0.0)
        assert np.array_equal(y[:n // 2], 0.5 * (x[0::2] + x[1::2]))


@pytest.mark.parametrize("w", [1, 2, 3])
def test_perfect_reconstruction_and_linearity(oracle, w):
    rng = np.random.default_rng(w)
    for n in (8, 16, 32):
        T = np.array([oracle.row(e, w, True) for e in np.eye(n)]).T   # dense matrix by unit vectors
        Ti = np.array([oracle.row(e, w, False) for e in np.eye(n)]).T
        assert np.allclose(Ti @ T, np.eye(n), atol=1e-12)
        x = rng.standard_normal(n)
        assert np.allclose(oracle.row(x, w, True), T @ x, atol=1e-12)
    blk = rng.standard_normal((32, 32, 32))
    assert np.allclose(oracle.iwt(oracle.fwt(blk, w), w), blk, atol=1e-10)


def test_3d_schedule_is_axis_passes_with_piecewise_permutation(oracle):
    """The reference's in-place swaps (WaveletsOnInterval.h:205-276) equal 1D
    passes along fixed axes (level 32: x,y,z; 16: z,x,y; 8: y,z,x) followed by
    the piecewise axis permutation the GPU kernels use for mask/stream
    positions.  Brute force on one block."""
    rng = np.random.default_rng(7)
    blk = rng.standard_normal((32, 32, 32)).astype(np.float32)
    for w in (1, 3):
        o = blk.copy()
        for bs, axes in ((32, (2, 1, 0)), (16, (0, 2, 1)), (8, (1, 0, 2))):
            for ax in axes:
                sub = np.moveaxis(o[:bs, :bs, :bs], ax, -1).copy()
                for idx in np.ndindex(bs, bs):
                    sub[idx] = oracle.row(sub[idx], w, True)
                o[:bs, :bs, :bs] = np.moveaxis(sub, -1, ax)
        phys = oracle.fwt(blk, w).reshape(-1)
        p = np.arange(32768)
        a0, a1, a2 = p >> 10, (p >> 5) & 31, p & 31
        m = np.maximum(np.maximum(a0, a1), a2)
        z = np.where(m >= 16, a2, np.where(m >= 8, a1, a0))
        y = np.where(m >= 16, a0, np.where(m >= 8, a2, a1))
        x = np.where(m >= 16, a1, np.where(m >= 8, a0, a2))
        assert np.array_equal(o[z, y, x].view(np.uint32), phys.view(np.uint32))


def _stream_positions():
    """Physical positions in stream order (FullWaveletTransform.h:59-96,163-176)."""
    pos = [a0 * 1024 + a1 * 32 + a2 for a0 in range(4) for a1 in range(4) for a2 in range(4)]
    for bs in (8, 16, 32):
        h = bs // 2
        for code in range(1, 8):
            o2, o1, o0 = h * (code & 1), h * ((code >> 1) & 1), h * ((code >> 2) & 1)
            pos += [(o0 + i0) * 1024 + (o1 + i1) * 32 + o2 + i2
                    for i0 in range(h) for i1 in range(h) for i2 in range(h)]
    return np.array(pos)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_mask_and_stream_order_closed_form(oracle, prec):
    dt = np.float32 if prec == "f32" else np.float64
    rng = np.random.default_rng(3)
    blk = (rng.standard_normal((32, 32, 32)) * np.logspace(-5, 0, 32)).astype(dt)
    coef = oracle.fwt(blk, 3).reshape(-1)
    pos = _stream_positions()
    assert np.array_equal(np.sort(pos), np.arange(32768))
    eps = 1e-3
    keep = np.abs(coef[pos]) > dt(eps)
    keep[:64] = True
    pay = oracle.compress(blk, eps, 3)
    mask = np.unpackbits(pay[:4096], bitorder="little").astype(bool)
    want_mask = np.zeros(32768, bool)
    want_mask[pos[keep]] = True
    assert np.array_equal(mask, want_mask)
    surv = np.frombuffer(pay[4096:].tobytes(), dt)
    assert np.array_equal(surv, coef[pos[keep]])
    assert surv.size == mask.sum() >= 64


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_zerobits_and_shuffle_closed_form(oracle, prec):
    dt = np.float32 if prec == "f32" else np.float64
    s = np.dtype(dt).itemsize
    rng = np.random.default_rng(4)
    blk = rng.standard_normal((32, 32, 32)).astype(dt)
    plain = oracle.compress(blk, 1e-1, 1)
    S = (plain.size - 4096) // s
    for zb in (4, 8, 12, 16):
        z = oracle.compress(blk, 1e-1, 1, zb)
        words = plain[4096:].view(np.uint32).reshape(S, s // 4).copy()
        words[:, 0] &= np.uint32((0xFFFFFFFF << zb) & 0xFFFFFFFF)
        assert np.array_equal(z[4096:], words.view(np.uint8).reshape(-1))
        zs = oracle.compress(blk, 1e-1, 1, zb, True)
        assert np.array_equal(zs[4096:], words.view(np.uint8).reshape(S, s).T.reshape(-1))
        assert np.array_equal(zs[:4096], plain[:4096])


@pytest.mark.parametrize("w", [1, 2, 3])
def test_eps_zero_decode_is_iwt_of_fwt(oracle, w):
    rng = np.random.default_rng(5)
    blk = rng.standard_normal((32, 32, 32)).astype(np.float32)
    for sh in (False, True):
        dec = oracle.decompress(oracle.compress(blk, 0.0, w, 0, sh), w, "f32", sh)
        assert np.array_equal(dec, oracle.iwt(oracle.fwt(blk, w), w))


def test_dp_eps_is_float_rounded(oracle):
    """The double build compares |c| > (double)(float)eps (WaveletCompressor.h:61)."""
    blk = np.zeros((32, 32, 32))
    blk[5, 6, 7] = 1.0
    coef = oracle.fwt(blk, 1).reshape(-1)
    c = np.abs(coef[np.abs(coef) > 0]).min()   # a coefficient magnitude
    e32 = np.float32(c)
    if float(e32) < c:          # eps just below c in double but rounding to >= c as float
        eps = np.nextafter(c, 0)
        pay = oracle.compress(blk, eps, 1)
        n_keep_float = int(np.sum(np.abs(coef) > float(np.float32(eps))))
        assert (pay.size - 4096) // 8 >= 64
        assert np.unpackbits(pay[:4096]).sum() >= n_keep_float


def test_decoder_rejects_bad_sizes(oracle):
    rng = np.random.default_rng(6)
    pay = oracle.compress(rng.standard_normal((32, 32, 32)).astype(np.float32), 1e-2, 3)
    with pytest.raises(ValueError):
        oracle.decompress(pay[:-4], 3, "f32")
    bad = pay.copy()
    bad[200] ^= 0xFF
    with pytest.raises(ValueError):
        oracle.decompress(bad, 3, "f32")


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_oracle_matches_compiled_reference_random(oracle, prec):
    if not oracle.ref_available(prec, 4, True):
        pytest.skip("oracle/_ref not built here")
    dt = np.float32 if prec == "f32" else np.float64
    rng = np.random.default_rng(8)
    for _ in range(3):
        blk = (rng.standard_normal((32, 32, 32)) * rng.uniform(1e-3, 1e3)).astype(dt)
        for w in (1, 2, 3):
            for zb, sh in ((0, False), (4, True)):
                p = oracle.compress(blk, 1e-2, w, zb, sh)
                assert np.array_equal(p, oracle.ref_compress(blk, 1e-2, w, zb, sh))
                assert np.array_equal(oracle.decompress(p, w, prec, sh).view(np.uint8),
                                      oracle.ref_decompress(p, w, prec, zb, sh).view(np.uint8))


def special_block(dt, rng, kind):
    """A smooth block with one kind of special value planted: subnormals,
    signed zeros, values whose predictors round past the largest finite Real,
    infinities or NaNs."""
    i = np.arange(32, dtype=np.float64)
    blk = 1.0 + 0.3 * np.sin(i[:, None, None] / 5.0) * np.cos(i[None, :, None] / 7.0) + 0.01 * i[None, None, :]
    blk = blk.astype(dt)
    fi = np.finfo(dt)
    pos = tuple(rng.integers(0, 32, (3, 40)))
    if kind == "subnormal":
        blk *= fi.tiny  # the whole block below the normal range
        blk[pos] = fi.smallest_subnormal * rng.integers(1, 50, 40)
    elif kind == "zeros":
        blk[:, :, :16] = 0.0
        blk[pos] = -0.0
    elif kind == "huge":
        blk[pos] = fi.max * rng.choice([-0.9, 0.9], 40)
    elif kind == "inf":
        blk[pos] = rng.choice([-np.inf, np.inf], 40)
    elif kind == "nan":
        blk[pos] = np.nan
    return blk


def nan_aware_equal(a, b, dt):
    """Equal bit for bit except that a NaN matches any NaN (NaN payloads are
    not pinned across hardware)."""
    a, b = np.asarray(a).view(dt), np.asarray(b).view(dt)
    na, nb = np.isnan(a), np.isnan(b)
    it = np.uint32 if dt == np.float32 else np.uint64
    return bool(np.array_equal(na, nb) and np.array_equal(a[~na].view(it), b[~nb].view(it)))


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("kind", ["subnormal", "zeros", "huge", "inf", "nan"])
def test_oracle_matches_compiled_reference_special_values(oracle, prec, kind):
    """Subnormals, signed zeros, overflowing predictors, infinities and NaNs:
    the oracle keeps the reference's masks and values (NaN payloads aside)."""
    if not oracle.ref_available(prec, 0, False):
        pytest.skip("oracle/_ref not built here")
    dt = np.float32 if prec == "f32" else np.float64
    rng = np.random.default_rng(21)
    blk = special_block(dt, rng, kind)
    eps = 0.0 if kind == "subnormal" else 1e-3
    for w in (1, 2, 3):
        p = oracle.compress(blk, eps, w)
        r = oracle.ref_compress(blk, eps, w)
        assert len(p) == len(r) and np.array_equal(p[:4096], r[:4096]), (kind, w)
        assert nan_aware_equal(p[4096:], r[4096:], dt), (kind, w)
        d = oracle.decompress(r, w, prec)
        assert nan_aware_equal(d, oracle.ref_decompress(r, w, prec), dt), (kind, w)
