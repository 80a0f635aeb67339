"""The O-orientation transform schedule and the physical coefficient layout,
pinned on the CPU against the oracle (itself pinned by the golden vectors of
the compiled reference).

The GPU transform kernels keep a block in its own orientation O[z][y][x] and
change the pass axis instead of swapping axes (WaveletsOnInterval.h:205-276).
Here the same schedule is run with the oracle's 1D row transform on numpy
arrays; mapped through cubismz_amd.autograd.to_physical it must equal the
oracle's whole-block fwt bit for bit, and the inverse must invert it.  The
mask of threshold_mask, mapped the same way, must equal the mask bytes of the
oracle's payload.
"""
import numpy as np
import pytest

from cubismz_amd import autograd as A

AX = {"z": 0, "y": 1, "x": 2}
FWD = [(32, "xyz"), (16, "zxy"), (8, "yzx")]
INV = [(8, "xzy"), (16, "yxz"), (32, "zyx")]


def _pass(oracle, b, n, ax, wtype, forward):
    v = np.moveaxis(b[:n, :n, :n], AX[ax], -1).copy()
    flat = v.reshape(-1, n)
    for r in range(flat.shape[0]):
        flat[r] = oracle.row(flat[r], wtype, forward)
    b[:n, :n, :n] = np.moveaxis(flat.reshape(v.shape), -1, AX[ax])


def o_fwt(oracle, block, wtype):
    b = block.copy()
    for n, axes in FWD:
        for ax in axes:
            _pass(oracle, b, n, ax, wtype, True)
    return b


def o_iwt(oracle, coeffs, wtype):
    b = coeffs.copy()
    for n, axes in INV:
        for ax in axes:
            _pass(oracle, b, n, ax, wtype, False)
    return b


def _block(seed, dt):
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(*(np.linspace(0, 1, 32),) * 3, indexing="ij")
    f = 1.0 + 0.3 * np.tanh((x + 0.5 * y - 0.7 * z) * 8) + 0.05 * np.sin(9 * x * y + 4 * z)
    return (f + 1e-3 * rng.standard_normal(f.shape)).astype(dt)


def test_physical_index_is_permutation():
    idx = A.physical_index()
    assert np.array_equal(np.sort(idx), np.arange(32768))
    c = np.arange(32768, dtype=np.float64).reshape(32, 32, 32)
    assert np.array_equal(A.from_physical(A.to_physical(c)), c)
    # the 4^3 coarse corner and the [0,8)^3 level stay in place
    assert np.array_equal(A.to_physical(c)[:8, :8, :8], c[:8, :8, :8])


@pytest.mark.parametrize("wtype", [1, 2, 3])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_o_schedule_matches_oracle_fwt_and_iwt(oracle, wtype, dt):
    blk = _block(wtype, dt)
    co = o_fwt(oracle, blk, wtype)
    ref = oracle.fwt(blk, wtype)
    assert np.array_equal(A.to_physical(co).view(np.uint8), ref.view(np.uint8))
    back = o_iwt(oracle, co, wtype)
    ref_back = oracle.iwt(ref, wtype)
    assert np.array_equal(back.view(np.uint8), ref_back.view(np.uint8))


@pytest.mark.parametrize("wtype", [1, 3])
def test_threshold_mask_matches_payload_mask(oracle, wtype):
    import torch

    blk = _block(10 + wtype, np.float32)
    eps = 1e-3
    co = o_fwt(oracle, blk, wtype)
    m = A.threshold_mask(torch.from_numpy(co), eps).numpy()
    bits = np.unpackbits(oracle.compress(blk, eps, wtype)[:4096], bitorder="little").astype(bool)
    assert np.array_equal(A.to_physical(m).reshape(-1), bits)
    assert 64 <= m.sum() < 32768
