"""Differentiable wavelet inverse with a frozen survivor mask.

The decompressor of CubismZ reconstructs a block as ``iwt(m * c)``: the
coefficients ``c`` that survived thresholding (``fabs(c) > eps``,
FullWaveletTransform.h:84, plus the 4^3 coarse scalings always kept,
:166-173), zeros elsewhere (:100-150), then the inverse transform (:51-56).
With the mask ``m`` frozen this is linear in ``c``; :class:`MaskedIWT` makes
it a ``torch.autograd.Function`` whose backward is ``m * iwt^T(g)``, the
transpose computed by the hand-written adjoint kernel (op 2 of
``czw_transform_blocks``), not by autograd tracing.

Coefficient tensors are kept in the block's own orientation O[z][y][x]
(the orientation the GPU kernels transform in).  The reference stores them in
a piecewise axis-permuted layout; :func:`to_physical` / :func:`from_physical`
convert between the two:

    region of O                 physical (a0, a1, a2) of O (z, y, x)
    outside [0,16)^3            (y, x, z)
    [0,16)^3 minus [0,8)^3      (x, z, y)
    [0,8)^3                     (z, y, x)

with physical index ``a2 + 32*(a1 + 32*a0)``.
"""
import numpy as np

from . import czw

BLOCK = czw.BLOCK


def _perm_index():
    """Flat physical index of every O position, as an int64 array [32768]."""
    z, y, x = np.meshgrid(np.arange(BLOCK), np.arange(BLOCK), np.arange(BLOCK), indexing="ij")
    r8 = (z < 8) & (y < 8) & (x < 8)
    r16 = (z < 16) & (y < 16) & (x < 16) & ~r8
    a0 = np.where(r8, z, np.where(r16, x, y))
    a1 = np.where(r8, y, np.where(r16, z, x))
    a2 = np.where(r8, x, np.where(r16, y, z))
    return (a2 + BLOCK * (a1 + BLOCK * a0)).reshape(-1).astype(np.int64)


_PHYS = None


def physical_index():
    global _PHYS
    if _PHYS is None:
        _PHYS = _perm_index()
    return _PHYS


def to_physical(coeffs):
    """O-orientation coefficients [..., 32,32,32] -> the reference's layout."""
    idx = physical_index()
    flat = coeffs.reshape(*coeffs.shape[:-3], BLOCK ** 3)
    if isinstance(coeffs, np.ndarray):
        out = np.empty_like(flat)
        out[..., idx] = flat
        return out.reshape(coeffs.shape)
    import torch

    out = torch.empty_like(flat)
    out[..., torch.as_tensor(idx, device=coeffs.device)] = flat
    return out.reshape(coeffs.shape)


def from_physical(coeffs):
    """Inverse of :func:`to_physical`."""
    idx = physical_index()
    flat = coeffs.reshape(*coeffs.shape[:-3], BLOCK ** 3)
    if isinstance(coeffs, np.ndarray):
        return flat[..., idx].reshape(coeffs.shape)
    import torch

    return flat[..., torch.as_tensor(idx, device=coeffs.device)].reshape(coeffs.shape)


def field_to_blocks(field):
    """Field [Z,Y,X] (x fastest) -> blocks [nb,32,32,32], block index x fastest."""
    Z, Y, X = field.shape
    if Z % BLOCK or Y % BLOCK or X % BLOCK:
        raise czw.CZWError("field extents must be multiples of 32")
    b = field.reshape(Z // BLOCK, BLOCK, Y // BLOCK, BLOCK, X // BLOCK, BLOCK)
    return b.permute(0, 2, 4, 1, 3, 5).reshape(-1, BLOCK, BLOCK, BLOCK).contiguous()


def blocks_to_field(blocks, shape):
    Z, Y, X = shape
    b = blocks.reshape(Z // BLOCK, Y // BLOCK, X // BLOCK, BLOCK, BLOCK, BLOCK)
    return b.permute(0, 3, 1, 4, 2, 5).reshape(Z, Y, X).contiguous()


def fwt(blocks, wtype):
    """Forward transform of a copy of ``blocks`` (O orientation output)."""
    return czw.transform_blocks(blocks.contiguous().clone(), wtype, "fwt")


def iwt(coeffs, wtype):
    return czw.transform_blocks(coeffs.contiguous().clone(), wtype, "iwt")


def iwt_adjoint(grad, wtype):
    return czw.transform_blocks(grad.contiguous().clone(), wtype, "iwt_adjoint")


def threshold_mask(coeffs, eps):
    """Survivor mask of O-orientation coefficients, as the reference decides it.

    ``fabs(c) > eps`` in the coefficient type with eps rounded to float first
    (FullWaveletTransform.h:84, WaveletCompressor.cpp:153-160), and the 4^3
    coarse scalings always kept (:166-173).
    """
    import torch

    e = float(np.float32(eps))
    m = coeffs.abs() > torch.tensor(e, dtype=coeffs.dtype, device=coeffs.device)
    m[..., :4, :4, :4] = True
    return m


class MaskedIWT:
    """``y = iwt(m * c)`` with ``m`` frozen; ``dL/dc = m * iwt^T(dL/dy)``.

    Use :func:`masked_iwt`.  Inputs: coefficients ``c`` [nb,32,32,32] (f32 or
    f64, O orientation, on the GPU), boolean mask of the same shape, wavelet
    type.
    """

    _fn = None

    @classmethod
    def fn(cls):
        if cls._fn is None:
            import torch

            class _F(torch.autograd.Function):
                @staticmethod
                def forward(ctx, c, mask, wtype):
                    ctx.save_for_backward(mask)
                    ctx.wtype = int(wtype)
                    # where() makes a fresh contiguous tensor, transformed in place
                    return czw.transform_blocks(torch.where(mask, c, 0.0), wtype, "iwt")

                @staticmethod
                def backward(ctx, g):
                    (mask,) = ctx.saved_tensors
                    ga = czw.transform_blocks(g.contiguous().clone(), ctx.wtype, "iwt_adjoint")
                    return torch.where(mask, ga, 0.0), None, None

            cls._fn = _F
        return cls._fn


def masked_iwt(coeffs, mask, wtype):
    return MaskedIWT.fn().apply(coeffs, mask, wtype)
