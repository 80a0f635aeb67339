"""GPU tests of the whole-block transforms and the differentiable inverse.

* fwt / iwt kernels (ops 0/1 of czw_transform_blocks) against the oracle, bit
  for bit, in binary32 and binary64, through the physical layout map pinned
  on the CPU (test_transform_layout.py).
* iwt(m * fwt(x)) equals the decompressor's output bit for bit.
* The adjoint kernel (op 2) is the transpose of the inverse: the dot-product
  identity in binary64, and the full dense 32768 x 32768 matrix of the
  inverse against the full matrix of the adjoint (A^T == B).
* MaskedIWT's gradient against binary64 central finite differences
  (h = 1e-6, relative error <= 1e-6) on 32^3 blocks, with loss = MSE to a
  second-seed field.
"""
import numpy as np
import pytest

from cubismz_amd import autograd as A
from cubismz_amd import czw, fields

pytestmark = pytest.mark.gpu


def _blocks(torch, n, seed, dtype):
    shape = (64, 64, 32 * ((n + 3) // 4)) if n > 1 else (32, 32, 32)
    f = fields.bubble_field(shape, seed=seed, nbubbles=6, dtype=dtype, device="cuda")
    return A.field_to_blocks(f)[:n].contiguous()


@pytest.mark.parametrize("wtype", [1, 2, 3])
@pytest.mark.parametrize("dt", ["f32", "f64"])
def test_fwt_iwt_bitexact_vs_oracle(torch_gpu, oracle, wtype, dt):
    torch = torch_gpu
    dtype = torch.float32 if dt == "f32" else torch.float64
    print(f"[case] fwt/iwt wtype={wtype} {dt}", flush=True)
    b = _blocks(torch, 4, 3 + wtype, dtype)
    co = A.fwt(b, wtype)
    bh = b.cpu().numpy()
    coh = co.cpu().numpy()
    for k in range(b.shape[0]):
        ref = oracle.fwt(bh[k], wtype)
        assert np.array_equal(A.to_physical(coh[k]).view(np.uint8), ref.view(np.uint8)), k
    back = A.iwt(co, wtype).cpu().numpy()
    for k in range(b.shape[0]):
        ref = oracle.iwt(A.to_physical(coh[k]), wtype)
        assert np.array_equal(back[k].view(np.uint8), ref.view(np.uint8)), k


@pytest.mark.parametrize("wtype", [1, 2, 3])
def test_masked_iwt_equals_decompressor(torch_gpu, wtype):
    torch = torch_gpu
    print(f"[case] masked iwt vs decompress wtype={wtype}", flush=True)
    b = _blocks(torch, 8, 20 + wtype, torch.float32)
    eps = 1e-3
    rec, off = czw.compress_blocks(b, eps, wtype)
    dec = czw.decompress_blocks(rec, off, b.shape[0], torch.float32, wtype)
    co = A.fwt(b, wtype)
    m = A.threshold_mask(co, eps)
    y = A.masked_iwt(co, m, wtype)
    assert torch.equal(y.view(torch.int32), dec.view(torch.int32))


@pytest.mark.parametrize("wtype", [1, 2, 3])
def test_adjoint_dot_product_f64(torch_gpu, wtype):
    torch = torch_gpu
    print(f"[case] adjoint dot product wtype={wtype}", flush=True)
    g = torch.Generator(device="cuda").manual_seed(wtype)
    u = torch.randn(16, 32, 32, 32, dtype=torch.float64, device="cuda", generator=g)
    v = torch.randn(16, 32, 32, 32, dtype=torch.float64, device="cuda", generator=g)
    lhs = float((A.iwt(u, wtype) * v).sum())
    rhs = float((u * A.iwt_adjoint(v, wtype)).sum())
    assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + abs(rhs) + 1.0)
    # f32 adjoint agrees with the f64 one
    a32 = A.iwt_adjoint(v.float(), wtype).double()
    a64 = A.iwt_adjoint(v, wtype)
    assert float((a32 - a64).abs().max()) <= 1e-5 * float(a64.abs().max())


@pytest.mark.parametrize("wtype", [1, 2, 3])
def test_adjoint_dense_transpose_f64(torch_gpu, wtype):
    """Full dense 32768 x 32768 matrices of iwt and of its adjoint (8 GiB each)."""
    torch = torch_gpu
    print(f"[case] dense transpose wtype={wtype}", flush=True)
    n = 32768
    a = torch.zeros(n, n, dtype=torch.float64, device="cuda")
    a.view(-1)[:: n + 1] = 1.0
    b = a.clone()
    A.czw.transform_blocks(a.view(n, 32, 32, 32), wtype, "iwt")          # a[i] = iwt(e_i) = column i
    A.czw.transform_blocks(b.view(n, 32, 32, 32), wtype, "iwt_adjoint")  # b[j] = iwt^T(e_j) = row j
    scale = float(a.abs().max())
    worst = 0.0
    for r0 in range(0, n, 4096):
        worst = max(worst, float((b[r0:r0 + 4096] - a[:, r0:r0 + 4096].t()).abs().max()))
    nnz = int((a != 0).sum())
    del a, b
    torch.cuda.empty_cache()
    assert worst <= 1e-13 * scale, (worst, scale)
    assert nnz > n  # a genuinely dense-ish operator, not the identity


@pytest.mark.parametrize("wtype", [1, 3])
def test_gradient_vs_central_fd_f64(torch_gpu, wtype):
    torch = torch_gpu
    print(f"[case] FD gradient wtype={wtype}", flush=True)
    x = _blocks(torch, 2, 40 + wtype, torch.float64)
    target = _blocks(torch, 2, 90 + wtype, torch.float64)
    co = A.fwt(x, wtype)
    m = A.threshold_mask(co, 1e-3)
    c = co.clone().requires_grad_(True)
    loss = ((A.masked_iwt(c, m, wtype) - target) ** 2).mean()
    loss.backward()
    grad = c.grad.detach()
    assert torch.equal(grad[~m], torch.zeros_like(grad[~m]))

    def dL(cp, cm):
        # central difference of the loss, summed elementwise: the terms outside
        # the perturbed coefficient's support cancel exactly
        rp = A.masked_iwt(cp, m, wtype) - target
        rm = A.masked_iwt(cm, m, wtype) - target
        return float(((rp - rm) * (rp + rm)).mean())

    rng = np.random.default_rng(wtype)
    on = torch.nonzero(m.view(-1)).view(-1).cpu().numpy()
    picks = list(rng.choice(on, 24, replace=False)) + list(range(0, 64, 9))
    h = 1e-6
    for p in picks:
        cp = co.clone().view(-1)
        cm = co.clone().view(-1)
        cp[p] += h
        cm[p] -= h
        fd = dL(cp.view_as(co), cm.view_as(co)) / (2 * h)
        an = float(grad.view(-1)[p])
        assert abs(fd - an) <= 1e-6 * max(abs(an), 1e-3 * float(grad.abs().max())), (p, fd, an)


def test_masked_iwt_f32_gradient_matches_f64(torch_gpu):
    torch = torch_gpu
    wtype = 1
    x = _blocks(torch, 8, 7, torch.float32)
    t = _blocks(torch, 8, 8, torch.float32)
    co = A.fwt(x, wtype)
    m = A.threshold_mask(co, 1e-3)
    grads = []
    for dt in (torch.float32, torch.float64):
        c = co.detach().to(dt).clone().requires_grad_(True)
        ((A.masked_iwt(c, m, wtype) - t.to(dt)) ** 2).sum().backward()
        grads.append(c.grad.double())
    assert float((grads[0] - grads[1]).abs().max()) <= 1e-5 * float(grads[1].abs().max())
