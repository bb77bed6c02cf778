"""Host-side checks of the chi^2 plumbing in lofar_amd.autograd (CPU tensors, no device).

- ``chi2`` is one autograd function: its value matches sum_i w_i |V_i - m_i|^2 evaluated
  elementwise in float64, and its gradient has the bits of the elementwise graph's (the
  factor -2 is a power of two, so folding it into the weight changes no rounding).
- ``ImagingOperator.adjoint`` hands float32 cotangents and float32 weights straight to the
  imaging call in FP32C mode, where the library forms the float product weight*vis. That
  equals rounding the double product to float, which the float64 path computes: a product of
  two floats is exact in double (24 + 24 significant bits), and float multiplication rounds
  that exact product once.
"""
import numpy as np
import pytest


def _elementwise(m_re, m_im, vr, vi, w):
    import torch
    dr = vr.to(torch.float64) - m_re.to(torch.float64)
    di = vi.to(torch.float64) - m_im.to(torch.float64)
    return (w.to(torch.float64) * (dr * dr + di * di)).sum()


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_chi2_value_and_gradient_match_the_elementwise_graph(dtype):
    import torch
    from lofar_amd.autograd import chi2
    dt = getattr(torch, dtype)
    g = torch.Generator().manual_seed(11)
    n = 50000
    base_re = torch.randn(n, generator=g).to(dt)
    base_im = torch.randn(n, generator=g).to(dt)
    vr = torch.randn(n, generator=g).float()
    vi = torch.randn(n, generator=g).float()
    w = (torch.rand(n, generator=g) + 0.1).float()
    a_re, a_im = base_re.clone().requires_grad_(True), base_im.clone().requires_grad_(True)
    b_re, b_im = base_re.clone().requires_grad_(True), base_im.clone().requires_grad_(True)
    la = chi2(a_re, a_im, vr, vi, w)
    lb = _elementwise(b_re, b_im, vr, vi, w)
    la.backward()
    lb.backward()
    assert abs(float(la.detach()) - float(lb.detach())) <= 1e-13 * float(lb.detach())
    assert a_re.grad.dtype == dt and a_im.grad.dtype == dt
    assert torch.equal(a_re.grad, b_re.grad) and torch.equal(a_im.grad, b_im.grad)


def test_chi2_gradient_scales_with_the_upstream_gradient():
    """d(c * chi2)/dm = c * dchi2/dm, bit for bit for a power-of-two c."""
    import torch
    from lofar_amd.autograd import chi2
    g = torch.Generator().manual_seed(12)
    n = 1000
    m_re = torch.randn(n, generator=g, dtype=torch.float64).requires_grad_(True)
    m_im = torch.randn(n, generator=g, dtype=torch.float64).requires_grad_(True)
    vr, vi = torch.randn(n, generator=g).float(), torch.randn(n, generator=g).float()
    w = torch.rand(n, generator=g).float()
    chi2(m_re, m_im, vr, vi, w).backward()
    g1 = m_re.grad.clone()
    m_re.grad = None
    (4.0 * chi2(m_re, m_im, vr, vi, w)).backward()
    assert torch.equal(m_re.grad, 4.0 * g1)


def test_float_product_equals_rounded_double_product():
    """The identity the FP32C adjoint relies on, over normal, tiny and huge magnitudes."""
    rng = np.random.default_rng(13)
    n = 1 << 20
    a = (rng.standard_normal(n) * np.exp2(rng.integers(-60, 60, n))).astype(np.float32)
    b = (rng.random(n) * np.exp2(rng.integers(-60, 20, n))).astype(np.float32)
    via_double = (a.astype(np.float64) * b.astype(np.float64)).astype(np.float32)
    assert np.array_equal((a * b).view(np.uint32), via_double.view(np.uint32))
