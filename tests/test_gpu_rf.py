"""RF-DT (recursive-filter domain transform; not in the reference) on the GPU against
the FP64 oracle (oracle/rf_oracle.c, pinned in test_oracle_rf.py): forward, the input
and guide adjoints, the exact-transpose identity, tape and argument errors, and a
test-only torch.autograd.Function over the device-pointer entry points."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S(gpu_lib):
    from domain_transform_solver_amd import solver
    return solver


def _frame(S, oracle_mod, w, h, cs, seed):
    rng = np.random.default_rng(seed)
    rgba = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    # smooth regions as well as noise, so weights span (0, 1)
    rgba[: h // 2] = (np.arange(w)[None, :, None] // 7 * 40 % 256).astype(np.uint8)
    g = S.DomainTransformSolver(w, h)
    g.InitFrame(cs, rgba, np.zeros((h, w), np.float32))
    guide = oracle_mod.convert_color(rgba, cs).astype(np.float64)
    return g, guide, rng


def _rel(a, b):
    return float(np.abs(a - b).max() / max(1e-30, np.abs(b).max()))


@pytest.mark.parametrize("shape", [(181, 133), (64, 1), (1, 64), (2100, 5), (5, 2100), (3, 3)])
@pytest.mark.parametrize("n", [1, 3])
def test_rf_forward_matches_oracle(S, oracle_mod, shape, n):
    w, h = shape
    g, guide, rng = _frame(S, oracle_mod, w, h, 0, 100 + w + n)
    x = rng.normal(size=(h, w)).astype(np.float32)
    fp = S.DomainFilterParams(20.0, 15.0, 0.4)
    out = g.RFFilter(fp, x, n)
    ref = oracle_mod.rf_filter(guide, 20.0, 15.0, 0.4, n, x)
    err = _rel(out, ref)
    print(f"RF {shape} N={n}: max-abs / max {err:.2e}")
    assert err <= 1e-5


def test_rf_ycbcr_guide_and_constant_input(S, oracle_mod):
    w, h = 97, 61
    g, guide, rng = _frame(S, oracle_mod, w, h, 1, 7)
    fp = S.DomainFilterParams(30.0, 30.0, 0.2)
    x = rng.uniform(-3, 3, (h, w)).astype(np.float32)
    assert _rel(g.RFFilter(fp, x, 3), oracle_mod.rf_filter(guide, 30.0, 30.0, 0.2, 3, x)) <= 1e-5
    c = g.RFFilter(fp, np.full((h, w), 1.75, np.float32), 3)
    assert np.abs(c - 1.75).max() <= 1e-5


@pytest.mark.parametrize("shape,n", [((131, 77), 3), ((40, 1), 2), ((1, 40), 2), ((2100, 4), 1)])
def test_rf_adjoint_matches_oracle(S, oracle_mod, shape, n):
    w, h = shape
    g, guide, rng = _frame(S, oracle_mod, w, h, 0, 200 + w)
    x = rng.normal(size=(h, w)).astype(np.float32)
    yb = rng.normal(size=(h, w)).astype(np.float32)
    fp = S.DomainFilterParams(12.0, 9.0, 0.5)
    y = g.RFFilter(fp, x, n, keep_tape=True)
    xb, gb = g.RFFilterAdjoint(yb)
    rxb, rgb = oracle_mod.rf_filter_adjoint(guide, 12.0, 9.0, 0.5, n, x, yb)
    ex, eg = _rel(xb, rxb), _rel(gb, rgb)
    print(f"RF adjoint {shape} N={n}: xbar {ex:.2e}, gbar {eg:.2e} (max |gbar| {np.abs(rgb).max():.3e})")
    assert ex <= 1e-5
    assert eg <= 1e-5
    # <ybar, F x> = <F^T ybar, x>: the input adjoint is the exact transpose
    lhs = float(np.vdot(yb.astype(np.float64), y.astype(np.float64)))
    rhs = float(np.vdot(xb.astype(np.float64), x.astype(np.float64)))
    assert abs(lhs - rhs) <= 1e-5 * (abs(lhs) + np.abs(yb).sum() * 1e-3)


def test_rf_tape_and_argument_errors(S, oracle_mod):
    from domain_transform_solver_amd._lib import DtsError
    w, h = 32, 24
    g, guide, rng = _frame(S, oracle_mod, w, h, 0, 9)
    x = rng.normal(size=(h, w)).astype(np.float32)
    fp = S.DomainFilterParams(10.0, 10.0, 0.5)
    with pytest.raises(DtsError):
        g.RFFilterAdjoint(x)                       # no tape yet
    g.RFFilter(fp, x, 2, keep_tape=True)
    g.RFFilterAdjoint(x)
    g.InitFrame(0, np.zeros((h, w, 4), np.uint8), x)
    with pytest.raises(DtsError):
        g.RFFilterAdjoint(x)                       # new frame: tape invalid
    with pytest.raises(DtsError):
        g.RFFilter(fp, x, 0)
    with pytest.raises(DtsError):
        g.RFFilter(S.DomainFilterParams(10.0, 10.0, 0.0), x, 2)


def test_rf_torch_autograd_function(S, oracle_mod):
    """Test-only torch.autograd.Function: forward = RFFilter on device pointers
    (keeping the tape), backward = the adjoint; the input gradient of
    sum(out * m) equals the oracle adjoint for ybar = m."""
    import ctypes
    import torch
    from domain_transform_solver_amd import _lib
    w, h, n = 80, 56, 3
    g, guide, rng = _frame(S, oracle_mod, w, h, 0, 11)
    lib = _lib.load()
    _lib.check(lib.dts_set_stream(g.handle, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    fp = S.DomainFilterParams(16.0, 16.0, 0.3)

    class RF(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x):
            x = x.contiguous()
            y = torch.empty_like(x)
            g.RFFilterDevice(fp, x.data_ptr(), y.data_ptr(), n, keep_tape=True)
            return y

        @staticmethod
        def backward(ctx, yb):
            yb = yb.contiguous()
            xb = torch.empty_like(yb)
            g.RFFilterAdjointDevice(yb.data_ptr(), xb.data_ptr(), None)
            return xb

    x_np = rng.normal(size=(h, w)).astype(np.float32)
    m_np = rng.normal(size=(h, w)).astype(np.float32)
    x = torch.from_numpy(x_np).cuda().requires_grad_(True)
    m = torch.from_numpy(m_np).cuda()
    (RF.apply(x) * m).sum().backward()
    torch.cuda.synchronize()
    _lib.check(lib.dts_set_stream(g.handle, None))
    rxb, _ = oracle_mod.rf_filter_adjoint(guide, 16.0, 16.0, 0.3, n, x_np, m_np)
    assert _rel(x.grad.cpu().numpy(), rxb) <= 1e-5
