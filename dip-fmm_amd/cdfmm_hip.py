"""Python front-end of the HipFull backend (the reference's ``cdfmm`` interface).

The extension ``_cdfmm_hip`` links the HIP runtime.  PyTorch-ROCm ships its
own copy of that runtime; two copies in one process cannot share the GPU, so
torch (when installed) is imported first and the extension then binds to the
runtime already loaded.
"""

try:  # one HIP runtime per process
    import torch  # noqa: F401
except ImportError:  # pragma: no cover - torch is optional for host-memory use
    torch = None

from _cdfmm_hip import *  # noqa: F401,F403
from _cdfmm_hip import _laplace_derivatives, _m2l_class_matrix  # noqa: F401


# ---------------------------------------------------------------------------
# torch.autograd caller: the field map H = A m of one plan as a differentiable
# function of the moments.  The backward is the plan's exact transpose
# (UniformFmm.evaluate_adjoint_device), itself differentiable, so higher-order
# derivatives compose.  Positions are constants of the plan.

_MAX_BATCH = 64  # moment states per device pass


def _plan_dtype(fmm):
    return torch.float32 if fmm.precision == StaticPrecision.FLOAT32 else torch.float64  # noqa: F405


def _apply_device(fmm, x, n_out, adjoint):
    """Run the forward (or adjoint) device pass over (K, n, 3) in chunks of 64 states."""
    if not x.is_cuda:
        raise ValueError("the HipFull autograd path takes device tensors")
    x = x.to(_plan_dtype(fmm)).contiguous()
    out = torch.empty((x.shape[0], n_out, 3), dtype=x.dtype, device=x.device)
    stream = torch.cuda.current_stream(x.device).cuda_stream
    run = fmm.evaluate_adjoint_device if adjoint else fmm.evaluate_device
    for k0 in range(0, x.shape[0], _MAX_BATCH):
        k1 = min(k0 + _MAX_BATCH, x.shape[0])
        run(x[k0:k1].data_ptr(), out[k0:k1].data_ptr(), k1 - k0, stream)
    return out


if torch is not None:

    class _FieldMap(torch.autograd.Function):
        @staticmethod
        def forward(ctx, moments, fmm, n_sources, n_targets):
            ctx.fmm, ctx.n_sources, ctx.n_targets = fmm, n_sources, n_targets
            return _apply_device(fmm, moments, n_targets, adjoint=False)

        @staticmethod
        def backward(ctx, cotangents):
            return _AdjointMap.apply(cotangents, ctx.fmm, ctx.n_sources, ctx.n_targets), None, None, None

    class _AdjointMap(torch.autograd.Function):
        @staticmethod
        def forward(ctx, cotangents, fmm, n_sources, n_targets):
            ctx.fmm, ctx.n_sources, ctx.n_targets = fmm, n_sources, n_targets
            return _apply_device(fmm, cotangents, n_sources, adjoint=True)

        @staticmethod
        def backward(ctx, grad):
            return _FieldMap.apply(grad, ctx.fmm, ctx.n_sources, ctx.n_targets), None, None, None


def dipole_field(fmm, moments):
    """Fields of a HipFull plan as a differentiable torch function of the moments.

    moments: (n_sources, 3) or (K, n_sources, 3) device tensor in physical
    units, user source order; returns (n_targets, 3) or (K, n_targets, 3) in
    the plan precision.  Gradients with respect to the moments come from the
    plan's adjoint pass (the exact transpose of the same FMM operator).
    """
    if torch is None:
        raise RuntimeError("dipole_field needs torch")
    ns, nt = fmm.source_count, fmm.target_count
    single = moments.dim() == 2
    m = moments.unsqueeze(0) if single else moments
    if m.dim() != 3 or m.shape[1] != ns or m.shape[2] != 3:
        raise ValueError("moments must have shape (n_sources, 3) or (batch, n_sources, 3)")
    h = _FieldMap.apply(m, fmm, ns, nt)
    return h[0] if single else h


# ---------------------------------------------------------------------------
# Gradients with respect to positions.  For the loss L = sum_k sum_t
# g_kt . H_k(x_t) of a plan, dL/dx_t = G_k(t) g_kt (the field gradient G is
# symmetric) and, since the dipole kernel is even and symmetric,
# L = sum_k sum_s m_ks . H'_k(x_s) with H' the field of the dipoles g at the
# targets, so dL/dx_s = G'_k(s) m_ks.  Both are exact derivatives of the
# plan's own FMM output with the tree held fixed.

_GRAD_PAIRS = ((0, 1, 2), (1, 3, 4), (2, 4, 5))  # rows of G from its six stored entries


def _apply_gradient_device(fmm, x, n_out, adjoint):
    x = x.to(_plan_dtype(fmm)).contiguous()
    out = torch.empty((x.shape[0], n_out, 6), dtype=x.dtype, device=x.device)
    stream = torch.cuda.current_stream(x.device).cuda_stream
    run = fmm.evaluate_field_gradient_adjoint_device if adjoint else fmm.evaluate_field_gradient_device
    for k0 in range(0, x.shape[0], _MAX_BATCH):
        k1 = min(k0 + _MAX_BATCH, x.shape[0])
        run(x[k0:k1].data_ptr(), out[k0:k1].data_ptr(), k1 - k0, stream)
    return out


def _contract(grad6, v):
    """G v for G stored as (..., 6) xx xy xz yy yz zz and v (..., 3)."""
    return torch.stack([sum(grad6[..., j] * v[..., i] for i, j in enumerate(row)) for row in _GRAD_PAIRS], dim=-1)


def field_gradient(fmm, moments):
    """Field gradient dH_c/dx_j at the targets: (K, n_sources, 3) device
    moments -> (K, n_targets, 6) as xx xy xz yy yz zz, in the plan precision."""
    if torch is None:
        raise RuntimeError("field_gradient needs torch")
    if not moments.is_cuda or moments.dim() != 3 or moments.shape[1:] != (fmm.source_count, 3):
        raise ValueError("moments must be a (batch, n_sources, 3) device tensor")
    return _apply_gradient_device(fmm, moments, fmm.target_count, adjoint=False)


def position_gradients(fmm, moments, cotangents):
    """dL/d(target positions) (n_targets, 3) and dL/d(source positions)
    (n_sources, 3) of L = sum over states and targets of cotangents . fields.

    moments: (K, n_sources, 3) and cotangents: (K, n_targets, 3) device
    tensors in user order.  Where a target and a source are one point (the
    self-identity map of the plan), its position gradient is the sum of both.
    """
    if torch is None:
        raise RuntimeError("position_gradients needs torch")
    ns, nt = fmm.source_count, fmm.target_count
    if not moments.is_cuda or moments.dim() != 3 or moments.shape[1:] != (ns, 3):
        raise ValueError("moments must be a (batch, n_sources, 3) device tensor")
    if not cotangents.is_cuda or cotangents.shape != (moments.shape[0], nt, 3):
        raise ValueError("cotangents must be a (batch, n_targets, 3) device tensor")
    dt = _plan_dtype(fmm)
    g_t = _apply_gradient_device(fmm, moments, nt, adjoint=False)
    g_s = _apply_gradient_device(fmm, cotangents, ns, adjoint=True)
    d_targets = _contract(g_t, cotangents.to(dt)).sum(dim=0)
    d_sources = _contract(g_s, moments.to(dt)).sum(dim=0)
    return d_targets, d_sources
