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
