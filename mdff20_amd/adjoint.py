"""torch.autograd.Function over the batched force-matching adjoint.

ForceMatchLoss.apply(theta, batch, spec, fref) returns the loss
sum |F - F_ref|^2 of every configuration of `batch` at the parameters
spec.params(theta); its backward is the hand-written adjoint of
mdff_batch_adjoint (no autograd tape, no stored SCF iterations).  PyTorch is
used only for this wrapper: theta is a small CPU float64 vector.
"""
import numpy as np

try:
    import torch
except ImportError:  # pragma: no cover - torch is part of the image
    torch = None

__all__ = ["ForceMatchLoss", "fit"]

if torch is not None:
    class ForceMatchLoss(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta, batch, spec, fref):
            th = theta.detach().cpu().numpy().astype(np.float64)
            batch.set_params(spec.params(th))
            res = batch.adjoint(fref)
            g = spec.gather(res, th)
            ctx.save_for_backward(torch.as_tensor(g, dtype=theta.dtype, device=theta.device))
            ctx.info = res
            return theta.new_tensor(res["loss"])

        @staticmethod
        def backward(ctx, gout):
            (g,) = ctx.saved_tensors
            return gout * g, None, None, None


def fit(batch, spec, fref, steps=50, lr=0.02, theta0=None, log=None):
    """Adam on theta (log parameters by default) through ForceMatchLoss;
    returns theta and the loss history"""
    if torch is None:
        raise RuntimeError("torch is required for fit()")
    th0 = spec.values() if theta0 is None else np.asarray(theta0, dtype=np.float64)
    theta = torch.tensor(th0, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([theta], lr=lr)
    hist = []
    for k in range(steps):
        opt.zero_grad()
        loss = ForceMatchLoss.apply(theta, batch, spec, fref)
        loss.backward()
        opt.step()
        hist.append(float(loss.item()))
        if log:
            log(k, hist[-1])
    return theta.detach().numpy(), hist
