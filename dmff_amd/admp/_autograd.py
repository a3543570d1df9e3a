"""torch.autograd plumbing shared by the drop-in calculators: an energy whose
gradients the library computed in the same call (value and gradient in one
pass, as the reference's jax.value_and_grad provides them)."""
import numpy as np
import torch

F64 = torch.float64


def as_t(x, dtype=F64):
    """A tensor view of x (tensors pass through unchanged, keeping autograd)."""
    return x if isinstance(x, torch.Tensor) else torch.as_tensor(np.asarray(x), dtype=dtype)


class EnergyWithGrads(torch.autograd.Function):
    """apply(run, *inputs): run(want_grad) -> (E scalar tensor, tuple of dE/d input
    per input, None where not computed).  Backward scales them by the incoming
    gradient; nothing is recomputed."""

    @staticmethod
    def forward(ctx, run, *inputs):
        need = any(ctx.needs_input_grad[1:])
        E, grads = run(need)
        ctx.grads = grads
        ctx.shapes = [x.shape if x is not None else None for x in inputs]
        ctx.devices = [x.device if x is not None else None for x in inputs]
        return E

    @staticmethod
    def backward(ctx, gE):
        out = [None]
        for k, (g, shp, dev) in enumerate(zip(ctx.grads, ctx.shapes, ctx.devices)):
            if g is None or shp is None or not ctx.needs_input_grad[k + 1]:
                out.append(None)
            else:
                out.append((gE.reshape(()).to(g.device) * g).reshape(shp).to(dev))
        return tuple(out)
