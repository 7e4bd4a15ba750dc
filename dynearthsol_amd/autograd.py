"""torch.autograd binding of the differentiated run (des_adj_forward /
des_adj_backward in include/des_amd.h).

    topo = SurfaceTopography(model, nsteps=2000, window=100)
    uz = topo(log10_visc, cohesion)          # top-surface vertical displacement
    loss = ((uz - target) ** 2).sum()
    loss.backward()                           # gradients on both inputs

The forward restarts the model from its initial state with the given
per-element parameters, runs nsteps fused FP64 steps on the device with a
checkpoint every `window` steps, and returns the vertical displacement of the
top-surface nodes (bcflag & BOUNDZ1).  The backward recomputes each window
from its checkpoint (window=0: keeps every step's records from the forward
and recomputes nothing) and runs the hand-written reverse kernels.  Torch tensors
carry the parameters and gradients; the model state stays in the library's
device buffers.
"""
import math

import numpy as np
import torch

from .lib import DesModel

BOUNDZ1 = 32


class _Run(torch.autograd.Function):
    @staticmethod
    def forward(ctx, log_visc, cohesion, topo):
        g = topo.g
        visc = 10.0 ** log_visc.detach().double().cpu().numpy()
        coh = cohesion.detach().double().cpu().numpy()
        g.reset()
        g.set_elem_params(elem_visc=visc, elem_cohesion=coh)
        g.adj_forward(topo.nsteps, topo.window)
        topo.runs += 1
        z = g.download(["coord"])["coord"][topo.top, -1]
        uz = z - topo.z0
        ctx.topo = topo
        ctx.run = topo.runs
        ctx.visc = visc
        ctx.device = log_visc.device
        return torch.as_tensor(uz, dtype=torch.float64, device=log_visc.device)

    @staticmethod
    def backward(ctx, grad_uz):
        topo = ctx.topo
        if ctx.run != topo.runs:
            # the device holds the checkpoints (or tape) of the latest forward only
            raise RuntimeError("SurfaceTopography: another forward ran after this graph's; its device "
                               "checkpoints are gone (run the backward before the next forward, or use "
                               "one SurfaceTopography per graph)")
        cb = np.zeros((topo.g.N, topo.g.nd))
        cb[topo.top, -1] = grad_uz.detach().double().cpu().numpy()
        vb, hb = topo.g.adj_backward(cb)
        g_log = torch.as_tensor(vb * ctx.visc * math.log(10.0), dtype=torch.float64, device=ctx.device)
        g_coh = torch.as_tensor(hb, dtype=torch.float64, device=ctx.device)
        return g_log, g_coh, None


class SurfaceTopography:
    """Differentiable map (log10 viscosity, cohesion) per element -> vertical
    displacement of the top-surface nodes after nsteps steps."""

    def __init__(self, model, nsteps, window=100):
        self.model = model
        self.g = DesModel(model, init=False)
        self.nsteps = nsteps
        self.window = window
        self.top = np.nonzero(model.bcflag & BOUNDZ1)[0]
        self.z0 = model.coord[self.top, -1].copy()
        self.runs = 0          # forwards so far; a graph's backward needs the latest one

    def __call__(self, log_visc, cohesion):
        return _Run.apply(log_visc, cohesion, self)
