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
and recomputes nothing) and runs the hand-written reverse kernels.

With tensors on the GPU the parameters, the final coordinates, the loss
weights and the gradients pass between torch and the library as device
pointers (des_upload_fields_device, des_download_fields_device,
des_adj_backward_device): nothing crosses to the host.  CPU tensors go
through host copies.  The model state stays in the library's device buffers.
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
        dev = log_visc.device
        g.reset()
        if log_visc.is_cuda:
            visc = torch.pow(10.0, log_visc.detach().double()).contiguous()
            coh = cohesion.detach().double().contiguous()
            torch.cuda.current_stream(dev).synchronize()          # the inputs are written
            g.upload_device(elem_visc=visc.data_ptr(), elem_cohesion=coh.data_ptr())
        else:
            visc = 10.0 ** log_visc.detach().double()
            coh = cohesion.detach().double()
            g.set_elem_params(elem_visc=visc.numpy(), elem_cohesion=coh.numpy())
        g.adj_forward(topo.nsteps, topo.window)
        topo.runs += 1
        if log_visc.is_cuda:
            coord = torch.empty((g.N, g.nd), dtype=torch.float64, device=dev)
            g.download_device(coord=coord.data_ptr())
            top = topo.top_t(dev)
            uz = coord[top, -1] - topo.z0_t(dev)
        else:
            z = g.download(["coord"])["coord"][topo.top, -1]
            uz = torch.as_tensor(z - topo.z0, dtype=torch.float64)
        ctx.topo = topo
        ctx.run = topo.runs
        ctx.visc = visc
        return uz

    @staticmethod
    def backward(ctx, grad_uz):
        topo = ctx.topo
        if ctx.run != topo.runs:
            # the device holds the checkpoints (or tape) of the latest forward only
            raise RuntimeError("SurfaceTopography: another forward ran after this graph's; its device "
                               "checkpoints are gone (run the backward before the next forward, or use "
                               "one SurfaceTopography per graph)")
        visc = ctx.visc
        g = topo.g
        if grad_uz.is_cuda:
            dev = grad_uz.device
            cb = torch.zeros((g.N, g.nd), dtype=torch.float64, device=dev)
            cb[topo.top_t(dev), -1] = grad_uz.detach().double()
            vb = torch.empty(g.E, dtype=torch.float64, device=dev)
            hb = torch.empty(g.E, dtype=torch.float64, device=dev)
            torch.cuda.current_stream(dev).synchronize()
            g.adj_backward_device(cb.data_ptr(), vb.data_ptr(), hb.data_ptr())
        else:
            cb = np.zeros((g.N, g.nd))
            cb[topo.top, -1] = grad_uz.detach().double().numpy()
            vbn, hbn = g.adj_backward(cb)
            vb, hb = torch.as_tensor(vbn), torch.as_tensor(hbn)
        return vb * visc * math.log(10.0), hb, None


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
        self._dev = {}

    def top_t(self, dev):
        if (dev, "top") not in self._dev:
            self._dev[(dev, "top")] = torch.as_tensor(self.top, dtype=torch.long, device=dev)
        return self._dev[(dev, "top")]

    def z0_t(self, dev):
        if (dev, "z0") not in self._dev:
            self._dev[(dev, "z0")] = torch.as_tensor(self.z0, dtype=torch.float64, device=dev)
        return self._dev[(dev, "z0")]

    def __call__(self, log_visc, cohesion):
        return _Run.apply(log_visc, cohesion, self)
