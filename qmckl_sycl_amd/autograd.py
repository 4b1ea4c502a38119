"""torch.autograd.Function harness around the C-ABI adjoint.

``mo_vgl(nucl_coord, exponent, mo_coef, system)`` returns mo_vgl [P][5][M]
(FP64) for ``system`` with its nuclear coordinates [N][3], primitive
exponents [Np] and MO coefficients [M][A] replaced by the given tensors; its
backward calls ``qmckl_get_mo_basis_mo_vgl_adjoint_device`` with the
incoming gradient as the weight array, so ``torch.autograd.grad`` of any
scalar built from mo_vgl yields dL/dR, dL/dalpha (at fixed normalised
contraction coefficients) and dL/dC from the hand-written adjoint kernels.

The forward builds a fresh context from the tensors' values (the qmckl
nucleus setters are set-once); the context is kept for the backward pass.
"""
import dataclasses

import numpy as np
import torch

from .context import Context


class MoVgl(torch.autograd.Function):
    @staticmethod
    def forward(ctx, nucl_coord, exponent, mo_coef, system, device_id=0):
        s = dataclasses.replace(
            system,
            nucl_coord=np.ascontiguousarray(nucl_coord.detach().cpu().numpy().T, dtype=np.float64),
            exponent=np.ascontiguousarray(exponent.detach().cpu().numpy(), dtype=np.float64),
            mo_coef=np.ascontiguousarray(mo_coef.detach().cpu().numpy(), dtype=np.float64))
        q = Context(device_id)
        q.set_system(s)
        mo = torch.from_numpy(q.mo_vgl(s.point_num))
        ctx.qmckl = q
        ctx.shapes = (tuple(nucl_coord.shape), tuple(exponent.shape), tuple(mo_coef.shape))
        ctx.devices = (nucl_coord.device, exponent.device, mo_coef.device)
        ctx.device_id = device_id
        return mo.to(nucl_coord.device)

    @staticmethod
    def backward(ctx, grad_out):
        q = ctx.qmckl
        dev = torch.device("cuda", ctx.device_id)
        w = grad_out.detach().to(device=dev, dtype=torch.float64).contiguous()
        shp_r, shp_a, shp_c = ctx.shapes
        dR = torch.empty(shp_r, dtype=torch.float64, device=dev)
        dA = torch.empty(shp_a, dtype=torch.float64, device=dev)
        dC = torch.empty(shp_c, dtype=torch.float64, device=dev)
        torch.cuda.synchronize(dev)   # w is complete before the library's stream reads it
        q.call("qmckl_get_mo_basis_mo_vgl_adjoint_device", w.data_ptr(), w.numel(),
               dR.data_ptr(), dA.data_ptr(), dC.data_ptr())
        q.destroy()
        ctx.qmckl = None
        d0, d1, d2 = ctx.devices
        return dR.to(d0), dA.to(d1), dC.to(d2), None, None


def mo_vgl(nucl_coord, exponent, mo_coef, system, device_id=0):
    return MoVgl.apply(nucl_coord, exponent, mo_coef, system, device_id)
