"""PyTorch face of the engine's adjoint (gpe_engine_adjoint_forward/_backward).

`OverlapLoss` is a torch.autograd.Function: its forward sets V_ext on an
engine, runs nt real-time steps from psi0 and returns
L = 1 - |<psi_t|psi(T)>|^2 / (|psi_t|^2 |psi(0)|^2); its backward runs the
hand-written discrete adjoint of the split step on the GPU (sqrt(nt)
checkpoints) and returns dL/dV_ext.  `gaussian_bumps` builds a potential from
bump amplitudes with torch ops, so autograd carries dL/dV_ext on to the
amplitudes.  A potential tensor on the GPU crosses the C ABI as device
memory (gpe_engine_set_external_potential_device and
gpe_engine_adjoint_backward_device, after torch's stream has finished
producing it), a CPU tensor as a host array.
"""
import numpy as np
import torch

from . import engine as E


class OverlapLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, V, eng, psi0, psi_target, nt, ckpt=0):
        if V.is_cuda:
            v = V.detach().to(torch.float64).contiguous()
            torch.cuda.current_stream(v.device).synchronize()
            eng.set_external_potential_device(v.data_ptr())
        else:
            eng.set_external_potential(V.detach().to(torch.float64).contiguous().numpy())
        eng.set_psi(psi0)
        loss = eng.adjoint_forward(nt, psi_target, ckpt)
        ctx.eng = eng
        ctx.generation = eng.adjoint_generation
        return torch.tensor(loss, dtype=V.dtype, device=V.device)

    @staticmethod
    def backward(ctx, grad_out):
        # the engine holds the checkpoints of its last forward run only
        if ctx.eng.adjoint_generation != ctx.generation:
            raise RuntimeError("OverlapLoss.backward: the engine ran another adjoint forward after this loss was "
                               "computed; use one engine per live loss")
        if grad_out.is_cuda:
            g = torch.empty(ctx.eng.shape, dtype=torch.float64, device=grad_out.device)
            torch.cuda.current_stream(g.device).synchronize()
            ctx.eng.adjoint_backward_device(g.data_ptr())
            g = g.to(grad_out.dtype)
        else:
            g = torch.from_numpy(ctx.eng.adjoint_backward()).to(dtype=grad_out.dtype)
        return grad_out * g, None, None, None, None, None


def overlap_loss(V, eng, psi0, psi_target, nt, ckpt=0):
    return OverlapLoss.apply(V, eng, psi0, psi_target, nt, ckpt)


def bump_centres(shape, count, seed):
    """Seeded centres (lattice coordinates) of `count` bumps."""
    rng = np.random.default_rng(seed)
    return np.stack([rng.uniform(0, n, count) for n in shape], axis=1)


def gaussian_bumps(amp, centres, width, shape):
    """V(x) = sum_j amp_j exp(-|x - x_j|^2 / (2 width^2)) on the periodic
    lattice (minimum-image distance per axis), separable per bump."""
    dev, dt = amp.device, amp.dtype
    factors = []
    for a, n in enumerate(shape):
        x = torch.arange(n, device=dev, dtype=dt)
        c = torch.as_tensor(centres[:, a], device=dev, dtype=dt)
        d = x[None, :] - c[:, None]
        d = d - n * torch.round(d / n)
        factors.append(torch.exp(-d * d / (2.0 * width * width)))
    gx, gy, gz = factors
    # sum_j a_j gx[j, ix] gy[j, iy] gz[j, iz] as two matrix products
    t = (amp[:, None] * gx).T @ (gy[:, :, None] * gz[:, None, :]).reshape(len(amp), -1)
    return t.reshape(shape)
