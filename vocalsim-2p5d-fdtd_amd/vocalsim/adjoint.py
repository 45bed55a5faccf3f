"""Differentiable tube simulation for PyTorch.

:class:`TubeProbe` is a ``torch.autograd.Function`` whose forward runs the
differentiable forward pass of a batch of tubes on the GPU
(``vs_adjoint_forward``: the FDTD time step of FDTDEngineCPU.cpp:316-410 in
flux form, with a state checkpoint every ``checkpoint_every`` steps) and whose
backward runs the hand-written reverse-time kernels
(``vs_adjoint_backward``): each checkpointed segment is recomputed, reversed
step by step, and the per-cell coefficient gradients are chained to the
per-section inputs.

Inputs are ``[B, S]`` tensors of section areas [cm^2], depth scales and wall
admittances (mu3d); the output is the ``[B, N]`` probe pressure of every tube
(all tubes of one call share S sections and N steps).  The library stores the
inputs as float32 (the reference's float32 area function,
TubeGeometry.cu:83-88), so the gradient is taken at the float32-rounded
values.  The integer geometry (diameters in cells) is held fixed.
"""
from __future__ import annotations

from typing import Sequence

import numpy as np
import torch

from .solver import Solver
from .tubes import TubeSpec


_last_backward_timing = (0.0, 0.0)


def last_backward_timing():
    """(device reverse pass ms, host chain ms) of the last TubeProbe backward."""
    return _last_backward_timing


class TubeProbe(torch.autograd.Function):
    @staticmethod
    def forward(ctx, area, depth_scale, mu3d, specs: Sequence[TubeSpec], precision: int = 64,
                checkpoint_every: int = 512, threads: int = 1024):
        if area.dim() != 2 or area.shape != depth_scale.shape or area.shape != mu3d.shape:
            raise ValueError("area, depth_scale and mu3d must be [B, S] tensors of one shape")
        B, S = area.shape
        if len(specs) != B:
            raise ValueError(f"need {B} tube specs, got {len(specs)}")
        a = area.detach().to("cpu", torch.float64).numpy()
        d = depth_scale.detach().to("cpu", torch.float64).numpy()
        m = mu3d.detach().to("cpu", torch.float64).numpy()
        run = []
        for b, sp in enumerate(specs):
            if sp.num_sections != S:
                raise ValueError(f"tube {b} has {sp.num_sections} sections, expected {S}")
            run.append(sp.replace(areas_cm2=a[b].astype(np.float32), depth_scale=d[b].astype(np.float32),
                                  mu3d=m[b].astype(np.float32)))
        solver = Solver(run, precision=precision, threads=threads)
        steps = {d["num_steps"] for d in solver.dims}
        if len(steps) != 1:
            solver.close()
            raise ValueError("all tubes of one call must have the same number of steps")
        N = steps.pop()
        ctx.solver = solver
        ctx.meta = (area.device, area.dtype)
        if area.is_cuda:
            # probe traces written by the library straight into device memory
            out = torch.empty((B, N), dtype=torch.float64, device=area.device)
            solver.adjoint_forward_to(out, checkpoint_every)
            return out
        return torch.from_numpy(np.stack(solver.adjoint_forward(checkpoint_every)))

    @staticmethod
    def backward(ctx, gprobe):
        device, dtype = ctx.meta
        g = gprobe.detach().to(torch.float64).contiguous()
        if g.is_cuda:
            ga, gs, gm = ctx.solver.adjoint_backward_from(g)
        else:
            g = g.numpy()
            ga, gs, gm = ctx.solver.adjoint_backward([g[b] for b in range(g.shape[0])])
        global _last_backward_timing
        _last_backward_timing = ctx.solver.adjoint_timing()
        ctx.solver.close()
        ctx.solver = None
        to = lambda x: torch.from_numpy(np.stack(x)).to(device=device, dtype=dtype)
        return to(ga), to(gs), to(gm), None, None, None, None


def tube_probe(area: torch.Tensor, depth_scale: torch.Tensor, mu3d: torch.Tensor, specs: Sequence[TubeSpec],
               precision: int = 64, checkpoint_every: int = 512, threads: int = 1024) -> torch.Tensor:
    """Probe pressure [B, N] of the tubes ``specs`` with their area function,
    depth scale and wall admittance replaced by the given [B, S] tensors;
    differentiable with respect to all three."""
    return TubeProbe.apply(area, depth_scale, mu3d, list(specs), precision, checkpoint_every, threads)


def dft_basis(n: int, fs: float, freqs, device="cuda"):
    """cos / sin bases [n, F] in float64 for X(f) = sum_k x[k] exp(-i 2 pi f k / fs).
    The phase f k is reduced modulo fs (one cycle) before the trigonometric
    call; the reduction is exact when f is an integer number of Hz."""
    f = torch.as_tensor(np.asarray(freqs, dtype=np.float64), device=device)
    k = torch.arange(n, dtype=torch.float64, device=device)
    ph = torch.remainder(k[:, None] * f[None, :], fs) * (2.0 * np.pi / fs)
    return torch.cos(ph), torch.sin(ph)


def log_magnitude(x: torch.Tensor, basis) -> torch.Tensor:
    """ln |X(f)| of signals x [B, n] at the frequencies of ``basis`` (dft_basis)."""
    c, s = basis
    re = x @ c
    im = x @ s
    return 0.5 * torch.log(re * re + im * im)
