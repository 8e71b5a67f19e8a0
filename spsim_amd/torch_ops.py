"""PyTorch autograd front end of the structure factor and its adjoint.

`intensity(pos, atom_b, ...)` returns I = |F|^2 on a q list, by the direct
sum (spsim_hip_sf_pattern) or, with method="nufft", by the non-uniform FFT
(spsim_hip_sf_nfft, per-atom B by Gaussian spreading) to a stated accuracy.
Its backward is the hand-written adjoint kernel (spsim_hip_sf_adjoint), which
recomputes phases per pixel chunk instead of storing atom x pixel terms and
takes F from the forward.
`poisson_nll(I, y, K)` is the Poisson negative log-likelihood
sum(K I - y log(K I)) whose forward and gradient come from
spsim_hip_poisson_nll. All tensors live on the GPU; only pointers and the
stream cross into libspsim_amd.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from . import device as dv
from ._lib import check


class SpeciesLayout:
    """Species sort of a fixed list of atomic numbers: perm[packed] = original,
    segments of the packed order, and the amplitude per packed atom."""

    def __init__(self, Z, device, amp=None, delta_atoms=False):
        Z = np.ascontiguousarray(Z, np.int32)
        n = Z.shape[0]
        packed = dv.pack_atoms(Z, np.zeros((n, 3), np.float32), amp=amp, atom_b=None, delta_atoms=delta_atoms)
        self.natoms = packed.natoms
        self.segs_host = np.ascontiguousarray(packed.segs, np.int32)
        self.nsegs = int(self.segs_host.shape[0])
        self.segs = torch.from_numpy(self.segs_host if self.nsegs else np.zeros((1, 4), np.int32)).to(device)
        self.perm = torch.from_numpy(packed.perm.astype(np.int64)).to(device)
        self.amp = torch.from_numpy(packed.atoms[:, 3].copy() if self.natoms else np.zeros(0, np.float32)).to(device)


def _sp():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _p(t):
    return None if t is None else C.c_void_p(t.data_ptr())


class _Intensity(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, atom_b, hkl, layout, ff, b_global, tile, method, eps):
        packed_pos = pos.detach().float()[layout.perm].contiguous()
        atoms = torch.cat([packed_pos, layout.amp[:, None]], dim=1).contiguous()
        pb = None if atom_b is None else atom_b.detach().float()[layout.perm].contiguous()
        npix = int(hkl.shape[0])
        F = torch.empty((npix, 2), dtype=torch.float32, device=hkl.device)
        I = torch.empty((npix,), dtype=torch.float32, device=hkl.device)
        if method == "nufft":
            check(_lib.load().spsim_hip_sf_nfft(_p(hkl), npix, _p(atoms), _p(pb), _p(layout.segs), layout.nsegs,
                                                layout.natoms, _p(ff), float(b_global), None, float(eps), _p(F),
                                                _p(I), None, _sp()),
                  "spsim_hip_sf_nfft")
        elif method == "direct":
            check(_lib.load().spsim_hip_sf_pattern(_p(hkl), npix, _p(atoms), _p(pb), _p(layout.segs), layout.nsegs,
                                                   layout.natoms, _p(ff), float(b_global), None, _p(F), _p(I),
                                                   C.byref(tile) if tile is not None else None, _sp()),
                  "spsim_hip_sf_pattern")
        else:
            raise ValueError(f"method must be 'direct' or 'nufft', not {method!r}")
        ctx.save_for_backward(atoms, pb if pb is not None else torch.empty(0, device=hkl.device), F)
        ctx.has_b = pb is not None
        ctx.want_b = atom_b is not None and atom_b.requires_grad
        ctx.hkl, ctx.layout, ctx.ff, ctx.b_global = hkl, layout, ff, b_global
        ctx.n_orig = int(pos.shape[0])
        return I

    @staticmethod
    def backward(ctx, gI):
        atoms, pb, F = ctx.saved_tensors
        layout = ctx.layout
        gI = gI.contiguous().float()
        n = layout.natoms
        gpos_p = torch.empty((max(n, 1), 3), dtype=torch.float32, device=atoms.device)
        gb_p = torch.empty((max(n, 1),), dtype=torch.float32, device=atoms.device) if ctx.want_b else None
        segs = layout.segs_host
        check(_lib.load().spsim_hip_sf_adjoint(_p(ctx.hkl), int(ctx.hkl.shape[0]), _p(atoms),
                                               _p(pb) if ctx.has_b else None,
                                               segs.ctypes.data_as(C.c_void_p), layout.nsegs, n, _p(ctx.ff),
                                               float(ctx.b_global), _p(F), _p(gI), _p(gpos_p), _p(gb_p), _sp()),
              "spsim_hip_sf_adjoint")
        gpos = torch.zeros((ctx.n_orig, 3), dtype=torch.float32, device=atoms.device)
        if n:
            gpos[layout.perm] = gpos_p[:n]
        gb = None
        if ctx.want_b:
            gb = torch.zeros((ctx.n_orig,), dtype=torch.float32, device=atoms.device)
            if n:
                gb[layout.perm] = gb_p[:n]
        return gpos, gb, None, None, None, None, None, None, None


def intensity(pos: torch.Tensor, hkl: torch.Tensor, layout: SpeciesLayout, ff: torch.Tensor,
              atom_b: torch.Tensor | None = None, b_global: float = 0.0, tile=None, method: str = "direct",
              eps: float = 1e-6) -> torch.Tensor:
    """I = |F|^2 on hkl for atoms at pos [N, 3] (metres), differentiable in pos and
    atom_b. method "direct" sums every term; "nufft" evaluates F by the
    non-uniform FFT at accuracy eps (B factors must then be >= 0), and the
    backward uses that F."""
    return _Intensity.apply(pos, atom_b, hkl, layout, ff, b_global, tile, method, eps)


class _PoissonNLL(torch.autograd.Function):
    @staticmethod
    def forward(ctx, I, y, K):
        I = I.contiguous().float()
        npix = int(I.shape[0])
        loss = torch.empty((1,), dtype=torch.float64, device=I.device)
        gI = torch.empty_like(I)
        check(_lib.load().spsim_hip_poisson_nll(_p(I), _p(y), _p(K), npix, _p(loss), _p(gI), _sp()),
              "spsim_hip_poisson_nll")
        ctx.save_for_backward(gI)
        return loss[0]

    @staticmethod
    def backward(ctx, g):
        (gI,) = ctx.saved_tensors
        return gI * g.to(gI.dtype), None, None


def poisson_nll(I: torch.Tensor, y: torch.Tensor, K: torch.Tensor | None = None) -> torch.Tensor:
    """sum_i (K_i I_i - y_i log(K_i I_i)) in FP64 (K = 1 when None)."""
    y = y.contiguous().float()
    K = None if K is None else K.contiguous().float()
    return _PoissonNLL.apply(I, y, K)
