"""Backscatter yield and exit-energy histogram as differentiable functions.

``backscatter(params, spec)`` returns the backscatter yield Y (fraction of
electrons with outcome 'bks') and the histogram H (fraction of electrons that
backscatter with exit energy in each of ``spec.n_bins`` equal bins over
[0, E0]) of one transport configuration, as torch tensors, for
``params = [Wc, s_tilt, J, density_scale]``.  The forward pass is the
transport kernel itself (``mcd_forward``); ``backward`` calls the hand-written
replay kernel (``mcd_backward``) with the upstream gradients, which
regenerates every electron from its Philox counters and returns
dL/d(Wc, s_tilt, J, density_scale).

Parameter meaning (the reference's run parameters, fileTools.readInput):
Wc [eV] the Moller cut-off (DS only); s_tilt [degree] the sample tilt; J [eV]
the Bethe mean ionisation energy (continuous classical/JL only); density_scale
multiplies the tabulated mass density, moving n, E_F and E_pl together as the
reference's material density does (material.py:127-131).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

from . import _abi
from .engine import Transport, backscatter_stats, check_guard
from .material import Material, bethe_J


@dataclass
class BackscatterSpec:
    """A differentiable transport workload (everything but the parameters)."""

    material: str = "Cu"
    E0: float = 20000.0
    Emin: float = 1000.0
    mode: str = "DS"
    bethe_model: str = "JL"
    n: int = 1 << 24
    seed: int = 4
    n_bins: int = 64
    tol_E: float = 1e-3
    tol_W: float = 1e-3
    device: int = 0
    chunk: int = 1 << 22          # electrons per forward launch (bounds the record buffer)
    rel_step: float = 0.0         # threshold-term parameter step (0: library default 1e-2)
    graze: float = 0.0            # grazing cutoff (0: library default 0.005)
    _cache: dict = field(default_factory=dict, repr=False)

    def params(self, Wc=50.0, s_tilt=70.0, J=None, density_scale=1.0, dtype=torch.float64):
        J = bethe_J(Material(self.material).Z) if J is None else J
        return torch.tensor([Wc, s_tilt, J, density_scale], dtype=dtype)

    def transport(self, p) -> Transport:
        """The Transport (device tables included) at parameters p, cached."""
        key = tuple(float(v) for v in p)
        tr = self._cache.get(key)
        if tr is None:
            Wc, s_tilt, J, scale = key
            mat = Material(self.material, density_scale=scale)
            tr = Transport(mat, self.E0, self.Emin, Wc=Wc, s_tilt=s_tilt, mode=self.mode,
                           bethe_model=self.bethe_model, tol_E=self.tol_E, tol_W=self.tol_W, J=J,
                           device=self.device)
            if len(self._cache) >= 8:
                self._cache.pop(next(iter(self._cache)))
            self._cache[key] = tr
        return tr

    def forward_stats(self, p, n=None, first=0):
        """Exact transport of electrons [first, first+n): (backscatter count,
        histogram counts) as device int64 tensors."""
        n = self.n if n is None else n
        tr = self.transport(p)
        dev = torch.device("cuda", self.device)
        m = min(self.chunk, n)
        els = torch.empty(m * _abi.ELECTRON_DTYPE.itemsize, dtype=torch.uint8, device=dev)
        cnt = torch.zeros(_abi.NCOUNTERS, dtype=torch.int64, device=dev)
        total = torch.zeros((), dtype=torch.int64, device=dev)
        hist = torch.zeros(self.n_bins, dtype=torch.int64, device=dev)
        done = 0
        while done < n:
            k = min(m, n - done)
            tr.launch(k, els, cnt, seed=self.seed, first=first + done)
            c, h = backscatter_stats(els, k, self.E0, self.n_bins)
            total += c
            hist += h
            done += k
        check_guard(cnt.cpu().numpy().view(np.uint64), "forward_stats")  # one check over every chunk
        return total, hist


class BackscatterFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, spec):
        p = params.detach().to("cpu", torch.float64).tolist()
        total, hist = spec.forward_stats(p)
        ctx.spec = spec
        ctx.p = p
        ctx.save_for_backward(params)
        Y = (total.double() / spec.n).to(device=params.device, dtype=params.dtype)
        H = (hist.double() / spec.n).to(device=params.device, dtype=params.dtype)
        return Y, H

    @staticmethod
    def backward(ctx, gY, gH):
        (params,) = ctx.saved_tensors
        spec, p = ctx.spec, ctx.p
        g = torch.zeros(1 + spec.n_bins, dtype=torch.float32)
        if gY is not None:
            g[0] = float(gY)
        if gH is not None:
            g[1:] = gH.detach().to("cpu", torch.float32).reshape(-1)
        tr = spec.transport(p)
        out = tr.vjp(spec.n, g, seed=spec.seed, rel_step=spec.rel_step, graze=spec.graze)
        grads = np.array(out[1:], dtype=np.float64)
        grads[_abi.GRAD_LNRHO] /= p[3]  # d/d scale = d/d ln(rho) / scale
        return torch.as_tensor(grads, dtype=params.dtype, device=params.device), None


def backscatter(params: torch.Tensor, spec: BackscatterSpec):
    """(Y, H) of the configuration at params = [Wc, s_tilt, J, density_scale]."""
    return BackscatterFunction.apply(params, spec)


def jacobian(spec: BackscatterSpec, params, n=None, first=0):
    """Replay estimate of d(Y, H)/d(Wc, s_tilt, J, density_scale) per electron:
    (values[1 + n_bins], jac[1 + n_bins, 4]); row 0 is Y, row 1+b is H_b."""
    p = [float(v) for v in params]
    tr = spec.transport(p)
    n = spec.n if n is None else n
    J = tr.jacobian(n, seed=spec.seed, first=first, n_bins=spec.n_bins, rel_step=spec.rel_step,
                    graze=spec.graze)
    jac = J[:, 1:].copy()
    jac[:, _abi.GRAD_LNRHO] /= p[3]
    return J[:, 0].copy(), jac


__all__ = ["BackscatterSpec", "BackscatterFunction", "backscatter", "jacobian"]
