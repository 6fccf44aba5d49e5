"""torch.autograd over the C ABI: P_l(k) of a periodic catalogue as a
differentiable function of the particle positions, line-of-sight velocities
and weights (used for checking the adjoint; the product is the C ABI).

The raw bins (capow_power) and their backward (capow_backward) run on the
GPU; the host-side normalisation of the reference (src/main.c:103-148 for
shot noise and N^2/V, src/main.c:1014-1056 for the (2l+1)/(N_modes norm)
scaling, shot subtraction and the zeroed last bin) is restated here in torch
so that its dependence on the weights (N = sum w, moments sum w v^n) is
differentiated by autograd.  Every particle is taken as accepted by the main
deposit (in-box catalogues with LOS = 0).
"""
import numpy as np
import torch

from . import api


class _RawBins(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xyzv, w, ps):
        # the forward keeps its sort and transform: a backward that follows
        # with nothing run on `ps` in between reuses them
        ps.keep_forward(True)
        ps.set_particles(xyzv.detach().contiguous(), None if w is None else w.detach().contiguous())
        ctx.generation = ps.generation
        res = ps.power()
        ns = len(res["codes"])
        raw = np.stack([res[c]["pk"] for c in res["codes"]])            # [ns][5][nk]
        ctx.ps = ps
        ctx.save_for_backward(xyzv, w if w is not None else torch.empty(0, device=xyzv.device))
        ctx.has_w = w is not None
        ctx.nmodes = res["nmodes"].copy()
        out = torch.from_numpy(raw).to(torch.float64)
        return out, torch.from_numpy(ctx.nmodes.astype(np.float64))

    @staticmethod
    def backward(ctx, graw, _gnm):
        xyzv, w = ctx.saved_tensors
        ps = ctx.ps
        if ps.generation != ctx.generation:
            # other particles were set since: recompute on these
            ps.set_particles(xyzv.detach().contiguous(), w.detach().contiguous() if ctx.has_w else None)
        gx, gw = ps.backward(graw.detach().cpu().numpy(), want_w=True)
        if not ctx.has_w:
            # unit weights: the gradient with respect to them is still returned
            # through the weight-free path (no tensor to attach it to)
            gw = None
        return gx.to(xyzv.dtype), gw, None


def raw_bins(ps, xyzv, w=None):
    """Raw binned sums [n_spectra][5][nk] (float64, CPU) and mode counts."""
    return _RawBins.apply(xyzv, w, ps)


def multipoles(ps, xyzv, w=None):
    """Normalised P_l(k) per spectrum code: dict code -> [5][nk] tensor
    (float64, CPU), differentiable in xyzv and w."""
    p = ps.p
    raw, nm = raw_bins(ps, xyzv, w)
    ww = w if w is not None else torch.ones(xyzv.shape[0], dtype=xyzv.dtype, device=xyzv.device)
    v = xyzv[:, 3].to(torch.float64)
    ww64 = ww.to(torch.float64)
    N = ww64.sum().cpu()
    vol = (p.xmax - p.xmin) * (p.ymax - p.ymin) * (p.zmax - p.zmin)
    norm = N * N / vol
    mom = [(ww64 * v ** k).sum().cpu() / N for k in range(5)]
    out = {}
    nmv = nm
    filled = (nmv > 0)
    last = int(torch.nonzero(filled).max()) if filled.any() else -1
    for s, code in enumerate(ps.codes):
        if code == 0:
            shot = N
        else:
            e = {1: 1, 2: 2, 3: 2, 4: 3, 6: 3}.get(code, 4)
            shot = mom[e] * N
        nmc = torch.where(filled, nmv, torch.ones_like(nmv))
        P = []
        for l, f in ((0, 1.0), (1, 3.0), (2, 5.0), (3, 7.0), (4, 9.0)):
            r = raw[s, l]
            if l == 0:
                val = (r - nmc * shot) / (nmc * norm)
            else:
                val = f * r / (nmc * norm)
            val = torch.where(filled, val, torch.zeros_like(val))
            if last >= 0:
                keep = torch.ones_like(val)
                keep[last] = 0.0
                val = val * keep
            P.append(val)
        out[code] = torch.stack(P)
    return out
