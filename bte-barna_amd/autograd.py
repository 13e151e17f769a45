"""torch.autograd bridge for the specularity gradient of RTAMC2D.

forward runs one GPU RTAMC2D run with the edge specularities p (4 sides:
left, bottom, top, right) and returns the temperature T(t, x); backward
contracts the upstream gradient with the score-function tallies dT/dp that
the same run accumulated (only sides with 0 < p < 1 get a derivative)."""
import numpy as np
import torch


class _SpecularityT(torch.autograd.Function):
    @staticmethod
    def forward(fctx, p, model):
        T, dTdp = model.run(p.detach().to("cpu", torch.float64).numpy())
        fctx.save_for_backward(torch.from_numpy(dTdp).to(p.device))
        return torch.from_numpy(T).to(p.device)

    @staticmethod
    def backward(fctx, gT):
        (dTdp,) = fctx.saved_tensors
        gp = (gT.unsqueeze(0).to(dTdp.dtype) * dTdp).sum(dim=(1, 2))
        return gp, None


class SpecularityModel:
    """T(p) for one geometry and mode table; call with a (4,) tensor p."""

    def __init__(self, pkg, ctx, modes, geom, T0, nparticles, thickness, dt, maxtime, seed=1, sigma_bounce=0.25,
                 ballistic=False):
        self.pkg, self.ctx, self.modes, self.geom = pkg, ctx, modes, geom
        self.kw = dict(thickness=thickness, dt=dt, maxtime=maxtime, sigma_bounce=sigma_bounce)
        self.run_kw = dict(geom=geom, T0=T0, nparticles=nparticles, seed=seed, ballistic=ballistic)
        self.last = None

    def run(self, p):
        p = np.asarray(p, dtype=np.float64)
        b = self.pkg.RtaBatch(self.ctx, self.modes, [dict(self.run_kw, specularity=tuple(p), grad=True)],
                              **self.kw)
        try:
            (out,), rep = b.execute(grad=True)
        finally:
            b.close()
        self.last = rep
        return out["T"], out["dTdp"]

    def __call__(self, p):
        return _SpecularityT.apply(p, self)
