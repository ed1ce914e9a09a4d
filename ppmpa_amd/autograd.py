"""torch.autograd.Function over the PPM-PA passes: forward renders the mean
image of K passes for given Lambertian albedos and light scale; backward
replays the passes through the adjoint kernels (ppm_session_backward).

  render = Renderer(ctx, dscene, cfg, n_pass, n_params=32)
  img = render(albedo, light_scale)      # albedo: (n_params, 3) float64 tensor
  loss = ((img - target) ** 2).mean(); loss.backward()

The parameters are materials 0..n_params-1 of the scene (their c1); every
call rewinds the renderer's session and runs it with the same seed, so
forward and backward see the same random numbers.  The reference has no counterpart (it is not
differentiable); this is the forward/backward pair of the differentiable
form.  Tensors live on the host: the image and the gradients cross the
C ABI as double arrays.
"""
import numpy as np
import torch

from .tracer import Session


class _RenderFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, albedo, light_scale, renderer):
        r = renderer
        a = albedo.detach().cpu().numpy().astype(np.float64)
        s = float(light_scale.detach().cpu().item())
        r.dscene.set_params(albedo=a, survival=r.survival(a), light_scale=s)
        sess = r.session()
        sess.run(r.n_pass)
        mean, _, _ = sess.image(with_var=False)
        ctx.renderer = r
        ctx.generation = r.generation
        return torch.from_numpy(mean.copy())

    @staticmethod
    def backward(ctx, grad_out):
        r = ctx.renderer
        if ctx.generation != r.generation:
            raise RuntimeError("backward of a render whose session has since rendered again")
        ga, gl = r._session.backward(grad_out.detach().cpu().numpy().astype(np.float64))
        n = r.n_params
        return torch.from_numpy(ga[:n].copy()), torch.tensor(gl, dtype=torch.float64), None


class Renderer:
    """The forward/backward pair for one scene and pass schedule.

    survival: None (q = the current albedo: the reference's estimator) or an
    (n_params, 3) array held fixed, which makes the rendered image a smooth
    function of the albedos for fixed random numbers (for finite-difference
    checks)."""

    def __init__(self, ctx, dscene, cfg, n_pass, n_params=32, survival=None):
        self.ctx, self.dscene, self.cfg, self.n_pass, self.n_params = ctx, dscene, cfg, int(n_pass), int(n_params)
        self.fixed_q = None if survival is None else np.asarray(survival, np.float64).reshape(n_params, 3)
        self._session = None
        self.generation = 0

    def session(self):
        """The renderer's session, rewound (one set of device buffers for
        every forward; a backward must come before the next forward)."""
        if self._session is None:
            self._session = Session(self.ctx, self.dscene, self.cfg)
        else:
            self._session.reset()
        self.generation += 1
        return self._session

    def close(self):
        if self._session is not None:
            self._session.close()
            self._session = None

    def survival(self, albedo):
        if self.fixed_q is not None:
            return self.fixed_q
        return np.clip(albedo, 1e-6, 1.0)

    def __call__(self, albedo, light_scale=None):
        if light_scale is None:
            light_scale = torch.tensor(1.0, dtype=torch.float64)
        return _RenderFn.apply(albedo, light_scale, self)
