"""The adjoint of the bake's bounces over the C-ABI (lb_adjoint_*), and a torch.autograd.Function
on it: the light atlas after the bounces (crates/lightbaker/src/bake.rs:300-330) as a function of
the materials' albedo, the sky (zenith, horizon, ground: trace.wgsl:63-66) and the sun's
irradiance, with visibility held fixed (every gather ray's hit recorded once, then replayed).

    sess = BakeSession(scene, gpu); sess.execute()
    adj = Adjoint(sess)
    light = baked_light(adj, albedo, sky, sun)      # torch, on the GPU; differentiable
    loss = ((light[..., :3] - target[..., :3]) ** 2 * target[..., 3:]).sum()
    loss.backward()

The reference has no gradients; the forward is its bake, replayed.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _native as N
from .bake import BakeSession


class Adjoint:
    """lb_adjoint over a bake session that has run lb_bake_execute (it must stay open)."""

    def __init__(self, sess: BakeSession):
        self.lib = N.lib()
        self.sess = sess
        self.gpu = sess.gpu
        h = C.c_void_p()
        N.check(self.lib.lb_adjoint_create(sess.handle, C.byref(h)))
        self.handle = h
        N.own(self, N.ADJOINT)
        atlas, params, items, rays = (C.c_size_t() for _ in range(4))
        N.check(self.lib.lb_adjoint_info(h, C.byref(atlas), C.byref(params), C.byref(items),
                                         C.byref(rays)))
        P = sess.array(N.BAKE_PARAMS, N.PARAMS)
        self.shape = (int(P["atlas"][0][1]), int(P["atlas"][0][0]), 4)
        assert self.shape[0] * self.shape[1] == atlas.value
        self.n_params = params.value
        self.items, self.rays = items.value, rays.value
        self.materials = (self.n_params - 24) // 3
        self.forwards = 0  # forwards run; the backward differentiates the last

    def records(self) -> np.ndarray:
        """(rays, items, 2) u32: each gather ray's word and tag."""
        out = np.zeros((self.rays, self.items, 2), np.uint32)
        N.check(self.lib.lb_adjoint_records(self.handle, out.ctypes.data, out.size))
        return out

    def _params(self, albedo, sky, sun, sky_sh=None):
        alb = np.ascontiguousarray(np.asarray(albedo, np.float32).reshape(-1, 3))
        sky = np.asarray(sky, np.float32).reshape(3, 3)
        p = N.LightParams()
        p.albedo = alb.ctypes.data
        p.n_materials = len(alb)
        p.zenith[:] = [float(v) for v in sky[0]]
        p.horizon[:] = [float(v) for v in sky[1]]
        p.ground[:] = [float(v) for v in sky[2]]
        p.sun[:] = [float(v) for v in np.asarray(sun, np.float32).reshape(3)]
        if sky_sh is not None:
            p.sky_sh[:] = [float(v) for v in np.asarray(sky_sh, np.float32).reshape(12)]
        return p, alb

    def forward(self, albedo, sky, sun, sky_sh=None) -> torch.Tensor:
        """The final light atlas, (ah, aw, 4) float32 on the GPU; sky rows zenith, horizon,
        ground; sky_sh (4, 3) the sky's L1 coefficients c0..c3 (None: zero)."""
        p, keep = self._params(albedo, sky, sun, sky_sh)
        light = torch.empty(self.shape, dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        N.check(self.lib.lb_adjoint_forward(self.handle, C.byref(p), C.c_void_p(light.data_ptr())))
        self.forwards += 1
        del keep
        return light

    def backward(self, dlight: torch.Tensor) -> torch.Tensor:
        """The gradient by (albedo rgb per material, zenith, horizon, ground, the sky's L1
        coefficients c0..c3, sun; rgb each) of a loss whose derivative by the last forward's atlas
        is `dlight` ((ah, aw, 4), rgb used)."""
        d = dlight.detach().to(device="cuda", dtype=torch.float32).contiguous()
        if tuple(d.shape) != self.shape:
            raise ValueError(f"dlight shape {tuple(d.shape)}, the atlas is {self.shape}")
        grad = torch.empty(self.n_params, dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        N.check(self.lib.lb_adjoint_backward(self.handle, C.c_void_p(d.data_ptr()),
                                             C.c_void_p(grad.data_ptr())))
        return grad

    def close(self) -> None:
        if self.handle:
            self.lib.lb_adjoint_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BakedLight(torch.autograd.Function):
    """light = bake(albedo (M, 3), sky (3, 3), sun (3,), sky_sh (4, 3)) with visibility fixed."""

    @staticmethod
    def forward(ctx, adj: Adjoint, albedo, sky, sun, sky_sh):
        args = [t.detach().cpu().numpy() for t in (albedo, sky, sun, sky_sh)]
        light = adj.forward(*args)
        ctx.adj = adj
        ctx.forward_id = adj.forwards
        ctx.save_for_backward(*(t.detach().clone() for t in (albedo, sky, sun, sky_sh)))
        return light

    @staticmethod
    def backward(ctx, grad_light):
        adj: Adjoint = ctx.adj
        albedo, sky, sun, sky_sh = ctx.saved_tensors
        if adj.forwards != ctx.forward_id:
            # another forward ran since: restore this one's bounces before differentiating
            adj.forward(*(t.cpu().numpy() for t in (albedo, sky, sun, sky_sh)))
            ctx.forward_id = adj.forwards
        g = adj.backward(grad_light)
        M = albedo.shape[0]
        ga = g[:M * 3].view(M, 3).to(albedo.device)
        gs = g[M * 3:M * 3 + 9].view(3, 3).to(sky.device)
        gh = g[M * 3 + 9:M * 3 + 21].view(4, 3).to(sky_sh.device)
        gu = g[M * 3 + 21:].view(3).to(sun.device)
        return None, ga, gs, gu, gh


def baked_light(adj: Adjoint, albedo: torch.Tensor, sky: torch.Tensor, sun: torch.Tensor,
                sky_sh: torch.Tensor | None = None) -> torch.Tensor:
    if sky_sh is None:
        sky_sh = torch.zeros(4, 3, dtype=torch.float32)
    return BakedLight.apply(adj, albedo, sky, sun, sky_sh)


def scene_light_params(scene):
    """A scene's own (albedo (M, 3), sky (3, 3), sun (3,), sky_sh (4, 3)) as float32 arrays."""
    alb = np.array([m.albedo for m in scene.materials], np.float32)
    sky = np.array([scene.zenith, scene.horizon, scene.ground], np.float32)
    sh = (np.zeros((4, 3), np.float32) if scene.sky_sh is None
          else np.asarray(scene.sky_sh, np.float32).reshape(4, 3))
    return alb, sky, np.asarray(scene.sun_irradiance, np.float32), sh


def l2_loss(light: torch.Tensor, target: torch.Tensor) -> torch.Tensor:
    """Sum over the atlas texels the target knows (w > 0.5) of the rgb's squared error."""
    w = (target[..., 3:] > 0.5).to(light.dtype)
    return (((light[..., :3] - target[..., :3]) ** 2) * w).sum()


def fit(adj: Adjoint, target: torch.Tensor, albedo0, sky0, sun0, sky_sh0=None, steps: int = 50,
        lr: float = 0.02, fit_sun: bool = False, fit_sky: str = "gradient", log=None) -> dict:
    """Adam on the L2 loss to `target`, visibility fixed, over the albedo (kept in [0, 1]) and the
    sky: `fit_sky` "gradient" (zenith, horizon, ground, kept non-negative), "sh" (the L1
    coefficients c0..c3) or "both"; and the sun's irradiance when `fit_sun`."""
    albedo = torch.tensor(np.asarray(albedo0, np.float32), requires_grad=True)
    grad_sky = fit_sky in ("gradient", "both")
    sh_sky = fit_sky in ("sh", "both")
    sky = torch.tensor(np.asarray(sky0, np.float32), requires_grad=grad_sky)
    sh0 = np.zeros((4, 3), np.float32) if sky_sh0 is None else np.asarray(sky_sh0, np.float32)
    sky_sh = torch.tensor(sh0.reshape(4, 3), requires_grad=sh_sky)
    sun = torch.tensor(np.asarray(sun0, np.float32), requires_grad=fit_sun)
    params = ([albedo] + ([sky] if grad_sky else []) + ([sky_sh] if sh_sky else [])
              + ([sun] if fit_sun else []))
    opt = torch.optim.Adam(params, lr=lr)
    history = []
    for k in range(steps):
        opt.zero_grad()
        loss = l2_loss(baked_light(adj, albedo, sky, sun, sky_sh), target)
        loss.backward()
        opt.step()
        with torch.no_grad():
            albedo.clamp_(0.0, 1.0)
            sky.clamp_(min=0.0)
            sun.clamp_(min=0.0)
        history.append(float(loss.detach()))
        if log:
            log(f"step {k}: loss {history[-1]:.6g}")
    return {"albedo": albedo.detach().numpy(), "sky": sky.detach().numpy(),
            "sky_sh": sky_sh.detach().numpy(), "sun": sun.detach().numpy(), "loss": history}
