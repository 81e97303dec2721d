"""torch.autograd binding of the differentiable renderer.

``render(scene, albedo, tau, air)`` is a ``torch.autograd.Function``: the
forward pass restarts the view's accumulation and renders one progressive
pass on the GPU (``PathTracer.encode``); the backward pass runs the HIP
adjoint of that same pass (``PathTracer.adjoint``, same draws) and maps its
gradient vector onto the three parameter tensors:

* ``albedo`` (B, 3): ``sg_sphere.color`` of the bodies in ``scene.albedo_bodies``;
* ``tau`` (512,): the tabulated ring profile of ``scene.ring_body``;
* ``air`` (5,): Rayleigh rgb (``atmo_rayleigh.xyz``), Mie (``atmo_rayleigh.w``)
  and Mie g (``atmo_params.x``), in the frame's units (per AU);
* ``density`` (16,): the tabulated relative air density per altitude slice
  (when ``scene.density`` is set; otherwise exp(-h / H) and no gradient).

Tensors live on the CPU: the GPU work happens in the library, and each
call moves one image and one gradient vector.

The adjoint replays the draws of the pass it differentiates, so ``backward``
must run before the scene's tracer renders again: the forward pass records
the tracer's pass identity (``PathTracer.pass_id``) and ``backward`` raises
``StalePassError`` when another pass, a profile change or a restart has
replaced it.
"""
from __future__ import annotations

import numpy as np
import torch

from .layouts import DENSITY_BINS, GRAD_AIR, GRAD_DENSITY, GRAD_RING, RING_PROFILE_BINS, grad_albedo
from .pathtracer import Frame, Options, PathTracer


class StalePassError(RuntimeError):
    """``backward`` of a pass the tracer no longer holds."""


class DiffScene:
    """A frame, its differentiable parameters and the tracer that renders it."""

    def __init__(self, frame: Frame, size, spp: int, bounces: int, albedo_bodies=(), ring_body=None,
                 density: bool = False):
        self.frame = Frame(frame.globals.copy(), frame.bodies.copy(), frame.ground.copy(), frame.scene_time)
        self.size = (int(size[0]), int(size[1]))
        self.tracer = PathTracer(self.size, options=Options(samples_per_frame=spp, max_bounces=bounces,
                                                            milky_way=0.0))
        self.tracer.enable_adjoint(True)
        self.albedo_bodies = list(albedo_bodies)
        self.ring_body = ring_body
        self.density = density

    def initial(self):
        """Parameter tensors holding the frame's current values (tau: the
        banded profile's nominal depth in every bin)."""
        albedo = torch.tensor(np.stack([self.frame.bodies["color"][b] for b in self.albedo_bodies])
                              if self.albedo_bodies else np.zeros((0, 3)), dtype=torch.float32)
        tau = None
        if self.ring_body is not None:
            tau = torch.full((RING_PROFILE_BINS,), float(self.frame.bodies["ring_params"][self.ring_body][1]))
        g = self.frame.globals
        air = torch.tensor([*g["atmo_rayleigh"], g["atmo_params"][0]], dtype=torch.float32)
        top, scale = float(g["atmo_params"][2]), float(g["atmo_params"][1])
        density = torch.exp(-(torch.arange(DENSITY_BINS) + 0.5) * top / DENSITY_BINS / scale).float()
        return albedo, tau, air, density

    def frame_with(self, albedo: np.ndarray, air: np.ndarray) -> Frame:
        bodies = self.frame.bodies.copy()
        for row, b in enumerate(self.albedo_bodies):
            bodies["color"][b] = albedo[row]
        g = self.frame.globals.copy()
        g["atmo_rayleigh"] = air[:4]
        p = g["atmo_params"].copy()
        p[0] = air[4]
        g["atmo_params"] = p
        return Frame(g, bodies, self.frame.ground, self.frame.scene_time)

    def close(self):
        self.tracer.close()


class _Render(torch.autograd.Function):
    @staticmethod
    def forward(ctx, albedo, tau, air, density, scene: DiffScene):
        frame = scene.frame_with(albedo.detach().cpu().numpy().astype(np.float32),
                                 air.detach().cpu().numpy().astype(np.float32))
        if scene.ring_body is not None:
            scene.tracer.set_ring_profile(scene.ring_body, tau.detach().cpu().numpy().astype(np.float32))
        if scene.density:
            scene.tracer.set_density_profile(density.detach().cpu().numpy().astype(np.float32))
        scene.tracer.restart()
        scene.tracer.encode(frame)
        image = scene.tracer.read_accumulation()[..., :3].copy()
        ctx.scene, ctx.frame = scene, frame
        ctx.pass_id = scene.tracer.pass_id()
        ctx.has_tau = scene.ring_body is not None
        return torch.from_numpy(image)

    @staticmethod
    def backward(ctx, grad_image):
        scene = ctx.scene
        if scene.tracer.pass_id() != ctx.pass_id:
            raise StalePassError("backward of a pass this scene's tracer has since replaced: differentiate each "
                                 "render before rendering the scene again")
        g = scene.tracer.adjoint(ctx.frame, grad_image.detach().cpu().numpy().astype(np.float32))
        g = torch.from_numpy(g)
        albedo = torch.stack([g[grad_albedo(b, 0):grad_albedo(b, 0) + 3] for b in scene.albedo_bodies]) \
            if scene.albedo_bodies else torch.zeros((0, 3))
        tau = g[GRAD_RING:GRAD_RING + RING_PROFILE_BINS].clone() if ctx.has_tau else None
        air = g[GRAD_AIR:GRAD_AIR + 5].clone()
        density = g[GRAD_DENSITY:GRAD_DENSITY + DENSITY_BINS].clone() if scene.density else None
        return albedo, tau, air, density, None


def render(scene: DiffScene, albedo: torch.Tensor, tau: torch.Tensor | None, air: torch.Tensor,
           density: torch.Tensor | None = None) -> torch.Tensor:
    """Pass mean radiance (H, W, 3) of one fresh pass; differentiable in every
    parameter tensor the scene uses."""
    if tau is None:
        tau = torch.zeros(RING_PROFILE_BINS)
    if density is None:
        if scene.density:
            raise ValueError("the scene renders with a density profile: pass `density`")
        density = torch.zeros(DENSITY_BINS)
    return _Render.apply(albedo, tau, air, density, scene)
