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


PARAMETER_GROUPS = ("albedo", "tau", "air", "density")


def render_converged(scene: DiffScene, values: dict, passes: int, noise_seed: int = 0x7A29E7) -> torch.Tensor:
    """Mean radiance (H, W, 3) of ``passes`` progressive passes of the scene
    at parameter ``values`` (numpy arrays keyed as in ``recover``), on a
    tracer of its own with its own noise seed: a target image for
    ``recover`` whose noise does not repeat the scene's draws."""
    spp = scene.tracer.options.samples_per_frame
    t = PathTracer(scene.size, options=Options(noise_seed=noise_seed, samples_per_frame=spp,
                                               max_bounces=scene.tracer.options.max_bounces, milky_way=0.0))
    try:
        if scene.ring_body is not None and "tau" in values:
            t.set_ring_profile(scene.ring_body, np.asarray(values["tau"], np.float32))
        if scene.density:
            t.set_density_profile(np.asarray(values["density"], np.float32))
        frame = scene.frame_with(np.asarray(values["albedo"], np.float32), np.asarray(values["air"], np.float32))
        for _ in range(passes):
            t.encode(frame)
        return torch.from_numpy(t.read_accumulation()[..., :3].copy())
    finally:
        t.close()


def recover(scene: DiffScene, target: torch.Tensor, start: dict, steps: int = 200, lr: float = 0.01,
            final_lr: float = 0.001, log=None) -> tuple[dict, list]:
    """Fit the scene's parameters to a target image with Adam.

    ``start`` holds the starting values (numpy arrays ``albedo`` (B, 3),
    ``tau`` (512,), ``air`` (5,) and, when the scene renders with a density
    table, ``density`` (16,)). Every parameter is optimised in relative
    coordinates, value = start * (1 + u), so one learning rate suits
    albedos near 0.5 and air coefficients near 1e6 per AU alike. The
    learning rate falls linearly from ``lr`` to ``final_lr``.

    Each step renders two fresh passes (new draws). The first gives the
    residual to ``target`` and the step's loss (mean squared error); the
    adjoint of the second, weighted by that residual, gives the gradient.
    With one pass for both, the residual would carry the pass's own noise
    and the gradient would also descend the variance of the estimate.
    Returns the fitted values (numpy, same keys) and the loss of every
    step."""
    names = [n for n in PARAMETER_GROUPS if n in start and (n != "density" or scene.density)
             and (n != "tau" or scene.ring_body is not None)]
    base = {n: torch.tensor(np.asarray(start[n], np.float32)) for n in names}
    u = {n: torch.zeros_like(base[n], requires_grad=True) for n in names}
    opt = torch.optim.Adam(list(u.values()), lr=lr)

    def values():
        v = {n: base[n] * (1.0 + u[n]) for n in names}
        v["albedo"] = v["albedo"].clamp(0.0, 1.0)
        if "tau" in v:
            v["tau"] = v["tau"].clamp(min=0.0)
        if "density" in v:
            v["density"] = v["density"].clamp(min=0.0)
        air = v["air"].clamp(min=0.0)
        v["air"] = torch.cat([air[:4], air[4:].clamp(max=0.99)])
        return v

    losses = []
    for it in range(steps):
        for group in opt.param_groups:
            group["lr"] = lr + (final_lr - lr) * it / max(1, steps - 1)
        opt.zero_grad()
        v = values()
        with torch.no_grad():
            residual = render(scene, v["albedo"], v.get("tau"), v["air"], v.get("density")) - target
        image = render(scene, v["albedo"], v.get("tau"), v["air"], v.get("density"))
        (image * (2.0 / residual.numel()) * residual).sum().backward()
        opt.step()
        losses.append(float((residual ** 2).mean()))
        if log is not None and (it % 50 == 0 or it == steps - 1):
            log(it, losses[-1])
    with torch.no_grad():
        fitted = {n: t.numpy().copy() for n, t in values().items()}
    return fitted, losses


def parameter_errors(values: dict, target: dict, segments: int = 16) -> dict:
    """Mean relative error of a parameter set against a target set, per
    quantity the image determines:

    * ``albedo``: every albedo channel;
    * ``ring_bins``: every ring-profile bin, and ``ring_segments``: the
      profile averaged over ``segments`` equal runs of bins (512 bins over
      the ring's radial span resolve finer than a pixel at 512 x 512);
    * ``extinction``: every product of a coefficient (Rayleigh rgb, Mie)
      and a density slice, which is what a tabulated density enters the
      image through (scaling the table and dividing the coefficients by
      the same factor leaves the image unchanged);
    * ``mie_g``: the phase asymmetry."""

    def rel(a, b):
        a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
        return float(np.mean(np.abs(a - b) / np.abs(b)))

    out = {"albedo": rel(values["albedo"], target["albedo"])}
    if "tau" in values and "tau" in target:
        out["ring_bins"] = rel(values["tau"], target["tau"])
        seg = lambda t: np.asarray(t, np.float64).reshape(segments, -1).mean(axis=1)  # noqa: E731
        out["ring_segments"] = rel(seg(values["tau"]), seg(target["tau"]))
    if "density" in values and "density" in target:
        prod = lambda v: np.outer(np.asarray(v["air"][:4], np.float64), np.asarray(v["density"], np.float64))  # noqa: E731
        out["extinction"] = rel(prod(values), prod(target))
    else:
        out["extinction"] = rel(values["air"][:4], target["air"][:4])
    out["mie_g"] = rel(values["air"][4:], target["air"][4:])
    return out
