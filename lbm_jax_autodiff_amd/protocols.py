"""Run protocols of the reference's drivers on the fused HIP engine.

* channel flow         examples/01_channel_flow.py:22-89
* sphere Cd            examples/10_cd_convergence.py:46-121 (time-averaged Cd)
* sphere dCd/dr, Adam  examples/11_optimisation_v2.py:56-204
* sphere geometry at a world centre on the box [-4,4]x[-2,2]x[-2,2]; a centre
  given in lattice units is read as the continuous coordinate X (node i at
  X = i + 1/2), i.e. world = X / s - (4, 2, 2) with s = nx/8.

Usage: python -m lbm_jax_autodiff_amd.protocols {c1,c2,c3,c4,c5} [options]
prints one JSON line per result.
"""

from __future__ import annotations

import argparse
import json
import math
import sys
import time
from typing import Optional

import torch

from . import _field as F
from .adjoint import drag_series, rollout
from .core.ops import compute_density, compute_velocity, equilibrium, projected_area
from .engine import Lattice
from .geometry import compute_q_from_solid, create_sphere, smooth_sphere_geometry


def _dev():
    return torch.device("cuda", torch.cuda.current_device())


def _dtype(storage):
    return torch.float64 if storage == "f64" else torch.float32


def world_center(nx, ny, nz, center_lattice):
    return (center_lattice[0] / (nx / 8.0) - 4.0, center_lattice[1] / (ny / 4.0) - 2.0,
            center_lattice[2] / (nz / 4.0) - 2.0)


def tau_for(u, D, re):
    return 3.0 * (u * D / re) + 0.5


# ---------------------------------------------------------------- channel --

def channel_flow(nx=64, ny=21, nz=21, tau=0.8, u_in=0.05, n_steps=5000, storage="f32"):
    """Poiseuille channel: halfway walls at y = 0, ny-1, BGK, Zou-He, f0 = w, no guard."""
    dev = _dev()
    dt = _dtype(storage)
    solid = torch.zeros((nx, ny, nz), dtype=torch.int32, device=dev)
    solid[:, 0, :] = 1
    solid[:, ny - 1, :] = 1
    q = compute_q_from_solid(solid, nx, ny, nz, 0.5, dtype=dt)
    f0 = equilibrium(torch.ones((nx, ny, nz), dtype=dt, device=dev), torch.zeros((3, nx, ny, nz), dtype=dt,
                                                                                  device=dev))
    lat = Lattice(nx, ny, nz, tau=tau, inlet_velocity=(u_in, 0.0, 0.0), collision="bgk", solid=solid, q=q,
                  storage=storage, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s, _, _ = lat.run(f0, n_steps, forces=False)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    s = s.to(torch.float64)
    rho = compute_density(s)
    ux = compute_velocity(s, rho)[0]
    prof = ux[nx // 2].mean(-1).cpu()
    H = float(ny - 2)
    umax = float(prof[1:-1].max())
    ana = torch.tensor([0.0 if j in (0, ny - 1) else umax * 4.0 * (j - 0.5) * (H - (j - 0.5)) / (H * H)
                        for j in range(ny)], dtype=torch.float64)
    fluid = slice(1, ny - 1)
    l2 = float(torch.linalg.norm(prof[fluid] - ana[fluid]) / torch.linalg.norm(ana[fluid]))
    return {"config": "channel", "grid": [nx, ny, nz], "steps": n_steps, "storage": storage, "tau": tau,
            "profile": [round(float(v), 8) for v in prof], "rel_l2_vs_parabola": l2,
            "mlups": nx * ny * nz * n_steps / wall / 1e6, "wall_s": wall}


# ----------------------------------------------------------------- sphere --

def sphere_lattice(nx, ny, nz, center_world, radius=0.5, *, re=100.0, u_in=0.05, collision="mrt",
                   mode="hard", sharpness=10.0, smagorinsky_cs=None, guard="pre", storage="f32",
                   offset="inlet"):
    dev = _dev()
    D = 2.0 * radius * nx / 8.0
    tau = tau_for(u_in, D, re)
    if mode == "hard":
        solid, q = create_sphere(nx, ny, nz, center_world, radius, device=dev)
        kw = dict(solid=solid, q=q.to(_dtype(storage)))
    else:
        por, solid, q = smooth_sphere_geometry(nx, ny, nz, center_world, radius, sharpness, device=dev,
                                               out_dtype=_dtype(storage))
        kw = dict(porosity=por, q=q)
    lat = Lattice(nx, ny, nz, tau=tau, inlet_velocity=(u_in, 0.0, 0.0), collision=collision,
                  smagorinsky_cs=smagorinsky_cs, guard=guard, guard_velocity=(u_in, 0.0, 0.0), storage=storage,
                  offset=offset, device=dev, **kw)
    area = float(projected_area(solid, 0))
    return lat, solid, tau, area


def tail_stats(cd, frac=0.3):
    cd = cd.to(torch.float64)
    n = cd.numel()
    k = max(1, int(round(frac * n)))
    tail = cd[n - k:]
    return float(tail.mean()), float(tail.std(unbiased=False))


def snapshot_stats(cd, stride=500, frac=0.3):
    """examples/10_cd_convergence.py:88-106: a sample every `stride` steps,
    mean and population std of the last max(3, int(0.3 * #samples))."""
    samples = cd[stride - 1::stride].to(torch.float64).cpu()
    if samples.numel() == 0:
        return float("nan"), float("nan")
    k = max(3, int(samples.numel() * frac))
    tail = samples[-k:]
    return float(tail.mean()), float(tail.std(unbiased=False))


def sphere_cd(nx, ny, nz, n_steps, center_world, radius=0.5, *, init_u=0.0, storage="f32", **kw):
    lat, solid, tau, area = sphere_lattice(nx, ny, nz, center_world, radius, storage=storage, **kw)
    dev = lat.device
    dt = _dtype(storage)
    u = torch.zeros((3, nx, ny, nz), dtype=dt, device=dev)
    u[0] = init_u
    f0 = equilibrium(torch.ones((nx, ny, nz), dtype=dt, device=dev), u)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, _, forces = lat.run(f0, n_steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    u_in = lat.inlet_velocity[0]
    cd = drag_series(forces, u_in, area)
    mean, std = tail_stats(cd)
    smean, sstd = snapshot_stats(cd)
    return {"grid": [nx, ny, nz], "steps": n_steps, "storage": storage, "tau": tau, "area": area,
            "center_world": list(center_world), "radius_world": radius, "links": lat.links.n,
            "cd_tail30_mean": mean, "cd_tail30_std": std, "cd_snap500_mean": smean, "cd_snap500_std": sstd,
            "cd_last": float(cd[-1]), "mlups": nx * ny * nz * n_steps / wall / 1e6, "wall_s": wall}, cd


def sphere_loss_and_grad(nx, ny, nz, n_steps, center_world, radius=0.5, *, avg_last=600, q_mode="bouzidi",
                         porosity_mode="hard", sharpness=20.0, collision="mrt", re=100.0, u_in=0.05,
                         storage="f32", checkpoint="all", smagorinsky_cs=None, guard="none", init_u=0.05,
                         offset="inlet", geometry_dtype=None, r_ref=0.5, adjoint_storage=None,
                         rounding="stochastic"):
    """loss = mean Cd over the last `avg_last` steps and dloss/dr (world units)
    through q(r) (Bouzidi) and porosity(r) (soft), examples/11:56-167.  The
    initial state is f_eq(1, (init_u, 0, 0)) (examples/11:147-149 use u_inlet).  The
    viscosity comes from the nominal diameter 2 r_ref (examples/11:228-230),
    so tau does not move with r."""
    dev = _dev()
    torch.cuda.reset_peak_memory_stats(dev)
    dt = _dtype(storage)
    gdt = geometry_dtype or (torch.float64 if storage == "f64" else torch.float32)
    r = torch.tensor(float(radius), dtype=gdt, device=dev, requires_grad=True)
    por_soft, solid, q = smooth_sphere_geometry(nx, ny, nz, center_world, r, sharpness, device=dev,
                                                coord_dtype=gdt, out_dtype=dt)
    if q_mode == "halfway":
        q = compute_q_from_solid(solid, nx, ny, nz, 0.5, dtype=dt)
    # "hard" in examples/11:62-65 is the soft blend fed porosity = solid (0/1,
    # zero gradient).  With p in {0, 1} the blend equals the hard-mask
    # reversal bit for bit and its dL/dp is multiplied by zero, so the hard
    # mask gives the same loss and dCd/dr without the per-cell blend and the
    # float64 dL/dp accumulation.
    porosity = por_soft if porosity_mode == "soft" else None
    hard_solid = solid if porosity_mode == "hard" else None
    if q_mode == "halfway" and porosity_mode == "hard":
        # nothing else depends on r here: keep the reference's 0 * porosity(r)
        # term so the adjoint runs and returns its exact zero
        porosity, hard_solid = solid.to(dt) + 0.0 * por_soft, None
    D = 2.0 * float(r_ref) * nx / 8.0
    tau = tau_for(u_in, D, re)
    u = torch.zeros((3, nx, ny, nz), dtype=dt, device=dev)
    u[0] = init_u
    f0 = equilibrium(torch.ones((nx, ny, nz), dtype=dt, device=dev), u)
    area = torch.clamp(projected_area(solid, 0).to(torch.float64), min=1.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, _, forces = rollout(f0, n_steps, tau=tau, q=q, porosity=porosity, solid=hard_solid,
                           inlet_velocity=(u_in, 0.0, 0.0),
                           collision=collision, smagorinsky_cs=smagorinsky_cs, guard=guard,
                           guard_velocity=(u_in, 0.0, 0.0), storage=storage, offset=offset, checkpoint=checkpoint,
                           adjoint_storage=adjoint_storage, rounding=rounding)
    cd = drag_series(forces, u_in, area)
    loss = cd[-avg_last:].mean()
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    (g,) = torch.autograd.grad(loss, r, allow_unused=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    g = 0.0 if g is None else float(g)
    return {"grid": [nx, ny, nz], "steps": n_steps, "storage": storage, "q_mode": q_mode,
            "porosity_mode": porosity_mode, "tau": tau, "radius_world": float(radius), "loss": float(loss.detach()),
            "dloss_dr_world": g, "dloss_dr_lattice": g / (nx / 8.0), "checkpoint": str(checkpoint),
            "adjoint_storage": adjoint_storage or ("f16" if storage in ("f16", "bf16") else storage),
            "fwd_s": t_fwd, "fwd_adj_s": wall, "fwd_adj_mlups": nx * ny * nz * n_steps / wall / 1e6,
            "peak_hbm_gib": torch.cuda.max_memory_allocated(dev) / 2.0 ** 30}


def adam(loss_and_grad, r0, n_iters, lr, r_min=0.3, r_max=0.8, b1=0.9, b2=0.999, eps=1e-8):
    """optax.adam(lr) + clip to [r_min, r_max] on a scalar (examples/11:182-194)."""
    r, m, v = float(r0), 0.0, 0.0
    hist = []
    for t in range(1, n_iters + 1):
        loss, g = loss_and_grad(r)
        hist.append({"iter": t - 1, "r": r, "loss": loss, "grad": g})
        m = b1 * m + (1.0 - b1) * g
        v = b2 * v + (1.0 - b2) * g * g
        mh = m / (1.0 - b1 ** t)
        vh = v / (1.0 - b2 ** t)
        r = r - lr * mh / (math.sqrt(vh) + eps)
        r = min(max(r, r_min), r_max)
    return r, hist


# ------------------------------------------- published MRT-path studies --

SHARPNESS_ALPHAS = (5.0, 10.0, 20.0, 40.0, 80.0)


def sharpness_case(alpha, *, storage="f32", radius=0.5, n_steps=2880, fd_eps=None, nx=96, ny=48, nz=48,
                   geometry_dtype=torch.float32):
    """examples/12_sharpness_sensitivity.py:53-105: 96x48x48, Re = 50 (tau from
    D = nx/8 = 12), MRT, soft porosity sigma(alpha sdf) with ray q(r), centre
    world 0, f_0 = f_eq(1, (U, 0, 0)), no guard, 2880 steps; the loss is the
    Cd of the last (f_post, f_final) pair, and dCd/dr is in world units.
    ``fd_eps`` adds the script's central difference (r +- eps) of the same
    forward.  The geometry is computed from float32 coordinates for every
    storage by default, as the script's (jnp.arange float32, smooth.py:62-71),
    so a float64 run differs from the float32 one in the populations only;
    ``geometry_dtype`` float64 takes r and the coordinates in float64 (for
    central differences in r at small steps)."""
    res = sphere_loss_and_grad(nx, ny, nz, n_steps, (0.0, 0.0, 0.0), radius, avg_last=1, q_mode="bouzidi",
                               porosity_mode="soft", sharpness=float(alpha), collision="mrt", re=50.0,
                               storage=storage, checkpoint="all", guard="none", init_u=0.05,
                               geometry_dtype=geometry_dtype)
    out = {"alpha": float(alpha), "cd": res["loss"], "grad_ad": res["dloss_dr_world"], "tau": res["tau"],
           "fwd_adj_s": res["fwd_adj_s"]}
    if fd_eps:
        kw = dict(avg_last=1, q_mode="bouzidi", porosity_mode="soft", sharpness=float(alpha), collision="mrt",
                  re=50.0, storage=storage, checkpoint="all", guard="none", init_u=0.05,
                  geometry_dtype=geometry_dtype)
        lp = sphere_loss_and_grad(nx, ny, nz, n_steps, (0.0, 0.0, 0.0), radius + fd_eps, **kw)["loss"]
        lm = sphere_loss_and_grad(nx, ny, nz, n_steps, (0.0, 0.0, 0.0), radius - fd_eps, **kw)["loss"]
        out["grad_fd"] = (lp - lm) / (2.0 * fd_eps)
        out["fd_eps"] = fd_eps
    return out


FOUR_MODES = (("bouzidi", "soft"), ("halfway", "soft"), ("bouzidi", "hard"), ("halfway", "hard"))


def four_mode_case(q_mode, porosity_mode, *, storage="f32", radius=0.5, n_steps=2000, iters=0, lr=3e-3):
    """examples/11_optimisation_v2.py:56-204: 96x48x48, Re = 50, MRT, sharpness
    20, q ray (bouzidi) or 1/2 (halfway), porosity soft or the 0/1 mask
    (hard), f_0 = f_eq(1, (U, 0, 0)), 2000 steps, loss = Cd of the last
    (f_post, f_final) pair.  ``iters`` > 0 runs the script's Adam (optax.adam
    update, r clipped to [0.3, 0.8]) and reports base, best and final values."""

    def lg(r):
        out = sphere_loss_and_grad(96, 48, 48, n_steps, (0.0, 0.0, 0.0), r, avg_last=1, q_mode=q_mode,
                                   porosity_mode=porosity_mode, sharpness=20.0, collision="mrt", re=50.0,
                                   storage=storage, checkpoint="all", guard="none", init_u=0.05,
                                   geometry_dtype=torch.float32)
        return out["loss"], out["dloss_dr_world"]

    cd0, g0 = lg(radius)
    out = {"mode": f"{q_mode}_{porosity_mode}", "base_cd": cd0, "base_grad": g0}
    if iters > 0:
        r_end, hist = adam(lg, radius, iters, lr)
        cds = [h["loss"] for h in hist]
        out.update({"iters": iters, "lr": lr, "best_cd": min(cds), "final_r": hist[-1]["r"], "r_after": r_end,
                    "avg_abs_grad": sum(abs(h["grad"]) for h in hist) / len(hist),
                    "history": [(round(h["r"], 6), round(h["loss"], 6), round(h["grad"], 6)) for h in hist]})
    return out


# -------------------------------------------------------------------- CLI --

def _emit(d):
    print(json.dumps(d), flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("config", choices=["c1", "c2", "c3", "c3fd", "c3fdh", "c3adam", "c4", "c5", "ex10", "ex11",
                                       "ex12"])
    ap.add_argument("--grid", default="64")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--storage", default="f32")
    ap.add_argument("--checkpoint", default=None)
    ap.add_argument("--adjoint-storage", default=None, choices=["f32", "f16"])
    ap.add_argument("--mode", default="hard")
    ap.add_argument("--repeat", type=int, default=1, help="c3/c3fd: gradient evaluations in one process")
    ap.add_argument("--iters", type=int, default=0, help="ex11: Adam iterations (the script runs 60)")
    ap.add_argument("--offset", default="inlet",
                    help="c4/c5 half storage: 'inlet' (f_eq(1, U)), 'weights' (w_i) or a number k (f_eq(1, k U))")
    ap.add_argument("--fd-eps", type=float, default=0.0, help="ex12: central-difference step (the script: 1e-3)")
    ap.add_argument("--rounding", default="stochastic", choices=["nearest", "stochastic"],
                    help="c4/c5 16-bit storage rounding (default stochastic)")
    args = ap.parse_args(argv)
    offset = args.offset
    if offset not in ("inlet", "weights"):
        from .engine import equilibrium_values
        offset = equilibrium_values(1.0, (float(offset) * 0.05, 0.0, 0.0))
    if args.config == "c1":
        _emit(channel_flow(n_steps=args.steps or 5000, storage=args.storage))
    elif args.config == "c2":
        nx, ny, nz = 64, 32, 32
        res, _ = sphere_cd(nx, ny, nz, args.steps or 8000, world_center(nx, ny, nz, (16, 16, 16)), 0.5,
                           storage=args.storage)
        res["config"] = "c2"
        _emit(res)
    elif args.config in ("c3", "c3fd"):
        nx, ny, nz = 96, 48, 48
        c = world_center(nx, ny, nz, (24, 24, 24))
        steps = args.steps or 2000
        # a repeated call reuses the stored states' memory from torch's caching
        # allocator, so from the second call on the wall time is the pipeline's
        for k in range(max(1, args.repeat)):
            res = sphere_loss_and_grad(nx, ny, nz, steps, c, 0.5, storage=args.storage)
            res["config"] = "c3"
            res["call"] = k + 1
            _emit(res)
        if args.config == "c3fd":
            h = 1e-3
            lp = sphere_loss_and_grad(nx, ny, nz, steps, c, 0.5 + h, storage="f64")["loss"]
            lm = sphere_loss_and_grad(nx, ny, nz, steps, c, 0.5 - h, storage="f64")["loss"]
            g64 = sphere_loss_and_grad(nx, ny, nz, steps, c, 0.5, storage="f64")
            _emit({"config": "c3", "fd_f64_h1e-3": (lp - lm) / (2 * h), "adjoint_f64": g64["dloss_dr_world"],
                   "adjoint_" + args.storage: res["dloss_dr_world"]})
            hw = sphere_loss_and_grad(nx, ny, nz, steps, c, 0.5, storage=args.storage, q_mode="halfway")
            hw["config"] = "c3-halfway-hard"
            _emit(hw)
    elif args.config == "c3fdh":
        # central differences of the FP64 kernels at shrinking h
        nx, ny, nz = 96, 48, 48
        c = world_center(nx, ny, nz, (24, 24, 24))
        steps = args.steps or 2000
        ad = sphere_loss_and_grad(nx, ny, nz, steps, c, 0.5, storage="f64")
        out = {"config": "c3-fd-sweep", "adjoint_f64": ad["dloss_dr_world"], "fd": {}}
        for h in (1e-3, 3e-4, 1e-4, 3e-5, 1e-5, 1e-6):
            lp = sphere_loss_and_grad(nx, ny, nz, steps, c, 0.5 + h, storage="f64")["loss"]
            lm = sphere_loss_and_grad(nx, ny, nz, steps, c, 0.5 - h, storage="f64")["loss"]
            out["fd"][str(h)] = (lp - lm) / (2 * h)
        _emit(out)
    elif args.config == "ex10":
        # examples/10_cd_convergence.py:46-121: centre world 0, r = 0.5, init u = U,
        # pre-collision guard to eq(1, U), Cd snapshot every 500 steps
        sizes = {"64": (64, 32, 32, 8000), "128": (128, 64, 64, 16000), "256": (256, 128, 128, 28000),
                 "384": (384, 192, 192, 40000)}
        nx, ny, nz, steps = sizes[args.grid]
        res, cd = sphere_cd(nx, ny, nz, args.steps or steps, (0.0, 0.0, 0.0), 0.5, init_u=0.05,
                            storage=args.storage, guard="pre")
        res["config"] = "ex10-" + args.grid
        res["cd_snapshots"] = [round(float(v), 6) for v in cd[499::500].cpu()]
        _emit(res)
    elif args.config == "ex12":
        for a in SHARPNESS_ALPHAS:
            res = sharpness_case(a, storage=args.storage, n_steps=args.steps or 2880, fd_eps=args.fd_eps or None)
            res["config"] = "ex12"
            res["storage"] = args.storage
            _emit(res)
    elif args.config == "ex11":
        for qm, pm in FOUR_MODES:
            res = four_mode_case(qm, pm, storage=args.storage, n_steps=args.steps or 2000,
                                 iters=0 if (qm, pm) == ("halfway", "hard") else args.iters)
            res["config"] = "ex11"
            res["storage"] = args.storage
            _emit(res)
    elif args.config == "c3adam":
        nx, ny, nz = 96, 48, 48
        c = world_center(nx, ny, nz, (24, 24, 24))
        steps = args.steps or 2000

        def lg(r):
            out = sphere_loss_and_grad(nx, ny, nz, steps, c, r, storage=args.storage)
            return out["loss"], out["dloss_dr_world"]
        r, hist = adam(lg, 0.5, 20, 0.05)
        _emit({"config": "c3-adam", "lr": 0.05, "iters": 20, "final_r": r, "history": hist})
    elif args.config == "c4":
        nx, ny, nz = 256, 128, 128
        steps = args.steps or 28000
        # every 100 steps; float64 states (610 MiB each) take two levels
        ck = (200, 14) if args.storage == "f64" else 100
        if args.checkpoint:
            parts = [int(v) for v in args.checkpoint.split(",")]
            ck = parts[0] if len(parts) == 1 else tuple(parts)
        res = sphere_loss_and_grad(nx, ny, nz, steps, (0.0, 0.0, 0.0), 0.5, storage=args.storage,
                                   checkpoint=ck, avg_last=int(0.3 * steps), guard="pre",
                                   adjoint_storage=args.adjoint_storage, offset=offset,
                                   rounding=args.rounding)
        res["config"] = "c4"
        res["rounding"] = args.rounding if args.storage in ("f16", "bf16") else None
        _emit(res)
    elif args.config == "c5":
        nx, ny, nz = 384, 192, 192
        steps = args.steps or 40000
        soft = args.mode == "soft"
        # segments of 200 steps with sub-checkpoints every 14 for half storage
        # (123 GB); fewer, longer segments for the 1.08 GB float32 and 2.15 GB
        # float64 states: about N/K1 + K1/K2 + K2 states are held at once
        ck = {"f32": (400, 20), "f64": (1200, 35)}.get(args.storage, (200, 14))
        if args.checkpoint:
            ck = tuple(int(v) for v in args.checkpoint.split(","))
        rounding = args.rounding
        res = sphere_loss_and_grad(nx, ny, nz, steps, (0.0, 0.0, 0.0), 0.5, storage=args.storage,
                                   checkpoint=ck, avg_last=int(0.3 * steps), guard="pre",
                                   porosity_mode="soft" if soft else "hard", sharpness=10.0,
                                   smagorinsky_cs=0.1 if soft else None, adjoint_storage=args.adjoint_storage,
                                   offset=offset, rounding=rounding)
        res["config"] = "c5-" + args.mode
        res["offset"] = args.offset
        res["rounding"] = rounding if args.storage in ("f16", "bf16") else None
        _emit(res)


if __name__ == "__main__":
    sys.exit(main())
