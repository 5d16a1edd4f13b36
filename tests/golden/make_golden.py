"""Writes the golden fixtures under tests/golden/ (run on the CPU, no GPU).

    python tests/golden/make_golden.py kats   # published known-answer values
    python tests/golden/make_golden.py c1     # 64x21x21 channel, 5000 steps (~10 min)
    python tests/golden/make_golden.py c2     # 64x32x32 sphere Cd, 8000 steps (~15 min)
    python tests/golden/make_golden.py c3     # 96x48x48 dCd/dr, 2000 steps (~40 min)
    python tests/golden/make_golden.py c2p    # c2 in float32 from 3 perturbed starts (~30 min)
    python tests/golden/make_golden.py ex12   # examples/12 sharpness sweep forward, float32 (~10 min)

The reference (jax_lbm) cannot be imported here: jax, jaxlib and optax are not
installed.  ``kats`` therefore holds the values the reference publishes (its
README and paper); the other fixtures are outputs of the CPU oracle
(oracle/lbm_oracle.py), whose agreement with those published values is tested
in tests/test_oracle.py.  The GPU tests compare the HIP kernels with these
files, so the oracle does not have to rerun the long workloads on the GPU host.

Every fixture is plain JSON (inputs described, outputs as numbers).
"""

from __future__ import annotations

import json
import os
import sys
import time

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import lbm_oracle as O  # noqa: E402

sys.path.insert(0, os.path.dirname(HERE))
from _helpers import perturbed_start  # noqa: E402


def _threads():
    torch.set_num_threads(int(os.environ.get("GOLDEN_THREADS", os.cpu_count() or 1)))


def _write(name, obj):
    path = os.path.join(HERE, name)
    with open(path, "w") as fh:
        json.dump(obj, fh, indent=1)
        fh.write("\n")
    print(f"wrote {path}", flush=True)


def kats():
    """Values published by the reference for this path (data, with sources)."""
    _write("kats.json", {
        "poiseuille_60x20x4": {
            "source": "README.md:59-69; docs/getting_started/first_simulation.md:101-109; paper/paper.tex:375-393",
            "setup": "60x20x4 D3Q19 BGK tau=0.8, halfway walls (q=1/2) at y=0 and y=19, periodic z, Zou-He "
                     "velocity inlet u_x=0.05 at x=0 and pressure outlet rho=1 at x=59, f0 = w, 3000 steps, "
                     "float32; u_x at x=30 averaged over z",
            "ux_by_y": {"0": 0.000200, "1": 0.008247, "5": 0.057873, "9": 0.076879, "10": 0.076879,
                        "15": 0.048352, "19": 0.000200},
            "decimals": 6,
        },
        "d_sum_feq_d_rho": {"source": "paper/paper.tex:501; tests/test_gradients.py:45-48",
                             "setup": "4x4x4, d(sum f_eq(rho, u=0))/d rho", "value": 64.0},
        "d_mean_rho_d_u_inlet": {"source": "paper/paper.tex:502; tests/test_gradients_f64.py:95-127",
                                  "setup": "8x4x4 BGK tau=0.8, no walls, Zou-He, 5 steps, float64, "
                                           "mean rho over x in [2, nx-2)", "value": 0.611238, "decimals": 6},
        "d_cd_d_r_bouzidi_soft": {"source": "paper/paper.tex:503; tests/test_gradients_f64.py:129-217",
                                   "setup": "32x16x16 BGK tau=3*(0.05*8/50)+1/2, soft porosity sigmoid(20 sdf), "
                                            "ray q(r), r=0.5 world, Zou-He, init eq(1, U), 10 steps, float64, "
                                            "Cd from the last (f_post, f) pair, area = max(A, 1)",
                                   "value": 93.33275, "decimals": 5},
        "drag_coefficient": {"source": "tests/test_forces.py:86-91",
                             "setup": "F = (0.01, 0, 0), U = 0.05, A = 100, rho = 1 -> |F_x| / (0.5 rho U^2 A + 1e-10)",
                             "force_x": 0.01, "u": 0.05, "area": 100.0, "value": 0.08},
        "projected_area_box": {"source": "tests/test_forces.py:62-67",
                               "setup": "16x8x8 mask with ones at [5:8, 2:6, 2:6], axis 0",
                               "shape": [16, 8, 8], "box": [[5, 8], [2, 6], [2, 6]], "value": 16},
        "projected_area_ground": {"source": "tests/test_forces.py:75-80",
                                  "setup": "16x8x8 mask with 2 (ground) on the y=0 plane, axis 0", "value": 0},
        "halfway_hard_d_cd_d_r": {"source": "README.md:35; paper/paper.tex:651-662",
                                  "setup": "q = 1/2 everywhere (compute_q_from_solid) with a hard mask: nothing "
                                           "depends on r", "value": 0.0},
        "sharpness_sweep_96x48x48": {
            "source": "paper/make_convergence_figures.py:86-89 (A100 runs of examples/12_sharpness_sensitivity.py)",
            "setup": "examples/12_sharpness_sensitivity.py:53-139: 96x48x48 MRT, Re=50 (tau=float32(0.536)), soft "
                     "porosity sigmoid(alpha sdf) + ray q(r), centre world 0, r=0.5, f0 = f_eq(1,(0.05,0,0)), no "
                     "guard, 2880 steps, Cd of the last (f_post, f_final) pair, jax.grad and FD (eps=1e-3) in r; "
                     "float32",
            "alpha": [5.0, 10.0, 20.0, 40.0, 80.0], "cd": [3.7555, 1.3538, 0.8745, 0.3283, 0.5414],
            "grad_ad": [15.1634, 3.6649, 1.8524, 9.4889, 2.4426],
            "grad_fd": [15.4666, 3.7023, 2.2820, 10.3985, 3.3549], "decimals": 4},
        "four_mode_96x48x48": {
            "source": "README.md:31-39; paper/make_convergence_figures.py:131-134 (A100 runs of "
                      "examples/11_optimisation_v2.py)",
            "setup": "examples/11_optimisation_v2.py:56-240: 96x48x48 MRT, Re=50, sharpness 20, q ray (bouzidi) or "
                     "1/2 (halfway), porosity soft or the 0/1 mask (hard), f0 = f_eq(1,(0.05,0,0)), 2000 steps, "
                     "Cd of the last (f_post, f_final) pair; optax.adam lr=3e-3 from r=0.5, r clipped to "
                     "[0.3, 0.8], 60 iterations (halfway_hard: one evaluation); float32",
            "modes": ["halfway_hard", "bouzidi_hard", "halfway_soft", "bouzidi_soft"],
            "base_cd": [1.85, 1.77, 0.43, 0.45], "best_cd": [None, 1.66, 0.003, 0.002],
            "avg_abs_grad": [0.0, 5.55, 6.30, 3.19], "final_r": [0.500, 0.314, 0.652, 0.663]},
        "cd_convergence_ex10": {
            "source": "paper/make_convergence_figures.py:52-54; README.md:45-52 (A100 runs of "
                      "examples/10_cd_convergence.py)",
            "setup": "MRT + hard sphere with ray q, Re=100, centre world 0, r=0.5, f0 = f_eq(1,(0.05,0,0)), "
                     "pre-collision guard to f_eq(1,(0.05,0,0)), Cd sampled every 500 steps, mean and population "
                     "std of the last max(3, int(0.3 * samples)); float32",
            "grids": [[64, 32, 32], [128, 64, 64], [256, 128, 128], [384, 192, 192]],
            "steps": [8000, 16000, 28000, 40000], "cd_mean": [1.2903, 1.1863, 1.1854, 1.1679],
            "cd_std": [0.0190, 0.0406, 0.0705, 0.0769]},
    })


def c1():
    """Channel flow of examples/01_channel_flow.py at 64x21x21, 5000 steps."""
    _threads()
    out = {"setup": "64x21x21 D3Q19 BGK tau=0.8, halfway walls (solid rows y=0 and y=20, q=1/2 links), "
                    "periodic z, Zou-He velocity inlet u_x=0.05 at x=0 and pressure outlet rho=1 at x=63, "
                    "f0 = w, no guard, 5000 steps; u_x(y) at x=32 averaged over z",
           "generator": "oracle.lbm_oracle.channel_flow", "profiles": {}}
    for name, dt in (("f64", torch.float64), ("f32", torch.float32)):
        t0 = time.perf_counter()
        prof, f = O.channel_flow(nx=64, ny=21, nz=21, tau=0.8, u_in=0.05, n_steps=5000, dtype=dt)
        out["profiles"][name] = [float(v) for v in prof]
        print(f"c1 {name}: {time.perf_counter() - t0:.1f} s, u_max {float(prof.max()):.8f}", flush=True)
    _write("c1_channel_64x21x21.json", out)


def _sphere_setup(nx, ny, nz, center_lattice, dt):
    c = O.sphere_center_world(nx, ny, nz, center_lattice)
    D = 2.0 * 0.5 * nx / 8.0
    tau = 3.0 * (0.05 * D / 100.0) + 0.5
    return c, tau


def c2(kinds=("f64", "f32")):
    """Sphere drag at 64x32x32: MRT + hard mask with ray q, centre X=(16,16,16)
    (world (-2,0,0)), r=0.5 world (D=8), tau=0.512, f0 = w, pre-collision
    guard to eq(1, U), 8000 steps; Cd of every step, by the float64 and the
    float32 oracle (``kinds``; an existing fixture keeps the other kind)."""
    _threads()
    nx, ny, nz, T = 64, 32, 32, 8000
    path = os.path.join(HERE, "c2_sphere_cd_64x32x32.json")
    old = {}
    if os.path.exists(path):
        with open(path) as fh:
            old = json.load(fh)
    out = {"setup": "64x32x32 D3Q19 MRT tau=0.512 (nu = 0.05*8/100), hard sphere dist^2 <= r^2 with ray q, "
                    "centre world (-2,0,0) (continuous lattice X=(16,16,16)), r=0.5 world, periodic y/z, Zou-He "
                    "inlet u=0.05 / outlet rho=1, f0 = w, pre-collision guard to f_eq(1,(0.05,0,0)), 8000 steps; "
                    "Cd_t = |F_x| / (0.5 U^2 A + 1e-10), F from momentum exchange of (f_post_t, f_t+1)",
           "generator": "oracle.lbm_oracle.hot_step + momentum_exchange", "steps": T}
    for name in ("f64", "f32"):
        if name not in kinds and name in old:
            out[name] = old[name]
    for name, dt in (("f64", torch.float64), ("f32", torch.float32)):
        if name not in kinds:
            continue
        c, tau = _sphere_setup(nx, ny, nz, (16, 16, 16), dt)
        solid, q = O.create_sphere(nx, ny, nz, c, 0.5)
        q = q.to(dt)
        area = float(O.projected_area(solid))
        rho = torch.ones((nx, ny, nz), dtype=dt)
        u0 = torch.zeros((3, nx, ny, nz), dtype=dt)
        uref = u0.clone()
        uref[0] = 0.05
        f = O.equilibrium(rho, u0)
        kw = dict(tau=torch.tensor(tau, dtype=dt), velocity=torch.tensor([0.05, 0.0, 0.0], dtype=dt), q=q,
                  solid=solid, collision="mrt", guard="pre", f_eq_ref=O.equilibrium(rho, uref))
        cd = []
        t0 = time.perf_counter()
        with torch.no_grad():
            for t in range(T):
                f, fp = O.hot_step(f, **kw)
                fx = O.momentum_exchange(fp, f, q)
                cd.append(float(O.drag_coefficient(fx, 0.05, area)))
                if (t + 1) % 1000 == 0:
                    print(f"c2 {name}: step {t + 1}, Cd {cd[-1]:.6f}, {time.perf_counter() - t0:.0f} s", flush=True)
        out[name] = {"area": area, "tau": tau, "cd": cd}
    _write("c2_sphere_cd_64x32x32.json", out)


def c3(steps=2000, avg_last=600, seg=5):
    """dCd/dr (world units) by torch autograd of the float64 oracle for the
    96x48x48 sphere of examples/11: MRT, hard mask + ray q(r) from the smooth
    geometry (sharpness 20, float64 coordinates), centre world (-2,0,0), r=0.5,
    tau from the nominal D=12, f0 = f_eq(1, U), no guard; loss = mean Cd over
    the last `avg_last` steps.  Segments of `seg` steps are recomputed in the
    backward (torch.utils.checkpoint) to bound memory."""
    from torch.utils.checkpoint import checkpoint
    _threads()
    nx, ny, nz = 96, 48, 48
    dt = torch.float64
    c, tau = _sphere_setup(nx, ny, nz, (24, 24, 24), dt)
    r = torch.tensor(0.5, dtype=dt, requires_grad=True)
    _, solid, q = O.smooth_sphere_geometry(nx, ny, nz, c, r, 20.0, coord_dtype=dt)
    area = torch.clamp(O.projected_area(solid).to(dt), min=1.0)
    u = torch.zeros((3, nx, ny, nz), dtype=dt)
    u[0] = 0.05
    f = O.equilibrium(torch.ones((nx, ny, nz), dtype=dt), u)
    kw = dict(tau=torch.tensor(tau, dtype=dt), velocity=torch.tensor([0.05, 0.0, 0.0], dtype=dt), solid=solid,
              collision="mrt")

    def segment(f_in, q_in, n):
        fx = []
        ff = f_in
        for _ in range(n):
            ff, fp = O.hot_step(ff, q=q_in, **kw)
            fx.append(O.momentum_exchange(fp, ff, q_in)[0])
        return ff, torch.stack(fx)

    t0 = time.perf_counter()
    fxs = []
    done = 0
    while done < steps:
        n = min(seg, steps - done)
        # reentrant: the segment's forward keeps no graph (the non-reentrant
        # form held about 650 MB of activations per step here)
        f, fx = checkpoint(segment, f, q, n, use_reentrant=True)
        fxs.append(fx)
        done += n
        if done % 200 == 0:
            print(f"c3 forward: step {done}, {time.perf_counter() - t0:.0f} s", flush=True)
    fx = torch.cat(fxs)
    cd = fx.abs() / (0.5 * 0.05 * 0.05 * area + 1e-10)
    loss = cd[-avg_last:].mean()
    loss.backward()
    g = r.grad
    loss = loss.detach()
    print(f"c3: loss {float(loss):.10f} dCd/dr {float(g):.10f} ({time.perf_counter() - t0:.0f} s)", flush=True)
    _write("c3_sphere_dcd_dr_96x48x48.json", {
        "setup": "96x48x48 D3Q19 MRT tau=0.518 (nominal D=12), hard mask solid = sigmoid(20 sdf) > 1/2 and ray "
                 "q(r) from smooth_sphere_geometry (float64 coordinates, disc floor 1e-10, denominator eps 1e-20), "
                 "centre world (-2,0,0), r=0.5 world (6 lattice units), periodic y/z, Zou-He inlet u=0.05 / outlet "
                 "rho=1, f0 = f_eq(1,(0.05,0,0)), no guard, 2000 steps; loss = mean Cd over the last 600 steps, "
                 "area = max(A, 1); gradient in world units (divide by 12 for lattice units)",
        "generator": "torch.autograd through oracle.lbm_oracle (float64), segment recompute every 5 steps",
        "steps": steps, "avg_last": avg_last, "area": float(area), "loss": float(loss), "dloss_dr_world": float(g),
        "cd_tail": [float(v) for v in cd[-avg_last:].detach()]})


C2P_SEEDS = (1, 2, 3)


def c2p(seeds=C2P_SEEDS, merge=True):
    """The 64x32x32 sphere of ``c2`` in float32 from three initial states
    perturbed by 1e-8 (``perturbed_start``): how far float32 rounding alone
    moves the time-averaged Cd of this flow started from rest."""
    _threads()
    nx, ny, nz, T = 64, 32, 32, 8000
    dt = torch.float32
    c, tau = _sphere_setup(nx, ny, nz, (16, 16, 16), dt)
    solid, q = O.create_sphere(nx, ny, nz, c, 0.5)
    q = q.to(dt)
    area = float(O.projected_area(solid))
    rho = torch.ones((nx, ny, nz), dtype=dt)
    uref = torch.zeros((3, nx, ny, nz), dtype=dt)
    uref[0] = 0.05
    kw = dict(tau=torch.tensor(tau, dtype=dt), velocity=torch.tensor([0.05, 0.0, 0.0], dtype=dt), q=q,
              solid=solid, collision="mrt", guard="pre", f_eq_ref=O.equilibrium(rho, uref))
    path = os.path.join(HERE, "c2_sphere_cd_64x32x32_perturbed_f32.json")
    runs = {}
    if merge and os.path.exists(path):
        with open(path) as fh:
            runs = json.load(fh)["runs"]
    k = max(1, int(round(0.3 * T)))
    for seed in seeds:
        f = perturbed_start(nx, ny, nz, seed).to(dt)
        cd = []
        t0 = time.perf_counter()
        with torch.no_grad():
            for t in range(T):
                f, fp = O.hot_step(f, **kw)
                cd.append(float(O.drag_coefficient(O.momentum_exchange(fp, f, q), 0.05, area)))
        runs[str(seed)] = {"cd_tail30_mean": sum(cd[T - k:]) / k, "cd_every_500": cd[499::500]}
        print(f"c2p seed {seed}: tail mean {runs[str(seed)]['cd_tail30_mean']:.9f} "
              f"({time.perf_counter() - t0:.0f} s)", flush=True)
    _write("c2_sphere_cd_64x32x32_perturbed_f32.json", {
        "setup": "the c2 workload (c2_sphere_cd_64x32x32.json) in float32 with f0 = w + 1e-8 N(0,1), noise from "
                 "torch.Generator().manual_seed(seed) in float64, shape (19, nx, ny, nz), cast to float32",
        "generator": "oracle.lbm_oracle.hot_step + momentum_exchange (float32)", "steps": T, "amp": 1e-8,
        "tail_steps": k, "runs": runs})


SHARPNESS_ALPHAS = (5.0, 10.0, 20.0, 40.0, 80.0)


def ex12(alphas=SHARPNESS_ALPHAS, n_steps=2880):
    """examples/12_sharpness_sensitivity.py:53-79 forward in float32: 96x48x48,
    Re = 50 (tau = float32(3 * 0.05 * 12 / 50 + 1/2)), MRT, soft porosity
    sigmoid(alpha sdf) with ray q(r) (float32 coordinates), centre world 0,
    r = 0.5, f0 = f_eq(1, (0.05, 0, 0)), no guard, 2880 steps; Cd of the last
    (f_post, f_final) pair with area = max(A, 1)."""
    _threads()
    nx, ny, nz = 96, 48, 48
    dt = torch.float32
    u_in = 0.05
    tau = torch.tensor(3.0 * (u_in * (nx * 0.5 / 4.0) / 50.0) + 0.5, dtype=dt)
    out = {"setup": "96x48x48 D3Q19 MRT tau=float32(0.536) (Re=50, nominal D=12), soft porosity sigmoid(alpha sdf) "
                    "with ray q(r) from smooth_sphere_geometry (float32 coordinates), centre world 0, r=0.5 world, "
                    "periodic y/z, Zou-He inlet u=0.05 / outlet rho=1, f0 = f_eq(1,(0.05,0,0)), no guard, 2880 "
                    "steps; Cd of the last (f_post, f_final) pair, area = max(A, 1)",
           "generator": "oracle.lbm_oracle (float32)", "steps": n_steps, "cd": {}}
    for a in alphas:
        por, solid, q = O.smooth_sphere_geometry(nx, ny, nz, (0.0, 0.0, 0.0), torch.tensor(0.5, dtype=dt), a)
        u = torch.zeros((3, nx, ny, nz), dtype=dt)
        u[0] = u_in
        f = O.equilibrium(torch.ones((nx, ny, nz), dtype=dt), u)
        vel = torch.tensor([u_in, 0.0, 0.0], dtype=dt)
        area = torch.clamp(O.projected_area(solid, 0), min=1.0)
        t0 = time.perf_counter()
        with torch.no_grad():
            fp = f
            for _ in range(n_steps):
                fc = O.soft_bounce_back(O.mrt(f, tau), por)
                fp = fc
                f = O.zou_he_outlet(O.zou_he_inlet(O.stream(fc, q, True), vel), 1.0)
        cd = float(O.drag_coefficient(O.momentum_exchange(fp, f, q), u_in, area))
        out["cd"][str(a)] = cd
        print(f"ex12 alpha {a}: Cd {cd:.9f} ({time.perf_counter() - t0:.0f} s)", flush=True)
    _write("ex12_sharpness_oracle_f32.json", out)


if __name__ == "__main__":
    which = sys.argv[1:] or ["kats"]
    for w in which:
        if w.startswith("c2:"):
            c2(tuple(w[3:].split(",")))  # e.g. c2:f32 adds the float32 series
        elif w.startswith("c2p:"):
            c2p(tuple(int(v) for v in w[4:].split(",")))  # e.g. c2p:4,5 adds those seeds
        else:
            {"kats": kats, "c1": c1, "c2": c2, "c3": c3, "c2p": c2p, "ex12": ex12}[w]()
