"""Fused-step parity: Lattice.run (one kernel launch per step) against the
oracle's unfused rollout (oracle.hot_step), for every variant of the step."""

import itertools
import json
import os

import numpy as np
import pytest
import torch

from lbm_jax_autodiff_amd.engine import Lattice
from oracle import lbm_oracle as O
from _helpers import near_eq_state, random_q, cpu, maxabs

pytestmark = pytest.mark.gpu

# nx = 20: at 4 (2) cells per thread a row is 5 (10) lanes, so rows straddle
# wavefront boundaries and both the lane-shuffle and the wave-edge loads run
NX, NY, NZ = 20, 9, 6


def _oracle(f0, steps, **kw):
    with torch.no_grad():
        return O.rollout(f0, steps, forces=True, **kw)


def _run(dev, f0, steps, storage, collision, bounce, guard, smag, periodic, q, solid, phi, tau=0.6,
         vel=(0.05, 0.0, 0.0), uref=None, vec=0):
    lat = Lattice(NX, NY, NZ, tau=tau, inlet_velocity=vel, collision=collision, cells_per_thread=vec,
                  smagorinsky_cs=0.1 if smag else None,
                  solid=solid.to(dev) if (bounce == "hard") else None,
                  porosity=phi.to(dev) if bounce == "soft" else None,
                  q=q.to(dev) if q is not None else None, guard=guard, guard_velocity=uref,
                  periodic_yz=periodic, storage=storage, device=dev)
    dt = torch.float64 if storage == "f64" else torch.float32
    return lat.run(f0.to(dt).to(dev), steps)


CASES = list(itertools.product(["bgk", "mrt"], ["none", "hard", "soft"], ["none", "post", "pre"], [False, True],
                               [True, False], [0.1, 0.0], [1, 2, 4]))


@pytest.mark.parametrize("collision,bounce,guard,smag,periodic,p_wall,vec", CASES)
def test_fused_matches_oracle_f64(dev, collision, bounce, guard, smag, periodic, p_wall, vec):
    steps = 6
    f0 = near_eq_state(NX, NY, NZ, amp=2e-3)
    q = random_q(NX, NY, NZ, seed=4, p_wall=p_wall)
    g = torch.Generator().manual_seed(2)
    solid = (torch.rand((NX, NY, NZ), generator=g) < 0.1).to(torch.int32)
    phi = 0.3 * torch.rand((NX, NY, NZ), generator=g, dtype=torch.float64)
    vel = torch.tensor([0.05, 0.0, 0.0], dtype=torch.float64)
    uref = (0.05, 0.0, 0.0)
    feq_ref = O.equilibrium(torch.ones((NX, NY, NZ), dtype=torch.float64),
                            torch.stack([torch.full((NX, NY, NZ), 0.05, dtype=torch.float64),
                                         torch.zeros((NX, NY, NZ), dtype=torch.float64),
                                         torch.zeros((NX, NY, NZ), dtype=torch.float64)]))
    s_ref, p_ref, f_ref = _oracle(
        f0, steps, tau=torch.tensor(0.6, dtype=torch.float64), velocity=vel, q=q,
        solid=solid if bounce == "hard" else None, porosity=phi if bounce == "soft" else None,
        collision=collision, smagorinsky_cs=0.1 if smag else None, periodic_yz=periodic, guard=guard,
        f_eq_ref=feq_ref)
    s, p, fh = _run(dev, f0, steps, "f64", collision, bounce, guard, smag, periodic, q, solid, phi, uref=uref,
                    vec=vec)
    assert maxabs(s, s_ref) < 1e-12
    assert maxabs(p, p_ref) < 1e-12
    assert maxabs(fh, f_ref) < 1e-11


@pytest.mark.parametrize("vec", [1, 4])
@pytest.mark.parametrize("collision", ["bgk", "mrt"])
def test_fused_f32_close_to_f64_oracle(dev, collision, vec):
    steps = 20
    f0 = near_eq_state(NX, NY, NZ, amp=2e-3)
    q = random_q(NX, NY, NZ, seed=4, p_wall=0.1)
    vel = torch.tensor([0.05, 0.0, 0.0], dtype=torch.float64)
    s_ref, p_ref, f_ref = _oracle(f0, steps, tau=torch.tensor(0.6, dtype=torch.float64), velocity=vel, q=q,
                                  collision=collision)
    s, p, fh = _run(dev, f0, steps, "f32", collision, "none", "none", False, True, q, None, None, vec=vec)
    assert maxabs(s, s_ref) < 5e-6
    assert maxabs(p, p_ref) < 5e-6
    assert maxabs(fh, f_ref) < 1e-4


@pytest.mark.parametrize("vec", [1, 2, 4])
@pytest.mark.parametrize("storage,tol", [("f16", 3e-4), ("bf16", 3e-3)])
def test_half_storage_close_to_f32(dev, storage, tol, vec):
    steps = 10
    f0 = near_eq_state(NX, NY, NZ, amp=2e-3)
    q = random_q(NX, NY, NZ, seed=4, p_wall=0.02)
    s32, p32, f32 = _run(dev, f0, steps, "f32", "mrt", "none", "none", False, True, q, None, None, vec=1)
    sh, ph, fh = _run(dev, f0, steps, storage, "mrt", "none", "none", False, True, q, None, None, vec=vec)
    assert maxabs(sh, s32) < tol
    assert maxabs(ph, p32) < tol


def test_guard_resets_only_bad_cells(dev):
    """Forced NaN / out-of-range cells are reset, the others untouched (state.py:158-169)."""
    f0 = near_eq_state(NX, NY, NZ, amp=1e-3)
    bad = [(3, 2, 1), (7, 4, 3), (10, 8, 5)]
    f0[:, 3, 2, 1] = float("nan")
    f0[:, 7, 4, 3] *= 3.0   # rho ~ 3 > 2
    fast = O.equilibrium(torch.ones((1,), dtype=torch.float64),
                         torch.tensor([[0.5], [0.0], [0.0]], dtype=torch.float64))
    f0[:, 10, 8, 5] = fast[:, 0]  # rho = 1, |u| = 0.5 > 0.4
    vel = torch.tensor([0.05, 0.0, 0.0], dtype=torch.float64)
    s_ref, p_ref, _ = _oracle(f0, 1, tau=torch.tensor(0.6, dtype=torch.float64), velocity=vel,
                              q=-torch.ones((19, NX, NY, NZ), dtype=torch.float64), collision="bgk", guard="pre",
                              f_eq_ref=O.equilibrium(torch.ones((NX, NY, NZ), dtype=torch.float64),
                                                     torch.zeros((3, NX, NY, NZ), dtype=torch.float64)))
    lat = Lattice(NX, NY, NZ, tau=0.6, collision="bgk", guard="pre", guard_velocity=(0, 0, 0), storage="f64",
                  device=dev)
    s, p, _ = lat.run(f0.to(dev), 1)
    p = cpu(p)
    assert torch.isfinite(p).all()
    assert maxabs(p, p_ref) < 1e-12
    w = O.weights(torch.float64)
    for (x, y, z) in bad:
        # reset to eq(1, 0) = w before collision; BGK then moves it only by the
        # float32 rounding of the weights (their sum is 1 + O(1e-8))
        assert float((p[:, x, y, z] - w).abs().max()) < 1e-7
    good = torch.ones((NX, NY, NZ), dtype=torch.bool)
    for (x, y, z) in bad:
        good[x, y, z] = False
    dev_from_w = (p - w.reshape(19, 1, 1, 1)).abs().amax(0)
    assert bool((dev_from_w[good] > 1e-6).all())  # no other cell was reset


def test_poiseuille_kat(dev):
    """60x20x4, tau=0.8, 3000 BGK steps: the reference's published profile to 6
    decimals (README.md:59-69, docs/getting_started/first_simulation.md:101-109)."""
    nx, ny, nz = 60, 20, 4
    solid = O.channel_walls(nx, ny, nz)
    q = O.compute_q_from_solid(solid, nx, ny, nz, 0.5)
    f0 = O.equilibrium(torch.ones((nx, ny, nz), dtype=torch.float32), torch.zeros((3, nx, ny, nz)))
    lat = Lattice(nx, ny, nz, tau=0.8, collision="bgk", solid=solid.to(dev), q=q.to(dev), storage="f32",
                  device=dev)
    s, _, _ = lat.run(f0.to(dev), 3000, forces=False)
    s = cpu(s).to(torch.float64)
    rho = s.sum(0)
    ux = (s * O.velocities(torch.float64)[:, 0].reshape(19, 1, 1, 1)).sum(0) / rho
    prof = ux[nx // 2].mean(-1)
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "kats.json")) as fh:
        published = json.load(fh)["poiseuille_60x20x4"]["ux_by_y"]
    for y, v in published.items():
        assert abs(float(prof[int(y)]) - v) <= 6e-7, (y, float(prof[int(y)]), v)


@pytest.mark.parametrize("vec", [1, 4])
@pytest.mark.parametrize("storage", ["f32", "f16", "f64"])
def test_state_layout_does_not_change_results(dev, storage, vec):
    """The row-interleaved layout (default, also with padded rows), the padded
    population stride and the dense (19, nz, ny, nx) array give bit-identical
    states and forces."""
    steps = 5
    f0 = near_eq_state(NX, NY, NZ, amp=2e-3)
    q = random_q(NX, NY, NZ, seed=4, p_wall=0.1)
    g = torch.Generator().manual_seed(2)
    solid = (torch.rand((NX, NY, NZ), generator=g) < 0.1).to(torch.int32)
    dt = torch.float64 if storage == "f64" else torch.float32
    outs = []
    for layout, stride in (("rows", None), ("rows", NX + 4), ("apart", None), ("apart", NX * NY * NZ),
                           ("apart", NX * NY * NZ + 4 * 37)):
        lat = Lattice(NX, NY, NZ, tau=0.6, collision="mrt", solid=solid.to(dev), q=q.to(dev), guard="pre",
                      storage=storage, device=dev, cells_per_thread=vec, pop_stride=stride, layout=layout)
        outs.append([cpu(t) for t in lat.run(f0.to(dt).to(dev), steps)])
    assert outs[0][0].shape == (19, NX, NY, NZ)
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert torch.equal(a, b)


@pytest.mark.parametrize("storage,layout", [("f16", "rows"), ("f32", "apart"), ("f64", "rows")])
def test_states_from_one_allocation(dev, storage, layout):
    """Lattice.new_states carves n state buffers from one allocation: each
    starts on a 256-byte boundary, none overlaps another, and a chain through
    them gives the same states and forces as separately allocated buffers."""
    nx, ny, nz = 13, 7, 5  # an extent that is not a multiple of 256 bytes
    f0 = near_eq_state(nx, ny, nz, amp=2e-3)
    q = random_q(nx, ny, nz, seed=4, p_wall=0.1)
    dt = torch.float64 if storage == "f64" else torch.float32
    lat = Lattice(nx, ny, nz, tau=0.6, collision="mrt", q=q.to(dev), guard="pre", storage=storage, device=dev,
                  layout=layout)
    n = 6
    bufs = lat.new_states(n)
    esize = bufs[0].element_size()
    spans = []
    for b in bufs:
        assert b.data_ptr() % 256 == 0
        assert b.shape == (19, nz, ny, nx) and b.stride() == (lat.P, ny * lat.RS, lat.RS, 1)
        spans.append((b.data_ptr(), b.data_ptr() + lat.extent * esize))
    spans.sort()
    assert all(e0 <= s1 for (_, e0), (s1, _) in zip(spans, spans[1:]))
    s0 = lat.to_state(f0.to(dt).to(dev))
    out = []
    for states in (bufs, [lat.new_state() for _ in range(n)]):
        states[0].copy_(s0)
        fp = torch.zeros((n - 1, max(lat.slots, 1), 3), dtype=torch.float64, device=dev)
        lat.chain(states, fp if lat.slots else None, step=0)
        out.append([cpu(lat.from_state(states[-1])), cpu(fp)])
    for a, b in zip(*out):
        assert torch.equal(a, b)


def test_step_loop_reuses_the_lattice_until_the_geometry_changes(dev):
    """A loop of step() calls compiles the link list once and equals
    simulate() bit for bit; an in-place write to q, or another q tensor,
    builds a new lattice whose result equals a freshly built one."""
    import lbm_jax_autodiff_amd as lbm
    from lbm_jax_autodiff_amd import state as state_mod
    nx, ny, nz = 32, 16, 16
    f0 = near_eq_state(nx, ny, nz, amp=2e-3).to(dev, torch.float32)
    q = random_q(nx, ny, nz, seed=9, p_wall=0.05).to(dev, torch.float32)
    g = torch.Generator().manual_seed(3)
    solid = (torch.rand((nx, ny, nz), generator=g) < 0.05).to(torch.int32).to(dev)
    params = lbm.LBMParams(0.6, (0.05, 0.0, 0.0), 1)
    st = lbm.LBMState(f0, solid, q)
    st = lbm.step(st, params)
    lat = state_mod._last_lattice[-1]
    for _ in range(4):
        st = lbm.step(st, params)
        assert state_mod._last_lattice[-1] is lat
    ref = lbm.simulate(lbm.LBMState(f0, solid, q), params, 5)
    assert torch.equal(st.f, ref.f)

    def fresh(qq):
        lt = Lattice(nx, ny, nz, tau=0.6, inlet_velocity=(0.05, 0.0, 0.0), collision="mrt", solid=solid, q=qq,
                     guard="post", device=dev)
        return lt.run(f0, 1, forces=False)[0]

    q[q >= 0] = 0.5  # in place: same tensor, new version
    one = lbm.step(lbm.LBMState(f0, solid, q), params)
    assert state_mod._last_lattice[-1] is not lat
    assert torch.equal(one.f, fresh(q))
    lat2 = state_mod._last_lattice[-1]
    q2 = random_q(nx, ny, nz, seed=10, p_wall=0.05).to(dev, torch.float32)
    two = lbm.step(lbm.LBMState(f0, solid, q2), params)
    assert state_mod._last_lattice[-1] is not lat2
    assert torch.equal(two.f, fresh(q2))


ROW_CASES = list(itertools.product(["f16", "bf16"], ["bgk", "mrt"], ["none", "hard", "soft"],
                                   ["none", "post", "pre"], [True, False]))


@pytest.mark.parametrize("dims", [(64, 7, 5), (128, 5, 3)])
@pytest.mark.parametrize("storage,collision,bounce,guard,periodic", ROW_CASES)
def test_row_sweep_matches_per_lane_addressing(dev, storage, collision, bounce, guard, periodic, dims):
    """At nx = 64 the 16-bit sweep keeps each wave's row addresses in scalar
    registers (and reads the wrapped neighbour where Zou-He overwrites the
    pulled value); cells_per_thread=1 forces per-lane addresses.  The two
    kernels are compiled separately (fused multiply-adds may round
    differently), so states and forces agree to one 16-bit rounding of a
    deviation, with Smagorinsky, links and clamped y/z faces."""
    nx, ny, nz = dims  # nx = 128: two waves per row, two rows per workgroup
    steps = 6
    f0 = near_eq_state(nx, ny, nz, amp=2e-3)
    q = random_q(nx, ny, nz, seed=4, p_wall=0.1)
    g = torch.Generator().manual_seed(2)
    solid = (torch.rand((nx, ny, nz), generator=g) < 0.1).to(torch.int32)
    phi = 0.3 * torch.rand((nx, ny, nz), generator=g, dtype=torch.float32)
    outs = []
    for cpt in (0, 1):
        lat = Lattice(nx, ny, nz, tau=0.6, inlet_velocity=(0.05, 0.0, 0.0), collision=collision,
                      smagorinsky_cs=0.1, solid=solid.to(dev) if bounce == "hard" else None,
                      porosity=phi.to(dev) if bounce == "soft" else None, q=q.to(dev), guard=guard,
                      guard_velocity=(0.05, 0.0, 0.0), periodic_yz=periodic, storage=storage, device=dev,
                      offset="inlet", cells_per_thread=cpt)
        outs.append([cpu(t).double() for t in lat.run(f0.to(torch.float32).to(dev), steps)])
    tol = 5e-5 if storage == "f16" else 5e-4
    for a, b in zip(*outs):
        assert float((a - b).abs().max()) <= tol * max(1.0, float(a.abs().max()))


@pytest.mark.parametrize("storage,tdt", [("f16", torch.float16), ("bf16", torch.bfloat16)])
def test_half_storage_rounds_to_nearest_even(dev, storage, tdt):
    """Encoding a state rounds to nearest, ties to even, exactly as torch's
    conversion does: exact ties, values just above and below them, and
    random values over a wide range (offsets 0, so the stored value is f)."""
    nx, ny, nz = 64, 4, 4
    n = 19 * nx * ny * nz
    g = torch.Generator().manual_seed(7)
    if storage == "bf16":
        hi = torch.randint(0, 0x7f00, (n,), generator=g, dtype=torch.int64)  # finite magnitudes
        low = torch.tensor([0x8000, 0x8001, 0x7fff, 0x0000], dtype=torch.int64)[torch.arange(n) % 4]
        bits = (hi << 16) | low
        sign = torch.randint(0, 2, (n,), generator=g, dtype=torch.int64) << 31
        f = torch.tensor(((bits | sign) & 0xffffffff).to(torch.int32).numpy().view("float32"))
    else:
        f = torch.randn(n, generator=g, dtype=torch.float64) * torch.exp2(torch.randint(-20, 12, (n,), generator=g))
        f = f.to(torch.float32)
        # exact ties: midpoints of neighbouring half values
        h = torch.randn(n // 4, generator=g).to(torch.float16)
        up = torch.nextafter(h, torch.full_like(h, float("inf")))
        f[: n // 4] = ((h.to(torch.float64) + up.to(torch.float64)) / 2).to(torch.float32)
    f = f.reshape(19, nx, ny, nz)
    lat = Lattice(nx, ny, nz, tau=0.6, storage=storage, offset=[0.0] * 19, device=dev, rounding="nearest")
    back = cpu(lat.from_state(lat.to_state(f.to(dev))))
    assert torch.equal(back, f.to(tdt).to(torch.float32))


def test_half_storage_deviation_rounded_once(dev):
    """With non-zero offsets the binary16 state holds f - off_i rounded once
    to nearest even (the exact difference, as numpy's float64 -> float16
    conversion rounds it), and decoding returns h + off_i rounded once to
    float32.  Deviations span the subnormal and normal half range."""
    nx, ny, nz = 64, 4, 4
    n = nx * ny * nz
    g = torch.Generator().manual_seed(11)
    off = (torch.rand(19, generator=g, dtype=torch.float64) * 0.4 + 0.01).to(torch.float32)
    dev_ = torch.randn(19, n, generator=g, dtype=torch.float64) * torch.exp2(
        torch.randint(-26, -1, (19, n), generator=g).to(torch.float64))
    f = (off.to(torch.float64)[:, None] + dev_).to(torch.float32)
    lat = Lattice(nx, ny, nz, tau=0.6, storage="f16", offset=off.tolist(), device=dev, rounding="nearest")
    fs = f.reshape(19, nx, ny, nz)
    buf = lat.to_state(fs.to(dev))
    raw = cpu(buf.permute(0, 3, 2, 1)).reshape(19, n).numpy()
    exact = f.to(torch.float64).numpy() - off.to(torch.float64).numpy()[:, None]
    want = exact.astype(np.float16)
    assert np.array_equal(raw.view(np.uint16), want.view(np.uint16))
    back = cpu(lat.from_state(buf)).reshape(19, n).numpy()
    want_back = (want.astype(np.float64) + off.to(torch.float64).numpy()[:, None]).astype(np.float32)
    assert np.array_equal(back, want_back)


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_stochastic_rounding_is_unbiased(dev, storage):
    """With rounding="stochastic" each stored deviation is one of the two
    16-bit neighbours of the exact deviation, the upper one with probability
    (d - lo) / (hi - lo): over 256 time indices the mean stored value matches
    the exact value (no bias over all values to a small fraction of an ulp),
    each value's frequency agrees with its probability, the dither is a
    function of (seed, step) only, and a different seed draws differently.
    Deviations span the binary16 subnormals (float16) and lie well inside
    their binades, so the float32 input holds each one to far below a 16-bit
    unit in the last place."""
    nx, ny, nz = 64, 4, 4
    n = nx * ny * nz
    g = torch.Generator().manual_seed(5)
    off = (torch.rand(19, generator=g, dtype=torch.float64) * 0.3 + 0.02).to(torch.float32)
    lowest = -22 if storage == "f16" else -12
    mag = torch.exp2(torch.randint(lowest, -3, (19, n), generator=g).to(torch.float64))
    d = mag * (1.05 + 0.9 * torch.rand(19, n, generator=g, dtype=torch.float64))  # inside [2^e, 2^(e+1))
    d = d * torch.where(torch.rand(19, n, generator=g) < 0.5, -1.0, 1.0).to(torch.float64)
    f = (off.to(torch.float64)[:, None] + d).to(torch.float32)
    exact = f.to(torch.float64).numpy() - off.to(torch.float64).numpy()[:, None]
    lat = Lattice(nx, ny, nz, tau=0.6, storage=storage, offset=off.tolist(), device=dev, rounding="stochastic",
                  seed=123)
    fs = f.reshape(19, nx, ny, nz).to(dev)

    def stored(step, latt=lat):
        buf = latt.to_state(fs, step=step)
        raw = cpu(buf.permute(0, 3, 2, 1)).reshape(19, n)
        return raw.to(torch.float64).numpy()

    K = 256
    draws = np.stack([stored(k) for k in range(K)])
    # neighbours of the exact deviation in the 16-bit format
    tdt = torch.float16 if storage == "f16" else torch.bfloat16
    near = torch.tensor(exact).to(tdt)
    towards = torch.where(torch.tensor(exact) > near.to(torch.float64), torch.full_like(near, float("inf")),
                          torch.full_like(near, float("-inf")))
    other = torch.nextafter(near, towards)
    lo = torch.minimum(near, other).to(torch.float64).numpy()
    hi = torch.maximum(near, other).to(torch.float64).numpy()
    assert np.all((draws == lo[None]) | (draws == hi[None]))
    p = (exact - lo) / (hi - lo)
    freq = (draws == hi[None]).mean(0)
    # per value: binomial spread of K draws
    assert np.abs(freq - p).max() < 6.0 * np.sqrt(0.25 / K)
    # over all values: the mean error in ulps has no bias
    bias = ((draws.mean(0) - exact) / (hi - lo)).mean()
    assert abs(bias) < 6.0 * np.sqrt(0.25 / K / exact.size), bias
    assert np.array_equal(stored(7), draws[7])
    lat2 = Lattice(nx, ny, nz, tau=0.6, storage=storage, offset=off.tolist(), device=dev, rounding="stochastic",
                   seed=124)
    assert not np.array_equal(stored(7, lat2), draws[7])


def test_default_state_layouts(dev):
    lat = Lattice(256, 128, 128, tau=0.6, device=dev)
    assert lat.layout == "rows" and lat.P == 256 and lat.RS == 19 * 256
    buf = lat.new_state()
    assert buf.stride() == (256, 128 * 19 * 256, 19 * 256, 1) and buf.untyped_storage().nbytes() == 19 * lat.N * 4
    lat = Lattice(256, 128, 128, tau=0.6, device=dev, layout="apart")
    assert lat.P > lat.N and lat.P % 8192 == 1024
    buf = lat.new_state()
    assert buf.stride() == (lat.P, 256 * 128, 256, 1)


def _bench_like_lattice(dev, storage="f32"):
    solid = torch.zeros((NX, NY, NZ), dtype=torch.int32)
    solid[8:12, 3:6, 2:4] = 1
    q = O.compute_q_from_solid(solid, NX, NY, NZ, 0.4)
    return Lattice(NX, NY, NZ, tau=0.6, collision="mrt", solid=solid.to(dev), q=q.to(dev), guard="pre",
                   storage=storage, device=dev)


def test_step_on_user_stream_matches_default_stream(dev):
    """The link launches fork from and join back into the caller's stream."""
    f0 = near_eq_state(NX, NY, NZ, amp=2e-3).to(torch.float32)
    lat = _bench_like_lattice(dev)
    assert lat.links.n > 0
    ref = [cpu(t) for t in lat.run(f0.to(dev), 7)]
    side = torch.cuda.Stream(device=dev)
    with torch.cuda.stream(side):
        out = lat.run(f0.to(dev), 7)
        got = [t.to("cpu") for t in out]
    torch.cuda.synchronize()
    for a, b in zip(ref, got):
        assert torch.equal(a, b)


def test_step_graph_replays_match_eager(dev):
    """One fused step captured in a graph (fork/join included) replays
    identically; each replay is checked against eager steps."""
    f0 = near_eq_state(NX, NY, NZ, amp=2e-3).to(torch.float32)
    lat = _bench_like_lattice(dev)
    s0 = lat.to_state(f0.to(dev))
    a, b = lat.new_state(), lat.new_state()
    lat.collide(s0, a)
    fp = torch.zeros((max(lat.slots, 1), 3), dtype=torch.float64, device=dev)
    ea, eb = lat.clone_state(a), lat.new_state()
    efp = torch.zeros_like(fp)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        lat.step(a, b, fp)
    for rep in range(4):
        g.replay()
        lat.step(ea, eb, efp)  # eager reference of the same step
        torch.cuda.synchronize()
        assert torch.equal(cpu(b), cpu(eb)), rep
        assert torch.equal(cpu(fp), cpu(efp)), rep
        a.copy_(b)
        ea.copy_(eb)
        torch.cuda.synchronize()
