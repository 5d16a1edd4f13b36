"""Host logic of the adjoint, checked on the CPU.

1. The segment-recompute checkpoint schedules yield every P_t once, in
   reverse order, for store-all, one-level and two-level policies.
2. The transpose of streaming decomposes into the kernel's owner writes
   (entry (x - e_i)_i gets mu(x)_i for a normal pull, 0 at x-face fallbacks
   and Bouzidi links) plus the scatter plan: checked against torch autograd
   of the oracle's stream() on a small grid with random Bouzidi links.
"""

import types
import weakref

import numpy as np
import pytest
import torch

from lbm_jax_autodiff_amd import adjoint as A
from oracle import lbm_oracle as O
from _helpers import random_q


class _CountingLattice:
    """Stand-in lattice whose state is a 1-element tensor holding t."""

    def __init__(self):
        self.steps = 0

    def new_state(self):
        return torch.zeros(1)

    def new_states(self, n):
        return [self.new_state() for _ in range(n)]

    def clone_state(self, buf):
        return buf.clone()

    def step(self, src, dst):
        self.steps += 1
        dst.copy_(src + 1)

    def chain(self, states, fpart=None, step=0):
        # the time index handed down must be that of the state (the
        # stochastic-rounding dither of a recomputed step depends on it)
        for k in range(len(states) - 1):
            assert float(states[k]) == float(step + k)
            self.step(states[k], states[k + 1])

    def pingpong(self, cur, spare_a, spare_b, n, fpart=None, step=0):
        first = spare_a if cur is not spare_a else spare_b
        second = spare_b if first is spare_a else spare_a
        seq = [cur] + [first if k % 2 == 0 else second for k in range(n)]
        self.chain(seq, step=step)
        return seq[-1]


@pytest.mark.parametrize("policy,n", [("all", 7), (3, 10), (4, 4), ((7, 3), 23), ((200, 14), 431), ((5, 5), 12)])
def test_checkpoint_schedules(policy, n):
    lat = _CountingLattice()
    ck = A._Checkpoints(lat, policy, n)
    cur = torch.zeros(1)
    for t in range(n):
        ck.record(t, cur)
        if t < n - 1:
            nxt = torch.zeros(1)
            nxt.copy_(cur + 1)
            cur = nxt
    seen = []
    for t, p in ck.descending():
        assert float(p) == float(t)
        seen.append(t)
    assert seen == list(range(n - 1, -1, -1))


class _TrackingLattice(_CountingLattice):
    """Counting lattice that also tracks how many state buffers are alive."""

    def __init__(self):
        super().__init__()
        self.refs = []
        self.peak = 0

    def _track(self, t):
        self.refs.append(weakref.ref(t))
        self.peak = max(self.peak, self.live())
        return t

    def live(self):
        return sum(1 for r in self.refs if r() is not None)

    def new_state(self):
        return self._track(torch.zeros(1))

    def clone_state(self, buf):
        return self._track(buf.clone())


@pytest.mark.parametrize("policy,n", [(100, 2800), ((200, 14), 4000), ((600, 25), 12000)])
def test_checkpoint_peak_states(policy, n):
    """At most about N/K1 + K1/K2 + K2 states are alive while the backward
    walks the segments, with the consumer's loop variable still holding the
    previous segment (as the rollout's backward loop does)."""
    lat = _TrackingLattice()
    ck = A._Checkpoints(lat, policy, n)
    a, b = lat.new_state(), lat.new_state()
    cur, t = a, 0
    while t < n - 1:
        ck.record(t, cur)
        end = min(n - 1, (t // ck.stride + 1) * ck.stride)
        cur = lat.pingpong(cur, a, b, end - t, step=t)
        t = end
    ck.record(n - 1, cur)
    expect_t = n - 1
    for t0, t1, have in ck.segments():
        for t in range(t1 - 1, t0 - 1, -1):
            assert float(have[t]) == float(t) and t == expect_t
            expect_t -= 1
    assert expect_t == -1
    if isinstance(policy, int):
        bound = -(-n // policy) + policy
    else:
        k1, k2 = policy
        bound = -(-n // k1) + -(-k1 // k2) + k2
    # + the two ping-pong buffers of each recompute level and the forward's pair
    assert lat.peak <= bound + 6, (lat.peak, bound)


def _owner_transpose(mu, q, nx, ny, nz):
    """Kernel owner writes, emulated: lam(x - e_i)_i = mu(x)_i unless x keeps
    its own value (x faces) or direction i is a Bouzidi link at x."""
    lam = torch.zeros_like(mu)
    for i in range(19):
        ex, ey, ez = O.VELOCITIES[i]
        for x in range(nx):
            fallback = (ex == 1 and x == 0) or (ex == -1 and x == nx - 1)
            link = (q[O.OPPOSITE[i], x] >= 0) if i else torch.zeros((ny, nz), dtype=torch.bool)
            val = torch.where(link, torch.zeros_like(mu[i, x]), mu[i, x])
            if fallback:
                val = torch.zeros_like(val)
            tx = (x - ex) % nx
            lam[i, tx] = torch.roll(val, shifts=(-ey, -ez), dims=(0, 1))
    return lam


def test_stream_transpose_decomposition():
    nx, ny, nz = 7, 5, 4
    g = torch.Generator().manual_seed(0)
    P = torch.rand((19, nx, ny, nz), generator=g, dtype=torch.float64, requires_grad=True)
    q = random_q(nx, ny, nz, seed=3, p_wall=0.2)
    S = O.stream(P, q, True)
    mu = torch.randn(S.shape, generator=g, dtype=torch.float64)
    # Zou-He overwrites the x-face fallback populations, so their cotangent is 0
    for i in range(19):
        ex = O.VELOCITIES[i][0]
        if ex == 1:
            mu[i, 0] = 0.0
        if ex == -1:
            mu[i, nx - 1] = 0.0
    (ref,) = torch.autograd.grad(S, P, mu)

    lam = _owner_transpose(mu, q, nx, ny, nz)
    # scatter plan on a stand-in lattice carrying the compiled links
    N = nx * ny * nz
    qs = q.permute(0, 3, 2, 1).contiguous().reshape(19, N)  # lattice storage (x fastest)
    has = (qs[1:] >= 0).any(0)
    cells = torch.nonzero(has).reshape(-1)
    links = types.SimpleNamespace(n=int(cells.numel()), cell=cells.to(torch.int32), q=qs[1:, cells].clone())
    lat = types.SimpleNamespace(links=links, nx=nx, ny=ny, nz=nz, N=N, P=N, RS=nx, device=torch.device("cpu"))
    tgt, start, src, coef = A.scatter_plan(lat)
    mu_s = mu.permute(0, 3, 2, 1).contiguous().reshape(19, N)
    mu_link = torch.zeros((19, links.n), dtype=torch.float64)
    for i in range(1, 19):
        link_i = qs[O.OPPOSITE[i], cells] >= 0
        mu_link[i] = torch.where(link_i, mu_s[i, cells], torch.zeros(()))
    lam_s = lam.permute(0, 3, 2, 1).contiguous().reshape(-1)
    mu_flat = mu_link.reshape(-1)
    for k in range(tgt.numel()):
        acc = 0.0
        for e in range(int(start[k]), int(start[k + 1])):
            s = int(src[e])
            if s >= 0:
                acc += float(coef[e]) * float(mu_flat[s])
        lam_s[int(tgt[k])] += acc
    lam = lam_s.reshape(19, nz, ny, nx).permute(0, 3, 2, 1)
    assert float((lam - ref).abs().max()) < 1e-13


def test_scatter_plan_is_sorted_and_complete():
    nx, ny, nz = 6, 5, 4
    q = random_q(nx, ny, nz, seed=5, p_wall=0.3)
    N = nx * ny * nz
    qs = q.permute(0, 3, 2, 1).contiguous().reshape(19, N)
    cells = torch.nonzero((qs[1:] >= 0).any(0)).reshape(-1)
    links = types.SimpleNamespace(n=int(cells.numel()), cell=cells.to(torch.int32), q=qs[1:, cells].clone())
    lat = types.SimpleNamespace(links=links, nx=nx, ny=ny, nz=nz, N=N, P=N, RS=nx, device=torch.device("cpu"))
    tgt, start, src, coef = A.scatter_plan(lat)
    t = tgt.numpy()
    assert np.all(np.diff(t) > 0)
    n_links = int((qs[1:, cells] >= 0).sum())
    n_force = int((src < 0).sum())
    assert n_force == n_links
    n_hi = int((qs[1:, cells][qs[1:, cells] >= 0] >= 0.5).sum())
    assert int(src.numel()) == n_force + n_links + n_links  # own + (own i | neighbour) + force
    assert n_hi <= n_links
    # entry (x, j) gets at most four additions: the own term of link (x, opp j), the high-branch
    # and force terms of link (x, j), the low-branch neighbour term of link (x + e_j, opp j);
    # one load batch of the scatter kernel covers it
    per_target = (start[1:] - start[:-1]).numpy()
    assert per_target.min() >= 1 and per_target.max() <= 4


def _plan_by_entry(plan, decode):
    """{(population, cell): sorted [(src, coef)]} of a scatter plan."""
    tgt, start, src, coef = plan
    out = {}
    for k in range(tgt.numel()):
        ent = [(int(src[e]), float(coef[e])) for e in range(int(start[k]), int(start[k + 1]))]
        out[decode(int(tgt[k]))] = sorted(ent)
    return out


def test_scatter_plan_follows_the_state_layout():
    """Targets address population i of cell (x, y, z) at
    i * P + (z * ny + y) * RS + x: the dense array, a padded population stride
    and the row-interleaved layout give the same entries."""
    nx, ny, nz = 6, 5, 4
    q = random_q(nx, ny, nz, seed=5, p_wall=0.3)
    N = nx * ny * nz
    qs = q.permute(0, 3, 2, 1).contiguous().reshape(19, N)
    cells = torch.nonzero((qs[1:] >= 0).any(0)).reshape(-1)
    links = types.SimpleNamespace(n=int(cells.numel()), cell=cells.to(torch.int32), q=qs[1:, cells].clone())

    def lattice(P, RS):
        return types.SimpleNamespace(links=links, nx=nx, ny=ny, nz=nz, N=N, P=P, RS=RS, device=torch.device("cpu"))

    def decoder(P, RS):
        def dec(t):
            if RS == nx:
                return t // P, t % P
            row, rem = divmod(t, RS)
            return rem // P, row * nx + rem % P
        return dec

    ref = _plan_by_entry(A.scatter_plan(lattice(N, nx)), decoder(N, nx))
    for P, RS in ((N + 37, nx), (nx, 19 * nx), (nx + 2, 19 * (nx + 2) + 5)):
        plan = A.scatter_plan(lattice(P, RS))
        assert np.all(np.diff(plan[0].numpy()) > 0)  # sorted, unique targets
        assert _plan_by_entry(plan, decoder(P, RS)) == ref
