// Reverse-mode products (VJPs) of the per-operator surface.
//
// One kernel per operator, one thread per cell, dense (19, nz, ny, nx)
// arrays as in the forward per-operator kernels.  Each takes the forward
// inputs it needs and the cotangent of the forward output, and writes the
// cotangents of the inputs.  Every output entry is written by exactly one
// thread (gathers, no atomics), so the results are deterministic.  The
// conventions are JAX's transposition of the reference graph: `where` routes
// the cotangent to the taken branch, `clip` passes it inside its bounds only,
// comparisons carry none.
#include <cmath>

#include "lbm_adjoint_physics.hpp"
#include "lbm_geometry.hpp"

namespace lbm {

constexpr int kVjpBlock = 256;

inline unsigned vjp_grid(size_t n) { return static_cast<unsigned>((n + kVjpBlock - 1) / kVjpBlock); }

inline bool vjp_dims_ok(int nx, int ny, int nz) {
  if (nx <= 0 || ny <= 0 || nz <= 0) return false;
  return 1LL * nx * ny * nz * Q < (1LL << 31);
}

template <typename R>
__device__ __forceinline__ void vload(const R* __restrict__ f, size_t N, size_t c, R* v) {
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    v[i] = f[i * N + c];
  });
}

template <typename R>
__device__ __forceinline__ void vstore(R* __restrict__ f, size_t N, size_t c, const R* v) {
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    f[i * N + c] = v[i];
  });
}

// f_eq(rho, u) (equilibrium.py:90-151): (g_rho, g_u) from g on f_eq.
template <typename R>
__global__ void k_equilibrium_vjp(const R* __restrict__ rho, const R* __restrict__ u, const R* __restrict__ gout,
                                  R* __restrict__ grho, R* __restrict__ gu, size_t N) {
  const size_t c = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (c >= N) return;
  Macro<R> m;
  m.rho = rho[c];
  m.ux = u[c];
  m.uy = u[N + c];
  m.uz = u[2 * N + c];
  m.usq = m.ux * m.ux + m.uy * m.uy + m.uz * m.uz;
  m.inv = R(1) / m.rho;
  R g[Q];
  vload(gout, N, c, g);
  const Pairs<R> pg = to_pairs(g);
  R Gr, Gx, Gy, Gz;
  feq_vjp<true>(pg.f0, pg.s, pg.d, m, Gr, Gx, Gy, Gz);
  grho[c] = Gr;
  gu[c] = Gx;
  gu[N + c] = Gy;
  gu[2 * N + c] = Gz;
}

// rho = sum_i f_i (equilibrium.py:22-48): g_f_i = g_rho.
template <typename R>
__global__ void k_density_vjp(const R* __restrict__ grho, R* __restrict__ gf, size_t N) {
  const size_t c = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (c >= N) return;
  const R g = grho[c];
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    gf[i * N + c] = g;
  });
}

// u = (sum_i f_i e_i) / rho with rho an input (equilibrium.py:51-87):
// g_f_i = (g_u . e_i) / rho, g_rho = -(g_u . u) / rho.
template <typename R>
__global__ void k_velocity_vjp(const R* __restrict__ f, const R* __restrict__ rho, const R* __restrict__ gu,
                               R* __restrict__ gf, R* __restrict__ grho, size_t N) {
  const size_t c = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (c >= N) return;
  const R inv = R(1) / rho[c];
  const R gx = gu[c] * inv, gy = gu[N + c] * inv, gz = gu[2 * N + c] * inv;
  R jx = R(0), jy = R(0), jz = R(0);
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    const R fi = f[i * N + c];
    add_scaled<EX[i]>(jx, fi);
    add_scaled<EY[i]>(jy, fi);
    add_scaled<EZ[i]>(jz, fi);
    R v = R(0);
    add_scaled<EX[i]>(v, gx);
    add_scaled<EY[i]>(v, gy);
    add_scaled<EZ[i]>(v, gz);
    gf[i * N + c] = v;
  });
  if (grho) grho[c] = -(gx * jx + gy * jy + gz * jz) * inv;
}

// BGK / MRT collision with tau scalar or per cell (collision.py:46-184):
// g_f = J^T g and, if gtau, the per-cell dL/dtau = -dL/domega / tau^2.
template <typename R, bool MRT>
__global__ void k_collide_vjp(const R* __restrict__ fin, R tau, const R* __restrict__ tau_field,
                              const R* __restrict__ gout, R* __restrict__ gf, R* __restrict__ gtau, size_t N) {
  const size_t c = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (c >= N) return;
  R f[Q], g[Q];
  vload(fin, N, c, f);
  vload(gout, N, c, g);
  const Pairs<R> ps = to_pairs(f);
  const Macro<R> m = macro(ps);
  const R omega = R(1) / (tau_field ? tau_field[c] : tau);
  const R gom = MRT ? mrt_vjp<true>(ps, m, g, omega) : bgk_vjp<true>(ps, m, g, omega);
  vstore(gf, N, c, g);
  if (gtau) gtau[c] = -gom * omega * omega;
}

// tau_eff = (tau + sqrt(tau^2 + 18 cs^2 |Pi| / (rho + 1e-10))) / 2
// (collision.py:187-270): g_f from g_teff; if gtau, the per-cell
// dL/dtau = g_teff (1 + tau / sqrt(...)) / 2.  Subgradient 0 at |Pi| = 0.
template <typename R>
__global__ void k_smag_vjp(const R* __restrict__ fin, R tau, R cs, const R* __restrict__ gteff,
                           R* __restrict__ gf, R* __restrict__ gtau, size_t N) {
  const size_t c = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (c >= N) return;
  R f[Q], g[Q];
  vload(fin, N, c, f);
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    g[i] = R(0);
  });
  const Pairs<R> ps = to_pairs(f);
  const Macro<R> m = macro(ps);
  const R gt = gteff[c];
  smag_vjp(ps, m, g, tau, cs, gt);
  vstore(gf, N, c, g);
  if (gtau) {
    const Stress<R> st = neq_stress(ps, m);
    const R pi = sqrt(st.xx * st.xx + st.yy * st.yy + st.zz * st.zz +
                      R(2) * (st.xy * st.xy + st.xz * st.xz + st.yz * st.yz));
    const R root = sqrt(tau * tau + R(18) * cs * cs * pi / (m.rho + R(1e-10)));
    gtau[c] = gt * R(0.5) * (R(1) + tau / root);
  }
}

// (1 - p) f + p f[opp] (smooth.py:118-155): g_f, and g_p = sum g_i (f_opp - f_i).
template <typename R>
__global__ void k_soft_bounce_vjp(const R* __restrict__ fin, const R* __restrict__ phi, const R* __restrict__ gout,
                                  R* __restrict__ gf, R* __restrict__ gphi, size_t N) {
  const size_t c = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (c >= N) return;
  R f[Q], g[Q];
  vload(fin, N, c, f);
  vload(gout, N, c, g);
  const R gp = blend_vjp(f, g, phi[c]);
  vstore(gf, N, c, g);
  if (gphi) gphi[c] = gp;
}

// f[opp] where solid >= 1 (state.py:146-149): the permutation is its own transpose.
template <typename R>
__global__ void k_hard_bounce_vjp(const int32_t* __restrict__ solid, const R* __restrict__ gout, R* __restrict__ gf,
                                  size_t N) {
  const size_t c = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (c >= N) return;
  R g[Q];
  vload(gout, N, c, g);
  if (solid[c] >= 1) reverse_all(g);
  vstore(gf, N, c, g);
}

// Stability guard (state.py:158-169): where(unstable, f_eq(1, 0), f) sends the
// cotangent to the reset branch at unstable cells, and that branch is a
// constant there (where(unstable, 1, rho), where(unstable, 0, u)), so those
// cells pass none; the predicate is recomputed from the guard's input.
template <typename R>
__global__ void k_guard_vjp(const R* __restrict__ fin, const R* __restrict__ gout, R* __restrict__ gf, size_t N) {
  const size_t c = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (c >= N) return;
  R f[Q];
  vload(fin, N, c, f);
  const bool bad = guard_bad(f);
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    gf[i * N + c] = bad ? R(0) : gout[i * N + c];
  });
}

// Bouzidi weights of the two branches (streaming.py:155-178), with their
// clips: A = 1 / (2 clip(q, 1/2, 1)) for q >= 1/2, B = 2 clip(q, 0, 1/2) below.
template <typename R>
__device__ __forceinline__ R bouzidi_a(R q) {
  return R(1) / (R(2) * fmin(fmax(q, R(0.5)), R(1)));
}
template <typename R>
__device__ __forceinline__ R bouzidi_b(R q) {
  return R(2) * fmin(fmax(q, R(0)), R(0.5));
}

// Transpose of pull streaming with Bouzidi links (streaming.py:99-183),
// gathered per input entry (y, j).  Its consumers are
//   (a) the normal pull of cell y + e_j (y/z wrapped when periodic; none past
//       an x face or, when not periodic, past a y/z face) where that cell has
//       no wall link along j;
//   (b) y itself where it keeps its own value (x faces, y/z faces when not
//       periodic) and has no wall link along j;
//   (c) y's own Bouzidi value in direction i = opp(j) (wall along j at y),
//       weight A or B of q_j(y);
//   (d) y's high-branch Bouzidi value in direction j (q_i(y) >= 1/2), weight 1 - A;
//   (e) the low-branch value in direction i of cell y + e_j (all axes
//       wrapped) when 0 <= q_j(y + e_j) < 1/2, weight 1 - B.
// With gq, also dL/dq_j(y) = g_i(y) dS_i(y)/dq_j on the taken branch:
//   -(f_j(y) - f_i(y)) / (2 q^2) for 1/2 <= q <= 1, 2 (f_j(y) - f_j(y + e_i)) for 0 <= q < 1/2.
template <typename R>
__global__ void k_stream_vjp(const R* __restrict__ f, const R* __restrict__ q, int periodic,
                             const R* __restrict__ gout, R* __restrict__ gf, R* __restrict__ gq, int nx, int ny,
                             int nz) {
  const size_t N = size_t(nx) * ny * nz;
  const size_t c = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (c >= N) return;
  const int x = int(c % nx);
  const int y = int((c / nx) % ny);
  const int z = int(c / (size_t(nx) * ny));
  auto at = [&](int cx, int cy, int cz) { return (size_t(cz) * ny + cy) * nx + cx; };
  gf[c] = gout[c];
  if (gq) gq[c] = R(0);
  sfor<Q>([&](auto J) {
    constexpr int j = J;
    if constexpr (j != 0) {
      constexpr int i = opp(j);
      R acc = R(0);
      // (a) normal pull of y + e_j
      {
        int sx = x + EX[j], sy = y + EY[j], sz = z + EZ[j];
        bool ok = sx >= 0 && sx < nx;
        if (periodic) {
          sy = (sy + ny) % ny;
          sz = (sz + nz) % nz;
        } else {
          ok = ok && sy >= 0 && sy < ny && sz >= 0 && sz < nz;
        }
        if (ok) {
          const size_t s = at(sx, sy, sz);
          if (q[i * N + s] < R(0)) acc += gout[j * N + s];
        }
      }
      // (b) y keeps its own value
      {
        bool own = (EX[j] == 1 && x == 0) || (EX[j] == -1 && x == nx - 1);
        if (!periodic)
          own = own || (EY[j] == 1 && y == 0) || (EY[j] == -1 && y == ny - 1) || (EZ[j] == 1 && z == 0) ||
                (EZ[j] == -1 && z == nz - 1);
        if (own && q[i * N + c] < R(0)) acc += gout[j * N + c];
      }
      // (c) own Bouzidi value in direction i
      const R qj = q[j * N + c];
      if (qj >= R(0)) acc += (qj >= R(0.5) ? bouzidi_a(qj) : bouzidi_b(qj)) * gout[i * N + c];
      // (d) high branch in direction j
      const R qi = q[i * N + c];
      if (qi >= R(0.5)) acc += (R(1) - bouzidi_a(qi)) * gout[j * N + c];
      // (e) low branch of cell y + e_j, all axes wrapped
      {
        const int fx = (x + EX[j] + nx) % nx, fy = (y + EY[j] + ny) % ny, fz = (z + EZ[j] + nz) % nz;
        const size_t s = at(fx, fy, fz);
        const R qs = q[j * N + s];
        if (qs >= R(0) && qs < R(0.5)) acc += (R(1) - bouzidi_b(qs)) * gout[i * N + s];
      }
      gf[j * N + c] = acc;
      if (gq) {
        R d = R(0);
        if (qj >= R(0.5)) {
          if (qj <= R(1)) d = -(f[j * N + c] - f[i * N + c]) / (R(2) * qj * qj);
        } else if (qj >= R(0)) {
          const int fx = (x + EX[i] + nx) % nx, fy = (y + EY[i] + ny) % ny, fz = (z + EZ[i] + nz) % nz;
          d = R(2) * (f[j * N + c] - f[j * N + at(fx, fy, fz)]);
        }
        gq[j * N + c] = gout[i * N + c] * d;
      }
    }
  });
}

// Zou-He velocity inlet at x = 0 (boundary.py:35-124), in place on g (a copy
// of the output cotangent).  Per row, gu_rows receives dL/du (3 doubles) and,
// with a density target, grt_rows dL/drho_target.
template <typename R>
__global__ void k_zou_he_in_vjp(const R* __restrict__ f, R ux, R uy, R uz, int has_rt, R rt, R* __restrict__ g,
                                double* __restrict__ gu_rows, double* __restrict__ grt_rows, int nx, int ny, int nz) {
  const size_t N = size_t(nx) * ny * nz;
  const size_t p = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (p >= size_t(ny) * nz) return;
  const size_t c = p * nx;
  R s[Q], h[Q];
  vload(f, N, c, s);
  vload(g, N, c, h);
  double gu[3] = {0.0, 0.0, 0.0};
  if (!has_rt) {
    zou_he_in_vjp(zou_he_in_rho(s, ux), h, ux, uy, uz, gu);
  } else {
    // rho is the target: the written values depend on f only through
    // f2, f8, f10, f12, f14
    const R g1 = h[1], g7 = h[7], g9 = h[9], g11 = h[11], g13 = h[13];
    const R grho = g1 * ux * R(1.0 / 3.0) +
                   (g7 * (ux + uy) + g9 * (ux - uy) + g11 * (ux + uz) + g13 * (ux - uz)) * R(1.0 / 6.0);
    gu[0] = double(g1 * rt * R(1.0 / 3.0) + (g7 + g9 + g11 + g13) * rt * R(1.0 / 6.0));
    gu[1] = double((g7 - g9) * rt * R(1.0 / 6.0));
    gu[2] = double((g11 - g13) * rt * R(1.0 / 6.0));
    if (grt_rows) grt_rows[p] = double(grho);
    h[1] = h[7] = h[9] = h[11] = h[13] = R(0);
    h[2] += g1;
    h[10] += g7;
    h[8] += g9;
    h[14] += g11;
    h[12] += g13;
  }
  vstore(g, N, c, h);
  if (gu_rows) {
    gu_rows[3 * p] = gu[0];
    gu_rows[3 * p + 1] = gu[1];
    gu_rows[3 * p + 2] = gu[2];
  }
}

// Zou-He pressure outlet at x = nx - 1 (boundary.py:127-200), in place on g;
// grho_rows receives dL/drho_out per row.
template <typename R>
__global__ void k_zou_he_out_vjp(R rho_out, R* __restrict__ g, double* __restrict__ grho_rows, int nx, int ny,
                                 int nz) {
  const size_t N = size_t(nx) * ny * nz;
  const size_t p = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (p >= size_t(ny) * nz) return;
  const size_t c = p * nx + (nx - 1);
  R h[Q];
  vload(g, N, c, h);
  if (grho_rows) grho_rows[p] = double(h[2] * R(1.0 / 3.0) + (h[10] + h[8] + h[14] + h[12]) * R(1.0 / 6.0));
  zou_he_out_vjp(h, rho_out);
  vstore(g, N, c, h);
}

// Momentum exchange (forces.py:37-196): F = sum over links (q_i >= 0) of
// e_i (P_i + S_opp(i)); the pressure part sum e_i (feq_i + feq_opp(i)) of
// f_eq(rho(P), u(P)).  gF = dL/dF (3), gFp = dL/dF_pressure (3, optional).
template <typename R>
__global__ void k_mex_vjp(const R* __restrict__ fp, const R* __restrict__ q, const double* __restrict__ gF,
                          const double* __restrict__ gFp, R* __restrict__ gP, R* __restrict__ gS, size_t N) {
  const size_t c = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (c >= N) return;
  const R ax = R(gF[0]), ay = R(gF[1]), az = R(gF[2]);
  R gp[Q], gs[Q], h[Q];
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    gp[i] = R(0);
    gs[i] = R(0);
    h[i] = R(0);
  });
  R bx = R(0), by = R(0), bz = R(0);
  if (gFp) {
    bx = R(gFp[0]);
    by = R(gFp[1]);
    bz = R(gFp[2]);
  }
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    if constexpr (i != 0) {
      constexpr int io = opp(i);
      if (q[i * N + c] >= R(0)) {
        R t = R(0);
        add_scaled<EX[i]>(t, ax);
        add_scaled<EY[i]>(t, ay);
        add_scaled<EZ[i]>(t, az);
        gp[i] += t;
        gs[io] += t;
        R tp = R(0);
        add_scaled<EX[i]>(tp, bx);
        add_scaled<EY[i]>(tp, by);
        add_scaled<EZ[i]>(tp, bz);
        h[i] += tp;
        h[io] += tp;
      }
    }
  });
  if (gFp) {
    R f[Q];
    vload(fp, N, c, f);
    const Macro<R> m = macro(f);
    const Pairs<R> ph = to_pairs(h);
    R Gr, Gx, Gy, Gz;
    feq_vjp<true>(ph.f0, ph.s, ph.d, m, Gr, Gx, Gy, Gz);
    add_macro_pullback(gp, m, Gr, Gx, Gy, Gz);
  }
  vstore(gP, N, c, gp);
  vstore(gS, N, c, gs);
}

// Cd_t = |F_t,x| / (0.5 rho U^2 A + 1e-10) (forces.py:237-276):
// dL/dF_t = (sign(F_t,x) g_t / den, 0, 0), sign(0) = 0.
__global__ void k_drag_vjp(const double* __restrict__ force, const double* __restrict__ gcd, int T, double U,
                           const double* __restrict__ area, double rho, double* __restrict__ gforce) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  const double den = 0.5 * rho * U * U * area[0] + 1e-10;
  const double fx = force[3 * t];
  const double s = fx > 0.0 ? 1.0 : (fx < 0.0 ? -1.0 : 0.0);
  gforce[3 * t] = s * gcd[t] / den;
  gforce[3 * t + 1] = 0.0;
  gforce[3 * t + 2] = 0.0;
}

// Sphere geometry (jax_lbm/geometry/smooth.py:158-248, primitives.py:56-174):
// dL/d(r, cx, cy, cz) from the cotangents of the porosity and of q, in
// float64, per workgroup (grid-stride, fixed order) into part[blockIdx][4].
//   porosity = sigmoid(alpha (r - |x - c|)):  d/dr = alpha p (1 - p),
//     d/dc = alpha p (1 - p) (x - c) / |x - c|  (0 at |x - c| = 0, where JAX gives NaN);
//   q = clip(t, 0.01, 0.99) on links with t = (-b - sqrt(max(disc, floor))) / (2a + eps),
//     b = 2 o.d, cc = |o|^2 - r^2, disc = b^2 - 4 a cc, o = x - c, d = e_i / s:
//     dt/dr = -(4 a r / sqrt(disc)) / (2a + eps),
//     dt/dc = (2 d + (2 b d - 4 a o) / sqrt(disc)) / (2a + eps),
//   only where the ray hit is valid, disc > floor and 0.01 < t < 0.99
//   (the where, max and clip then pass the cotangent), else 0.
// The sweep covers the box [box[0], box[1]) x [box[2], box[3]) x [box[4],
// box[5]) of the grid: all of it when the porosity term is summed, else the
// sphere's surroundings, the only cells with links.
template <typename C, typename O>
__global__ void k_sphere_vjp(int mode, C cx, C cy, C cz, C r, C alpha, C disc_floor, C denom_eps,
                             const int32_t* __restrict__ solid, const O* __restrict__ gphi, const O* __restrict__ gq,
                             double* __restrict__ part, int nx, int ny, int nz, int4 lo, int4 ext) {
  const size_t N = size_t(nx) * ny * nz;
  const size_t M = size_t(ext.x) * ext.y * ext.z;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (size_t k = blockIdx.x * size_t(blockDim.x) + threadIdx.x; k < M; k += size_t(gridDim.x) * blockDim.x) {
    const int x = lo.x + int(k % ext.x), y = lo.y + int((k / ext.x) % ext.y), z = lo.z + int(k / (size_t(ext.x) * ext.y));
    const size_t c = (size_t(z) * ny + y) * nx + x;
    C wx, wy, wz;
    world_xyz(x, y, z, nx, ny, nz, wx, wy, wz);
    const C ox = wx - cx, oy = wy - cy, oz = wz - cz;
    if (gphi && mode == 1) {
      const C dist = sqrt(ox * ox + oy * oy + oz * oz);
      const C phi = C(1) / (C(1) + exp(-(alpha * (r - dist))));
      const C k = C(gphi[c]) * alpha * phi * (C(1) - phi);
      acc[0] += double(k);
      if (dist > C(0)) {
        acc[1] += double(k * ox / dist);
        acc[2] += double(k * oy / dist);
        acc[3] += double(k * oz / dist);
      }
    }
    if (gq) {
      for (int i = 1; i < Q; ++i) {
        if (!link_from(solid, x, y, z, i, nx, ny, nz)) continue;
        const C g = C(gq[i * N + c]);
        if (g == C(0)) continue;
        const C dx = C(EX[i]) / C(nx / 8.0), dy = C(EY[i]) / C(ny / 4.0), dz = C(EZ[i]) / C(nz / 4.0);
        const C a = dx * dx + dy * dy + dz * dz;
        const C b = C(2) * (ox * dx + oy * dy + oz * dz);
        const C cc = ox * ox + oy * oy + oz * oz - r * r;
        const C disc = b * b - C(4) * a * cc;
        const C sq = sqrt(fmax(disc, disc_floor));
        const C den = C(2) * a + denom_eps;
        const C t = (-b - sq) / den;
        const bool valid = (disc >= C(0)) && (t > C(0)) && (t <= C(1));
        if (!(valid && disc > disc_floor && t > C(0.01) && t < C(0.99))) continue;
        acc[0] += double(g * (-(C(4) * a * r / sq) / den));
        acc[1] += double(g * (C(2) * dx + (C(2) * b * dx - C(4) * a * ox) / sq) / den);
        acc[2] += double(g * (C(2) * dy + (C(2) * b * dy - C(4) * a * oy) / sq) / den);
        acc[3] += double(g * (C(2) * dz + (C(2) * b * dz - C(4) * a * oz) / sq) / den);
      }
    }
  }
  __shared__ double sh[4][kVjpBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double v = wave_sum(acc[k]);
    if (lane == 0) sh[k][wid] = v;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double s = 0.0;
    for (int w = 0; w < kVjpBlock / 64; ++w) s += sh[threadIdx.x][w];
    part[blockIdx.x * 4 + threadIdx.x] = s;
  }
}

// out[k] = sum over the rows of part[rows][4], one thread per k, fixed order.
constexpr int kGeomVjpBlocks = 512;
constexpr int kSum4Threads = 256;

// out[k] = sum over the kGeomVjpBlocks rows of part[rows][4] in a fixed
// order: each thread adds its rows t, t + 256, ..., then a halving tree in LDS.
__global__ __launch_bounds__(kSum4Threads) void k_sum4(const double* __restrict__ part, double* __restrict__ out) {
  static_assert(kGeomVjpBlocks % kSum4Threads == 0, "whole rows per thread");
  __shared__ double sh[4][kSum4Threads];
  const int t = threadIdx.x;
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  for (int r = t; r < kGeomVjpBlocks; r += kSum4Threads) {
#pragma unroll
    for (int k = 0; k < 4; ++k) s[k] += part[r * 4 + k];
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) sh[k][t] = s[k];
  __syncthreads();
  for (int h = kSum4Threads / 2; h > 0; h /= 2) {
    if (t < h) {
#pragma unroll
      for (int k = 0; k < 4; ++k) sh[k][t] += sh[k][t + h];
    }
    __syncthreads();
  }
  if (t < 4) out[t] = sh[t][0];
}

}  // namespace lbm

using namespace lbm;

#define LBM_VJP_DISPATCH(dtype, ...) \
  do {                               \
    if ((dtype) == LBM_F32) {        \
      using R = float;               \
      __VA_ARGS__;                   \
    } else if ((dtype) == LBM_F64) { \
      using R = double;              \
      __VA_ARGS__;                   \
    } else {                         \
      return LBM_E_DTYPE;            \
    }                                \
  } while (0)

static int vjp_status() { return hipGetLastError() == hipSuccess ? LBM_OK : LBM_E_LAUNCH; }

extern "C" {

int lbm_equilibrium_vjp(int dtype, const void* rho, const void* u, const void* g_feq, void* g_rho, void* g_u, int nx,
                        int ny, int nz, lbm_stream_t stream) {
  if (!vjp_dims_ok(nx, ny, nz) || !rho || !u || !g_feq || !g_rho || !g_u) return LBM_E_ARG;
  const size_t N = size_t(nx) * ny * nz;
  LBM_VJP_DISPATCH(dtype, hipLaunchKernelGGL(k_equilibrium_vjp<R>, vjp_grid(N), kVjpBlock, 0, (hipStream_t)stream,
                                             (const R*)rho, (const R*)u, (const R*)g_feq, (R*)g_rho, (R*)g_u, N));
  return vjp_status();
}

int lbm_density_vjp(int dtype, const void* g_rho, void* g_f, int nx, int ny, int nz, lbm_stream_t stream) {
  if (!vjp_dims_ok(nx, ny, nz) || !g_rho || !g_f) return LBM_E_ARG;
  const size_t N = size_t(nx) * ny * nz;
  LBM_VJP_DISPATCH(dtype, hipLaunchKernelGGL(k_density_vjp<R>, vjp_grid(N), kVjpBlock, 0, (hipStream_t)stream,
                                             (const R*)g_rho, (R*)g_f, N));
  return vjp_status();
}

int lbm_velocity_vjp(int dtype, const void* f, const void* rho, const void* g_u, void* g_f, void* g_rho, int nx,
                     int ny, int nz, lbm_stream_t stream) {
  if (!vjp_dims_ok(nx, ny, nz) || !f || !rho || !g_u || !g_f) return LBM_E_ARG;
  const size_t N = size_t(nx) * ny * nz;
  LBM_VJP_DISPATCH(dtype, hipLaunchKernelGGL(k_velocity_vjp<R>, vjp_grid(N), kVjpBlock, 0, (hipStream_t)stream,
                                             (const R*)f, (const R*)rho, (const R*)g_u, (R*)g_f, (R*)g_rho, N));
  return vjp_status();
}

int lbm_collide_vjp(int dtype, int collision, const void* f, double tau, const void* tau_field, const void* g_out,
                    void* g_f, void* g_tau, int nx, int ny, int nz, lbm_stream_t stream) {
  if (!vjp_dims_ok(nx, ny, nz) || !f || !g_out || !g_f) return LBM_E_ARG;
  if (collision != LBM_COLL_BGK && collision != LBM_COLL_MRT) return LBM_E_ARG;
  const size_t N = size_t(nx) * ny * nz;
  if (collision == LBM_COLL_MRT) {
    LBM_VJP_DISPATCH(dtype, hipLaunchKernelGGL((k_collide_vjp<R, true>), vjp_grid(N), kVjpBlock, 0,
                                               (hipStream_t)stream, (const R*)f, R(tau), (const R*)tau_field,
                                               (const R*)g_out, (R*)g_f, (R*)g_tau, N));
  } else {
    LBM_VJP_DISPATCH(dtype, hipLaunchKernelGGL((k_collide_vjp<R, false>), vjp_grid(N), kVjpBlock, 0,
                                               (hipStream_t)stream, (const R*)f, R(tau), (const R*)tau_field,
                                               (const R*)g_out, (R*)g_f, (R*)g_tau, N));
  }
  return vjp_status();
}

int lbm_smagorinsky_tau_vjp(int dtype, const void* f, double tau, double cs, const void* g_tau_eff, void* g_f,
                            void* g_tau, int nx, int ny, int nz, lbm_stream_t stream) {
  if (!vjp_dims_ok(nx, ny, nz) || !f || !g_tau_eff || !g_f) return LBM_E_ARG;
  const size_t N = size_t(nx) * ny * nz;
  LBM_VJP_DISPATCH(dtype, hipLaunchKernelGGL(k_smag_vjp<R>, vjp_grid(N), kVjpBlock, 0, (hipStream_t)stream,
                                             (const R*)f, R(tau), R(cs), (const R*)g_tau_eff, (R*)g_f, (R*)g_tau,
                                             N));
  return vjp_status();
}

int lbm_soft_bounce_back_vjp(int dtype, const void* f, const void* porosity, const void* g_out, void* g_f,
                             void* g_porosity, int nx, int ny, int nz, lbm_stream_t stream) {
  if (!vjp_dims_ok(nx, ny, nz) || !f || !porosity || !g_out || !g_f) return LBM_E_ARG;
  const size_t N = size_t(nx) * ny * nz;
  LBM_VJP_DISPATCH(dtype, hipLaunchKernelGGL(k_soft_bounce_vjp<R>, vjp_grid(N), kVjpBlock, 0, (hipStream_t)stream,
                                             (const R*)f, (const R*)porosity, (const R*)g_out, (R*)g_f,
                                             (R*)g_porosity, N));
  return vjp_status();
}

int lbm_hard_bounce_vjp(int dtype, const int32_t* solid, const void* g_out, void* g_f, int nx, int ny, int nz,
                        lbm_stream_t stream) {
  if (!vjp_dims_ok(nx, ny, nz) || !solid || !g_out || !g_f) return LBM_E_ARG;
  const size_t N = size_t(nx) * ny * nz;
  LBM_VJP_DISPATCH(dtype, hipLaunchKernelGGL(k_hard_bounce_vjp<R>, vjp_grid(N), kVjpBlock, 0, (hipStream_t)stream,
                                             solid, (const R*)g_out, (R*)g_f, N));
  return vjp_status();
}

int lbm_stability_guard_vjp(int dtype, const void* f, const void* g_out, void* g_f, int nx, int ny, int nz,
                            lbm_stream_t stream) {
  if (!vjp_dims_ok(nx, ny, nz) || !f || !g_out || !g_f) return LBM_E_ARG;
  const size_t N = size_t(nx) * ny * nz;
  LBM_VJP_DISPATCH(dtype, hipLaunchKernelGGL(k_guard_vjp<R>, vjp_grid(N), kVjpBlock, 0, (hipStream_t)stream,
                                             (const R*)f, (const R*)g_out, (R*)g_f, N));
  return vjp_status();
}

int lbm_stream_vjp(int dtype, const void* f, const void* q, int periodic_yz, const void* g_out, void* g_f, void* g_q,
                   int nx, int ny, int nz, lbm_stream_t stream) {
  if (!vjp_dims_ok(nx, ny, nz) || !q || !g_out || !g_f || (g_q && !f)) return LBM_E_ARG;
  const size_t N = size_t(nx) * ny * nz;
  LBM_VJP_DISPATCH(dtype, hipLaunchKernelGGL(k_stream_vjp<R>, vjp_grid(N), kVjpBlock, 0, (hipStream_t)stream,
                                             (const R*)f, (const R*)q, periodic_yz, (const R*)g_out, (R*)g_f,
                                             (R*)g_q, nx, ny, nz));
  return vjp_status();
}

int lbm_zou_he_inlet_vjp(int dtype, const void* f, double ux, double uy, double uz, int has_rho_target,
                         double rho_target, void* g, double* g_u_rows, double* g_rho_target_rows, int nx, int ny,
                         int nz, lbm_stream_t stream) {
  if (!vjp_dims_ok(nx, ny, nz) || !f || !g) return LBM_E_ARG;
  const size_t P = size_t(ny) * nz;
  LBM_VJP_DISPATCH(dtype, hipLaunchKernelGGL(k_zou_he_in_vjp<R>, vjp_grid(P), kVjpBlock, 0, (hipStream_t)stream,
                                             (const R*)f, R(ux), R(uy), R(uz), has_rho_target, R(rho_target), (R*)g,
                                             g_u_rows, g_rho_target_rows, nx, ny, nz));
  return vjp_status();
}

int lbm_zou_he_outlet_vjp(int dtype, double rho_out, void* g, double* g_rho_out_rows, int nx, int ny, int nz,
                          lbm_stream_t stream) {
  if (!vjp_dims_ok(nx, ny, nz) || !g) return LBM_E_ARG;
  const size_t P = size_t(ny) * nz;
  LBM_VJP_DISPATCH(dtype, hipLaunchKernelGGL(k_zou_he_out_vjp<R>, vjp_grid(P), kVjpBlock, 0, (hipStream_t)stream,
                                             R(rho_out), (R*)g, g_rho_out_rows, nx, ny, nz));
  return vjp_status();
}

int lbm_momentum_exchange_vjp(int dtype, const void* f_post, const void* q, const double* g_force,
                              const double* g_pressure, void* g_post, void* g_streamed, int nx, int ny, int nz,
                              lbm_stream_t stream) {
  if (!vjp_dims_ok(nx, ny, nz) || !q || !g_force || !g_post || !g_streamed || (g_pressure && !f_post))
    return LBM_E_ARG;
  const size_t N = size_t(nx) * ny * nz;
  LBM_VJP_DISPATCH(dtype, hipLaunchKernelGGL(k_mex_vjp<R>, vjp_grid(N), kVjpBlock, 0, (hipStream_t)stream,
                                             (const R*)f_post, (const R*)q, g_force, g_pressure, (R*)g_post,
                                             (R*)g_streamed, N));
  return vjp_status();
}

size_t lbm_sphere_geometry_vjp_workspace_bytes(void) { return size_t(kGeomVjpBlocks) * 4 * sizeof(double); }

int lbm_sphere_geometry_vjp(int mode, int coord_dtype, int out_dtype, double cx, double cy, double cz, double radius,
                            double sharpness, double disc_floor, double denom_eps, const int32_t* solid,
                            const void* g_porosity, const void* g_q, double* out, double* workspace, int nx, int ny,
                            int nz, lbm_stream_t stream) {
  if (!vjp_dims_ok(nx, ny, nz) || !solid || !out || !workspace || (mode != 0 && mode != 1)) return LBM_E_ARG;
  hipStream_t s = (hipStream_t)stream;
  // Without the porosity term only link cells contribute: fluid cells next
  // to a body cell, and body cells lie within the radius of the centre.  The
  // sweep then covers the sphere's box in lattice indices, node i at world
  // (i + 1/2) / s - offset, widened by three cells for the link reach and
  // the rounding of the masks.
  int4 lo = make_int4(0, 0, 0, 0), ext = make_int4(nx, ny, nz, 0);
  if (!(g_porosity && mode == 1) && std::isfinite(cx) && std::isfinite(cy) && std::isfinite(cz) &&
      std::isfinite(radius)) {
    auto span = [&](double c, double off, double sc, int n, int& l, int& e) {
      const double ic = (c + off) * sc - 0.5, rr = std::fabs(radius) * sc;
      const double a = std::fmin(std::fmax(std::floor(ic - rr) - 3.0, 0.0), double(n));
      const double b = std::fmin(std::fmax(std::ceil(ic + rr) + 4.0, a), double(n));
      l = int(a);
      e = int(b) - int(a);
    };
    span(cx, 4.0, nx / 8.0, nx, lo.x, ext.x);
    span(cy, 2.0, ny / 4.0, ny, lo.y, ext.y);
    span(cz, 2.0, nz / 4.0, nz, lo.z, ext.z);
  }
#define LBM_GEOM_VJP(C, O)                                                                                          \
  hipLaunchKernelGGL((k_sphere_vjp<C, O>), kGeomVjpBlocks, kVjpBlock, 0, s, mode, C(cx), C(cy), C(cz), C(radius),   \
                     C(sharpness), C(mode == 0 ? 0.0 : disc_floor), C(denom_eps), solid, (const O*)g_porosity,       \
                     (const O*)g_q, workspace, nx, ny, nz, lo, ext)
  if (coord_dtype == LBM_F32 && out_dtype == LBM_F32) { LBM_GEOM_VJP(float, float); }
  else if (coord_dtype == LBM_F32 && out_dtype == LBM_F64) { LBM_GEOM_VJP(float, double); }
  else if (coord_dtype == LBM_F64 && out_dtype == LBM_F32) { LBM_GEOM_VJP(double, float); }
  else if (coord_dtype == LBM_F64 && out_dtype == LBM_F64) { LBM_GEOM_VJP(double, double); }
  else return LBM_E_DTYPE;
#undef LBM_GEOM_VJP
  hipLaunchKernelGGL(k_sum4, 1, kSum4Threads, 0, s, workspace, out);
  return vjp_status();
}

int lbm_drag_coefficient_vjp(const double* force, const double* g_cd, int T, double inlet_velocity,
                             const double* area, double rho, double* g_force, lbm_stream_t stream) {
  if (T <= 0 || !force || !g_cd || !area || !g_force) return LBM_E_ARG;
  hipLaunchKernelGGL(k_drag_vjp, (T + 255) / 256, 256, 0, (hipStream_t)stream, force, g_cd, T, inlet_velocity, area,
                     rho, g_force);
  return vjp_status();
}

}  // extern "C"
