// Cell-local reverse-mode (vector-Jacobian) products of the D3Q19 step.
//
// Each function takes the forward inputs of one operator and a cotangent g
// of its outputs, and overwrites g with the cotangent of its inputs.  The
// conventions follow JAX's transposition of the reference graph: a `where`
// sends the cotangent to the taken branch only, comparisons (guards, masks)
// carry none, and sqrt at 0 uses the subgradient 0 (JAX would give NaN there;
// this is the one documented deviation).
#pragma once

#include "lbm_physics.hpp"

namespace lbm {

// Pull back (G_rho|u, G_u) -- the cotangents of rho and u taken as
// independent variables -- to the populations through rho = sum f,
// u = j / rho:   g_i += G_rho - (G_u . u)/rho + (G_u . e_i)/rho.
template <typename R>
__device__ __forceinline__ void add_macro_pullback(R* g, const Macro<R>& m, R Grho, R Gux, R Guy, R Guz) {
  const R inv = m.inv;
  const R a = Grho - (Gux * m.ux + Guy * m.uy + Guz * m.uz) * inv;
  const R bx = Gux * inv, by = Guy * inv, bz = Guz * inv;
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    R v = a;
    add_scaled<EX[i]>(v, bx);
    add_scaled<EY[i]>(v, by);
    add_scaled<EZ[i]>(v, bz);
    g[i] += v;
  });
}

// Cotangents of (rho, u) from a cotangent h on f_eq(rho, u):
//   d feq_k / d rho|u = w_k (1 + 3 e.u + 4.5 (e.u)^2 - 1.5 u^2),
//   d feq_k / d u_a   = w_k rho (3 e_ka + 9 (e_k.u) e_ka - 3 u_a).
// h is given as its rest value and its pair sums hs / differences hd
// (hd is all zero when ODD is false); per pair the two directions give
//   w [hs (1 + 4.5 eu^2 - 1.5 u^2) + 3 eu hd]  and  w e_a (3 hd + 9 eu hs).
template <bool ODD, typename R>
__device__ __forceinline__ void feq_vjp(R h0, const R* hs, const R* hd, const Macro<R>& m, R& Grho, R& Gux,
                                        R& Guy, R& Guz) {
  constexpr float w0 = weight(0);
  Grho = h0 * R(w0) * (R(1) - R(1.5) * m.usq);
  R hw = h0 * R(w0);
  Gux = R(0);
  Guy = R(0);
  Guz = R(0);
  sfor<NP>([&](auto P) {
    constexpr int k = P;
    constexpr int a = PA[k];
    constexpr float w = weight(a);
    const R eu = edot<a>(m.ux, m.uy, m.uz);
    const R ws = R(w) * hs[k];
    R gr = ws * (R(1) + R(4.5) * eu * eu - R(1.5) * m.usq);
    R t = R(9) * eu * ws;
    if constexpr (ODD) {
      const R wd = R(w) * hd[k];
      gr += R(3) * eu * wd;
      t += R(3) * wd;
    }
    Grho += gr;
    hw += ws;
    add_scaled<EX[a]>(Gux, t);
    add_scaled<EY[a]>(Guy, t);
    add_scaled<EZ[a]>(Guz, t);
  });
  Gux = m.rho * (Gux - R(3) * m.ux * hw);
  Guy = m.rho * (Guy - R(3) * m.uy * hw);
  Guz = m.rho * (Guz - R(3) * m.uz * hw);
}

// BGK: c = s - omega (s - feq(s)).  g <- J^T g; returns dL/domega.
// J^T g = (1 - omega) g + omega (d feq / d s)^T g.
template <bool GOM, typename R>
__device__ __forceinline__ R bgk_vjp(const Pairs<R>& ps, const Macro<R>& m, R* g, R omega) {
  const Pairs<R> pg = to_pairs(g);
  // dL/domega = -sum_i g_i (s_i - feq_i), per pair
  //   = -(gs (ss - 2 sym) + gd (sd - 2 asym)) / 2
  // (only computed when a caller uses it: GOM)
  R gom = R(0);
  if constexpr (GOM) {
    gom = -pg.f0 * (ps.f0 - R(weight(0)) * m.rho * (R(1) - R(1.5) * m.usq));
    R acc = R(0);
    sfor<NP>([&](auto P) {
      constexpr int k = P;
      R sym, asym;
      feq_pair<k>(m, sym, asym);
      acc += pg.s[k] * (ps.s[k] - R(2) * sym) + pg.d[k] * (ps.d[k] - R(2) * asym);
    });
    gom -= R(0.5) * acc;
  }
  R Gr, Gx, Gy, Gz;
  feq_vjp<true>(pg.f0, pg.s, pg.d, m, Gr, Gx, Gy, Gz);
  const R k1 = R(1) - omega;
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    g[i] *= k1;
  });
  add_macro_pullback(g, m, omega * Gr, omega * Gx, omega * Gy, omega * Gz);
  return gom;
}

// Contraction of z with d meq / d(rho|u) and d meq / d u over the
// non-conserved rows (lattice.py:369-391).
template <typename R>
__device__ __forceinline__ void meq_vjp(const R* z, const Macro<R>& m, R& Grho, R& Gux, R& Guy, R& Guz) {
  const R r = m.rho, ux = m.ux, uy = m.uy, uz = m.uz, u2 = m.usq;
  const R dxx = R(2) * ux * ux - uy * uy - uz * uz;
  const R dww = uy * uy - uz * uz;
  const R c23 = R(-2.0 / 3.0);
  // rows 9/10 and 11/12 share their u-dependence up to the factor -1/2
  const R a = z[9] - R(0.5) * z[10];
  const R b = z[11] - R(0.5) * z[12];
  const R e = R(38) * z[1] - R(11) * z[2];
  Grho = z[1] * (R(-11) + R(19) * u2) + z[2] * (R(3) - R(5.5) * u2) + c23 * (z[4] * ux + z[6] * uy + z[8] * uz) +
         a * dxx + b * dww + z[13] * ux * uy + z[14] * uy * uz + z[15] * ux * uz;
  Gux = r * (e * ux + c23 * z[4] + R(4) * a * ux + z[13] * uy + z[15] * uz);
  Guy = r * (e * uy + c23 * z[6] - R(2) * a * uy + R(2) * b * uy + z[13] * ux + z[14] * uz);
  Guz = r * (e * uz + c23 * z[8] - R(2) * a * uz - R(2) * b * uz + z[14] * uy + z[15] * ux);
}

// MRT: c = s - M^T D (M s - meq(s)), D = diag(rate_k / ||M_k||^2).
// J^T g = g - M^T h + (d meq/d s)^T h with h = D M g.  Returns dL/domega.
template <bool GOM, typename R>
__device__ __forceinline__ R mrt_vjp(const Pairs<R>& ps, const Macro<R>& mac, R* g, R omega) {
  const Pairs<R> pg = to_pairs(g);
  R h[Q];
  R gom = R(0);
  sfor<Q>([&](auto K) {
    constexpr int k = K;
    if constexpr (mrt_is_conserved(k)) {
      h[k] = R(0);
    } else {
      constexpr double inv_n = 1.0 / double(mrt_norm(k));
      const R Mg = moment<k>(pg);
      if constexpr (mrt_is_stress(k)) {
        if constexpr (GOM) {
          const R dm = moment<k>(ps) - meq<k>(mac.rho, mac.ux, mac.uy, mac.uz, mac.usq);
          gom -= Mg * dm * R(inv_n);
        }
        h[k] = omega * Mg * R(inv_n);
      } else {
        h[k] = R(mrt_rate_fixed(k)) * Mg * R(inv_n);
      }
    }
  });
  sub_transposed(g, h);
  R Gr, Gx, Gy, Gz;
  meq_vjp(h, mac, Gr, Gx, Gy, Gz);
  add_macro_pullback(g, mac, Gr, Gx, Gy, Gz);
  return gom;
}

// Smagorinsky: tau_eff(s) = (tau + sqrt(tau^2 + 18 cs^2 |Pi| / (rho + 1e-10))) / 2,
// Pi_ab = sum_i e_ia e_ib (s_i - feq_i(s)).  Adds g_teff * d tau_eff / d s to g.
// The cotangent of the non-equilibrium part is even in e (the same for a
// and opp(a)) and zero at rest.
template <typename R>
__device__ __forceinline__ void smag_vjp(const Pairs<R>& ps, const Macro<R>& m, R* g, R tau, R cs, R g_teff) {
  const Stress<R> st = neq_stress(ps, m);
  const R pi = sqrt(st.xx * st.xx + st.yy * st.yy + st.zz * st.zz +
                    R(2) * (st.xy * st.xy + st.xz * st.xz + st.yz * st.yz));
  const R den = m.rho + R(1e-10);
  const R k18 = R(18) * cs * cs;
  const R root = sqrt(tau * tau + k18 * pi / den);
  // d tau_eff / d pi and the explicit d tau_eff / d rho
  const R dpi = R(0.25) * k18 / (den * root);
  const R drho = R(-0.25) * k18 * pi / (den * den * root);
  const R gpi = g_teff * dpi;
  // subgradient 0 at |Pi| = 0
  const R inv = pi > R(0) ? gpi / pi : R(0);
  const R gxx = inv * st.xx, gyy = inv * st.yy, gzz = inv * st.zz;
  const R gxy = R(2) * inv * st.xy, gxz = R(2) * inv * st.xz, gyz = R(2) * inv * st.yz;
  R hs[NP];
  sfor<NP>([&](auto P) {
    constexpr int k = P;
    constexpr int a = PA[k], o = opp(PA[k]);
    R v = R(0);
    add_scaled<EX[a] * EX[a]>(v, gxx);
    add_scaled<EY[a] * EY[a]>(v, gyy);
    add_scaled<EZ[a] * EZ[a]>(v, gzz);
    add_scaled<EX[a] * EY[a]>(v, gxy);
    add_scaled<EX[a] * EZ[a]>(v, gxz);
    add_scaled<EY[a] * EZ[a]>(v, gyz);
    g[a] += v;
    g[o] += v;
    hs[k] = R(2) * v;
  });
  R Gr, Gx, Gy, Gz;
  feq_vjp<false>(R(0), hs, hs, m, Gr, Gx, Gy, Gz);
  add_macro_pullback(g, m, g_teff * drho - Gr, -Gx, -Gy, -Gz);
}

// Density the Zou-He inlet computes from the pulled values it keeps
// (boundary.py:80-124): rho = (K0 + 2 K-) / (1 - u_x).
template <typename R>
__device__ __forceinline__ R zou_he_in_rho(const R* s, R ux) {
  const R k0 = s[0] + s[3] + s[4] + s[5] + s[6] + s[15] + s[16] + s[17] + s[18];
  const R km = s[2] + s[8] + s[10] + s[12] + s[14];
  return (k0 + R(2) * km) / (R(1) - ux);
}

// Zou-He inlet transpose.  rho = zou_he_in_rho of the pulled values; g the
// cotangent of the post-inlet values.  Accumulates dL/du_in into gu.
template <typename R>
__device__ __forceinline__ void zou_he_in_vjp(R rho, R* g, R ux, R uy, R uz, double* gu) {
  const R g1 = g[1], g7 = g[7], g9 = g[9], g11 = g[11], g13 = g[13];
  const R grho = g1 * ux * R(1.0 / 3.0) + (g7 * (ux + uy) + g9 * (ux - uy) + g11 * (ux + uz) + g13 * (ux - uz)) *
                 R(1.0 / 6.0);
  if (gu) {
    gu[0] += double(g1 * rho * R(1.0 / 3.0) + (g7 + g9 + g11 + g13) * rho * R(1.0 / 6.0) + grho * rho / (R(1) - ux));
    gu[1] += double((g7 - g9) * rho * R(1.0 / 6.0));
    gu[2] += double((g11 - g13) * rho * R(1.0 / 6.0));
  }
  const R a = grho / (R(1) - ux);
  g[1] = g[7] = g[9] = g[11] = g[13] = R(0);
  g[0] += a; g[3] += a; g[4] += a; g[5] += a; g[6] += a;
  g[15] += a; g[16] += a; g[17] += a; g[18] += a;
  g[2] += R(2) * a + g1;
  g[10] += R(2) * a + g7;
  g[8] += R(2) * a + g9;
  g[14] += R(2) * a + g11;
  g[12] += R(2) * a + g13;
}

// Zou-He outlet transpose (u_x = -1 + (K0 + 2 K+) / rho_out).
template <typename R>
__device__ __forceinline__ void zou_he_out_vjp(R* g, R rho_out) {
  const R g2 = g[2], g10 = g[10], g8 = g[8], g14 = g[14], g12 = g[12];
  const R gux = -rho_out * (g2 * R(1.0 / 3.0) + (g10 + g8 + g14 + g12) * R(1.0 / 6.0));
  const R a = gux / rho_out;
  g[2] = g[10] = g[8] = g[14] = g[12] = R(0);
  g[0] += a; g[3] += a; g[4] += a; g[5] += a; g[6] += a;
  g[15] += a; g[16] += a; g[17] += a; g[18] += a;
  g[1] += R(2) * a + g2;
  g[7] += R(2) * a + g10;
  g[9] += R(2) * a + g8;
  g[11] += R(2) * a + g14;
  g[13] += R(2) * a + g12;
}

// Soft-bounce transpose: returns dL/dp = sum_i g_i (c_opp(i) - c_i) and
// overwrites g with (1 - p) g + p g[opp].
template <typename R>
__device__ __forceinline__ R blend_vjp(const R* c, R* g, R p) {
  R gp = R(0);
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    constexpr int o = opp(i);
    gp += g[i] * (c[o] - c[i]);
  });
  blend_opposite(g, p);
  return gp;
}

}  // namespace lbm
