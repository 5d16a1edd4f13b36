// Cell-local D3Q19 physics shared by the per-op and the fused kernels.
// Everything works on 19 values held in registers.  Loops over the velocity
// set and the moment rows use sfor, so every lattice constant (velocity
// components, weights, opposites, integer moment-matrix entries) is a
// compile-time constant and zero terms are never emitted.
#pragma once

#include "lbm_common.hpp"

namespace lbm {

template <typename R>
struct Macro {
  R rho, ux, uy, uz, usq;
};

// acc += c * v for a compile-time integer c, without multiplications by +-1
// and with nothing emitted for c == 0.
template <int C, typename R>
__device__ __forceinline__ void add_scaled(R& acc, R v) {
  if constexpr (C == 1) acc += v;
  else if constexpr (C == -1) acc -= v;
  else if constexpr (C != 0) acc += R(C) * v;
}

// rho = sum f_i; u = sum f_i e_i / rho  (equilibrium.py:22-87)
template <typename R>
__device__ __forceinline__ Macro<R> macro(const R* f) {
  R rho = R(0), jx = R(0), jy = R(0), jz = R(0);
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    rho += f[i];
    add_scaled<EX[i]>(jx, f[i]);
    add_scaled<EY[i]>(jy, f[i]);
    add_scaled<EZ[i]>(jz, f[i]);
  });
  Macro<R> m;
  m.rho = rho;
  m.ux = jx / rho;
  m.uy = jy / rho;
  m.uz = jz / rho;
  m.usq = m.ux * m.ux + m.uy * m.uy + m.uz * m.uz;
  return m;
}

// e_i . u for a compile-time direction
template <int I, typename R>
__device__ __forceinline__ R edot(R ux, R uy, R uz) {
  R eu = R(0);
  add_scaled<EX[I]>(eu, ux);
  add_scaled<EY[I]>(eu, uy);
  add_scaled<EZ[I]>(eu, uz);
  return eu;
}

// f_eq_i = w_i rho (1 + e.u/cs2 + (e.u)^2/(2 cs2^2) - u.u/(2 cs2))
// (equilibrium.py:132-151) with cs2 = 1/3.
template <int I, typename R>
__device__ __forceinline__ R feq(R rho, R ux, R uy, R uz, R usq) {
  constexpr float w = weight(I);
  const R eu = edot<I>(ux, uy, uz);
  return R(w) * rho * (R(1) + R(3) * eu + R(4.5) * eu * eu - R(1.5) * usq);
}

// BGK relaxation (collision.py:89-100)
template <typename R>
__device__ __forceinline__ void collide_bgk(R* f, R omega) {
  const Macro<R> m = macro(f);
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    const R e = feq<i>(m.rho, m.ux, m.uy, m.uz, m.usq);
    f[i] = f[i] - omega * (f[i] - e);
  });
}

// Equilibrium moments (lattice.py:342-391) for the non-conserved rows.
template <int K, typename R>
__device__ __forceinline__ R meq(R rho, R ux, R uy, R uz, R usq) {
  if constexpr (K == 1) return R(-11) * rho + R(19) * rho * usq;
  else if constexpr (K == 2) return R(3) * rho - R(5.5) * rho * usq;
  else if constexpr (K == 4) return R(-2.0 / 3.0) * rho * ux;
  else if constexpr (K == 6) return R(-2.0 / 3.0) * rho * uy;
  else if constexpr (K == 8) return R(-2.0 / 3.0) * rho * uz;
  else if constexpr (K == 9) return rho * (R(2) * ux * ux - uy * uy - uz * uz);
  else if constexpr (K == 10) return R(-0.5) * rho * (R(2) * ux * ux - uy * uy - uz * uz);
  else if constexpr (K == 11) return rho * (uy * uy - uz * uz);
  else if constexpr (K == 12) return R(-0.5) * rho * (uy * uy - uz * uz);
  else if constexpr (K == 13) return rho * ux * uy;
  else if constexpr (K == 14) return rho * uy * uz;
  else if constexpr (K == 15) return rho * ux * uz;
  else return R(0);
}

// Moment K of f: sum_i M_Ki f_i (integer coefficients folded at compile time).
template <int K, typename R>
__device__ __forceinline__ R moment(const R* f) {
  R acc = R(0);
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    add_scaled<mrt_m(K, i)>(acc, f[i]);
  });
  return acc;
}

// acc_i -= sum_k M_ki d_k over the non-conserved rows (the transposed transform)
template <typename R>
__device__ __forceinline__ void sub_transposed(R* f, const R* d) {
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    R acc = R(0);
    sfor<Q>([&](auto K) {
      constexpr int k = K;
      if constexpr (!mrt_is_conserved(k)) add_scaled<mrt_m(k, i)>(acc, d[k]);
    });
    f[i] -= acc;
  });
}

// MRT relaxation (collision.py:149-184).  Written as
//   f <- f - M^T diag(1/||M_k||^2) S (M f - m_eq)
// which equals the reference's M^{-1}(m - S(m - m_eq)) since M^{-1} M = I.
// Conserved rows have rate 0 and drop out.
template <typename R>
__device__ __forceinline__ void collide_mrt(R* f, R omega) {
  const Macro<R> mac = macro(f);
  R d[Q];
  sfor<Q>([&](auto K) {
    constexpr int k = K;
    if constexpr (mrt_is_conserved(k)) {
      d[k] = R(0);
    } else {
      constexpr double inv_n = 1.0 / double(mrt_norm(k));
      const R rate = mrt_is_stress(k) ? omega : R(mrt_rate_fixed(k));
      d[k] = rate * R(inv_n) * (moment<k>(f) - meq<k>(mac.rho, mac.ux, mac.uy, mac.uz, mac.usq));
    }
  });
  sub_transposed(f, d);
}

// Non-equilibrium stress Pi_ab = sum_i e_ia e_ib (f_i - f_eq_i).
template <typename R>
struct Stress {
  R xx, yy, zz, xy, xz, yz;
};

template <typename R>
__device__ __forceinline__ Stress<R> neq_stress(const R* f, const Macro<R>& m) {
  Stress<R> s{R(0), R(0), R(0), R(0), R(0), R(0)};
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    if constexpr (i != 0) {
      const R n = f[i] - feq<i>(m.rho, m.ux, m.uy, m.uz, m.usq);
      add_scaled<EX[i] * EX[i]>(s.xx, n);
      add_scaled<EY[i] * EY[i]>(s.yy, n);
      add_scaled<EZ[i] * EZ[i]>(s.zz, n);
      add_scaled<EX[i] * EY[i]>(s.xy, n);
      add_scaled<EX[i] * EZ[i]>(s.xz, n);
      add_scaled<EY[i] * EZ[i]>(s.yz, n);
    }
  });
  return s;
}

// tau_eff = (tau + sqrt(tau^2 + 18 cs^2 |Pi| / (rho + 1e-10))) / 2
// (collision.py:241-270; |Pi| is the Frobenius norm of the full tensor).
template <typename R>
__device__ __forceinline__ R smag_tau(const R* f, R tau, R cs) {
  const Macro<R> m = macro(f);
  const Stress<R> s = neq_stress(f, m);
  const R pi = sqrt(s.xx * s.xx + s.yy * s.yy + s.zz * s.zz + R(2) * (s.xy * s.xy + s.xz * s.xz + s.yz * s.yz));
  return R(0.5) * (tau + sqrt(tau * tau + R(18) * cs * cs * pi / (m.rho + R(1e-10))));
}

// Zou-He velocity inlet at x = 0 (boundary.py:80-124): unknowns
// 1, 7, 9, 11, 13 from their opposites plus rho u corrections.
template <typename R>
__device__ __forceinline__ void zou_he_in(R* f, R ux, R uy, R uz, bool has_rt, R rt) {
  const R k0 = f[0] + f[3] + f[4] + f[5] + f[6] + f[15] + f[16] + f[17] + f[18];
  const R km = f[2] + f[8] + f[10] + f[12] + f[14];
  const R rho = has_rt ? rt : (k0 + R(2) * km) / (R(1) - ux);
  f[1] = f[2] + R(1.0 / 3.0) * rho * ux;
  f[7] = f[10] + R(1.0 / 6.0) * rho * (ux + uy);
  f[9] = f[8] + R(1.0 / 6.0) * rho * (ux - uy);
  f[11] = f[14] + R(1.0 / 6.0) * rho * (ux + uz);
  f[13] = f[12] + R(1.0 / 6.0) * rho * (ux - uz);
}

// Zou-He pressure outlet at x = nx-1 (boundary.py:168-200).
template <typename R>
__device__ __forceinline__ void zou_he_out(R* f, R rho_out) {
  const R k0 = f[0] + f[3] + f[4] + f[5] + f[6] + f[15] + f[16] + f[17] + f[18];
  const R kp = f[1] + f[7] + f[9] + f[11] + f[13];
  const R ux = R(-1) + (k0 + R(2) * kp) / rho_out;
  f[2] = f[1] - R(1.0 / 3.0) * rho_out * ux;
  f[10] = f[7] - R(1.0 / 6.0) * rho_out * ux;
  f[8] = f[9] - R(1.0 / 6.0) * rho_out * ux;
  f[14] = f[11] - R(1.0 / 6.0) * rho_out * ux;
  f[12] = f[13] - R(1.0 / 6.0) * rho_out * ux;
}

// Stability guard predicate (state.py:158-160; examples/10:71-74).
template <typename R>
__device__ __forceinline__ bool guard_bad(const R* f) {
  const Macro<R> m = macro(f);
  const R speed = sqrt(m.usq);
  return (m.rho < R(0.3)) | (m.rho > R(2.0)) | (speed > R(0.4)) | (m.rho != m.rho);
}

// f <- f[opp] (hard solid bounce, state.py:146-149)
template <typename R>
__device__ __forceinline__ void reverse_all(R* f) {
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    constexpr int o = opp(i);
    if constexpr (o > i) {
      const R t = f[i];
      f[i] = f[o];
      f[o] = t;
    }
  });
}

// f <- (1 - p) f + p f[opp] (soft bounce, smooth.py:151-155)
template <typename R>
__device__ __forceinline__ void blend_opposite(R* f, R p) {
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    constexpr int o = opp(i);
    if constexpr (o > i) {
      const R a = f[i], b = f[o];
      f[i] = (R(1) - p) * a + p * b;
      f[o] = (R(1) - p) * b + p * a;
    } else if constexpr (o == i) {
      // rest population: (1 - p) f0 + p f0
      f[i] = (R(1) - p) * f[i] + p * f[i];
    }
  });
}

// f_i = f_eq_i(1, u) (the guard reset targets)
template <typename R>
__device__ __forceinline__ void set_feq_unit(R* f, R ux, R uy, R uz) {
  const R uu = ux * ux + uy * uy + uz * uz;
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    f[i] = feq<i>(R(1), ux, uy, uz, uu);
  });
}

template <typename R>
__device__ __forceinline__ void set_weights(R* f) {
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    f[i] = R(weight(i));
  });
}

}  // namespace lbm
