// Cell-local D3Q19 physics shared by the per-op and the fused kernels.
//
// Everything works on 19 values held in registers.  Loops over the velocity
// set and the moment rows use sfor, so every lattice constant is a
// compile-time constant and zero terms are never emitted.
//
// The kernels are arithmetic-bound on gfx950 (packed FP32 is not used), so the
// collision works on opposite pairs (a, opp(a)): with s = f_a + f_opp and
// d = f_a - f_opp, the moments that are even in the lattice velocity (rho, e,
// eps, the stresses) depend on the nine sums only and the odd ones (j, q, the
// ghost m) on the nine differences only; the transposed transform splits the
// same way.  This halves the work of the moment transform against a dense
// 19 x 19 product and is exact algebra (float rounding differs by ulps).
#pragma once

#include "lbm_common.hpp"

namespace lbm {

// rho, u and u.u of a cell, with inv = 1 / rho kept for the adjoint's
// pull-back through u = j / rho.
template <typename R>
struct Macro {
  R rho, ux, uy, uz, usq, inv;
};

// 1 / rho: IEEE division, or with RCP (float only) the hardware reciprocal
// v_rcp_f32 (1 ulp), which the 16-bit storage sweeps use: their arithmetic is
// on the critical path and their populations carry 11-bit mantissas.
template <bool RCP, typename R>
__device__ __forceinline__ R recip(R x) {
  if constexpr (RCP && sizeof(R) == 4) return __builtin_amdgcn_rcpf(x);
  else return R(1) / x;
}

// The fused step and its adjoint take 1 / rho from the hardware reciprocal
// for the 16-bit storage formats (storage S), from IEEE division otherwise;
// the adjoint's recompute of a step makes the same choice as the step.
template <class S>
constexpr bool rcp_of = sizeof(typename S::T) == 2;

// acc += c * v for a compile-time integer c, without multiplications by +-1
// and with nothing emitted for c == 0.
template <int C, typename R>
__device__ __forceinline__ void add_scaled(R& acc, R v) {
  if constexpr (C == 1) acc += v;
  else if constexpr (C == -1) acc -= v;
  else if constexpr (C != 0) acc += R(C) * v;
}

// The nine pair representatives a (opp(a) > a).
constexpr int NP = 9;
constexpr int PA[NP] = {1, 3, 5, 7, 8, 11, 12, 15, 16};

__host__ __device__ constexpr bool row_even(int k) {
  return k == 0 || k == 1 || k == 2 || (k >= 9 && k <= 15);
}

// the parity property the pair form relies on, checked at compile time
__host__ __device__ constexpr bool mrt_parity_holds() {
  for (int k = 0; k < Q; ++k)
    for (int p = 0; p < NP; ++p) {
      const int a = PA[p], o = opp(PA[p]);
      if (row_even(k) && mrt_m(k, a) != mrt_m(k, o)) return false;
      if (!row_even(k) && (mrt_m(k, a) != -mrt_m(k, o) || mrt_m(k, 0) != 0)) return false;
    }
  for (int p = 0; p < NP; ++p)
    if (opp(PA[p]) <= PA[p]) return false;
  return true;
}
static_assert(mrt_parity_holds(), "D3Q19 moment rows must be even or odd under e -> -e");

template <typename R>
struct Pairs {
  R f0;
  R s[NP];  // f_a + f_opp(a)
  R d[NP];  // f_a - f_opp(a)
};

template <typename R>
__device__ __forceinline__ Pairs<R> to_pairs(const R* f) {
  Pairs<R> p;
  p.f0 = f[0];
  sfor<NP>([&](auto P) {
    constexpr int k = P;
    constexpr int a = PA[k], o = opp(PA[k]);
    p.s[k] = f[a] + f[o];
    p.d[k] = f[a] - f[o];
  });
  return p;
}

// rho = sum f; u = sum f e / rho (equilibrium.py:22-87), from the pairs:
// opposite directions contribute e_a (f_a - f_opp) to the momentum.
template <bool RCP = false, typename R>
__device__ __forceinline__ Macro<R> macro(const Pairs<R>& p) {
  R rho = p.f0, jx = R(0), jy = R(0), jz = R(0);
  sfor<NP>([&](auto P) {
    constexpr int k = P;
    constexpr int a = PA[k];
    rho += p.s[k];
    add_scaled<EX[a]>(jx, p.d[k]);
    add_scaled<EY[a]>(jy, p.d[k]);
    add_scaled<EZ[a]>(jz, p.d[k]);
  });
  Macro<R> m;
  m.rho = rho;
  const R inv = recip<RCP>(rho);
  m.inv = inv;
  m.ux = jx * inv;
  m.uy = jy * inv;
  m.uz = jz * inv;
  m.usq = m.ux * m.ux + m.uy * m.uy + m.uz * m.uz;
  return m;
}

template <bool RCP = false, typename R>
__device__ __forceinline__ Macro<R> macro(const R* f) {
  return macro<RCP>(to_pairs(f));
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

// The even and odd parts of f_eq for pair k: f_eq_a = sym + asym,
// f_eq_opp = sym - asym.
template <int K, typename R>
__device__ __forceinline__ void feq_pair(const Macro<R>& m, R& sym, R& asym) {
  constexpr int a = PA[K];
  constexpr float w = weight(a);
  const R eu = edot<a>(m.ux, m.uy, m.uz);
  const R wr = R(w) * m.rho;
  sym = wr * (R(1) + R(4.5) * eu * eu - R(1.5) * m.usq);
  asym = R(3) * wr * eu;
}

// BGK relaxation (collision.py:89-100) given the cell's macroscopic state.
template <typename R>
__device__ __forceinline__ void collide_bgk(R* f, R omega, const Macro<R>& m) {
  f[0] = f[0] - omega * (f[0] - R(weight(0)) * m.rho * (R(1) - R(1.5) * m.usq));
  sfor<NP>([&](auto P) {
    constexpr int k = P;
    constexpr int a = PA[k], o = opp(PA[k]);
    R sym, asym;
    feq_pair<k>(m, sym, asym);
    f[a] = f[a] - omega * (f[a] - (sym + asym));
    f[o] = f[o] - omega * (f[o] - (sym - asym));
  });
}

template <typename R>
__device__ __forceinline__ void collide_bgk(R* f, R omega) {
  collide_bgk(f, omega, macro(f));
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

// Moment K from the pairs: even rows use f0 and the sums, odd rows the differences.
template <int K, typename R>
__device__ __forceinline__ R moment(const Pairs<R>& p) {
  R acc = R(0);
  if constexpr (row_even(K)) {
    add_scaled<mrt_m(K, 0)>(acc, p.f0);
    sfor<NP>([&](auto P) {
      constexpr int k = P;
      add_scaled<mrt_m(K, PA[k])>(acc, p.s[k]);
    });
  } else {
    sfor<NP>([&](auto P) {
      constexpr int k = P;
      add_scaled<mrt_m(K, PA[k])>(acc, p.d[k]);
    });
  }
  return acc;
}

template <int K, typename R>
__device__ __forceinline__ R moment(const R* f) {
  return moment<K>(to_pairs(f));
}

// f_i -= (M^T h)_i over the non-conserved rows, split into the even part (the
// same for a and opp(a)) and the odd part (opposite signs).
template <typename R>
__device__ __forceinline__ void sub_transposed(R* f, const R* h) {
  R e0 = R(0);
  sfor<Q>([&](auto K) {
    constexpr int k = K;
    if constexpr (!mrt_is_conserved(k) && row_even(k)) add_scaled<mrt_m(k, 0)>(e0, h[k]);
  });
  f[0] -= e0;
  sfor<NP>([&](auto P) {
    constexpr int p = P;
    constexpr int a = PA[p], o = opp(PA[p]);
    R ev = R(0), od = R(0);
    sfor<Q>([&](auto K) {
      constexpr int k = K;
      if constexpr (!mrt_is_conserved(k)) {
        if constexpr (row_even(k)) add_scaled<mrt_m(k, a)>(ev, h[k]);
        else add_scaled<mrt_m(k, a)>(od, h[k]);
      }
    });
    f[a] -= ev + od;
    f[o] -= ev - od;
  });
}

// MRT relaxation (collision.py:149-184).  Written as
//   f <- f - M^T diag(1/||M_k||^2) S (M f - m_eq)
// which equals the reference's M^{-1}(m - S(m - m_eq)) since M^{-1} M = I.
// Conserved rows have rate 0 and drop out.
template <typename R>
__device__ __forceinline__ void collide_mrt(R* f, R omega, const Pairs<R>& pr, const Macro<R>& mac) {
  R h[Q];
  sfor<Q>([&](auto K) {
    constexpr int k = K;
    if constexpr (mrt_is_conserved(k)) {
      h[k] = R(0);
    } else {
      constexpr double inv_n = 1.0 / double(mrt_norm(k));
      const R rate = mrt_is_stress(k) ? omega : R(mrt_rate_fixed(k));
      h[k] = rate * R(inv_n) * (moment<k>(pr) - meq<k>(mac.rho, mac.ux, mac.uy, mac.uz, mac.usq));
    }
  });
  sub_transposed(f, h);
}

template <typename R>
__device__ __forceinline__ void collide_mrt(R* f, R omega) {
  const Pairs<R> pr = to_pairs(f);
  collide_mrt(f, omega, pr, macro(pr));
}

// Non-equilibrium stress Pi_ab = sum_i e_ia e_ib (f_i - f_eq_i): every
// e_a e_b is even, so only the pair sums and the even part of f_eq enter.
template <typename R>
struct Stress {
  R xx, yy, zz, xy, xz, yz;
};

template <typename R>
__device__ __forceinline__ Stress<R> neq_stress(const Pairs<R>& p, const Macro<R>& m) {
  Stress<R> s{R(0), R(0), R(0), R(0), R(0), R(0)};
  sfor<NP>([&](auto P) {
    constexpr int k = P;
    constexpr int a = PA[k];
    R sym, asym;
    feq_pair<k>(m, sym, asym);
    const R n = p.s[k] - R(2) * sym;
    add_scaled<EX[a] * EX[a]>(s.xx, n);
    add_scaled<EY[a] * EY[a]>(s.yy, n);
    add_scaled<EZ[a] * EZ[a]>(s.zz, n);
    add_scaled<EX[a] * EY[a]>(s.xy, n);
    add_scaled<EX[a] * EZ[a]>(s.xz, n);
    add_scaled<EY[a] * EZ[a]>(s.yz, n);
  });
  return s;
}

// tau_eff = (tau + sqrt(tau^2 + 18 cs^2 |Pi| / (rho + 1e-10))) / 2
// (collision.py:241-270; |Pi| is the Frobenius norm of the full tensor).
template <typename R>
__device__ __forceinline__ R smag_tau(const Pairs<R>& p, const Macro<R>& m, R tau, R cs) {
  const Stress<R> s = neq_stress(p, m);
  const R pi = sqrt(s.xx * s.xx + s.yy * s.yy + s.zz * s.zz + R(2) * (s.xy * s.xy + s.xz * s.xz + s.yz * s.yz));
  return R(0.5) * (tau + sqrt(tau * tau + R(18) * cs * cs * pi / (m.rho + R(1e-10))));
}

template <typename R>
__device__ __forceinline__ R smag_tau(const R* f, R tau, R cs) {
  const Pairs<R> p = to_pairs(f);
  return smag_tau(p, macro(p), tau, cs);
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

// Largest u.u whose correctly rounded square root is <= 0.4 in R: rounded
// sqrt is monotone, so sqrt(usq) > R(0.4) holds exactly when
// usq > guard_usq<R>() (also for inf and NaN), and the guard needs no sqrt.
template <typename R>
__host__ __device__ constexpr R guard_usq();
template <>
__host__ __device__ constexpr float guard_usq<float>() { return 0x1.47ae16p-3f; }
template <>
__host__ __device__ constexpr double guard_usq<double>() { return 0x1.47ae147ae147cp-3; }

// Stability guard predicate (state.py:158-160; examples/10:71-74):
// rho < 0.3, rho > 2, |u| > 0.4 or rho NaN.
template <typename R>
__device__ __forceinline__ bool guard_bad(const Macro<R>& m) {
  return (m.rho < R(0.3)) | (m.rho > R(2.0)) | (m.usq > guard_usq<R>()) | (m.rho != m.rho);
}

template <bool RCP = false, typename R>
__device__ __forceinline__ bool guard_bad(const R* f) {
  return guard_bad(macro<RCP>(f));
}

// f <- f[opp] (hard solid bounce, state.py:146-149).  For float the pairs
// are exchanged in place with v_swap_b32: a swap the compiler renames
// instead makes the cells that skip the (divergent) bounce copy 9 values at
// the join.
template <typename R>
__device__ __forceinline__ void reverse_all(R* f) {
  sfor<NP>([&](auto P) {
    constexpr int a = PA[P], o = opp(PA[P]);
    if constexpr (sizeof(R) == 4) {
      R x = f[a], y = f[o];
      asm("v_swap_b32 %0, %1" : "+v"(x), "+v"(y));
      f[a] = x;
      f[o] = y;
    } else {
      const R t = f[a];
      f[a] = f[o];
      f[o] = t;
    }
  });
}

// f <- (1 - p) f + p f[opp] (soft bounce, smooth.py:151-155)
template <typename R>
__device__ __forceinline__ void blend_opposite(R* f, R p) {
  f[0] = (R(1) - p) * f[0] + p * f[0];
  sfor<NP>([&](auto P) {
    constexpr int a = PA[P], o = opp(PA[P]);
    const R x = f[a], y = f[o];
    f[a] = (R(1) - p) * x + p * y;
    f[o] = (R(1) - p) * y + p * x;
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

// [pre guard reset] -> [Smagorinsky tau] -> BGK|MRT on one cell, sharing a
// single pass over the pairs and the macroscopic fields between the guard,
// the eddy viscosity and the collision (examples/10:71-78, state.py:133-145).
// omega = 1/tau is passed in for the constant-tau case.  Returns the omega
// used; sets `reset` when the guard fired.
template <bool MRT, bool RCP = false, typename R>
__device__ __forceinline__ R relax(R* f, R tau, R omega, bool pre_guard, R urx, R ury, R urz, bool smag, R cs,
                                   bool& reset) {
  Pairs<R> p = to_pairs(f);
  Macro<R> m = macro<RCP>(p);
  reset = pre_guard && guard_bad(m);
  if (reset) {
    set_feq_unit(f, urx, ury, urz);
    p = to_pairs(f);
    m = macro<RCP>(p);
  }
  if (smag) omega = R(1) / smag_tau(p, m, tau, cs);
  if constexpr (MRT) collide_mrt(f, omega, p, m);
  else collide_bgk(f, omega, m);
  return omega;
}

}  // namespace lbm
