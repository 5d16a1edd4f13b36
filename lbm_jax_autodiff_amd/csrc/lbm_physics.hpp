// Cell-local D3Q19 physics shared by the per-op and the fused kernels.
// Everything here works on 19 values held in registers; loops over the
// velocity set are fully unrolled so that lattice constants fold away.
#pragma once

#include "lbm_common.hpp"

namespace lbm {

template <typename R>
struct Macro {
  R rho, ux, uy, uz, usq;
};

// rho = sum f_i; u = sum f_i e_i / rho  (equilibrium.py:22-87)
template <typename R>
__device__ __forceinline__ Macro<R> macro(const R* f) {
  R rho = R(0), jx = R(0), jy = R(0), jz = R(0);
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    rho += f[i];
    if (EX[i]) jx += EX[i] > 0 ? f[i] : -f[i];
    if (EY[i]) jy += EY[i] > 0 ? f[i] : -f[i];
    if (EZ[i]) jz += EZ[i] > 0 ? f[i] : -f[i];
  }
  Macro<R> m;
  m.rho = rho;
  m.ux = jx / rho;
  m.uy = jy / rho;
  m.uz = jz / rho;
  m.usq = m.ux * m.ux + m.uy * m.uy + m.uz * m.uz;
  return m;
}

// f_eq_i = w_i rho (1 + e.u/cs2 + (e.u)^2/(2 cs2^2) - u.u/(2 cs2))
// (equilibrium.py:132-151) with cs2 = 1/3.
template <typename R>
__device__ __forceinline__ R feq(int i, R rho, R ux, R uy, R uz, R usq) {
  R eu = R(0);
  if (EX[i]) eu += EX[i] > 0 ? ux : -ux;
  if (EY[i]) eu += EY[i] > 0 ? uy : -uy;
  if (EZ[i]) eu += EZ[i] > 0 ? uz : -uz;
  return R(weight(i)) * rho * (R(1) + R(3) * eu + R(4.5) * eu * eu - R(1.5) * usq);
}

// BGK relaxation (collision.py:89-100)
template <typename R>
__device__ __forceinline__ void collide_bgk(R* f, R omega) {
  const Macro<R> m = macro(f);
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const R e = feq(i, m.rho, m.ux, m.uy, m.uz, m.usq);
    f[i] = f[i] - omega * (f[i] - e);
  }
}

// Equilibrium moments (lattice.py:342-391) for the non-conserved rows.
template <typename R>
__device__ __forceinline__ R meq(int k, R rho, R ux, R uy, R uz, R usq) {
  switch (k) {
    case 1: return R(-11) * rho + R(19) * rho * usq;
    case 2: return R(3) * rho - R(5.5) * rho * usq;
    case 4: return R(-2.0 / 3.0) * rho * ux;
    case 6: return R(-2.0 / 3.0) * rho * uy;
    case 8: return R(-2.0 / 3.0) * rho * uz;
    case 9: return rho * (R(2) * ux * ux - uy * uy - uz * uz);
    case 10: return R(-0.5) * rho * (R(2) * ux * ux - uy * uy - uz * uz);
    case 11: return rho * (uy * uy - uz * uz);
    case 12: return R(-0.5) * rho * (uy * uy - uz * uz);
    case 13: return rho * ux * uy;
    case 14: return rho * uy * uz;
    case 15: return rho * ux * uz;
    default: return R(0);
  }
}

// Moment k of f: sum_i M_ki f_i (integer coefficients fold at compile time).
template <typename R>
__device__ __forceinline__ R moment(int k, const R* f) {
  R acc = R(0);
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const int c = mrt_m(k, i);
    if (c == 1) acc += f[i];
    else if (c == -1) acc -= f[i];
    else if (c != 0) acc += R(c) * f[i];
  }
  return acc;
}

// MRT relaxation (collision.py:149-184).  Written as
//   f <- f - M^T diag(1/||M_k||^2) S (M f - m_eq)
// which equals the reference's M^{-1}(m - S(m - m_eq)) since M^{-1} M = I.
// Conserved rows have rate 0 and drop out.
template <typename R>
__device__ __forceinline__ void collide_mrt(R* f, R omega) {
  const Macro<R> mac = macro(f);
  R d[Q];
#pragma unroll
  for (int k = 0; k < Q; ++k) {
    if (mrt_is_conserved(k)) {
      d[k] = R(0);
      continue;
    }
    const R rate = mrt_is_stress(k) ? omega : R(mrt_rate_fixed(k));
    const R mk = moment(k, f);
    d[k] = rate * (mk - meq(k, mac.rho, mac.ux, mac.uy, mac.uz, mac.usq)) * R(1.0 / mrt_norm(k));
  }
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    R acc = R(0);
#pragma unroll
    for (int k = 0; k < Q; ++k) {
      if (mrt_is_conserved(k)) continue;
      const int c = mrt_m(k, i);
      if (c == 1) acc += d[k];
      else if (c == -1) acc -= d[k];
      else if (c != 0) acc += R(c) * d[k];
    }
    f[i] -= acc;
  }
}

// Non-equilibrium stress Pi_ab = sum_i e_ia e_ib (f_i - f_eq_i) and
// tau_eff = (tau + sqrt(tau^2 + 18 cs^2 |Pi| / (rho + 1e-10))) / 2
// (collision.py:241-270; |Pi| is the Frobenius norm of the full tensor).
template <typename R>
__device__ __forceinline__ R smag_tau(const R* f, R tau, R cs) {
  const Macro<R> m = macro(f);
  R sxx = 0, syy = 0, szz = 0, sxy = 0, sxz = 0, syz = 0;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const R n = f[i] - feq(i, m.rho, m.ux, m.uy, m.uz, m.usq);
    if (EX[i]) sxx += n;
    if (EY[i]) syy += n;
    if (EZ[i]) szz += n;
    if (EX[i] * EY[i]) sxy += EX[i] * EY[i] > 0 ? n : -n;
    if (EX[i] * EZ[i]) sxz += EX[i] * EZ[i] > 0 ? n : -n;
    if (EY[i] * EZ[i]) syz += EY[i] * EZ[i] > 0 ? n : -n;
  }
  const R pi = sqrt(sxx * sxx + syy * syy + szz * szz + R(2) * (sxy * sxy + sxz * sxz + syz * syz));
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

}  // namespace lbm
