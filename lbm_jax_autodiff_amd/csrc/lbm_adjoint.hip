// Hand-written adjoint of the fused D3Q19 step (reverse mode over one step).
//
// Given P_t and lambda_{t+1} = dL/dP_{t+1}, one launch recomputes each cell's
// forward step (pull, Bouzidi, Zou-He, guards, collision, bounce) and pulls
// the cotangent back to mu = dL/d(pulled values).  Pull streaming is a
// bijection once the x faces and the y/z axes wrap: entry (y, i) of P_t is
// consumed normally by exactly one cell x = y + e_i.  That "owner" writes
// lambda_t(y)_i with a plain store: mu_i for a normal pull, 0 where the
// reference keeps the cell's own value at the x faces (Zou-He overwrites
// those populations, so they carry no cotangent) or where a Bouzidi link
// replaces the pull.  Every entry of lambda_t is written exactly once, with
// no atomics.  The extra consumers -- Bouzidi links (P(x)_iopp, P(x)_i,
// P(x+e_i)_iopp) and the momentum-exchange force (P(x)_j) -- belong to the
// link cells: the P(x+e_i)_iopp term joins the link cell's own owner write of
// that entry, and the terms on the link cell's own populations are stored per
// link and added by the next launch as it loads them (lbm_adjoint_link_finish
// after a chain's last launch), or added by lbm_adjoint_scatter from a fixed
// per-entry plan.  Either way the result is deterministic.
#include "lbm_adjoint_physics.hpp"

namespace lbm {

constexpr int kAdjBlock = 256;
constexpr int kAdjBlockLarge = 512;

template <typename R>
struct AdjArgs {
  // forward configuration (same meaning as in the forward kernel)
  int nx, ny, nz;
  uint32_t N;   // cells
  uint32_t P;   // elements between consecutive populations (state and cotangent buffers)
  uint32_t RS;  // elements between consecutive (y, z) rows (state and cotangent buffers)
  FastDiv dnx, dny;
  int smag, bounce, guard, n_links, nlb, ndb;
  uint32_t tgroup;  // dense tiles per XCD group (xcd_group_tile)
  R tau, omega, cs;
  R uin[3];
  R rho_out;
  R uref[3];
  R off[Q];
  const uint8_t* flags;
  const R* phi;
  const int32_t* link_cell;
  const double* link_q;
  // adjoint outputs / scratch
  double* dq_acc;
  R* mu_link;
  double* phi_acc;
  double* u_acc;
  // folded link additions (lbm_adjoint_desc.fold_links): with `fold` the
  // link cells write their Bouzidi second-neighbour terms with their owner
  // writes and store, in mu_link, the additions to their own 19 entries
  // instead of their cotangents; add_prev (the previous launch's such
  // additions, null in a chain's first launch) is added to what they load
  int fold;
  const R* add_prev;
  // cotangent storage scale (half storage) and its inverse
  R lam_scale, lam_inv;
  // extents in bytes of a state and of a cotangent buffer (below 2^32 when
  // the row-uniform sweep is used)
  uint32_t state_bytes, lam_bytes;
};

// Storage of the cotangent (lambda) buffers: the arithmetic type itself, or
// IEEE binary16 holding lambda * lam_scale (a power of two chosen from the
// seeds' magnitude), rounded to nearest even.  The adjoint is linear in
// lambda, so the scale only moves the values into the half range.
template <typename R>
struct LamPlain {
  using T = R;
  __device__ static R ld(const T* p, size_t k, R) { return p[k]; }
  __device__ static void st(T* p, size_t k, R v, R) { p[k] = v; }
  // the value a store and a load would give back
  __device__ static R rt(R v, R, R) { return v; }
  __device__ static R bld(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so, R) {
    if constexpr (sizeof(R) == 4) return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, int(vo), int(so), 0));
    else return __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b64(r, int(vo), int(so), 0));
  }
  __device__ static void bst(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so, R v, R) {
    if constexpr (sizeof(R) == 4) __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, int(vo), int(so), 0);
    else __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, int(vo), int(so), 0);
  }
};
struct LamHalf {
  using T = unsigned short;
  __device__ static float ld(const T* p, size_t k, float inv_scale) { return f16_bits_to_float(p[k]) * inv_scale; }
  __device__ static void st(T* p, size_t k, float v, float scale) { p[k] = float_to_f16_bits(v * scale); }
  __device__ static float rt(float v, float scale, float inv_scale) {
    return f16_bits_to_float(float_to_f16_bits(v * scale)) * inv_scale;
  }
  __device__ static float bld(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so, float inv_scale) {
    return f16_bits_to_float(__builtin_amdgcn_raw_buffer_load_b16(r, int(vo), int(so), 0)) * inv_scale;
  }
  __device__ static void bst(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so, float v, float scale) {
    __builtin_amdgcn_raw_buffer_store_b16(float_to_f16_bits(v * scale), r, int(vo), int(so), 0);
  }
};

template <typename R>
__device__ __forceinline__ void adj_coords(const AdjArgs<R>& a, uint32_t c, int& x, int& y, int& z) {
  const uint32_t yz = fdiv(c, a.dnx);
  x = int(c - yz * uint32_t(a.nx));
  const uint32_t zz = fdiv(yz, a.dny);
  y = int(yz - zz * uint32_t(a.ny));
  z = int(zz);
}

// Element offset of cell (x, y, z) within a population (rows RS apart).
template <typename R>
__device__ __forceinline__ uint32_t adj_addr(const AdjArgs<R>& a, int x, int y, int z) {
  return (uint32_t(z) * uint32_t(a.ny) + uint32_t(y)) * a.RS + uint32_t(x);
}

// offset of cell x - e_I with all three axes wrapped: the entry whose normal
// consumer is x
template <int I, typename R>
__device__ __forceinline__ uint32_t owner_target(const AdjArgs<R>& a, int x, int y, int z) {
  int sx = x - EX[I], sy = y - EY[I], sz = z - EZ[I];
  sx = sx < 0 ? sx + a.nx : (sx >= a.nx ? sx - a.nx : sx);
  sy = sy < 0 ? sy + a.ny : (sy >= a.ny ? sy - a.ny : sy);
  sz = sz < 0 ? sz + a.nz : (sz >= a.nz ? sz - a.nz : sz);
  return adj_addr(a, sx, sy, sz);
}

// offset of cell x + e_I, all axes wrapped
template <int I, typename R>
__device__ __forceinline__ uint32_t next_cell(const AdjArgs<R>& a, int x, int y, int z) {
  int fx = x + EX[I], fy = y + EY[I], fz = z + EZ[I];
  fx = fx < 0 ? fx + a.nx : (fx >= a.nx ? fx - a.nx : fx);
  fy = fy < 0 ? fy + a.ny : (fy >= a.ny ? fy - a.ny : fy);
  fz = fz < 0 ? fz + a.nz : (fz >= a.nz ? fz - a.nz : fz);
  return adj_addr(a, fx, fy, fz);
}

template <int I>
__device__ __forceinline__ bool x_fallback(int x, int nx) {
  if constexpr (EX[I] == 1) return x == 0;
  else if constexpr (EX[I] == -1) return x == nx - 1;
  else return false;
}

// Pulled values of a non-link cell x at offset o (streaming.py:113-142):
// x - e_i with all axes wrapped, the cell itself at the x faces.
template <class S>
__device__ __forceinline__ void pull_dense(const AdjArgs<typename S::R>& a, const typename S::T* __restrict__ P,
                                           uint32_t o, int x, int y, int z, typename S::R* s) {
  const size_t N = a.P;
  s[0] = S::ld_c(P, o, a.off[0]);
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    if constexpr (i != 0) {
      const size_t k = i * N + (x_fallback<i>(x, a.nx) ? o : owner_target<i>(a, x, y, z));
      s[i] = S::ld_c(P, k, a.off[i]);
    }
  });
}

// The cotangent addition a link cell's Bouzidi pulls and force make to its
// own entry P_k(x) (streaming.py:155-178, forces.py:95-118), from the
// cotangents mu of its pulled values and the step's force seed g:
//   a wall in direction opp(k) (q_k = qk >= 0): c mu_opp(k) + e_k . g, with
//     c = 1/(2 q_k) (q_k >= 1/2) or 2 q_k (below);
//   a wall in direction k on the high branch (q_opp(k) = qko >= 1/2):
//     (1 - 1/(2 q_opp(k))) mu_k.
// The remaining term, (1 - 2q) mu_i into P_opp(i)(x + e_i), goes to an entry
// the cell owner-writes itself.  The coefficients are the forward's
// (storage-precision q).
template <int K, typename R>
__device__ __forceinline__ R own_link_addition(R qk, R qko, R muk, R muko, R gx, R gy, R gz) {
  R add = R(0);
  if (qk >= R(0)) {
    const R c = qk >= R(0.5) ? R(1) / (R(2) * fmin(qk, R(1))) : R(2) * qk;
    R f = R(0);
    add_scaled<EX[K]>(f, gx);
    add_scaled<EY[K]>(f, gy);
    add_scaled<EZ[K]>(f, gz);
    add = fma(c, muko, f);
  }
  if (qko >= R(0.5)) add = fma(R(1) - R(1) / (R(2) * fmin(qko, R(1))), muk, add);
  return add;
}

template <typename R>
__device__ __forceinline__ void apply_zou_he(const AdjArgs<R>& a, R* s, int x) {
  if (x == 0) zou_he_in(s, a.uin[0], a.uin[1], a.uin[2], false, R(0));
  if (x == a.nx - 1) zou_he_out(s, a.rho_out);
}

// g <- (pre-guard, collide, bounce)^T g at the collision input s (the cell's
// S_{t+1}); s is replaced by the pre-guard reset state where the guard fires.
// Accumulates dL/dporosity (the term times gunit, g's unit) unless `skip`.
// fl = the cell's flags.  Returns true when the pre guard fired.  SOFT: 1
// the caller knows the bounce is soft, 0 that it is not, -1 neither (read at
// run time); a sweep compiled for one case keeps the other branch out of its
// code (the soft path's new registers would make every other cell copy its
// 19 cotangents at the branch join).
template <bool RCP, int COLL, int SOFT = -1, typename R>
__device__ __forceinline__ bool collide_transpose(const AdjArgs<R>& a, R* s, R* g, R gunit, uint32_t c, uint32_t fl,
                                                  bool skip) {
  Pairs<R> ps = to_pairs(s);
  Macro<R> m = macro<RCP>(ps);
  const bool pre_reset = (a.guard == LBM_GUARD_PRE) && guard_bad(m);
  if (pre_reset) {
    set_feq_unit(s, a.uref[0], a.uref[1], a.uref[2]);
    ps = to_pairs(s);
    m = macro<RCP>(ps);
  }
  const R omega = a.smag ? R(1) / smag_tau(ps, m, a.tau, a.cs) : a.omega;
  if (SOFT != 1 && a.bounce == LBM_BOUNCE_HARD) {
    if (fl & 1) reverse_all(g);
  } else if (SOFT != 0 && a.bounce == LBM_BOUNCE_SOFT) {
    R cc[Q];
    sfor<Q>([&](auto I) {
      constexpr int i = I;
      cc[i] = s[i];
    });
    if constexpr (COLL == LBM_COLL_MRT) collide_mrt(cc, omega, ps, m);
    else collide_bgk(cc, omega, m);
    const R gp = blend_vjp(cc, g, a.phi[c]);
    if (a.phi_acc && !skip) a.phi_acc[c] += double(gp) * double(gunit);
  }
  // dL/domega is needed only for the Smagorinsky path (a uniform branch;
  // without it the transpose skips the stress moments' dm).  A cell the pre
  // guard reset passes no cotangent back: it is transposed like the others
  // (at its reset state) and then cleared, so the cleared case writes zeros
  // into the same registers instead of making every cell copy at a join.
  if (a.smag) {
    const R gom = (COLL == LBM_COLL_MRT) ? mrt_vjp<true>(ps, m, g, omega) : bgk_vjp<true>(ps, m, g, omega);
    smag_vjp(ps, m, g, a.tau, a.cs, -gom * omega * omega);
  } else if constexpr (COLL == LBM_COLL_MRT) {
    mrt_vjp<false>(ps, m, g, omega);
  } else {
    bgk_vjp<false>(ps, m, g, omega);
  }
  if (pre_reset) {
    sfor<Q>([&](auto I) {
      constexpr int i = I;
      g[i] = R(0);
    });
  }
  return pre_reset;
}

// The part of a cell's adjoint shared by the link, dense and row-uniform
// sweeps: Zou-He and the post guard on the pulled s, the transposed
// collision, the link cell's force seed, and the Zou-He transposes.  On entry
// g = dL/dP_{t+1}(x) (MODE FULL) or dL/dS_{t+1}(x) (MODE STREAM); on exit
// g = mu, the cotangent of the pulled values.  load_g(g) loads the incoming
// cotangent; it runs after the forward boundaries, which keeps the 19 values
// out of the registers while those are computed.  g may hold the cotangent
// times a constant (a storage scale): the map is linear in g, and gunit
// (1 / that constant) converts the accumulated parameter terms; the force
// seed (link cells) needs gunit = 1.  SOFT as for collide_transpose.
template <class S, int COLL, int MODE, bool LINK, int SOFT = -1, class LoadG>
__device__ __forceinline__ void cell_transpose(const AdjArgs<typename S::R>& a, typename S::R* s, typename S::R* g,
                                               typename S::R gunit, LoadG&& load_g, uint32_t c, int x, uint32_t fl,
                                               bool skip, uint32_t lmask, const double* __restrict__ gF) {
  using R = typename S::R;
  // the inlet's density depends on the pulled values it keeps
  R rho_in = R(0);
  if (x == 0) rho_in = zou_he_in_rho(s, a.uin[0]);
  apply_zou_he(a, s, x);
  // the reset stays under the uniform guard-mode test (if-converted outside
  // it, every cell would select its 19 values)
  bool post_reset = false;
  if (a.guard == LBM_GUARD_POST) {
    post_reset = guard_bad<rcp_of<S>>(s);
    if (post_reset) set_weights(s);
  }
  load_g(g);
  if constexpr (MODE == LBM_MODE_FULL) collide_transpose<rcp_of<S>, COLL, SOFT>(a, s, g, gunit, c, fl, skip);
  // momentum-exchange force uses S_{t+1}(x)_opp(j) for links j (forces.py:95-118)
  if constexpr (LINK) {
    if (gF) {
      const R gx = R(gF[0]), gy = R(gF[1]), gz = R(gF[2]);
      sfor<Q>([&](auto J) {
        constexpr int j = J;
        if constexpr (j != 0) {
          if (lmask & (1u << j)) {
            R v = R(0);
            add_scaled<EX[j]>(v, gx);
            add_scaled<EY[j]>(v, gy);
            add_scaled<EZ[j]>(v, gz);
            g[opp(j)] += v;
          }
        }
      });
    }
  }
  if (post_reset) {
    sfor<Q>([&](auto I) {
      constexpr int i = I;
      g[i] = R(0);
    });
  }
  // Zou-He^T (the outlet was applied last)
  if (x == a.nx - 1) zou_he_out_vjp(g, a.rho_out);
  if (x == 0) {
    double gu[3] = {0.0, 0.0, 0.0};
    zou_he_in_vjp(rho_in, g, a.uin[0], a.uin[1], a.uin[2], gu);
    if (a.u_acc && !skip) {
      const size_t row = size_t(c) / size_t(a.nx);
      a.u_acc[3 * row] += gu[0] * double(gunit);
      a.u_acc[3 * row + 1] += gu[1] * double(gunit);
      a.u_acc[3 * row + 2] += gu[2] * double(gunit);
    }
  }
}

// One dense-sweep cell: forward recompute + cotangent pull-back + owner
// writes.  MODE FULL: lam_in = dL/dP_{t+1}; MODE STREAM: lam_in = dL/dS_{t+1}.
// A thread on a link cell computes but writes nothing (the cell belongs to a
// link workgroup).
template <class S, class LM, int COLL, int MODE>
__device__ __forceinline__ void adjoint_cell(const AdjArgs<typename S::R>& a, const typename S::T* __restrict__ P,
                                             const typename LM::T* __restrict__ lam_in,
                                             typename LM::T* __restrict__ lam_out, uint32_t c, uint32_t fl) {
  using R = typename S::R;
  const size_t N = a.P;
  const bool skip = fl & 2u;
  int x, y, z;
  adj_coords(a, c, x, y, z);
  const uint32_t o = adj_addr(a, x, y, z);
  R s[Q], g[Q];
  pull_dense<S>(a, P, o, x, y, z, s);
  auto load_g = [&](R* gg) {
    sfor<Q>([&](auto I) {
      constexpr int i = I;
      gg[i] = LM::ld(lam_in, i * N + o, a.lam_inv);
    });
  };
  cell_transpose<S, COLL, MODE, false>(a, s, g, R(1), load_g, c, x, fl, skip, 0u, nullptr);
  if (skip) return;
  LM::st(lam_out, o, g[0], a.lam_scale);
  // the owner targets are the pull sources: recompute them here rather than
  // keep 18 addresses alive (or spilled) through the transposed collision
  int xw = x, yw = y, zw = z;
  asm volatile("" : "+v"(xw), "+v"(yw), "+v"(zw));
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    if constexpr (i != 0) {
      const R v = x_fallback<i>(xw, a.nx) ? R(0) : g[i];
      LM::st(lam_out, i * N + owner_target<i>(a, xw, yw, zw), v, a.lam_scale);
    }
  });
}

// LDS of an adjoint link workgroup: per direction and link cell (lane), the
// pulled value S_i (then the cotangent mu_i), dS_i/dq_opp(i), q_i, and the
// loaded cotangent of the cell's own P_i.
template <typename R>
struct LinkLds {
  R s[Q][kLinkCellsPerBlock];
  R dsdq[Q][kLinkCellsPerBlock];
  R q[Q][kLinkCellsPerBlock];
  R g[Q][kLinkCellsPerBlock];
};

// Phase 1 of an adjoint link workgroup, on the wave that owns the directions
// i with i % NW == W: the forward's Bouzidi pull of S_i (streaming.py:113-181)
// and dS_i/dq_opp(i) on the taken branch (0 outside the clip bounds), q_opp(i),
// and the cotangent of the cell's own P_i with the previous launch's folded
// addition (rounded to the cotangent storage as lbm_adjoint_link_finish would
// store it).  Straight-line code over the wave's few directions: two rounds
// of loads (q, then the populations and cotangents).  Sets bit opp(i) of mask
// where opp(i) is a wall link.
template <class S, class LM, int W, int NW>
__device__ __forceinline__ void adj_link_pull_part(const AdjArgs<typename S::R>& a, const typename S::T* __restrict__ P,
                                                   const typename LM::T* __restrict__ lam_in, int b, int x, int y,
                                                   int z, uint32_t o, LinkLds<typename S::R>& L, int lane,
                                                   uint32_t& mask) {
  using R = typename S::R;
  constexpr int ND = (Q - W + NW - 1) / NW;  // directions W, W + NW, ...
  const size_t N = a.P;
  R qo[ND], va[ND], vb[ND], lg[ND];
  sfor<ND>([&](auto K) {
    constexpr int i = W + int(K) * NW;
    if constexpr (i != 0) qo[K] = R(a.link_q[size_t(opp(i) - 1) * a.n_links + b]);
  });
  sfor<ND>([&](auto K) {
    constexpr int i = W + int(K) * NW;
    lg[K] = LM::ld(lam_in, i * N + o, a.lam_inv);
    if constexpr (i == 0) {
      va[K] = S::ld_c(P, o, a.off[0]);
    } else {
      constexpr int io = opp(i);
      const bool low = (qo[K] >= R(0)) & (qo[K] < R(0.5));
      const bool high = qo[K] >= R(0.5);
      const size_t normal = i * N + (x_fallback<i>(x, a.nx) ? o : owner_target<i>(a, x, y, z));
      const size_t ka = low ? io * N + next_cell<i>(a, x, y, z) : (high ? i * N + o : normal);
      va[K] = S::ld_cv(P, ka, low ? a.off[io] : a.off[i]);
      vb[K] = S::ld_c(P, io * N + o, a.off[io]);
      if (a.add_prev) lg[K] = LM::rt(lg[K] + a.add_prev[size_t(i) * a.n_links + b], a.lam_scale, a.lam_inv);
    }
  });
  sfor<ND>([&](auto K) {
    constexpr int i = W + int(K) * NW;
    L.g[i][lane] = lg[K];
    if constexpr (i == 0) {
      L.s[0][lane] = va[K];
      L.dsdq[0][lane] = R(0);
      L.q[0][lane] = R(-1);
    } else {
      constexpr int io = opp(i);
      const R q = qo[K];
      R si, d = R(0);
      if (q >= R(0.5)) {
        const R inv2q = R(1) / (R(2) * fmin(q, R(1)));
        si = inv2q * vb[K] + (R(1) - inv2q) * va[K];
        if (q <= R(1)) d = -(vb[K] - va[K]) / (R(2) * q * q);
      } else if (q >= R(0)) {
        const R twoq = R(2) * q;
        si = twoq * vb[K] + (R(1) - twoq) * va[K];
        d = R(2) * (vb[K] - va[K]);
      } else {
        si = va[K];
      }
      L.s[i][lane] = si;
      L.dsdq[i][lane] = d;
      L.q[io][lane] = q;
      if (q >= R(0)) mask |= 1u << io;
    }
  });
}

// Phase 3 of an adjoint link workgroup, on the wave that owns the directions
// i with i % NW == W (i != 0), from the cotangents mu of the cell's pulled
// values: the owner write of entry (x - e_i, i) -- mu_i for a normal pull, 0
// for a wall link or an x face, plus, folded, the low-branch second-neighbour
// term (1 - 2 q_i) mu_opp(i) of the wall direction opp(i) -- then dL/dq of the
// wall link opp(i) and mu_link (the link cotangent, or folded, the addition
// to the cell's own P_i for the next launch).
template <class S, class LM, int W, int NW>
__device__ __forceinline__ void adj_link_write_part(const AdjArgs<typename S::R>& a,
                                                    typename LM::T* __restrict__ lam_out,
                                                    const double* __restrict__ gF, int b, int x, int y, int z,
                                                    uint32_t mask, const LinkLds<typename S::R>& L, int lane) {
  using R = typename S::R;
  constexpr int ND = (Q - W + NW - 1) / NW;
  const size_t N = a.P;
  R gx = R(0), gy = R(0), gz = R(0);
  if (a.fold && gF) {
    gx = R(gF[0]);
    gy = R(gF[1]);
    gz = R(gF[2]);
  }
  sfor<ND>([&](auto K) {
    constexpr int i = W + int(K) * NW;
    if constexpr (i != 0) {
      constexpr int io = opp(i);
      const R gi = L.g[i][lane], gio = L.g[io][lane];
      const R qi = L.q[i][lane];
      const bool link = mask & (1u << io);
      const bool normal = !link && !x_fallback<i>(x, a.nx);
      R v = normal ? gi : R(0);
      if (a.fold && qi >= R(0) && qi < R(0.5)) v = fma(R(1) - R(2) * qi, gio, v);
      LM::st(lam_out, i * N + owner_target<i>(a, x, y, z), v, a.lam_scale);
      if (link && a.dq_acc) a.dq_acc[size_t(io - 1) * a.n_links + b] += double(gi) * double(L.dsdq[i][lane]);
      a.mu_link[size_t(i) * a.n_links + b] =
          a.fold ? own_link_addition<i>(qi, L.q[io][lane], gi, gio, gx, gy, gz) : (link ? gi : R(0));
    }
  });
}

// An adjoint link workgroup: kLinkCellsPerBlock link cells, lane l of every
// wave on link b = first + l.  Phase 1 (all waves): the Bouzidi pulls and the
// cotangent loads, spread over the waves by direction (adj_link_pull_part).
// Phase 2 (wave 0, one lane per cell): Zou-He, the transposed collision, the
// force seed and the Zou-He transposes (cell_transpose), leaving mu in LDS.
// Phase 3 (all waves): the owner writes and link bookkeeping by direction
// (adj_link_write_part).  A link cell so waits two rounds of loads before its
// transposed collision, and no wave runs all 18 directions' address work.
template <class S, class LM, int COLL, int MODE, int BS>
__device__ __forceinline__ void adjoint_link_block(const AdjArgs<typename S::R>& a,
                                                   const typename S::T* __restrict__ P,
                                                   const typename LM::T* __restrict__ lam_in,
                                                   typename LM::T* __restrict__ lam_out,
                                                   const double* __restrict__ gF, uint32_t bid) {
  using R = typename S::R;
  constexpr int NW = BS / 64;
  static_assert(kLinkCellsPerBlock == 64, "one link cell per lane");
  __shared__ LinkLds<R> L;
  __shared__ uint32_t lm[NW][kLinkCellsPerBlock];
  const int lane = int(threadIdx.x & 63u);
  const int w = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
  const int b = int(bid) * kLinkCellsPerBlock + lane;
  const bool live = b < a.n_links;
  uint32_t mask = 0, c = 0, fl = 0, o = 0;
  int x = 0, y = 0, z = 0;
  if (live) {
    c = uint32_t(a.link_cell[b]);
    adj_coords(a, c, x, y, z);
    o = adj_addr(a, x, y, z);
    if (w == 0) fl = a.flags[c];
    sfor<NW>([&](auto W) {
      if (w == int(W)) adj_link_pull_part<S, LM, int(W), NW>(a, P, lam_in, b, x, y, z, o, L, lane, mask);
    });
  }
  lm[w][lane] = mask;
  __syncthreads();
  mask = 0;
#pragma unroll
  for (int k = 0; k < NW; ++k) mask |= lm[k][lane];
  if (w == 0 && live) {
    R s[Q], g[Q];
    sfor<Q>([&](auto I) {
      constexpr int i = I;
      s[i] = L.s[i][lane];
    });
    auto load_g = [&](R* gg) {
      sfor<Q>([&](auto I) {
        constexpr int i = I;
        gg[i] = L.g[i][lane];
      });
    };
    cell_transpose<S, COLL, MODE, true>(a, s, g, R(1), load_g, c, x, fl, false, mask, gF);
    LM::st(lam_out, o, g[0], a.lam_scale);
    sfor<Q>([&](auto I) {
      constexpr int i = I;
      L.g[i][lane] = g[i];
    });
  }
  __syncthreads();
  if (live) {
    sfor<NW>([&](auto W) {
      if (w == int(W)) adj_link_write_part<S, LM, int(W), NW>(a, lam_out, gF, b, x, y, z, mask, L, lane);
    });
  }
}

// Dense cell when every wave's 64 cells share one (y, z) row (nx a multiple
// of 64, buffers below 4 GiB).  The pull source of direction i and its owner
// target are the same entry, x - e_i wrapped; its row is wave-uniform, so the
// state and cotangent accesses go through buffer resources with a scalar
// byte offset per direction and three per-lane x offsets.  At x = 0 the
// e_x = +1 pulls and at x = nx - 1 the e_x = -1 pulls read the wrapped
// neighbour instead of the cell's own value: Zou-He overwrites exactly those
// populations from the others (boundary.py:113-122, 190-198), so the recomputed
// step is the same.
template <class S, class LM, int COLL, int MODE, int SOFT>
__device__ __forceinline__ void adjoint_row(const AdjArgs<typename S::R>& a, const typename S::T* __restrict__ P,
                                            const typename LM::T* __restrict__ lam_in,
                                            typename LM::T* __restrict__ lam_out, uint32_t c) {
  using R = typename S::R;
  constexpr uint32_t B = sizeof(typename S::T), BL = sizeof(typename LM::T);
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t cw = __builtin_amdgcn_readfirstlane(c - lane);  // first cell of the wave
  const uint32_t yz = fdiv(cw, a.dnx);
  const uint32_t zz = fdiv(yz, a.dny);
  const int y = int(yz - zz * uint32_t(a.ny)), z = int(zz);
  const int x = int(cw - yz * uint32_t(a.nx) + lane);
  const uint32_t xm = uint32_t(x == 0 ? a.nx - 1 : x - 1), xp = uint32_t(x == a.nx - 1 ? 0 : x + 1);
  const uint32_t fl = a.flags ? uint32_t((a.flags + cw)[lane]) : 0u;
  const bool skip = fl & 2u;
  const __amdgpu_buffer_rsrc_t rp = buffer_rsrc(P, a.state_bytes);
  const __amdgpu_buffer_rsrc_t rli = buffer_rsrc(lam_in, a.lam_bytes);
  const __amdgpu_buffer_rsrc_t rlo = buffer_rsrc(lam_out, a.lam_bytes);
  // element offset of the row (y - dy, z - dz), periodic
  auto row_of = [&](int dy, int dz) -> uint32_t {
    int ry = y - dy, rz = z - dz;
    ry = ry < 0 ? ry + a.ny : (ry >= a.ny ? ry - a.ny : ry);
    rz = rz < 0 ? rz + a.nz : (rz >= a.nz ? rz - a.nz : rz);
    return (uint32_t(rz) * uint32_t(a.ny) + uint32_t(ry)) * a.RS;
  };
  const uint32_t own = yz * a.RS;
  R s[Q], g[Q];
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    constexpr int ex = EX[i];
    const uint32_t xs = ex == 1 ? xm : (ex == -1 ? xp : uint32_t(x));
    const uint32_t e = uint32_t(i) * a.P + row_of(EY[i], EZ[i]);
    s[i] = S::bld(rp, xs * B, e * B, a.off[i]);
  });
  // 16-bit sweeps load the cotangents together with the pulled values, so a
  // wave waits for one round of loads rather than two (FP16 forward + adjoint
  // at 256 x 128 x 128: 158 -> 153 us per step); the float32 sweep was 1%
  // slower this way
  constexpr bool early = sizeof(typename S::T) == 2;
  if constexpr (early) {
    sfor<Q>([&](auto I) {
      constexpr int i = I;
      g[i] = LM::bld(rli, uint32_t(x) * BL, (uint32_t(i) * a.P + own) * BL, R(1));
    });
  }
  auto load_g = [&](R* gg) {
    if constexpr (!early) {
      sfor<Q>([&](auto I) {
        constexpr int i = I;
        gg[i] = LM::bld(rli, uint32_t(x) * BL, (uint32_t(i) * a.P + own) * BL, R(1));
      });
    }
  };
  // g stays in the stored units (lambda times the storage scale, a power of
  // two): the transpose commutes with it exactly
  cell_transpose<S, COLL, MODE, false, SOFT>(a, s, g, a.lam_inv, load_g, c, x, fl, skip, 0u, nullptr);
  if (skip) return;
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    constexpr int ex = EX[i];
    const uint32_t xs = ex == 1 ? xm : (ex == -1 ? xp : uint32_t(x));
    const uint32_t e = uint32_t(i) * a.P + row_of(EY[i], EZ[i]);
    const bool normal = !x_fallback<i>(x, a.nx);
    LM::bst(rlo, xs * BL, e * BL, normal ? g[i] : R(0), R(1));
  });
}

// Occupancy target of the dense adjoint sweep: the FP32 MRT transpose fits
// 128 VGPRs (4 waves per SIMD) without spills; the other variants keep the
// compiler's allocation.
template <class S, int COLL>
constexpr int adj_waves_per_eu() {
  return (sizeof(typename S::R) == 4 && COLL == LBM_COLL_MRT) ? 4 : 1;
}

// One adjoint step: workgroups [0, nlb) own the link cells (Bouzidi transpose,
// mu_link, dq), the rest sweep the dense grid; dense threads on link cells
// compute but write nothing.
template <class S, class LM, int COLL, int MODE, bool ROW, int BS = kAdjBlock>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(adj_waves_per_eu<S, COLL>()))) void k_adjoint(
    AdjArgs<typename S::R> a, const typename S::T* __restrict__ P, const typename LM::T* __restrict__ lam_in,
    typename LM::T* __restrict__ lam_out, const double* __restrict__ gF) {
  using R = typename S::R;
  const uint32_t bid = blockIdx.x;
  if (int(bid) < a.nlb) {
    adjoint_link_block<S, LM, COLL, MODE, BS>(a, P, lam_in, lam_out, gF, bid);
    return;
  }
  // tiles in workgroup order (grouped per XCD where a tile holds part of a
  // row): consecutive rows run on all XCDs at once
  const uint32_t tile = xcd_group_tile(bid - uint32_t(a.nlb), uint32_t(a.ndb), a.tgroup);
  const uint32_t c = tile * BS + threadIdx.x;
  if (c >= a.N) return;
  if constexpr (ROW) {
    if (a.bounce == LBM_BOUNCE_SOFT) adjoint_row<S, LM, COLL, MODE, 1>(a, P, lam_in, lam_out, c);
    else adjoint_row<S, LM, COLL, MODE, 0>(a, P, lam_in, lam_out, c);
    return;
  }
  adjoint_cell<S, LM, COLL, MODE>(a, P, lam_in, lam_out, c, a.flags ? uint32_t(a.flags[c]) : 0u);
}

// Prologue transpose: dL/dS_0 = ([pre guard] collide bounce)^T lambda_0, cell-local.
template <class S, class LM, int COLL>
__global__ __launch_bounds__(kAdjBlock) void k_adjoint_collide(AdjArgs<typename S::R> a,
                                                               const typename S::T* __restrict__ S0,
                                                               const typename LM::T* __restrict__ lam_in,
                                                               typename S::R* __restrict__ g_out) {
  using R = typename S::R;
  const uint32_t c = blockIdx.x * kAdjBlock + threadIdx.x;
  if (c >= a.N) return;
  const size_t N = a.P;
  int x, y, z;
  adj_coords(a, c, x, y, z);
  const uint32_t o = adj_addr(a, x, y, z);
  R s[Q], g[Q];
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    s[i] = S::ld_c(S0, i * N + o, a.off[i]);
    g[i] = LM::ld(lam_in, i * N + o, a.lam_inv);
  });
  collide_transpose<rcp_of<S>, COLL>(a, s, g, R(1), c, a.flags ? uint32_t(a.flags[c]) : 0u, false);
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    g_out[i * N + o] = g[i];
  });
}

// Deterministic additions: lam[tgt[k]] += sum_c coef[c] * mu_link[src[c]]
// (+ gF . e_j for force terms, encoded src = -1 - j), summed in list order.
// The target's old value and a batch of SCATTER_BATCH entries are loaded
// before any of them is used, so a target costs about three dependent round
// trips instead of two per entry. Entry (x, j) receives at most four additions
// (the own term of link (x, opp j), the high-branch and force terms of link
// (x, j), the low-branch neighbour term of link (x + e_j, opp j)), so one batch
// covers every target; longer lists would take more batches.
constexpr int SCATTER_BATCH = 4;
// 64-thread workgroups: the few thousand targets spread over more CUs (96 x 48 x 48:
// 2000-step dCd/dr 73.5 -> 72.8 ms against 256-thread workgroups).
constexpr int kScatterBlock = 64;

template <typename R, class LM>
__global__ void k_adjoint_scatter(int n_targets, const int32_t* __restrict__ tgt, const int32_t* __restrict__ start,
                                  const int32_t* __restrict__ src, const double* __restrict__ coef,
                                  const R* __restrict__ mu_link, const double* __restrict__ gF,
                                  typename LM::T* __restrict__ lam, R lam_scale, R lam_inv) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_targets) return;
  const size_t t = size_t(tgt[k]);
  const R old = LM::ld(lam, t, lam_inv);
  const int e_end = start[k + 1];
  R gx = R(0), gy = R(0), gz = R(0);
  if (gF) {
    gx = R(gF[0]);
    gy = R(gF[1]);
    gz = R(gF[2]);
  }
  R acc = R(0);
  for (int e = start[k]; e < e_end; e += SCATTER_BATCH) {
    int sb[SCATTER_BATCH];
    double cb[SCATTER_BATCH];
    R mb[SCATTER_BATCH];
#pragma unroll
    for (int u = 0; u < SCATTER_BATCH; ++u) {
      const bool in = e + u < e_end;
      sb[u] = in ? src[e + u] : -1;
      cb[u] = in ? coef[e + u] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < SCATTER_BATCH; ++u) mb[u] = (e + u < e_end && sb[u] >= 0) ? mu_link[sb[u]] : R(0);
#pragma unroll
    for (int u = 0; u < SCATTER_BATCH; ++u) {
      if (e + u >= e_end) break;
      if (sb[u] >= 0) {
        acc += R(cb[u]) * mb[u];
      } else if (gF) {
        const int j = -1 - sb[u];
        R v = R(0);
        sfor<Q>([&](auto J) {
          constexpr int jj = J;
          if (j == jj) v = R(EX[jj]) * gx + R(EY[jj]) * gy + R(EZ[jj]) * gz;
        });
        acc += v;
      }
    }
  }
  LM::st(lam, t, old + acc, lam_scale);
}

// Folded link additions of one launch into its output: one thread per link
// cell adds the additions that launch stored (add, 19 x n_links) to the
// cell's own entries.  Each link cell's entries are its own, so no two
// threads touch the same entry.
template <typename R, class LM>
__global__ void k_link_finish(AdjArgs<R> a, const R* __restrict__ add, typename LM::T* __restrict__ lam) {
  const int b = int(blockIdx.x * blockDim.x + threadIdx.x);
  if (b >= a.n_links) return;
  int x, y, z;
  adj_coords(a, uint32_t(a.link_cell[b]), x, y, z);
  const uint32_t o = adj_addr(a, x, y, z);
  const size_t N = a.P;
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    if constexpr (i != 0) {
      const size_t e = size_t(i) * N + o;
      LM::st(lam, e, LM::ld(lam, e, a.lam_inv) + add[size_t(i) * a.n_links + b], a.lam_scale);
    }
  });
}

template <typename R>
int make_adj_args(const lbm_step_desc* d, const lbm_adjoint_desc* ad, int mode, AdjArgs<R>& a) {
  if (!d || !ad || d->nx < 2 || d->ny <= 0 || d->nz <= 0) return LBM_E_ARG;
  const long long n = 1LL * d->nx * d->ny * d->nz;
  long long pstride = 0, rstride = 0;
  if (!state_layout(d, pstride, rstride)) return LBM_E_ARG;
  if (!d->periodic_yz) return LBM_E_ARG;  // the owner-write scheme needs periodic y/z
  a.P = uint32_t(pstride);
  a.RS = uint32_t(rstride);
  a.nx = d->nx;
  a.ny = d->ny;
  a.nz = d->nz;
  a.N = uint32_t(n);
  a.dnx = make_fastdiv(uint32_t(d->nx));
  a.dny = make_fastdiv(uint32_t(d->ny));
  a.smag = d->smagorinsky;
  a.bounce = d->bounce;
  a.guard = d->guard;
  a.n_links = d->n_links;
  a.tau = R(d->tau);
  a.omega = R(1) / a.tau;
  a.cs = R(d->smag_cs);
  for (int k = 0; k < 3; ++k) {
    a.uin[k] = R(d->u_in[k]);
    a.uref[k] = R(d->u_ref[k]);
  }
  a.rho_out = R(d->rho_out);
  for (int i = 0; i < Q; ++i) a.off[i] = R(d->offset[i]);
  a.flags = d->flags;
  a.phi = static_cast<const R*>(d->porosity);
  a.link_cell = d->link_cell;
  a.link_q = d->link_q;
  a.dq_acc = ad->dq_acc;
  a.mu_link = static_cast<R*>(ad->mu_link);
  a.phi_acc = ad->phi_acc;
  a.u_acc = ad->u_acc;
  a.fold = ad->fold_links ? 1 : 0;
  a.add_prev = nullptr;
  const double scale = (ad->lam_storage == LBM_F16DEV && ad->lam_scale > 0.0) ? ad->lam_scale : 1.0;
  a.lam_scale = R(scale);
  a.lam_inv = R(1.0 / scale);
  if (a.bounce == LBM_BOUNCE_HARD && !a.flags) return LBM_E_ARG;
  if (a.bounce == LBM_BOUNCE_SOFT && !a.phi) return LBM_E_ARG;
  if (a.n_links < 0 || (a.n_links > 0 && (!a.link_cell || !a.link_q || !a.flags || !a.mu_link))) return LBM_E_ARG;
  a.nlb = (mode == LBM_MODE_COLLIDE) ? 0 : (a.n_links + kLinkCellsPerBlock - 1) / kLinkCellsPerBlock;
  a.ndb = int((n + kAdjBlock - 1) / kAdjBlock);
  a.tgroup = tile_group_size(uint32_t(a.nx), uint32_t(kAdjBlock));
  return LBM_OK;
}

// mu_cur (null: ad->mu_link) receives this launch's link-cell cotangents,
// or with fold_links their additions to the cells' own entries; add_prev is
// the previous launch's additions in a folded chain (else null).
template <class S, class LM>
int launch_adjoint(const lbm_step_desc* d, const lbm_adjoint_desc* ad, int mode, const void* state,
                   const void* lam_in, void* lam_out, const double* gF, hipStream_t st,
                   void* mu_cur = nullptr, const void* add_prev = nullptr) {
  using R = typename S::R;
  AdjArgs<R> a;
  const int rc = make_adj_args(d, ad, mode, a);
  if (rc) return rc;
  if (mu_cur) a.mu_link = static_cast<R*>(mu_cur);
  a.add_prev = static_cast<const R*>(add_prev);
  if (!state || !lam_in || !lam_out || lam_in == lam_out) return LBM_E_ARG;
  // the row-uniform sweep addresses both buffers through 32-bit byte offsets
  const long long extent = 1LL * (Q - 1) * a.P + (1LL * a.ny * a.nz - 1) * a.RS + a.nx;
  const long long sb = extent * (long long)sizeof(typename S::T), lb = extent * (long long)sizeof(typename LM::T);
  const bool row = d->cells_per_thread == 0 && a.nx % 64 == 0 && sb < (1LL << 32) && lb < (1LL << 32);
  // 512-thread workgroups for the float32 sweep of grids of 2^21 cells or
  // more (1.5% faster than 256 at 256 x 128 x 128); the 16-bit sweep is 8%
  // slower at 512 and keeps 256
  constexpr bool f32 = std::is_same<S, StoreF32>::value;
  const bool large = f32 && mode != LBM_MODE_COLLIDE && 1LL * a.N >= (1LL << 21);
  if (large) {
    a.nlb = (a.n_links + kLinkCellsPerBlock - 1) / kLinkCellsPerBlock;
    a.ndb = int((1LL * a.N + kAdjBlockLarge - 1) / kAdjBlockLarge);
    a.tgroup = tile_group_size(uint32_t(a.nx), uint32_t(kAdjBlockLarge));
  }
  a.state_bytes = row ? uint32_t(sb) : 0u;
  a.lam_bytes = row ? uint32_t(lb) : 0u;
  const auto* P = static_cast<const typename S::T*>(state);
  const auto* li = static_cast<const typename LM::T*>(lam_in);
  const bool mrt = d->collision == LBM_COLL_MRT;
  if (mode == LBM_MODE_COLLIDE) {
    // the prologue's output dL/dS_0 is in the arithmetic type
    auto* go = static_cast<R*>(lam_out);
    const unsigned grid = unsigned(a.ndb);
    if (mrt) hipLaunchKernelGGL((k_adjoint_collide<S, LM, LBM_COLL_MRT>), grid, kAdjBlock, 0, st, a, P, li, go);
    else hipLaunchKernelGGL((k_adjoint_collide<S, LM, LBM_COLL_BGK>), grid, kAdjBlock, 0, st, a, P, li, go);
  } else {
    auto* lo = static_cast<typename LM::T*>(lam_out);
    const unsigned grid = unsigned(a.nlb + a.ndb);
#define LBM_ADJ(COLL, MODE)                                                                                   \
  do {                                                                                                        \
    constexpr int BL = f32 ? kAdjBlockLarge : kAdjBlock;                                                      \
    if (large && row) hipLaunchKernelGGL((k_adjoint<S, LM, COLL, MODE, true, BL>), grid, BL, 0, st, a, P, li, lo, gF); \
    else if (large) hipLaunchKernelGGL((k_adjoint<S, LM, COLL, MODE, false, BL>), grid, BL, 0, st, a, P, li, lo, gF); \
    else if (row) hipLaunchKernelGGL((k_adjoint<S, LM, COLL, MODE, true>), grid, kAdjBlock, 0, st, a, P, li, lo, gF); \
    else hipLaunchKernelGGL((k_adjoint<S, LM, COLL, MODE, false>), grid, kAdjBlock, 0, st, a, P, li, lo, gF);      \
  } while (0)
    if (mode == LBM_MODE_FULL) {
      if (mrt) LBM_ADJ(LBM_COLL_MRT, LBM_MODE_FULL);
      else LBM_ADJ(LBM_COLL_BGK, LBM_MODE_FULL);
    } else if (mode == LBM_MODE_STREAM) {
      if (mrt) LBM_ADJ(LBM_COLL_MRT, LBM_MODE_STREAM);
      else LBM_ADJ(LBM_COLL_BGK, LBM_MODE_STREAM);
    } else {
      return LBM_E_ARG;
    }
#undef LBM_ADJ
  }
  return hipGetLastError() == hipSuccess ? LBM_OK : LBM_E_LAUNCH;
}

template <class S, class LM>
int launch_finish(const lbm_step_desc* d, const lbm_adjoint_desc* ad, const void* mu, void* lam, hipStream_t st) {
  using R = typename S::R;
  AdjArgs<R> a;
  const int rc = make_adj_args(d, ad, LBM_MODE_FULL, a);
  if (rc) return rc;
  if (a.n_links == 0) return LBM_OK;
  if (!mu || !lam) return LBM_E_ARG;
  const unsigned grid = unsigned((a.n_links + kScatterBlock - 1) / kScatterBlock);
  hipLaunchKernelGGL((k_link_finish<R, LM>), grid, kScatterBlock, 0, st, a, static_cast<const R*>(mu),
                     static_cast<typename LM::T*>(lam));
  return hipGetLastError() == hipSuccess ? LBM_OK : LBM_E_LAUNCH;
}

// Dispatch on the population storage and the cotangent storage.
template <template <class, class> class F, typename... A>
int adj_dispatch(const lbm_step_desc* d, const lbm_adjoint_desc* ad, A... args) {
  if (ad->lam_storage == LBM_F16DEV) {
    switch (d->storage) {
      case LBM_F32: return F<StoreF32, LamHalf>::run(d, ad, args...);
      case LBM_F16DEV: return F<StoreF16, LamHalf>::run(d, ad, args...);
      case LBM_BF16DEV: return F<StoreBF16, LamHalf>::run(d, ad, args...);
    }
    return LBM_E_DTYPE;
  }
  if (ad->lam_storage != 0) return LBM_E_DTYPE;
  switch (d->storage) {
    case LBM_F32: return F<StoreF32, LamPlain<float>>::run(d, ad, args...);
    case LBM_F64: return F<StoreF64, LamPlain<double>>::run(d, ad, args...);
    case LBM_F16DEV: return F<StoreF16, LamPlain<float>>::run(d, ad, args...);
    case LBM_BF16DEV: return F<StoreBF16, LamPlain<float>>::run(d, ad, args...);
  }
  return LBM_E_DTYPE;
}

template <class S, class LM>
struct AdjointLaunch {
  static int run(const lbm_step_desc* d, const lbm_adjoint_desc* ad, int mode, const void* state, const void* lam_in,
                 void* lam_out, const double* gF, hipStream_t st, void* mu_cur, const void* add_prev) {
    return launch_adjoint<S, LM>(d, ad, mode, state, lam_in, lam_out, gF, st, mu_cur, add_prev);
  }
};

template <class S, class LM>
struct FinishLaunch {
  static int run(const lbm_step_desc* d, const lbm_adjoint_desc* ad, const void* mu, void* lam, hipStream_t st) {
    return launch_finish<S, LM>(d, ad, mu, lam, st);
  }
};

}  // namespace lbm

using namespace lbm;

extern "C" {

int lbm_step_adjoint(const lbm_step_desc* d, const lbm_adjoint_desc* ad, int mode, const void* state,
                     const void* lam_in, void* lam_out, const double* gF, lbm_stream_t stream) {
  if (!d || !ad) return LBM_E_ARG;
  return adj_dispatch<AdjointLaunch>(d, ad, mode, state, lam_in, lam_out, gF, (hipStream_t)stream,
                                     static_cast<void*>(nullptr), static_cast<const void*>(nullptr));
}

int lbm_adjoint_link_finish(const lbm_step_desc* d, const lbm_adjoint_desc* ad, const void* mu, void* lam,
                            lbm_stream_t stream) {
  if (!d || !ad || !ad->fold_links) return LBM_E_ARG;
  return adj_dispatch<FinishLaunch>(d, ad, mu ? mu : static_cast<const void*>(ad->mu_link), lam,
                                    (hipStream_t)stream);
}

int lbm_adjoint_scatter(const lbm_step_desc* d, const lbm_adjoint_desc* ad, void* lam, const double* gF,
                        lbm_stream_t stream) {
  if (!d || !ad || !lam) return LBM_E_ARG;
  if (ad->n_targets <= 0) return LBM_OK;
  if (!ad->tgt_index || !ad->tgt_start || !ad->src || !ad->coef || !ad->mu_link) return LBM_E_ARG;
  const unsigned grid = unsigned((ad->n_targets + kScatterBlock - 1) / kScatterBlock);
  hipStream_t s = (hipStream_t)stream;
  if (ad->lam_storage == LBM_F16DEV) {
    if (d->storage == LBM_F64) return LBM_E_DTYPE;
    const float scale = ad->lam_scale > 0.0 ? float(ad->lam_scale) : 1.0f;
    hipLaunchKernelGGL((k_adjoint_scatter<float, LamHalf>), grid, kScatterBlock, 0, s, ad->n_targets, ad->tgt_index,
                       ad->tgt_start, ad->src, ad->coef, static_cast<const float*>(ad->mu_link), gF,
                       static_cast<unsigned short*>(lam), scale, 1.0f / scale);
  } else if (ad->lam_storage != 0) {
    return LBM_E_DTYPE;
  } else if (d->storage == LBM_F64) {
    hipLaunchKernelGGL((k_adjoint_scatter<double, LamPlain<double>>), grid, kScatterBlock, 0, s, ad->n_targets, ad->tgt_index,
                       ad->tgt_start, ad->src, ad->coef, static_cast<const double*>(ad->mu_link), gF,
                       static_cast<double*>(lam), 1.0, 1.0);
  } else {
    hipLaunchKernelGGL((k_adjoint_scatter<float, LamPlain<float>>), grid, kScatterBlock, 0, s, ad->n_targets, ad->tgt_index,
                       ad->tgt_start, ad->src, ad->coef, static_cast<const float*>(ad->mu_link), gF,
                       static_cast<float*>(lam), 1.0f, 1.0f);
  }
  return hipGetLastError() == hipSuccess ? LBM_OK : LBM_E_LAUNCH;
}


int lbm_adjoint_chain(const lbm_step_desc* d, const lbm_adjoint_desc* ad, int n, const void* const* states,
                      void* lam_a, void* lam_b, const double* gF, lbm_stream_t stream) {
  if (!d || !ad || n < 0 || (n > 0 && (!states || !lam_a || !lam_b)) || lam_a == lam_b) return LBM_E_ARG;
  if (ad->fold_links && d->n_links > 0 && (!ad->mu_link || !ad->mu_link_alt || ad->mu_link == ad->mu_link_alt))
    return LBM_E_ARG;
  hipStream_t st = (hipStream_t)stream;
  void* cur = lam_a;
  void* nxt = lam_b;
  // folded: each launch adds the previous launch's link additions to what it
  // loads and stores its own in the other mu buffer; the chain's last launch
  // is completed by lbm_adjoint_link_finish
  void* mu[2] = {ad->mu_link, ad->mu_link_alt};
  for (int k = n - 1; k >= 0; --k) {
    const double* g = gF ? gF + 3 * size_t(k) : nullptr;
    int rc;
    if (ad->fold_links) {
      const int w = (n - 1 - k) & 1;
      const bool first = (k == n - 1);
      rc = adj_dispatch<AdjointLaunch>(d, ad, int(LBM_MODE_FULL), states[k], static_cast<const void*>(cur), nxt, g, st,
                                       mu[w], first ? static_cast<const void*>(nullptr) : mu[w ^ 1]);
      if (rc) return rc;
      if (k == 0) {
        rc = lbm_adjoint_link_finish(d, ad, mu[w], nxt, stream);
        if (rc) return rc;
      }
    } else {
      rc = lbm_step_adjoint(d, ad, LBM_MODE_FULL, states[k], cur, nxt, g, stream);
      if (rc) return rc;
      rc = lbm_adjoint_scatter(d, ad, nxt, g, stream);
      if (rc) return rc;
    }
    void* t = cur;
    cur = nxt;
    nxt = t;
  }
  return LBM_OK;
}

}  // extern "C"
