// Fused D3Q19 timestep for gfx950.
//
// The carried state is the post-collision, post-bounce population P_t (the
// reference's f_post).  One step computes, for every cell x,
//   S_{t+1}(x) = guard(ZouHe(pull(P_t) with Bouzidi links))
//   P_{t+1}(x) = bounce(collide(S_{t+1}(x)))
// reading each population of P_t once (pull streaming is a bijection) and
// writing P_{t+1} once: 2 x 19 values per cell.
//
// One launch per step.  Its first workgroups own the sparse link cells
// (fluid cells with a Bouzidi link, listed in link_cell/link_q), 64 per
// workgroup, and produce the float64 momentum-exchange partials; the rest
// sweep the dense grid in workgroup order and skip link cells.  Hardware
// dispatch deals consecutive workgroups to the 8 XCDs in turn, so at any
// moment all XCDs update neighbouring rows and the row-interleaved state is
// read and written in a narrow address window (remapping tiles so that each
// XCD sweeps its own contiguous eighth of the grid was 6% slower).  A link
// workgroup spreads its cells' Bouzidi pulls over its waves by direction and
// then collides on one wave (link_block), so on small grids, where the step
// waits for the link cells, a link cell costs two dependent round trips and
// a fraction of the address work before its collision.  With 16-bit
// populations the sweep is bound by vector instructions rather than memory,
// and where nx is a multiple of 64 it takes each wave's row addresses from
// scalar registers (dense_row).
#include <mutex>

#include "lbm_physics.hpp"

namespace lbm {

// Threads per workgroup of a step launch.  256 for every sweep: the float32
// sweep of 256 x 128 x 128 runs 2% faster than with 512 (108.6 against
// 110.6 us per step, 384 x 192 x 192 380 against 384 us) since the force
// reduction left the link workgroups' collision path, the 16-bit sweep is 7%
// slower at 512, and small grids lose parallelism with larger workgroups.
constexpr int kStepBlock = 256;

template <typename R>
struct StepArgs {
  int nx, ny, nz;
  uint32_t N;   // cells
  uint32_t P;   // elements between consecutive populations of a state buffer
  uint32_t RS;  // elements between consecutive (y, z) rows of a state buffer
  FastDiv dnx, dny;
  int coll, smag, bounce, guard, periodic, n_links, nlb, ndb;
  uint32_t tgroup;      // dense tiles per XCD group (xcd_group_tile)
  int row_waves;        // nx is a multiple of 64: each wave's cells share one (y, z) row
  uint32_t state_bytes;  // extent of a state buffer in bytes (below 2^32 when row_waves)
  int sround;            // 16-bit storage rounds stochastically (dither hash seeded by `seed`)
  uint32_t seed;         // this launch's dither seed (from the call's seed, step and mode)
  R tau, omega, cs;
  R uin[3];
  R rho_out;
  R uref[3];
  R off[Q];
  const uint8_t* flags;
  const R* phi;
  const int32_t* link_cell;
  const double* link_q;
};

template <typename R>
__device__ __forceinline__ void cell_coords(const StepArgs<R>& a, uint32_t c, int& x, int& y, int& z) {
  const uint32_t yz = fdiv(c, a.dnx);
  x = int(c - yz * uint32_t(a.nx));
  const uint32_t zz = fdiv(yz, a.dny);
  y = int(yz - zz * uint32_t(a.ny));
  z = int(zz);
}

// Element offset of cell (x, y, z) within a population: rows are RS
// elements apart (RS = nx when the populations are whole arrays apart, 19 nx
// or more when the populations of a row are interleaved).
template <typename R>
__device__ __forceinline__ uint32_t cell_addr(const StepArgs<R>& a, int x, int y, int z) {
  return (uint32_t(z) * uint32_t(a.ny) + uint32_t(y)) * a.RS + uint32_t(x);
}
template <typename R>
__device__ __forceinline__ uint32_t cell_addr(const StepArgs<R>& a, uint32_t c) {
  if (a.RS == uint32_t(a.nx)) return c;
  return c + fdiv(c, a.dnx) * (a.RS - uint32_t(a.nx));
}

// Neighbour rows of a cell, wrapped (periodic) or flagged (clamped).
struct Rows {
  uint32_t ym, yp, zm, zp;  // element offsets of the rows y-1, y+1, z-1, z+1
  bool ym_ok, yp_ok, zm_ok, zp_ok;
};

template <typename R>
__device__ __forceinline__ Rows make_rows(const StepArgs<R>& a, int y, int z) {
  Rows r;
  const uint32_t nx = a.RS, plane = a.RS * uint32_t(a.ny);
  r.ym = (y == 0) ? (uint32_t(a.ny - 1) * nx) : uint32_t(0) - nx;
  r.yp = (y == a.ny - 1) ? uint32_t(0) - uint32_t(a.ny - 1) * nx : nx;
  r.zm = (z == 0) ? (uint32_t(a.nz - 1) * plane) : uint32_t(0) - plane;
  r.zp = (z == a.nz - 1) ? uint32_t(0) - uint32_t(a.nz - 1) * plane : plane;
  r.ym_ok = a.periodic || y > 0;
  r.yp_ok = a.periodic || y < a.ny - 1;
  r.zm_ok = a.periodic || z > 0;
  r.zp_ok = a.periodic || z < a.nz - 1;
  return r;
}

// Element offset of the source of the normal pull for direction I
// (streaming.py:113-142): x - e_I with y/z wrapped; the cell itself (c, its
// own offset) where the reference keeps the cell's own value (x faces, and
// y/z faces when not periodic).  Offsets are added modulo 2^32, so a wrapped
// row is a plain unsigned addition.
template <int I, typename R>
__device__ __forceinline__ uint32_t pull_src(const StepArgs<R>& a, const Rows& r, uint32_t c, int x) {
  constexpr int ex = EX[I], ey = EY[I], ez = EZ[I];
  if constexpr (ex == 1) {
    if (x == 0) return c;
  } else if constexpr (ex == -1) {
    if (x == a.nx - 1) return c;
  }
  uint32_t s = c - uint32_t(ex);
  if constexpr (ey == 1) {
    if (!r.ym_ok) return c;
    s += r.ym;
  } else if constexpr (ey == -1) {
    if (!r.yp_ok) return c;
    s += r.yp;
  }
  if constexpr (ez == 1) {
    if (!r.zm_ok) return c;
    s += r.zm;
  } else if constexpr (ez == -1) {
    if (!r.zp_ok) return c;
    s += r.zp;
  }
  return s;
}

// Offset of cell x + e_I wrapped periodically in all three axes
// (streaming.py:166-172).
template <int I, typename R>
__device__ __forceinline__ uint32_t wrap_next(const StepArgs<R>& a, int x, int y, int z) {
  int fx = x + EX[I], fy = y + EY[I], fz = z + EZ[I];
  fx = fx < 0 ? fx + a.nx : (fx >= a.nx ? fx - a.nx : fx);
  fy = fy < 0 ? fy + a.ny : (fy >= a.ny ? fy - a.ny : fy);
  fz = fz < 0 ? fz + a.nz : (fz >= a.nz ? fz - a.nz : fz);
  return cell_addr(a, fx, fy, fz);
}

// Zou-He planes and the post-collision guard on a pulled S (state.py:152-169).
template <bool RCP, typename R>
__device__ __forceinline__ void boundaries(const StepArgs<R>& a, R* s, int x) {
  if (x == 0) zou_he_in(s, a.uin[0], a.uin[1], a.uin[2], false, R(0));
  if (x == a.nx - 1) zou_he_out(s, a.rho_out);
  if (a.guard == LBM_GUARD_POST && guard_bad<RCP>(s)) set_weights(s);
}

// [pre guard] -> [Smagorinsky tau] -> BGK|MRT -> hard|soft bounce
// (examples/10:71-78, state.py:133-149, smooth.py:151-155).
// SOFT: 1 the caller knows the bounce is soft, 0 that it is not, -1 neither
// (the mode is read at run time).  A sweep compiled for one case keeps the
// other branch out of its code: a soft blend writing new registers would
// otherwise make every hard-mask cell copy its 19 values at the branch join.
template <bool RCP, int COLL, int SOFT = -1, typename R>
__device__ __forceinline__ void collide_cell(const StepArgs<R>& a, R* s, uint32_t c, uint32_t fl) {
  bool reset;
  relax<COLL == LBM_COLL_MRT, RCP>(s, a.tau, a.omega, a.guard == LBM_GUARD_PRE, a.uref[0], a.uref[1], a.uref[2],
                                   a.smag != 0, a.cs, reset);
  if (SOFT != 1 && a.bounce == LBM_BOUNCE_HARD) {
    if (fl & 1) reverse_all(s);
  } else if (SOFT != 0 && a.bounce == LBM_BOUNCE_SOFT) {
    blend_opposite(s, a.phi[c]);
  }
}

// Hash of cell c for this launch's stochastic rounding.
template <typename R>
__device__ __forceinline__ uint32_t cell_hash(const StepArgs<R>& a, uint32_t c) {
  return mix32(c ^ a.seed);
}

// s -> the 19 populations of cell c at cell offset o
template <class S>
__device__ __forceinline__ void store_cell(const StepArgs<typename S::R>& a, typename S::T* __restrict__ dst,
                                           uint32_t o, const typename S::R* s, uint32_t c) {
  const size_t N = a.P;
  if constexpr (is_half_storage<S>) {
    if (a.sround) {
      const uint32_t h = cell_hash(a, c);
      sfor<Q>([&](auto I) {
        constexpr int i = I;
        S::st_sr(dst, i * N + o, s[i], a.off[i], pop_bits<i>(h));
      });
      return;
    }
  }
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    S::st(dst, i * N + o, s[i], a.off[i]);
  });
}

// Dense sweep with V x-consecutive cells per thread (nx % V == 0, so the V
// cells share one row).  Every population is read with one aligned V-wide
// load from the source row; the one-cell x shift of the e_x = +-1 pulls is
// taken from the neighbouring lane's block (lane shuffle), with one
// predicated load at wave edges and the cell's own value at the x faces.
template <class S, int COLL, int MODE, int V>
__device__ __forceinline__ void dense_vec(const StepArgs<typename S::R>& a, const typename S::T* __restrict__ src,
                                          typename S::T* __restrict__ dst, uint32_t t) {
  using R = typename S::R;
  const size_t N = a.P;
  const uint32_t c0 = t * V;
  if (c0 >= a.N) return;
  int x0, y, z;
  cell_coords(a, c0, x0, y, z);
  const uint32_t o0 = cell_addr(a, x0, y, z);  // offset of cell c0 (its V cells are adjacent)
  const int lane = threadIdx.x & 63;
  const bool row_first = (x0 == 0), row_last = (x0 + V == a.nx);
  R s[V][Q];
  if constexpr (MODE == LBM_MODE_COLLIDE) {
    sfor<Q>([&](auto I) {
      constexpr int i = I;
      R blk[V];
      S::template ldv<V>(src, i * N + o0, a.off[i], blk);
#pragma unroll
      for (int j = 0; j < V; ++j) s[j][i] = blk[j];
    });
  } else {
    const Rows rows = make_rows(a, y, z);
    sfor<Q>([&](auto I) {
      constexpr int i = I;
      constexpr int ex = EX[i], ey = EY[i], ez = EZ[i];
      // aligned block of the source row (x not shifted), or the own row when
      // a clamped y/z face keeps the cell's own values
      uint32_t base = o0;
      bool row_ok = true;
      if constexpr (ey == 1) { row_ok = row_ok && rows.ym_ok; base += rows.ym; }
      else if constexpr (ey == -1) { row_ok = row_ok && rows.yp_ok; base += rows.yp; }
      if constexpr (ez == 1) { row_ok = row_ok && rows.zm_ok; base += rows.zm; }
      else if constexpr (ez == -1) { row_ok = row_ok && rows.zp_ok; base += rows.zp; }
      if (!row_ok) base = o0;
      R blk[V];
      S::template ldv<V>(src, i * N + base, a.off[i], blk);
      if constexpr (ex == 0) {
#pragma unroll
        for (int j = 0; j < V; ++j) s[j][i] = blk[j];
      } else if constexpr (ex == 1) {
        // cell x0 + j pulls x0 + j - 1
        R prev = __shfl_up(blk[V - 1], 1, 64);
        if (row_ok && !row_first && lane == 0) prev = S::ld_c(src, i * N + base - 1, a.off[i]);
        if (row_first && row_ok) prev = S::ld_c(src, i * N + o0, a.off[i]);  // x = 0 keeps its own value
        s[0][i] = row_ok ? prev : blk[0];
#pragma unroll
        for (int j = 1; j < V; ++j) s[j][i] = row_ok ? blk[j - 1] : blk[j];
      } else {
        // cell x0 + j pulls x0 + j + 1
        R next = __shfl_down(blk[0], 1, 64);
        if (row_ok && !row_last && lane == 63) next = S::ld_c(src, i * N + base + V, a.off[i]);
        if (row_last && row_ok) next = S::ld_c(src, i * N + o0 + V - 1, a.off[i]);  // x = nx-1 keeps its own
#pragma unroll
        for (int j = 0; j < V - 1; ++j) s[j][i] = row_ok ? blk[j + 1] : blk[j];
        s[V - 1][i] = row_ok ? next : blk[V - 1];
      }
    });
  }
  // per-cell flags (bit0 solid, bit1 link cell owned by a link workgroup)
  uint32_t fl = 0;
  if (a.flags) {
    if constexpr (V == 4) fl = *reinterpret_cast<const uint32_t*>(a.flags + c0);
    else if constexpr (V == 2) fl = *reinterpret_cast<const uint16_t*>(a.flags + c0);
    else fl = a.flags[c0];
  }
  R phi[V];
  if (a.bounce == LBM_BOUNCE_SOFT) {
#pragma unroll
    for (int j = 0; j < V; ++j) phi[j] = a.phi[c0 + j];
  }
#pragma unroll
  for (int j = 0; j < V; ++j) {
    const uint32_t fj = (fl >> (8 * j)) & 0xffu;
    if constexpr (MODE != LBM_MODE_COLLIDE) boundaries<rcp_of<S>>(a, s[j], x0 + j);
    if constexpr (MODE != LBM_MODE_STREAM) {
      bool reset;
      relax<COLL == LBM_COLL_MRT, rcp_of<S>>(s[j], a.tau, a.omega, a.guard == LBM_GUARD_PRE, a.uref[0], a.uref[1],
                                             a.uref[2], a.smag != 0, a.cs, reset);
      if (a.bounce == LBM_BOUNCE_HARD) {
        if (fj & 1u) reverse_all(s[j]);
      } else if (a.bounce == LBM_BOUNCE_SOFT) {
        blend_opposite(s[j], phi[j]);
      }
    }
  }
  // link cells are written by the link workgroups (dense MODE_COLLIDE writes all)
  const bool has_link = (MODE != LBM_MODE_COLLIDE) && (fl & 0x02020202u);
  if constexpr (is_half_storage<S>) {
    if (a.sround) {
#pragma unroll
      for (int j = 0; j < V; ++j) {
        if (MODE != LBM_MODE_COLLIDE && ((fl >> (8 * j)) & 2u)) continue;
        const uint32_t h = cell_hash(a, c0 + uint32_t(j));
        sfor<Q>([&](auto I) {
          constexpr int i = I;
          S::st_sr(dst, i * N + o0 + j, s[j][i], a.off[i], pop_bits<i>(h));
        });
      }
      return;
    }
  }
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    R o[V];
#pragma unroll
    for (int j = 0; j < V; ++j) o[j] = s[j][i];
    if (!has_link) {
      S::template stv<V>(dst, i * N + o0, o, a.off[i]);
    } else {
#pragma unroll
      for (int j = 0; j < V; ++j)
        if (!((fl >> (8 * j)) & 2u)) S::st(dst, i * N + o0 + j, o[j], a.off[i]);
    }
  });
}

// Dense sweep at one cell per thread when every wave's 64 cells share one
// (y, z) row (nx a multiple of 64).  The row, its neighbour rows and the base
// address of every pulled and stored population are then wave-uniform and
// live in scalar registers (buffer resource + scalar byte offset); a lane
// holds three byte offsets (x, and x - 1 and x + 1 wrapped), so no pull or
// store needs vector address arithmetic.  Used when a state buffer spans
// less than 4 GiB.  At
// x = 0 the e_x = +1 pulls and at x = nx - 1 the e_x = -1 pulls read the
// wrapped neighbour instead of the cell's own value: Zou-He then overwrites
// exactly those populations from the others (boundary.py:113-122, 190-198),
// so the step's result is the same.  PER: periodic y/z (no clamped rows);
// SOFT as for collide_cell.
template <class S, int COLL, int MODE, bool PER, int SOFT>
__device__ __forceinline__ void dense_row(const StepArgs<typename S::R>& a, const typename S::T* __restrict__ src,
                                          typename S::T* __restrict__ dst, uint32_t c) {
  using R = typename S::R;
  using T = typename S::T;
  constexpr uint32_t B = sizeof(T);
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t cw = __builtin_amdgcn_readfirstlane(c - lane);  // first cell of the wave
  const uint32_t yz = fdiv(cw, a.dnx);
  const uint32_t zz = fdiv(yz, a.dny);
  const int y = int(yz - zz * uint32_t(a.ny)), z = int(zz);
  const int x = int(cw - yz * uint32_t(a.nx) + lane);
  const uint32_t bx0 = uint32_t(x) * B;
  const uint32_t bxm = uint32_t(x == 0 ? a.nx - 1 : x - 1) * B;
  const uint32_t bxp = uint32_t(x == a.nx - 1 ? 0 : x + 1) * B;
  const uint32_t own = yz * a.RS;
  const Rows rows = make_rows(a, y, z);
  const uint32_t fl = a.flags ? uint32_t((a.flags + cw)[lane]) : 0u;
  const __amdgpu_buffer_rsrc_t rin = buffer_rsrc(src, a.state_bytes);
  R s[Q];
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    constexpr int ex = EX[i], ey = EY[i], ez = EZ[i];
    uint32_t row = own;
    bool ok = true;
    if constexpr (ey == 1) { ok = ok && rows.ym_ok; row += rows.ym; }
    else if constexpr (ey == -1) { ok = ok && rows.yp_ok; row += rows.yp; }
    if constexpr (ez == 1) { ok = ok && rows.zm_ok; row += rows.zm; }
    else if constexpr (ez == -1) { ok = ok && rows.zp_ok; row += rows.zp; }
    uint32_t bx = ex == 1 ? bxm : (ex == -1 ? bxp : bx0);
    if constexpr (!PER) {
      // a clamped y/z face keeps the cell's own value (streaming.py:134-142)
      if (!ok) {
        row = own;
        bx = bx0;
      }
    }
    s[i] = S::bld(rin, bx, (uint32_t(i) * a.P + row) * B, a.off[i]);
  });
  boundaries<rcp_of<S>>(a, s, x);
  if (MODE == LBM_MODE_FULL) collide_cell<rcp_of<S>, COLL, SOFT>(a, s, c, fl);
  if (fl & 2u) return;  // owned by a link workgroup
  const __amdgpu_buffer_rsrc_t rout = buffer_rsrc(dst, a.state_bytes);
  if constexpr (is_half_storage<S>) {
    if (a.sround) {
      const uint32_t h = cell_hash(a, c);
      sfor<Q>([&](auto I) {
        constexpr int i = I;
        S::bst_sr(rout, bx0, (uint32_t(i) * a.P + own) * B, s[i], a.off[i], pop_bits<i>(h));
      });
      return;
    }
  }
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    S::bst(rout, bx0, (uint32_t(i) * a.P + own) * B, s[i], a.off[i]);
  });
}

// Register budget of the step kernel at one cell per thread: at most 80 VGPRs
// (at least 6 waves per SIMD) with float arithmetic, 168 (3 waves) with FP64.
// The float32 step takes 61 VGPRs and runs 8 waves per SIMD; the 16-bit step
// 61 VGPRs and 103 SGPRs (its scalar row offsets), so 7 waves.  Asking for 8
// fits it in 78 SGPRs with lane spills and changed nothing (57.0 us either
// way at 256 x 128 x 128).
template <class S, int V>
constexpr int step_waves_per_eu() {
  return V > 1 ? 1 : (sizeof(typename S::R) == 8 ? 3 : 6);
}

// Phase 1 of a link workgroup, on the wave that owns the directions i with
// i % NW == W: for its link cell (streaming.py:113-181) it loads q_opp(i),
// then at once the cell's own P_opp(i) and the second value (the Bouzidi
// neighbour P(x + e_i)_opp below q = 1/2, the cell's own P_i above, the
// normal pull where opp(i) is no wall link), forms S_i into LDS, adds the own
// part e_opp(i) P_opp(i) of the momentum exchange (forces.py:95-118) to f and
// sets bit opp(i) of mask where opp(i) is a wall link.  With `pulled` (a cell
// that neither Zou-He nor the post guard changes) it also adds the pulled
// part e_opp(i) S_i.  Straight-line code over the wave's few directions, so
// its loads go out in two rounds (q, then the populations).
template <class S, int W, int NW>
__device__ __forceinline__ void link_pull_part(const StepArgs<typename S::R>& a, const typename S::T* __restrict__ src,
                                               int b, int x, int y, int z, uint32_t o, const Rows& rows,
                                               typename S::R (*sl)[kLinkCellsPerBlock], int lane, bool pulled,
                                               double* f, uint32_t& mask) {
  using R = typename S::R;
  constexpr int ND = (Q - W + NW - 1) / NW;  // directions W, W + NW, ...
  const size_t N = a.P;
  R qo[ND], va[ND], vb[ND];
  sfor<ND>([&](auto K) {
    constexpr int i = W + int(K) * NW;
    if constexpr (i != 0) qo[K] = R(a.link_q[size_t(opp(i) - 1) * a.n_links + b]);
  });
  sfor<ND>([&](auto K) {
    constexpr int i = W + int(K) * NW;
    if constexpr (i == 0) {
      va[K] = S::ld_c(src, o, a.off[0]);
    } else {
      constexpr int io = opp(i);
      const bool low = (qo[K] >= R(0)) & (qo[K] < R(0.5));
      const bool high = qo[K] >= R(0.5);
      const size_t ka = low ? io * N + wrap_next<i>(a, x, y, z)
                            : (high ? i * N + o : i * N + pull_src<i>(a, rows, o, x));
      va[K] = S::ld_cv(src, ka, low ? a.off[io] : a.off[i]);
      vb[K] = S::ld_c(src, io * N + o, a.off[io]);
    }
  });
  sfor<ND>([&](auto K) {
    constexpr int i = W + int(K) * NW;
    if constexpr (i == 0) {
      sl[0][lane] = va[K];
    } else {
      constexpr int io = opp(i);
      const R q = qo[K];
      R si;
      if (q >= R(0.5)) {
        const R inv2q = R(1) / (R(2) * fmin(q, R(1)));
        si = inv2q * vb[K] + (R(1) - inv2q) * va[K];
      } else if (q >= R(0)) {
        const R twoq = R(2) * q;
        si = twoq * vb[K] + (R(1) - twoq) * va[K];
      } else {
        si = va[K];
      }
      sl[i][lane] = si;
      if (q >= R(0)) {
        mask |= 1u << io;
        const double t = double(vb[K]);
        add_scaled<EX[io]>(f[0], t);
        add_scaled<EY[io]>(f[1], t);
        add_scaled<EZ[io]>(f[2], t);
        if (pulled) {
          const double u = double(si);
          add_scaled<EX[io]>(f[0], u);
          add_scaled<EY[io]>(f[1], u);
          add_scaled<EZ[io]>(f[2], u);
        }
      }
    }
  });
}

// A link workgroup: kLinkCellsPerBlock link cells, lane l of every wave on
// link b = first + l.  Phase 1 (all waves): the Bouzidi pulls and the
// momentum exchange, spread over the waves by direction (link_pull_part).
// Phase 2 (wave 0): per cell, Zou-He, the pulled part e_j S_opp(j) of the
// momentum exchange where Zou-He or the post guard changed S, collision and
// store.  The workgroup's force partial (float64) goes to its slot of
// fpart.  A link cell so waits two round trips before its collision, and no
// wave runs all 18 directions' address and branch work.
template <class S, int COLL, int MODE, int BS>
__device__ __forceinline__ void link_block(const StepArgs<typename S::R>& a, const typename S::T* __restrict__ src,
                                           typename S::T* __restrict__ dst, uint32_t bid,
                                           double* __restrict__ fpart) {
  using R = typename S::R;
  constexpr int NW = BS / 64;
  static_assert(kLinkCellsPerBlock == 64, "one link cell per lane");
  __shared__ R sl[Q][kLinkCellsPerBlock];
  __shared__ uint32_t lm[NW][kLinkCellsPerBlock];
  __shared__ double red[NW][3];
  __shared__ double fw[3][kLinkCellsPerBlock], gw[3][kLinkCellsPerBlock];
  const int lane = int(threadIdx.x & 63u);
  const int w = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
  const int b = int(bid) * kLinkCellsPerBlock + lane;
  const bool live = b < a.n_links;
  double f[3] = {0.0, 0.0, 0.0};
  uint32_t mask = 0, c = 0, fl = 0, o = 0;
  int x = 0, y = 0, z = 0;
  if (live) {
    c = uint32_t(a.link_cell[b]);
    cell_coords(a, c, x, y, z);
    o = cell_addr(a, x, y, z);
    if (w == 0 && MODE == LBM_MODE_FULL) fl = a.flags[c];
    const Rows rows = make_rows(a, y, z);
    const bool pulled = x != 0 && x != a.nx - 1 && a.guard != LBM_GUARD_POST;
    sfor<NW>([&](auto W) {
      if (w == int(W)) link_pull_part<S, int(W), NW>(a, src, b, x, y, z, o, rows, sl, lane, pulled, f, mask);
    });
  }
  lm[w][lane] = mask;
  __syncthreads();
  // The float64 force sums stay off the collision's path: waves 1.. reduce
  // their partials while wave 0 collides, and wave 0 parks its per-lane
  // partials in LDS and reduces them after its stores are issued.  With the
  // DPP reduction (wave_sum_dpp) this took 64 x 32 x 32 from 6.25 to
  // 5.9 us per step (5.75 us without forces).
  if (w != 0) {
    if (fpart) {
      const double fx = wave_sum_dpp(f[0]), fy = wave_sum_dpp(f[1]), fz = wave_sum_dpp(f[2]);
      if (lane == 0) {
        red[w][0] = fx;
        red[w][1] = fy;
        red[w][2] = fz;
      }
      __syncthreads();
    }
    return;
  }
  if (fpart) {
    fw[0][lane] = f[0];
    fw[1][lane] = f[1];
    fw[2][lane] = f[2];
  }
  R s[Q];
  double g[3] = {0.0, 0.0, 0.0};
  const bool face = live && (x == 0 || x == a.nx - 1 || a.guard == LBM_GUARD_POST);
  if (live) {
    sfor<Q>([&](auto I) {
      constexpr int i = I;
      s[i] = sl[i][lane];
    });
    boundaries<rcp_of<S>>(a, s, x);
  }
  if (face) {
#pragma unroll
    for (int k = 1; k < NW; ++k) mask |= lm[k][lane];
    sfor<Q>([&](auto J) {
      constexpr int j = J;
      if constexpr (j != 0) {
        if (mask & (1u << j)) {
          const double t = double(s[opp(j)]);
          add_scaled<EX[j]>(g[0], t);
          add_scaled<EY[j]>(g[1], t);
          add_scaled<EZ[j]>(g[2], t);
        }
      }
    });
  }
  // the pulled part of the x-face cells waits in LDS too: the collision
  // needs every register the dense budget leaves
  const bool any_face = __any(face);
  if (fpart && any_face) {
    gw[0][lane] = g[0];
    gw[1][lane] = g[1];
    gw[2][lane] = g[2];
  }
  if (live) {
    if (MODE == LBM_MODE_FULL) collide_cell<rcp_of<S>, COLL>(a, s, c, fl);
    store_cell<S>(a, dst, o, s, c);
  }
  if (fpart) {
    const double fx = wave_sum_dpp(fw[0][lane]), fy = wave_sum_dpp(fw[1][lane]), fz = wave_sum_dpp(fw[2][lane]);
    if (lane == 0) {
      red[0][0] = fx;
      red[0][1] = fy;
      red[0][2] = fz;
    }
    if (any_face) {
      const double gx = wave_sum_dpp(gw[0][lane]), gy = wave_sum_dpp(gw[1][lane]), gz = wave_sum_dpp(gw[2][lane]);
      if (lane == 0) {
        red[0][0] += gx;
        red[0][1] += gy;
        red[0][2] += gz;
      }
    }
    __syncthreads();
    if (lane < 3) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < NW; ++k) acc += red[k][lane];
      fpart[size_t(bid) * 3 + lane] = acc;
    }
  }
}

template <class S, int COLL, int MODE, int V, int BS = kStepBlock>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(step_waves_per_eu<S, V>()))) void k_step(
    StepArgs<typename S::R> a, const typename S::T* __restrict__ src, typename S::T* __restrict__ dst,
    double* __restrict__ fpart) {
  using R = typename S::R;
  const size_t N = a.P;
  const uint32_t bid = blockIdx.x;

  if (MODE != LBM_MODE_COLLIDE && int(bid) < a.nlb) {
    link_block<S, COLL, MODE, BS>(a, src, dst, bid, fpart);
    return;
  }

  // ---------------- dense sweep ----------------
  const uint32_t first = (MODE == LBM_MODE_COLLIDE) ? 0u : uint32_t(a.nlb);
  const uint32_t tile = xcd_group_tile(bid - first, uint32_t(a.ndb), a.tgroup);
  if constexpr (V > 1) {
    dense_vec<S, COLL, MODE, V>(a, src, dst, tile * BS + threadIdx.x);
    return;
  }
  const uint32_t c = tile * BS + threadIdx.x;
  if (c >= a.N) return;
  if constexpr (V == 0) {
    if (!a.periodic) dense_row<S, COLL, MODE, false, -1>(a, src, dst, c);
    else if (a.bounce == LBM_BOUNCE_SOFT) dense_row<S, COLL, MODE, true, 1>(a, src, dst, c);
    else dense_row<S, COLL, MODE, true, 0>(a, src, dst, c);
    return;
  }
  // The flags are loaded together with the populations; testing the
  // link-cell bit before the pulls would put a whole memory round trip in
  // front of them.  Link cells are computed and then not stored.
  const uint32_t fl = a.flags ? uint32_t(a.flags[c]) : 0u;
  R s[Q];
  uint32_t o;
  if (MODE == LBM_MODE_COLLIDE) {
    o = cell_addr(a, c);
    sfor<Q>([&](auto I) {
      constexpr int i = I;
      s[i] = S::ld(src, i * N + o, a.off[i]);
    });
    collide_cell<rcp_of<S>, COLL>(a, s, c, fl);
  } else {
    int x, y, z;
    cell_coords(a, c, x, y, z);
    o = cell_addr(a, x, y, z);
    const Rows rows = make_rows(a, y, z);
    sfor<Q>([&](auto I) {
      constexpr int i = I;
      s[i] = S::ld_c(src, i * N + pull_src<i>(a, rows, o, x), a.off[i]);
    });
    boundaries<rcp_of<S>>(a, s, x);
    if (MODE == LBM_MODE_FULL) collide_cell<rcp_of<S>, COLL>(a, s, c, fl);
    if (fl & 2u) return;  // owned by a link workgroup
  }
  store_cell<S>(a, dst, o, s, c);
}

__global__ void k_force_reduce(const double* __restrict__ part, int T, int slots, double* __restrict__ force) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  double fx = 0, fy = 0, fz = 0;
  const double* p = part + size_t(t) * slots * 3;
  for (int s = 0; s < slots; ++s) {
    fx += p[3 * s];
    fy += p[3 * s + 1];
    fz += p[3 * s + 2];
  }
  force[3 * t] = fx;
  force[3 * t + 1] = fy;
  force[3 * t + 2] = fz;
}

// One population per grid row (blockIdx.y), so the offset a half format is
// encoded against is uniform over the workgroup, as in the sweeps.
template <class SI, class SO>
__global__ void k_convert(const typename SI::T* __restrict__ in, typename SO::T* __restrict__ out,
                          StepArgs<float> a) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.N) return;
  const int i = int(blockIdx.y);
  const size_t e = size_t(i) * a.P + cell_addr(a, c);
  using RI = typename SI::R;
  using RO = typename SO::R;
  const float off = a.off[i];
  const RI v = SI::ld_c(in, e, RI(off));
  if constexpr (is_half_storage<SO>) {
    if (a.sround) {
      const uint32_t h = cell_hash(a, c);
      const uint32_t sh = uint32_t(5 * i) & 31u;
      SO::st_sr(out, e, RO(v), RO(off), sh ? __builtin_amdgcn_alignbit(h, h, sh) : h);
      return;
    }
  }
  SO::st(out, e, RO(v), RO(off));
}

template <typename R>
int make_args(const lbm_step_desc* d, int mode, StepArgs<R>& a) {
  if (!d || d->nx <= 0 || d->ny <= 0 || d->nz <= 0) return LBM_E_ARG;
  const long long n = 1LL * d->nx * d->ny * d->nz;
  long long pstride = 0, rstride = 0;
  if (!state_layout(d, pstride, rstride)) return LBM_E_ARG;
  a.P = uint32_t(pstride);
  a.RS = uint32_t(rstride);
  a.nx = d->nx;
  a.ny = d->ny;
  a.nz = d->nz;
  a.N = uint32_t(n);
  a.dnx = make_fastdiv(uint32_t(d->nx));
  a.dny = make_fastdiv(uint32_t(d->ny));
  a.coll = d->collision;
  a.smag = d->smagorinsky;
  a.bounce = d->bounce;
  a.guard = d->guard;
  a.periodic = d->periodic_yz;
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
  if (a.bounce == LBM_BOUNCE_HARD && !a.flags) return LBM_E_ARG;
  if (a.bounce == LBM_BOUNCE_SOFT && !a.phi) return LBM_E_ARG;
  if (a.n_links < 0 || (a.n_links > 0 && (!a.link_cell || !a.link_q || !a.flags))) return LBM_E_ARG;
  if (a.coll != LBM_COLL_BGK && a.coll != LBM_COLL_MRT) return LBM_E_ARG;
  if (a.guard < 0 || a.guard > 2 || a.bounce < 0 || a.bounce > 2) return LBM_E_ARG;
  a.nlb = (mode == LBM_MODE_COLLIDE) ? 0 : (a.n_links + kLinkCellsPerBlock - 1) / kLinkCellsPerBlock;
  a.ndb = int((n + kStepBlock - 1) / kStepBlock);
  a.tgroup = 1u;
  a.row_waves = 0;
  a.state_bytes = 0;
  a.sround = 0;
  a.seed = 0;
  if (d->rounding != LBM_ROUND_NEAREST && d->rounding != LBM_ROUND_STOCHASTIC) return LBM_E_ARG;
  return LBM_OK;
}

// Dither seed of one launch: a hash of the call's seed, the step index and
// the mode, so a recomputed step draws the same variates.
inline uint32_t launch_seed(uint32_t seed, int64_t step, int mode) {
  const uint64_t t = uint64_t(step);
  return mix32(seed ^ mix32(uint32_t(t) * 4u + uint32_t(mode)) ^ mix32(uint32_t(t >> 32) + 0x9e3779b9u));
}

// Cells per thread of the dense sweep: 2 or 4 when requested and nx allows,
// else 1 (the fastest on gfx950: the sweep is arithmetic-bound and wider
// threads only raise register pressure).
inline int choose_vec(const lbm_step_desc* d, uint32_t pstride, uint32_t rstride) {
  const int want = d->cells_per_thread;
  if (want == 2 || want == 4)
    return (d->nx % want == 0 && pstride % uint32_t(want) == 0 && rstride % uint32_t(want) == 0) ? want : 1;
  return 1;
}

template <class S, int COLL, int MODE>
int launch_step_cfg(const StepArgs<typename S::R>& a, int V, const typename S::T* in, typename S::T* out,
                    double* fpart, hipStream_t s) {
  const unsigned grid = unsigned(a.nlb + a.ndb);
  if (V == 4) hipLaunchKernelGGL((k_step<S, COLL, MODE, 4>), grid, kStepBlock, 0, s, a, in, out, fpart);
  else if (V == 2) hipLaunchKernelGGL((k_step<S, COLL, MODE, 2>), grid, kStepBlock, 0, s, a, in, out, fpart);
  else if (MODE != LBM_MODE_COLLIDE && a.row_waves)
    hipLaunchKernelGGL((k_step<S, COLL, MODE, 0>), grid, kStepBlock, 0, s, a, in, out, fpart);
  else hipLaunchKernelGGL((k_step<S, COLL, MODE, 1>), grid, kStepBlock, 0, s, a, in, out, fpart);
  return hipGetLastError() == hipSuccess ? LBM_OK : LBM_E_LAUNCH;
}

template <class S>
int launch_step(const lbm_step_desc* d, int mode, const void* src, void* dst, double* fpart, hipStream_t s,
                int64_t step) {
  using R = typename S::R;
  StepArgs<R> a;
  const int rc = make_args(d, mode, a);
  if (rc) return rc;
  if (is_half_storage<S> && d->rounding == LBM_ROUND_STOCHASTIC) {
    a.sround = 1;
    a.seed = launch_seed(d->seed, step, mode);
  }
  if (!src || !dst || src == dst) return LBM_E_ARG;
  const int V = choose_vec(d, a.P, a.RS);
  {
    // The row-uniform sweep (32-bit byte offsets) for 2-byte populations,
    // whose per-lane-addressed sweep was bound by vector instructions:
    // 62.2 -> 57.4 us per step at 256 x 128 x 128.  The float32 sweep is bound by memory and runs
    // 0.5% faster on the per-lane addresses, so it keeps them.
    const long long extent = 1LL * (Q - 1) * a.P + (1LL * a.ny * a.nz - 1) * a.RS + a.nx;
    const long long bytes = extent * (long long)sizeof(typename S::T);
    a.row_waves = (d->cells_per_thread == 0 && sizeof(typename S::T) == 2 && a.nx % 64 == 0 &&
                   bytes < (1LL << 32)) ? 1 : 0;
    a.state_bytes = a.row_waves ? uint32_t(bytes) : 0u;
  }
  a.nlb = (mode == LBM_MODE_COLLIDE) ? 0 : (a.n_links + kLinkCellsPerBlock - 1) / kLinkCellsPerBlock;
  const long long cells_per_block = 1LL * kStepBlock * V;
  a.ndb = int((1LL * a.N + cells_per_block - 1) / cells_per_block);
  a.tgroup = tile_group_size(uint32_t(a.nx), uint32_t(cells_per_block));
  const auto* in = static_cast<const typename S::T*>(src);
  auto* out = static_cast<typename S::T*>(dst);
  const bool mrt = a.coll == LBM_COLL_MRT;
  switch (mode) {
    case LBM_MODE_FULL:
      return mrt ? launch_step_cfg<S, LBM_COLL_MRT, LBM_MODE_FULL>(a, V, in, out, fpart, s)
                 : launch_step_cfg<S, LBM_COLL_BGK, LBM_MODE_FULL>(a, V, in, out, fpart, s);
    case LBM_MODE_COLLIDE:
      return mrt ? launch_step_cfg<S, LBM_COLL_MRT, LBM_MODE_COLLIDE>(a, V, in, out, fpart, s)
                 : launch_step_cfg<S, LBM_COLL_BGK, LBM_MODE_COLLIDE>(a, V, in, out, fpart, s);
    case LBM_MODE_STREAM:
      return mrt ? launch_step_cfg<S, LBM_COLL_MRT, LBM_MODE_STREAM>(a, V, in, out, fpart, s)
                 : launch_step_cfg<S, LBM_COLL_BGK, LBM_MODE_STREAM>(a, V, in, out, fpart, s);
  }
  return LBM_E_ARG;
}

// Force partials of lbm_step_chain_forces_noalloc: a file-scope device array
// (one copy per GPU, alive until the process ends) so that entry point
// allocates nothing.  Host code never passes `scratch` itself to a kernel:
// in host code the name is a stand-in, not a device address; the address of
// the current device's copy comes from hipGetSymbolAddress (noalloc_scratch).
constexpr size_t kNoAllocPartials = size_t(1) << 20;  // doubles (8 MiB)
static __device__ __align__(256) double scratch[kNoAllocPartials];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mutex;
static double* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's `scratch`, looked up on the first
// call on each device (with that device current) and kept.
inline int noalloc_scratch(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return int(e);
  if (dev < 0 || dev >= kMaxDevices) return LBM_E_ARG;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return int(e);
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return LBM_OK;
}

}  // namespace lbm

using namespace lbm;

extern "C" {

int64_t lbm_state_stride(const lbm_step_desc* d) {
  if (!d || d->nx <= 0 || d->ny <= 0 || d->nz <= 0) return 0;
  const int64_t n = int64_t(d->nx) * d->ny * d->nz;
  // N rounded up to a multiple of 8192 elements, plus 1024: consecutive
  // populations of one cell then sit 1024 elements apart modulo 8192 instead
  // of at a multiple of a large power of two, which spreads the 19 read and
  // 19 write streams over the HBM channels (measured: 5.3 -> 6.3 TB/s for
  // the 19-in/19-out float pattern at 256 x 128 x 128).
  return (n + 8191) / 8192 * 8192 + 1024;
}

int lbm_step_force_slots(const lbm_step_desc* d) {
  if (!d || d->n_links < 0) return 0;
  return (d->n_links + kLinkCellsPerBlock - 1) / kLinkCellsPerBlock;
}

}  // extern "C"

// One launch of step `step` (the dither index of stochastic rounding).
static int step_at(const lbm_step_desc* d, int mode, const void* src, void* dst, double* force_partial,
                   hipStream_t s, int64_t step) {
  switch (d->storage) {
    case LBM_F32: return launch_step<StoreF32>(d, mode, src, dst, force_partial, s, step);
    case LBM_F64: return launch_step<StoreF64>(d, mode, src, dst, force_partial, s, step);
    case LBM_F16DEV: return launch_step<StoreF16>(d, mode, src, dst, force_partial, s, step);
    case LBM_BF16DEV: return launch_step<StoreBF16>(d, mode, src, dst, force_partial, s, step);
  }
  return LBM_E_DTYPE;
}

extern "C" {

int lbm_step(const lbm_step_desc* d, int mode, const void* src, void* dst, double* force_partial,
             lbm_stream_t stream) {
  if (!d) return LBM_E_ARG;
  return step_at(d, mode, src, dst, force_partial, (hipStream_t)stream, d->step);
}

int lbm_step_chain(const lbm_step_desc* d, int n, void* const* states, double* force_partial,
                   int64_t partial_stride, lbm_stream_t stream) {
  if (!d || n < 0 || (n > 0 && !states) || partial_stride < 0) return LBM_E_ARG;
  for (int k = 0; k < n; ++k) {
    double* fp = force_partial ? force_partial + size_t(k) * size_t(partial_stride) : nullptr;
    const int rc = step_at(d, LBM_MODE_FULL, states[k], states[k + 1], fp, (hipStream_t)stream, d->step + k);
    if (rc) return rc;
  }
  return LBM_OK;
}

int lbm_force_reduce(const double* partials, int T, int slots, double* force, lbm_stream_t stream) {
  if (T <= 0 || !force || slots < 0 || (slots > 0 && !partials)) return LBM_E_ARG;
  hipLaunchKernelGGL(k_force_reduce, (T + 255) / 256, 256, 0, (hipStream_t)stream, partials, T, slots, force);
  return hipGetLastError() == hipSuccess ? LBM_OK : LBM_E_LAUNCH;
}

int lbm_step_chain_forces_noalloc(const lbm_step_desc* d, int n, void* const* states, double* force,
                                  lbm_stream_t stream) {
  if (!d || n <= 0 || !states || !force) return LBM_E_ARG;
  const int slots = lbm_step_force_slots(d);
  if (slots > 0 && size_t(n) * size_t(slots) * 3u > kNoAllocPartials) return LBM_E_ARG;
  double* part = nullptr;
  if (slots > 0) {
    const int rc = noalloc_scratch(&part);
    if (rc) return rc;
  }
  // every link workgroup writes its slot of every step, so the partials need
  // no clearing; without links the reduction writes zero forces
  int rc = lbm_step_chain(d, n, states, part, int64_t(slots) * 3, stream);
  if (rc) return rc;
  return lbm_force_reduce(part, n, slots, force, stream);
}

int lbm_convert_state(const lbm_step_desc* d, int src_dtype, const void* src, int dst_dtype, void* dst,
                      lbm_stream_t stream) {
  StepArgs<float> a;
  const int rc = make_args(d, LBM_MODE_COLLIDE, a);
  if (rc) return rc;
  if (!src || !dst) return LBM_E_ARG;
  if (d->rounding == LBM_ROUND_STOCHASTIC) {
    // encoding into a 16-bit format draws its dither like a step (mode 3)
    a.sround = 1;
    a.seed = launch_seed(d->seed, d->step, 3);
  }
  const dim3 grid((a.N + 255) / 256, Q);
  hipStream_t s = (hipStream_t)stream;
#define LBM_CONV(SI, SO) hipLaunchKernelGGL((k_convert<SI, SO>), grid, dim3(256), 0, s, \
                                            (const SI::T*)src, (SO::T*)dst, a)
  if (src_dtype == LBM_F32 && dst_dtype == LBM_F16DEV) LBM_CONV(StoreF32, StoreF16);
  else if (src_dtype == LBM_F32 && dst_dtype == LBM_BF16DEV) LBM_CONV(StoreF32, StoreBF16);
  else if (src_dtype == LBM_F16DEV && dst_dtype == LBM_F32) LBM_CONV(StoreF16, StoreF32);
  else if (src_dtype == LBM_BF16DEV && dst_dtype == LBM_F32) LBM_CONV(StoreBF16, StoreF32);
  else if (src_dtype == LBM_F32 && dst_dtype == LBM_F64) LBM_CONV(StoreF32, StoreF64);
  else if (src_dtype == LBM_F64 && dst_dtype == LBM_F32) LBM_CONV(StoreF64, StoreF32);
  else return LBM_E_DTYPE;
#undef LBM_CONV
  return hipGetLastError() == hipSuccess ? LBM_OK : LBM_E_LAUNCH;
}

int lbm_device_check(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return LBM_E_ARG;
  hipFuncAttributes attr;
  const hipError_t e = hipFuncGetAttributes(&attr, reinterpret_cast<const void*>(&k_force_reduce));
  return e == hipSuccess ? LBM_OK : int(e);
}

}  // extern "C"
