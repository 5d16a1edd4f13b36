// Fused D3Q19 timestep for gfx950: one launch per step.
//
// The carried state is the post-collision, post-bounce population P_t (the
// reference's f_post).  One launch computes, for every cell x,
//   S_{t+1}(x) = guard(ZouHe(pull(P_t) with Bouzidi links))
//   P_{t+1}(x) = bounce(collide(S_{t+1}(x)))
// reading each population of P_t once (pull streaming is a bijection) and
// writing P_{t+1} once: 2 x 19 values per cell.
//
// Workgroups [0, nlb) own the sparse link cells (fluid cells with a Bouzidi
// link, listed in link_cell/link_q) and also produce the float64
// momentum-exchange partials; workgroups [nlb, nlb + ndb) sweep the dense grid
// and skip link cells.  Dense workgroups are re-mapped so that consecutive
// x-rows share an XCD (and its L2).
#include "lbm_physics.hpp"

namespace lbm {

constexpr int kStepBlock = 256;

template <typename R>
struct StepArgs {
  int nx, ny, nz;
  uint32_t N;
  FastDiv dnx, dny;
  int coll, smag, bounce, guard, periodic, n_links, nlb, ndb;
  R tau, cs;
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

// Neighbour rows of a cell, wrapped (periodic) or flagged (clamped).
struct Rows {
  uint32_t ym, yp, zm, zp;  // row offsets (in cells) of y-1, y+1, z-1, z+1
  bool ym_ok, yp_ok, zm_ok, zp_ok;
};

template <typename R>
__device__ __forceinline__ Rows make_rows(const StepArgs<R>& a, int y, int z) {
  Rows r;
  const uint32_t nx = uint32_t(a.nx), plane = uint32_t(a.nx) * uint32_t(a.ny);
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

// Source cell of the normal pull for direction I (streaming.py:113-142):
// x - e_I with y/z wrapped; the cell itself where the reference keeps the
// cell's own value (x faces, and y/z faces when not periodic).  Offsets are
// added modulo 2^32, so a wrapped row is a plain unsigned addition.
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

// Cell x + e_I wrapped periodically in all three axes (streaming.py:166-172).
template <int I, typename R>
__device__ __forceinline__ uint32_t wrap_next(const StepArgs<R>& a, int x, int y, int z) {
  int fx = x + EX[I], fy = y + EY[I], fz = z + EZ[I];
  fx = fx < 0 ? fx + a.nx : (fx >= a.nx ? fx - a.nx : fx);
  fy = fy < 0 ? fy + a.ny : (fy >= a.ny ? fy - a.ny : fy);
  fz = fz < 0 ? fz + a.nz : (fz >= a.nz ? fz - a.nz : fz);
  return (uint32_t(fz) * uint32_t(a.ny) + uint32_t(fy)) * uint32_t(a.nx) + uint32_t(fx);
}

// Zou-He planes and the post-collision guard on a pulled S (state.py:152-169).
template <typename R>
__device__ __forceinline__ void boundaries(const StepArgs<R>& a, R* s, int x) {
  if (x == 0) zou_he_in(s, a.uin[0], a.uin[1], a.uin[2], false, R(0));
  if (x == a.nx - 1) zou_he_out(s, a.rho_out);
  if (a.guard == LBM_GUARD_POST && guard_bad(s)) set_weights(s);
}

// [pre guard] -> [Smagorinsky tau] -> BGK|MRT -> hard|soft bounce
// (examples/10:71-78, state.py:133-149, smooth.py:151-155).
template <typename R, int COLL>
__device__ __forceinline__ void collide_cell(const StepArgs<R>& a, R* s, uint32_t c) {
  if (a.guard == LBM_GUARD_PRE && guard_bad(s)) set_feq_unit(s, a.uref[0], a.uref[1], a.uref[2]);
  R tau = a.tau;
  if (a.smag) tau = smag_tau(s, a.tau, a.cs);
  const R omega = R(1) / tau;
  if constexpr (COLL == LBM_COLL_MRT) collide_mrt(s, omega);
  else collide_bgk(s, omega);
  if (a.bounce == LBM_BOUNCE_HARD) {
    if (a.flags[c] & 1) reverse_all(s);
  } else if (a.bounce == LBM_BOUNCE_SOFT) {
    blend_opposite(s, a.phi[c]);
  }
}

template <class S>
__device__ __forceinline__ void store_cell(const StepArgs<typename S::R>& a, typename S::T* __restrict__ dst,
                                           uint32_t c, const typename S::R* s) {
  const size_t N = a.N;
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    S::st(dst, i * N + c, s[i], a.off[i]);
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
  const size_t N = a.N;
  const uint32_t c0 = t * V;
  if (c0 >= a.N) return;
  int x0, y, z;
  cell_coords(a, c0, x0, y, z);
  const int lane = threadIdx.x & 63;
  const bool row_first = (x0 == 0), row_last = (x0 + V == a.nx);
  R s[V][Q];
  if constexpr (MODE == LBM_MODE_COLLIDE) {
    sfor<Q>([&](auto I) {
      constexpr int i = I;
      R blk[V];
      S::template ldv<V>(src, i * N + c0, a.off[i], blk);
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
      uint32_t base = c0;
      bool row_ok = true;
      if constexpr (ey == 1) { row_ok = row_ok && rows.ym_ok; base += rows.ym; }
      else if constexpr (ey == -1) { row_ok = row_ok && rows.yp_ok; base += rows.yp; }
      if constexpr (ez == 1) { row_ok = row_ok && rows.zm_ok; base += rows.zm; }
      else if constexpr (ez == -1) { row_ok = row_ok && rows.zp_ok; base += rows.zp; }
      if (!row_ok) base = c0;
      R blk[V];
      S::template ldv<V>(src, i * N + base, a.off[i], blk);
      if constexpr (ex == 0) {
#pragma unroll
        for (int j = 0; j < V; ++j) s[j][i] = blk[j];
      } else if constexpr (ex == 1) {
        // cell x0 + j pulls x0 + j - 1
        R prev = __shfl_up(blk[V - 1], 1, 64);
        if (row_ok && !row_first && lane == 0) prev = S::ld_c(src, i * N + base - 1, a.off[i]);
        if (row_first && row_ok) prev = S::ld_c(src, i * N + c0, a.off[i]);  // x = 0 keeps its own value
        s[0][i] = row_ok ? prev : blk[0];
#pragma unroll
        for (int j = 1; j < V; ++j) s[j][i] = row_ok ? blk[j - 1] : blk[j];
      } else {
        // cell x0 + j pulls x0 + j + 1
        R next = __shfl_down(blk[0], 1, 64);
        if (row_ok && !row_last && lane == 63) next = S::ld_c(src, i * N + base + V, a.off[i]);
        if (row_last && row_ok) next = S::ld_c(src, i * N + c0 + V - 1, a.off[i]);  // x = nx-1 keeps its own
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
    if constexpr (MODE != LBM_MODE_COLLIDE) boundaries(a, s[j], x0 + j);
    if constexpr (MODE != LBM_MODE_STREAM) {
      if (a.guard == LBM_GUARD_PRE && guard_bad(s[j])) set_feq_unit(s[j], a.uref[0], a.uref[1], a.uref[2]);
      R tau = a.tau;
      if (a.smag) tau = smag_tau(s[j], a.tau, a.cs);
      const R omega = R(1) / tau;
      if constexpr (COLL == LBM_COLL_MRT) collide_mrt(s[j], omega);
      else collide_bgk(s[j], omega);
      if (a.bounce == LBM_BOUNCE_HARD) {
        if (fj & 1u) reverse_all(s[j]);
      } else if (a.bounce == LBM_BOUNCE_SOFT) {
        blend_opposite(s[j], phi[j]);
      }
    }
  }
  // link cells are written by the link workgroups (dense MODE_COLLIDE writes all)
  const bool has_link = (MODE != LBM_MODE_COLLIDE) && (fl & 0x02020202u);
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    R o[V];
#pragma unroll
    for (int j = 0; j < V; ++j) o[j] = s[j][i];
    if (!has_link) {
      S::template stv<V>(dst, i * N + c0, o, a.off[i]);
    } else {
#pragma unroll
      for (int j = 0; j < V; ++j)
        if (!((fl >> (8 * j)) & 2u)) S::st(dst, i * N + c0 + j, o[j], a.off[i]);
    }
  });
}

template <class S, int COLL, int MODE, int V>
__global__ __launch_bounds__(kStepBlock) void k_step(StepArgs<typename S::R> a,
                                                     const typename S::T* __restrict__ src,
                                                     typename S::T* __restrict__ dst,
                                                     double* __restrict__ fpart) {
  using R = typename S::R;
  const size_t N = a.N;
  const uint32_t bid = blockIdx.x;

  if (MODE != LBM_MODE_COLLIDE && int(bid) < a.nlb) {
    // ---------------- link cells: Bouzidi + force partials ----------------
    const int b = int(bid) * kStepBlock + int(threadIdx.x);
    double fx = 0.0, fy = 0.0, fz = 0.0;
    if (b < a.n_links) {
      const uint32_t c = uint32_t(a.link_cell[b]);
      int x, y, z;
      cell_coords(a, c, x, y, z);
      const Rows rows = make_rows(a, y, z);
      R own[Q], s[Q];
      sfor<Q>([&](auto I) {
        constexpr int i = I;
        own[i] = S::ld_c(src, i * N + c, a.off[i]);
      });
      s[0] = own[0];
      sfor<Q>([&](auto I) {
        constexpr int i = I;
        if constexpr (i != 0) {
          constexpr int io = opp(i);
          const R qo = R(a.link_q[size_t(io - 1) * a.n_links + b]);
          if (qo >= R(0)) {
            if (qo >= R(0.5)) {
              const R inv2q = R(1) / (R(2) * fmin(fmax(qo, R(0.5)), R(1)));
              s[i] = inv2q * own[io] + (R(1) - inv2q) * own[i];
            } else {
              const R twoq = R(2) * fmin(fmax(qo, R(0)), R(0.5));
              const R ffv = S::ld_c(src, io * N + wrap_next<i>(a, x, y, z), a.off[io]);
              s[i] = twoq * own[io] + (R(1) - twoq) * ffv;
            }
          } else {
            s[i] = S::ld_c(src, i * N + pull_src<i>(a, rows, c, x), a.off[i]);
          }
        }
      });
      boundaries(a, s, x);
      // momentum exchange: links j with q_j(x) >= 0 (forces.py:95-118)
      sfor<Q>([&](auto J) {
        constexpr int j = J;
        if constexpr (j != 0) {
          if (a.link_q[size_t(j - 1) * a.n_links + b] >= 0.0) {
            const double t = double(own[j]) + double(s[opp(j)]);
            add_scaled<EX[j]>(fx, t);
            add_scaled<EY[j]>(fy, t);
            add_scaled<EZ[j]>(fz, t);
          }
        }
      });
      if (MODE == LBM_MODE_FULL) collide_cell<R, COLL>(a, s, c);
      store_cell<S>(a, dst, c, s);
    }
    if (fpart) {
      __shared__ double red[3][kStepBlock / 64];
      fx = wave_sum(fx);
      fy = wave_sum(fy);
      fz = wave_sum(fz);
      const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
      if (lane == 0) {
        red[0][w] = fx;
        red[1][w] = fy;
        red[2][w] = fz;
      }
      __syncthreads();
      if (threadIdx.x < 3) {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < kStepBlock / 64; ++k) acc += red[threadIdx.x][k];
        fpart[size_t(bid) * 3 + threadIdx.x] = acc;
      }
    }
    return;
  }

  // ---------------- dense sweep ----------------
  const uint32_t first = (MODE == LBM_MODE_COLLIDE) ? 0u : uint32_t(a.nlb);
  const uint32_t tile = xcd_swizzle(bid - first, uint32_t(a.ndb));
  if constexpr (V > 1) {
    dense_vec<S, COLL, MODE, V>(a, src, dst, tile * kStepBlock + threadIdx.x);
    return;
  }
  const uint32_t c = tile * kStepBlock + threadIdx.x;
  if (c >= a.N) return;
  R s[Q];
  if (MODE == LBM_MODE_COLLIDE) {
    sfor<Q>([&](auto I) {
      constexpr int i = I;
      s[i] = S::ld(src, i * N + c, a.off[i]);
    });
    collide_cell<R, COLL>(a, s, c);
  } else {
    if (a.flags && (a.flags[c] & 2)) return;  // owned by a link workgroup
    int x, y, z;
    cell_coords(a, c, x, y, z);
    const Rows rows = make_rows(a, y, z);
    sfor<Q>([&](auto I) {
      constexpr int i = I;
      s[i] = S::ld_c(src, i * N + pull_src<i>(a, rows, c, x), a.off[i]);
    });
    boundaries(a, s, x);
    if (MODE == LBM_MODE_FULL) collide_cell<R, COLL>(a, s, c);
  }
  store_cell<S>(a, dst, c, s);
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

template <class SI, class SO>
__global__ void k_convert(const typename SI::T* __restrict__ in, typename SO::T* __restrict__ out, StepArgs<float> a,
                          size_t total) {
  const size_t k = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (k >= total) return;
  const int i = int(k / a.N);
  using RI = typename SI::R;
  using RO = typename SO::R;
  const RI v = SI::ld_c(in, k, RI(a.off[i]));
  SO::st(out, k, RO(v), RO(a.off[i]));
}

template <typename R>
int make_args(const lbm_step_desc* d, int mode, StepArgs<R>& a) {
  if (!d || d->nx <= 0 || d->ny <= 0 || d->nz <= 0) return LBM_E_ARG;
  const long long n = 1LL * d->nx * d->ny * d->nz;
  if (n * Q >= (1LL << 31)) return LBM_E_ARG;
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
  a.nlb = (mode == LBM_MODE_COLLIDE) ? 0 : (a.n_links + kStepBlock - 1) / kStepBlock;
  a.ndb = int((n + kStepBlock - 1) / kStepBlock);
  return LBM_OK;
}

// Cells per thread of the dense sweep: the requested value if it divides nx,
// else the widest of 4, 2, 1 that does.
inline int choose_vec(const lbm_step_desc* d) {
  const int want = d->cells_per_thread;
  if (want == 1 || want == 2 || want == 4) return (d->nx % want == 0) ? want : 1;
  if (d->nx % 4 == 0) return 4;
  if (d->nx % 2 == 0) return 2;
  return 1;
}

template <class S>
int launch_step(const lbm_step_desc* d, int mode, const void* src, void* dst, double* fpart, hipStream_t s) {
  using R = typename S::R;
  StepArgs<R> a;
  const int rc = make_args(d, mode, a);
  if (rc) return rc;
  if (!src || !dst || src == dst) return LBM_E_ARG;
  const int V = choose_vec(d);
  const long long cells_per_block = 1LL * kStepBlock * V;
  a.ndb = int((1LL * a.N + cells_per_block - 1) / cells_per_block);
  const unsigned grid = unsigned(a.nlb + a.ndb);
  const auto* in = static_cast<const typename S::T*>(src);
  auto* out = static_cast<typename S::T*>(dst);
#define LBM_STEP_LAUNCH(COLL, MODE, VV) \
  hipLaunchKernelGGL((k_step<S, COLL, MODE, VV>), grid, kStepBlock, 0, s, a, in, out, fpart)
#define LBM_STEP_V(COLL, MODE)                                      \
  do {                                                              \
    if (V == 4) LBM_STEP_LAUNCH(COLL, MODE, 4);                     \
    else if (V == 2) LBM_STEP_LAUNCH(COLL, MODE, 2);                \
    else LBM_STEP_LAUNCH(COLL, MODE, 1);                            \
  } while (0)
  if (a.coll == LBM_COLL_MRT) {
    if (mode == LBM_MODE_FULL) LBM_STEP_V(LBM_COLL_MRT, LBM_MODE_FULL);
    else if (mode == LBM_MODE_COLLIDE) LBM_STEP_V(LBM_COLL_MRT, LBM_MODE_COLLIDE);
    else if (mode == LBM_MODE_STREAM) LBM_STEP_V(LBM_COLL_MRT, LBM_MODE_STREAM);
    else return LBM_E_ARG;
  } else {
    if (mode == LBM_MODE_FULL) LBM_STEP_V(LBM_COLL_BGK, LBM_MODE_FULL);
    else if (mode == LBM_MODE_COLLIDE) LBM_STEP_V(LBM_COLL_BGK, LBM_MODE_COLLIDE);
    else if (mode == LBM_MODE_STREAM) LBM_STEP_V(LBM_COLL_BGK, LBM_MODE_STREAM);
    else return LBM_E_ARG;
  }
#undef LBM_STEP_V
#undef LBM_STEP_LAUNCH
  return hipGetLastError() == hipSuccess ? LBM_OK : LBM_E_LAUNCH;
}

}  // namespace lbm

using namespace lbm;

extern "C" {

int lbm_step_force_slots(const lbm_step_desc* d) {
  if (!d || d->n_links < 0) return 0;
  return (d->n_links + kStepBlock - 1) / kStepBlock;
}

int lbm_step(const lbm_step_desc* d, int mode, const void* src, void* dst, double* force_partial,
             lbm_stream_t stream) {
  if (!d) return LBM_E_ARG;
  hipStream_t s = (hipStream_t)stream;
  switch (d->storage) {
    case LBM_F32: return launch_step<StoreF32>(d, mode, src, dst, force_partial, s);
    case LBM_F64: return launch_step<StoreF64>(d, mode, src, dst, force_partial, s);
    case LBM_F16DEV: return launch_step<StoreF16>(d, mode, src, dst, force_partial, s);
    case LBM_BF16DEV: return launch_step<StoreBF16>(d, mode, src, dst, force_partial, s);
  }
  return LBM_E_DTYPE;
}

int lbm_force_reduce(const double* partials, int T, int slots, double* force, lbm_stream_t stream) {
  if (T <= 0 || !force || slots < 0 || (slots > 0 && !partials)) return LBM_E_ARG;
  hipLaunchKernelGGL(k_force_reduce, (T + 255) / 256, 256, 0, (hipStream_t)stream, partials, T, slots, force);
  return hipGetLastError() == hipSuccess ? LBM_OK : LBM_E_LAUNCH;
}

int lbm_convert_state(const lbm_step_desc* d, int src_dtype, const void* src, int dst_dtype, void* dst,
                      lbm_stream_t stream) {
  StepArgs<float> a;
  const int rc = make_args(d, LBM_MODE_COLLIDE, a);
  if (rc) return rc;
  if (!src || !dst) return LBM_E_ARG;
  const size_t total = size_t(a.N) * Q;
  const unsigned grid = unsigned((total + 255) / 256);
  hipStream_t s = (hipStream_t)stream;
#define LBM_CONV(SI, SO) hipLaunchKernelGGL((k_convert<SI, SO>), grid, 256, 0, s, \
                                            (const SI::T*)src, (SO::T*)dst, a, total)
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
