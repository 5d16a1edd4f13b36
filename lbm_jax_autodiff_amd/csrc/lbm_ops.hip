// Per-operator kernels behind the jax_lbm.core-compatible surface.
// One thread per cell; populations are read with stride N (coalesced along x).
#include "lbm_geometry.hpp"

namespace lbm {

constexpr int kBlock = 256;

inline unsigned grid_for(size_t n) { return static_cast<unsigned>((n + kBlock - 1) / kBlock); }

inline int launch_status() {
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? LBM_OK : LBM_E_LAUNCH;
}

inline bool dims_ok(int nx, int ny, int nz) {
  if (nx <= 0 || ny <= 0 || nz <= 0) return false;
  const long long n = 1LL * nx * ny * nz;
  return n * Q < (1LL << 31);
}

template <typename R>
__global__ void k_equilibrium(const R* __restrict__ rho, const R* __restrict__ u, R* __restrict__ f,
                              size_t N) {
  const size_t c = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (c >= N) return;
  const R r = rho[c], ux = u[c], uy = u[N + c], uz = u[2 * N + c];
  const R usq = ux * ux + uy * uy + uz * uz;
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    f[i * N + c] = feq<i>(r, ux, uy, uz, usq);
  });
}

template <typename R>
__device__ __forceinline__ void load_cell(const R* __restrict__ f, size_t N, size_t c, R* v) {
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    v[i] = f[i * N + c];
  });
}

template <typename R>
__device__ __forceinline__ void store_cell(R* __restrict__ f, size_t N, size_t c, const R* v) {
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    f[i * N + c] = v[i];
  });
}

template <typename R>
__global__ void k_density(const R* __restrict__ f, R* __restrict__ rho, size_t N) {
  const size_t c = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (c >= N) return;
  R acc = R(0);
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    acc += f[i * N + c];
  });
  rho[c] = acc;
}

template <typename R>
__global__ void k_velocity(const R* __restrict__ f, const R* __restrict__ rho, R* __restrict__ u,
                           size_t N) {
  const size_t c = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (c >= N) return;
  R jx = 0, jy = 0, jz = 0;
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    const R v = f[i * N + c];
    add_scaled<EX[i]>(jx, v);
    add_scaled<EY[i]>(jy, v);
    add_scaled<EZ[i]>(jz, v);
  });
  const R r = rho[c];
  u[c] = jx / r;
  u[N + c] = jy / r;
  u[2 * N + c] = jz / r;
}

template <typename R, bool MRT>
__global__ void k_collide(const R* __restrict__ fin, R tau, const R* __restrict__ tau_field,
                          R* __restrict__ fout, size_t N) {
  const size_t c = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (c >= N) return;
  R f[Q];
  load_cell(fin, N, c, f);
  const R omega = R(1) / (tau_field ? tau_field[c] : tau);
  if constexpr (MRT) collide_mrt(f, omega);
  else collide_bgk(f, omega);
  store_cell(fout, N, c, f);
}

template <typename R>
__global__ void k_smag(const R* __restrict__ fin, R tau, R cs, R* __restrict__ tau_out, size_t N) {
  const size_t c = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (c >= N) return;
  R f[Q];
  load_cell(fin, N, c, f);
  tau_out[c] = smag_tau(f, tau, cs);
}

template <typename R>
__global__ void k_soft_bounce(const R* __restrict__ fin, const R* __restrict__ phi,
                              R* __restrict__ fout, size_t N) {
  const size_t c = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (c >= N) return;
  R f[Q];
  load_cell(fin, N, c, f);
  blend_opposite(f, phi[c]);
  store_cell(fout, N, c, f);
}

template <typename R>
__global__ void k_hard_bounce(const R* __restrict__ fin, const int32_t* __restrict__ solid,
                              R* __restrict__ fout, size_t N) {
  const size_t c = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (c >= N) return;
  R f[Q];
  load_cell(fin, N, c, f);
  if (solid[c] >= 1) reverse_all(f);
  store_cell(fout, N, c, f);
}

// Post-step stability guard (state.py:158-169): cells with rho < 0.3,
// rho > 2, |u| > 0.4 or rho = NaN become f_eq(1, 0) = w.
template <typename R>
__global__ void k_guard(const R* __restrict__ fin, R* __restrict__ fout, size_t N) {
  const size_t c = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (c >= N) return;
  R f[Q];
  load_cell(fin, N, c, f);
  if (guard_bad(f)) set_weights(f);
  store_cell(fout, N, c, f);
}

// Pull streaming + Bouzidi (streaming.py:99-183), dense q.
template <typename R>
__global__ void k_stream(const R* __restrict__ f, const R* __restrict__ q, int periodic,
                         R* __restrict__ out, int nx, int ny, int nz) {
  const size_t N = size_t(nx) * ny * nz;
  const size_t c = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (c >= N) return;
  const int x = int(c % nx);
  const int y = int((c / nx) % ny);
  const int z = int(c / (size_t(nx) * ny));
  out[c] = f[c];
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    if constexpr (i != 0) {
      constexpr int io = opp(i);
      // normal pull from x - e_i
      int sx = x - EX[i], sy = y - EY[i], sz = z - EZ[i];
      bool own = (sx < 0) || (sx >= nx);
      if (periodic) {
        sy = (sy + ny) % ny;
        sz = (sz + nz) % nz;
      } else {
        own = own || sy < 0 || sy >= ny || sz < 0 || sz >= nz;
      }
      R v = own ? f[i * N + c] : f[i * N + (size_t(sz) * ny + sy) * nx + sx];
      const R qo = q[io * N + c];
      if (qo >= R(0)) {
        const R a = f[io * N + c];
        if (qo >= R(0.5)) {
          const R inv2q = R(1) / (R(2) * fmin(fmax(qo, R(0.5)), R(1)));
          v = inv2q * a + (R(1) - inv2q) * f[i * N + c];
        } else {
          // x + e_i, periodic in all three axes (streaming.py:166-172)
          const int fx = (x + EX[i] + nx) % nx, fy = (y + EY[i] + ny) % ny, fz = (z + EZ[i] + nz) % nz;
          const R twoq = R(2) * fmin(fmax(qo, R(0)), R(0.5));
          v = twoq * a + (R(1) - twoq) * f[io * N + (size_t(fz) * ny + fy) * nx + fx];
        }
      }
      out[i * N + c] = v;
    }
  });
}

template <typename R>
__global__ void k_zou_he_in(R* __restrict__ f, R ux, R uy, R uz, int has_rt, R rt, int nx, int ny,
                            int nz) {
  const size_t N = size_t(nx) * ny * nz;
  const size_t p = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (p >= size_t(ny) * nz) return;
  const size_t c = p * nx;  // x = 0 of row p
  R g[Q];
  load_cell(f, N, c, g);
  zou_he_in(g, ux, uy, uz, has_rt != 0, rt);
  f[1 * N + c] = g[1];
  f[7 * N + c] = g[7];
  f[9 * N + c] = g[9];
  f[11 * N + c] = g[11];
  f[13 * N + c] = g[13];
}

template <typename R>
__global__ void k_zou_he_out(R* __restrict__ f, R rho_out, int nx, int ny, int nz) {
  const size_t N = size_t(nx) * ny * nz;
  const size_t p = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (p >= size_t(ny) * nz) return;
  const size_t c = p * nx + (nx - 1);
  R g[Q];
  load_cell(f, N, c, g);
  zou_he_out(g, rho_out);
  f[2 * N + c] = g[2];
  f[10 * N + c] = g[10];
  f[8 * N + c] = g[8];
  f[14 * N + c] = g[14];
  f[12 * N + c] = g[12];
}

// ---- deterministic float64 force reduction --------------------------------
constexpr int kRedBlocks = 512;

template <typename R>
__global__ void k_mex_partial(const R* __restrict__ fp, const R* __restrict__ fs, const R* __restrict__ q,
                              int want_p, double* __restrict__ part, size_t N) {
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (size_t c = blockIdx.x * size_t(blockDim.x) + threadIdx.x; c < N; c += size_t(gridDim.x) * blockDim.x) {
    R eqv[Q];
    if (want_p) {
      R f[Q];
      load_cell(fp, N, c, f);
      const Macro<R> m = macro(f);
      sfor<Q>([&](auto I) {
        constexpr int i = I;
        eqv[i] = feq<i>(m.rho, m.ux, m.uy, m.uz, m.usq);
      });
    }
    sfor<Q>([&](auto I) {
      constexpr int i = I;
      if constexpr (i != 0) {
        constexpr int io = opp(i);
        if (q[i * N + c] >= R(0)) {
          const double t = double(fp[i * N + c]) + double(fs[io * N + c]);
          add_scaled<EX[i]>(acc[0], t);
          add_scaled<EY[i]>(acc[1], t);
          add_scaled<EZ[i]>(acc[2], t);
          if (want_p) {
            const double pp = double(eqv[i]) + double(eqv[io]);
            add_scaled<EX[i]>(acc[3], pp);
            add_scaled<EY[i]>(acc[4], pp);
            add_scaled<EZ[i]>(acc[5], pp);
          }
        }
      }
    });
  }
  __shared__ double sh[6][kBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const double v = wave_sum(acc[k]);
    if (lane == 0) sh[k][wid] = v;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    double s = 0;
    for (int w = 0; w < kBlock / 64; ++w) s += sh[threadIdx.x][w];
    part[blockIdx.x * 6 + threadIdx.x] = s;
  }
}

__global__ void k_sum_rows(const double* __restrict__ part, int rows, int width, double* __restrict__ out) {
  // out[k] = sum_r part[r*width + k], one thread per k, fixed order
  const int k = threadIdx.x;
  if (k >= width) return;
  double s = 0;
  for (int r = 0; r < rows; ++r) s += part[size_t(r) * width + k];
  out[k] = s;
}

template <typename A, typename B>
__global__ void k_dot_partial(const A* __restrict__ a, const B* __restrict__ b, size_t n, double* __restrict__ part) {
  double acc = 0.0;
  for (size_t k = blockIdx.x * size_t(blockDim.x) + threadIdx.x; k < n; k += size_t(gridDim.x) * blockDim.x)
    acc += double(a[k]) * double(b[k]);
  __shared__ double sh[kBlock / 64];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0;
    for (int w = 0; w < kBlock / 64; ++w) s += sh[w];
    part[blockIdx.x] = s;
  }
}

__global__ void k_zero_u32(uint32_t* p) { *p = 0u; }

__global__ void k_area(const int32_t* __restrict__ solid, int axis, uint32_t* counter, int nx, int ny,
                       int nz) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  int n_a, n_b, len;
  if (axis == 0) { n_a = ny; n_b = nz; len = nx; }
  else if (axis == 1) { n_a = nx; n_b = nz; len = ny; }
  else { n_a = nx; n_b = ny; len = nz; }
  bool hit = false;
  if (t < n_a * n_b) {
    const int a = t % n_a, b = t / n_a;
    for (int s = 0; s < len && !hit; ++s) {
      int x, y, z;
      if (axis == 0) { x = s; y = a; z = b; }
      else if (axis == 1) { x = a; y = s; z = b; }
      else { x = a; y = b; z = s; }
      hit = solid[(size_t(z) * ny + y) * nx + x] == 1;
    }
  }
  // integer sum: order-independent
  const unsigned long long m = __ballot(hit);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(counter, static_cast<uint32_t>(__popcll(m)));
}

__global__ void k_u32_to_f32(const uint32_t* c, float* out) { *out = float(*c); }

__global__ void k_drag(const double* __restrict__ force, int T, double U, const double* __restrict__ area,
                       double rho, double* __restrict__ cd) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  cd[t] = fabs(force[3 * t]) / (0.5 * rho * U * U * area[0] + 1e-10);
}

// ---- geometry --------------------------------------------------------------
template <typename C, typename O>
__global__ void k_sphere_cells(int mode, C cx, C cy, C cz, C r, C alpha, int32_t* __restrict__ solid,
                               O* __restrict__ porosity, O* __restrict__ dphidr, int nx, int ny, int nz) {
  const size_t N = size_t(nx) * ny * nz;
  const size_t c = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (c >= N) return;
  const int x = int(c % nx), y = int((c / nx) % ny), z = int(c / (size_t(nx) * ny));
  C wx, wy, wz;
  world_xyz(x, y, z, nx, ny, nz, wx, wy, wz);
  const C dx = wx - cx, dy = wy - cy, dz = wz - cz;
  if (mode == 0) {
    solid[c] = (dx * dx + dy * dy + dz * dz <= r * r) ? 1 : 0;
    if (porosity) porosity[c] = O(solid[c]);
    if (dphidr) dphidr[c] = O(0);
  } else {
    const C dist = sqrt(dx * dx + dy * dy + dz * dz);
    const C phi = C(1) / (C(1) + exp(-(alpha * (r - dist))));
    solid[c] = phi > C(0.5) ? 1 : 0;
    if (porosity) porosity[c] = O(phi);
    // d sigmoid(alpha (r - dist)) / d r = alpha phi (1 - phi)
    if (dphidr) dphidr[c] = O(alpha * phi * (C(1) - phi));
  }
}

template <typename C, typename O>
__global__ void k_sphere_sdf(C cx, C cy, C cz, C r, O* __restrict__ sdf, int nx, int ny, int nz) {
  const size_t N = size_t(nx) * ny * nz;
  const size_t c = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (c >= N) return;
  const int x = int(c % nx), y = int((c / nx) % ny), z = int(c / (size_t(nx) * ny));
  C wx, wy, wz;
  world_xyz(x, y, z, nx, ny, nz, wx, wy, wz);
  const C dx = wx - cx, dy = wy - cy, dz = wz - cz;
  sdf[c] = O(r - sqrt(dx * dx + dy * dy + dz * dz));
}

template <typename R>
__global__ void k_sigmoid(const R* __restrict__ sdf, R alpha, R* __restrict__ out, size_t N) {
  const size_t c = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (c >= N) return;
  out[c] = R(1) / (R(1) + exp(-(alpha * sdf[c])));
}

template <typename C, typename O>
__global__ void k_sphere_q(C cx, C cy, C cz, C r, C disc_floor, C denom_eps,
                           const int32_t* __restrict__ solid, O* __restrict__ q, O* __restrict__ dqdr,
                           int nx, int ny, int nz) {
  const size_t N = size_t(nx) * ny * nz;
  const size_t c = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (c >= N) return;
  const int x = int(c % nx), y = int((c / nx) % ny), z = int(c / (size_t(nx) * ny));
  C wx, wy, wz;
  world_xyz(x, y, z, nx, ny, nz, wx, wy, wz);
  const C ox = wx - cx, oy = wy - cy, oz = wz - cz;
  q[c] = O(-1);
  if (dqdr) dqdr[c] = O(0);
  for (int i = 1; i < Q; ++i) {
    O qv = O(-1), dq = O(0);
    if (link_from(solid, x, y, z, i, nx, ny, nz)) {
      // ray-sphere hit along e_i in world units (primitives.py:115-138,
      // smooth.py:187-203)
      const C rdx = C(EX[i]) / C(nx / 8.0), rdy = C(EY[i]) / C(ny / 4.0), rdz = C(EZ[i]) / C(nz / 4.0);
      const C a = rdx * rdx + rdy * rdy + rdz * rdz;
      const C b = C(2) * (ox * rdx + oy * rdy + oz * rdz);
      const C cc = ox * ox + oy * oy + oz * oz - r * r;
      const C disc = b * b - C(4) * a * cc;
      const C sq = sqrt(fmax(disc, disc_floor));
      const C t = (-b - sq) / (C(2) * a + denom_eps);
      const bool valid = (disc >= C(0)) && (t > C(0)) && (t <= C(1));
      const C t0 = valid ? t : C(0.5);
      const C tq = fmin(fmax(t0, C(0.01)), C(0.99));
      qv = O(tq);
      // d t / d r = -(1/(2a+eps)) d sqrt(max(disc,floor))/dr, d disc/dr = 8 a r;
      // zero when the max, the where or the clip is not on the t branch
      if (valid && disc > disc_floor && t0 > C(0.01) && t0 < C(0.99))
        dq = O(-(C(4) * a * r / sq) / (C(2) * a + denom_eps));
    }
    q[i * N + c] = qv;
    if (dqdr) dqdr[i * N + c] = dq;
  }
}

template <typename O>
__global__ void k_q_from_solid(const int32_t* __restrict__ solid, O dq, O* __restrict__ q, int nx, int ny,
                               int nz) {
  const size_t N = size_t(nx) * ny * nz;
  const size_t c = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (c >= N) return;
  const int x = int(c % nx), y = int((c / nx) % ny), z = int(c / (size_t(nx) * ny));
  q[c] = O(-1);
  for (int i = 1; i < Q; ++i) q[i * N + c] = link_from(solid, x, y, z, i, nx, ny, nz) ? dq : O(-1);
}

// ---- link list -------------------------------------------------------------
template <typename R>
__device__ __forceinline__ bool has_link(const R* q, size_t c, size_t N) {
  bool any = false;
#pragma unroll
  for (int j = 1; j < Q; ++j) any |= q[j * N + c] >= R(0);
  return any;
}

template <typename R>
__global__ void k_links_count(const R* __restrict__ q, int32_t* __restrict__ counts, size_t N) {
  const size_t c = blockIdx.x * size_t(kBlock) + threadIdx.x;
  const bool l = c < N && has_link(q, c, N);
  __shared__ int sh[kBlock / 64];
  const unsigned long long m = __ballot(l);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = __popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int w = 0; w < kBlock / 64; ++w) s += sh[w];
    counts[blockIdx.x] = s;
  }
}

constexpr int kScanThreads = 1024;

// Exclusive prefix sum of the per-block link counts in one workgroup: each
// thread sums a contiguous chunk, the chunk totals are scanned in LDS, then
// each thread writes its chunk's offsets.  Runs once per geometry.
__global__ __launch_bounds__(kScanThreads) void k_exclusive_scan(const int32_t* __restrict__ counts,
                                                                 int32_t* __restrict__ offsets, int n,
                                                                 int32_t* __restrict__ total) {
  __shared__ int32_t part[kScanThreads];
  const int t = int(threadIdx.x);
  const int per = (n + kScanThreads - 1) / kScanThreads;
  const int b0 = min(n, t * per), b1 = min(n, b0 + per);
  int32_t s = 0;
  for (int b = b0; b < b1; ++b) s += counts[b];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    const int32_t v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int32_t run = t > 0 ? part[t - 1] : 0;
  for (int b = b0; b < b1; ++b) {
    offsets[b] = run;
    run += counts[b];
  }
  if (t == kScanThreads - 1) *total = part[t];
}

template <typename R>
__global__ void k_links_build(const R* __restrict__ q, const R* __restrict__ dqdr, const int32_t* __restrict__ solid,
                              const int32_t* __restrict__ offsets, int n_links, int32_t* __restrict__ link_cell,
                              double* __restrict__ link_q, double* __restrict__ link_dq,
                              uint8_t* __restrict__ flags, size_t N) {
  const size_t c = blockIdx.x * size_t(kBlock) + threadIdx.x;
  const bool l = c < N && has_link(q, c, N);
  __shared__ int sh[kBlock / 64];
  const unsigned long long m = __ballot(l);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = __popcll(m);
  __syncthreads();
  int base = offsets[blockIdx.x];
  for (int w = 0; w < wid; ++w) base += sh[w];
  const int rank = base + __popcll(m & ((1ull << lane) - 1ull));
  if (c < N) {
    uint8_t fl = 0;
    if (solid && solid[c] >= 1) fl |= 1;
    if (l) fl |= 2;
    flags[c] = fl;
  }
  if (l && rank < n_links) {
    link_cell[rank] = int32_t(c);
    for (int j = 1; j < Q; ++j) {
      link_q[size_t(j - 1) * n_links + rank] = double(q[j * N + c]);
      if (link_dq) link_dq[size_t(j - 1) * n_links + rank] = dqdr ? double(dqdr[j * N + c]) : 0.0;
    }
  }
}

}  // namespace lbm

using namespace lbm;

#define LBM_DISPATCH_RF(dtype, ...)                    \
  do {                                                 \
    if ((dtype) == LBM_F32) {                          \
      using R = float;                                 \
      __VA_ARGS__;                                     \
    } else if ((dtype) == LBM_F64) {                   \
      using R = double;                                \
      __VA_ARGS__;                                     \
    } else {                                           \
      return LBM_E_DTYPE;                              \
    }                                                  \
  } while (0)

extern "C" {

int lbm_version(void) { return 10000; }

int lbm_equilibrium(int dtype, const void* rho, const void* u, void* f_out, int nx, int ny, int nz,
                    lbm_stream_t stream) {
  if (!dims_ok(nx, ny, nz) || !rho || !u || !f_out) return LBM_E_ARG;
  const size_t N = size_t(nx) * ny * nz;
  LBM_DISPATCH_RF(dtype, hipLaunchKernelGGL(k_equilibrium<R>, grid_for(N), kBlock, 0, (hipStream_t)stream,
                                            (const R*)rho, (const R*)u, (R*)f_out, N));
  return launch_status();
}

int lbm_density(int dtype, const void* f, void* rho_out, int nx, int ny, int nz, lbm_stream_t stream) {
  if (!dims_ok(nx, ny, nz) || !f || !rho_out) return LBM_E_ARG;
  const size_t N = size_t(nx) * ny * nz;
  LBM_DISPATCH_RF(dtype, hipLaunchKernelGGL(k_density<R>, grid_for(N), kBlock, 0, (hipStream_t)stream,
                                            (const R*)f, (R*)rho_out, N));
  return launch_status();
}

int lbm_velocity(int dtype, const void* f, const void* rho, void* u_out, int nx, int ny, int nz,
                 lbm_stream_t stream) {
  if (!dims_ok(nx, ny, nz) || !f || !rho || !u_out) return LBM_E_ARG;
  const size_t N = size_t(nx) * ny * nz;
  LBM_DISPATCH_RF(dtype, hipLaunchKernelGGL(k_velocity<R>, grid_for(N), kBlock, 0, (hipStream_t)stream,
                                            (const R*)f, (const R*)rho, (R*)u_out, N));
  return launch_status();
}

int lbm_bgk(int dtype, const void* f, double tau, const void* tau_field, void* f_out, int nx, int ny, int nz,
            lbm_stream_t stream) {
  if (!dims_ok(nx, ny, nz) || !f || !f_out) return LBM_E_ARG;
  const size_t N = size_t(nx) * ny * nz;
  LBM_DISPATCH_RF(dtype, hipLaunchKernelGGL((k_collide<R, false>), grid_for(N), kBlock, 0, (hipStream_t)stream,
                                            (const R*)f, R(tau), (const R*)tau_field, (R*)f_out, N));
  return launch_status();
}

int lbm_mrt(int dtype, const void* f, double tau, const void* tau_field, void* f_out, int nx, int ny, int nz,
            lbm_stream_t stream) {
  if (!dims_ok(nx, ny, nz) || !f || !f_out) return LBM_E_ARG;
  const size_t N = size_t(nx) * ny * nz;
  LBM_DISPATCH_RF(dtype, hipLaunchKernelGGL((k_collide<R, true>), grid_for(N), kBlock, 0, (hipStream_t)stream,
                                            (const R*)f, R(tau), (const R*)tau_field, (R*)f_out, N));
  return launch_status();
}

int lbm_smagorinsky_tau(int dtype, const void* f, double tau, double cs, void* tau_out, int nx, int ny, int nz,
                        lbm_stream_t stream) {
  if (!dims_ok(nx, ny, nz) || !f || !tau_out) return LBM_E_ARG;
  const size_t N = size_t(nx) * ny * nz;
  LBM_DISPATCH_RF(dtype, hipLaunchKernelGGL(k_smag<R>, grid_for(N), kBlock, 0, (hipStream_t)stream,
                                            (const R*)f, R(tau), R(cs), (R*)tau_out, N));
  return launch_status();
}

int lbm_soft_bounce_back(int dtype, const void* f, const void* porosity, void* f_out, int nx, int ny, int nz,
                         lbm_stream_t stream) {
  if (!dims_ok(nx, ny, nz) || !f || !porosity || !f_out) return LBM_E_ARG;
  const size_t N = size_t(nx) * ny * nz;
  LBM_DISPATCH_RF(dtype, hipLaunchKernelGGL(k_soft_bounce<R>, grid_for(N), kBlock, 0, (hipStream_t)stream,
                                            (const R*)f, (const R*)porosity, (R*)f_out, N));
  return launch_status();
}

int lbm_hard_bounce(int dtype, const void* f, const int32_t* solid, void* f_out, int nx, int ny, int nz,
                    lbm_stream_t stream) {
  if (!dims_ok(nx, ny, nz) || !f || !solid || !f_out) return LBM_E_ARG;
  const size_t N = size_t(nx) * ny * nz;
  LBM_DISPATCH_RF(dtype, hipLaunchKernelGGL(k_hard_bounce<R>, grid_for(N), kBlock, 0, (hipStream_t)stream,
                                            (const R*)f, solid, (R*)f_out, N));
  return launch_status();
}

int lbm_stability_guard(int dtype, const void* f, void* f_out, int nx, int ny, int nz, lbm_stream_t stream) {
  if (!dims_ok(nx, ny, nz) || !f || !f_out) return LBM_E_ARG;
  const size_t N = size_t(nx) * ny * nz;
  LBM_DISPATCH_RF(dtype, hipLaunchKernelGGL(k_guard<R>, grid_for(N), kBlock, 0, (hipStream_t)stream,
                                            (const R*)f, (R*)f_out, N));
  return launch_status();
}

int lbm_stream(int dtype, const void* f, const void* q, int periodic_yz, void* f_out, int nx, int ny, int nz,
               lbm_stream_t stream) {
  if (!dims_ok(nx, ny, nz) || !f || !q || !f_out || f == f_out) return LBM_E_ARG;
  const size_t N = size_t(nx) * ny * nz;
  LBM_DISPATCH_RF(dtype, hipLaunchKernelGGL(k_stream<R>, grid_for(N), kBlock, 0, (hipStream_t)stream,
                                            (const R*)f, (const R*)q, periodic_yz, (R*)f_out, nx, ny, nz));
  return launch_status();
}

int lbm_zou_he_inlet(int dtype, void* f, double ux, double uy, double uz, int has_rho_target,
                     double rho_target, int nx, int ny, int nz, lbm_stream_t stream) {
  if (!dims_ok(nx, ny, nz) || !f) return LBM_E_ARG;
  const size_t P = size_t(ny) * nz;
  LBM_DISPATCH_RF(dtype, hipLaunchKernelGGL(k_zou_he_in<R>, grid_for(P), kBlock, 0, (hipStream_t)stream,
                                            (R*)f, R(ux), R(uy), R(uz), has_rho_target, R(rho_target), nx, ny, nz));
  return launch_status();
}

int lbm_zou_he_outlet(int dtype, void* f, double rho_out, int nx, int ny, int nz, lbm_stream_t stream) {
  if (!dims_ok(nx, ny, nz) || !f) return LBM_E_ARG;
  const size_t P = size_t(ny) * nz;
  LBM_DISPATCH_RF(dtype, hipLaunchKernelGGL(k_zou_he_out<R>, grid_for(P), kBlock, 0, (hipStream_t)stream,
                                            (R*)f, R(rho_out), nx, ny, nz));
  return launch_status();
}

size_t lbm_reduce_workspace_bytes(int, int, int) { return size_t(kRedBlocks) * 6 * sizeof(double); }

int lbm_momentum_exchange(int dtype, const void* f_post, const void* f_streamed, const void* q, int want_pressure,
                          double* out, void* workspace, int nx, int ny, int nz, lbm_stream_t stream) {
  if (!dims_ok(nx, ny, nz) || !f_post || !f_streamed || !q || !out || !workspace) return LBM_E_ARG;
  const size_t N = size_t(nx) * ny * nz;
  double* part = static_cast<double*>(workspace);
  LBM_DISPATCH_RF(dtype, hipLaunchKernelGGL(k_mex_partial<R>, kRedBlocks, kBlock, 0, (hipStream_t)stream,
                                            (const R*)f_post, (const R*)f_streamed, (const R*)q, want_pressure,
                                            part, N));
  hipLaunchKernelGGL(k_sum_rows, 1, 64, 0, (hipStream_t)stream, part, kRedBlocks, 6, out);
  return launch_status();
}

int lbm_dot(int a_dtype, const void* a, int b_dtype, const void* b, size_t n, double* out, void* workspace,
            lbm_stream_t stream) {
  if (!a || !b || !out || !workspace) return LBM_E_ARG;
  hipStream_t s = (hipStream_t)stream;
  double* part = static_cast<double*>(workspace);
#define LBM_DOT(A, B) hipLaunchKernelGGL((k_dot_partial<A, B>), kRedBlocks, kBlock, 0, s, (const A*)a, (const B*)b, n, part)
  if (a_dtype == LBM_F32 && b_dtype == LBM_F32) { LBM_DOT(float, float); }
  else if (a_dtype == LBM_F32 && b_dtype == LBM_F64) { LBM_DOT(float, double); }
  else if (a_dtype == LBM_F64 && b_dtype == LBM_F32) { LBM_DOT(double, float); }
  else if (a_dtype == LBM_F64 && b_dtype == LBM_F64) { LBM_DOT(double, double); }
  else return LBM_E_DTYPE;
#undef LBM_DOT
  hipLaunchKernelGGL(k_sum_rows, 1, 64, 0, s, part, kRedBlocks, 1, out);
  return launch_status();
}

int lbm_projected_area(const int32_t* solid, int axis, float* out, uint32_t* counter, int nx, int ny, int nz,
                       lbm_stream_t stream) {
  if (!dims_ok(nx, ny, nz) || !solid || !out || !counter || axis < 0 || axis > 2) return LBM_E_ARG;
  const int cols = axis == 0 ? ny * nz : (axis == 1 ? nx * nz : nx * ny);
  hipLaunchKernelGGL(k_zero_u32, 1, 1, 0, (hipStream_t)stream, counter);
  hipLaunchKernelGGL(k_area, grid_for(cols), kBlock, 0, (hipStream_t)stream, solid, axis, counter, nx, ny, nz);
  hipLaunchKernelGGL(k_u32_to_f32, 1, 1, 0, (hipStream_t)stream, counter, out);
  return launch_status();
}

int lbm_drag_coefficient(const double* force, int T, double inlet_velocity, const double* area, double rho,
                         double* cd, lbm_stream_t stream) {
  if (!force || !area || !cd || T <= 0) return LBM_E_ARG;
  hipLaunchKernelGGL(k_drag, grid_for(T), kBlock, 0, (hipStream_t)stream, force, T, inlet_velocity, area, rho, cd);
  return launch_status();
}

int lbm_sphere_geometry(int mode, int coord_dtype, int out_dtype, double cx, double cy, double cz, double radius,
                        double sharpness, double disc_floor, double denom_eps, int32_t* solid, void* porosity,
                        void* q, void* dqdr, void* dphidr, int nx, int ny, int nz, lbm_stream_t stream) {
  if (!dims_ok(nx, ny, nz) || !solid || !q || (mode != 0 && mode != 1)) return LBM_E_ARG;
  const size_t N = size_t(nx) * ny * nz;
  hipStream_t s = (hipStream_t)stream;
#define LBM_GEOM(C, O)                                                                                       \
  hipLaunchKernelGGL((k_sphere_cells<C, O>), grid_for(N), kBlock, 0, s, mode, C(cx), C(cy), C(cz), C(radius), \
                     C(sharpness), solid, (O*)porosity, (O*)dphidr, nx, ny, nz);                             \
  hipLaunchKernelGGL((k_sphere_q<C, O>), grid_for(N), kBlock, 0, s, C(cx), C(cy), C(cz), C(radius),           \
                     C(mode == 0 ? 0.0 : disc_floor), C(denom_eps), solid, (O*)q, (O*)dqdr, nx, ny, nz)
  if (coord_dtype == LBM_F32 && out_dtype == LBM_F32) { LBM_GEOM(float, float); }
  else if (coord_dtype == LBM_F32 && out_dtype == LBM_F64) { LBM_GEOM(float, double); }
  else if (coord_dtype == LBM_F64 && out_dtype == LBM_F32) { LBM_GEOM(double, float); }
  else if (coord_dtype == LBM_F64 && out_dtype == LBM_F64) { LBM_GEOM(double, double); }
  else return LBM_E_DTYPE;
#undef LBM_GEOM
  return launch_status();
}

int lbm_sphere_sdf(int coord_dtype, int out_dtype, double cx, double cy, double cz, double radius, void* sdf,
                   int nx, int ny, int nz, lbm_stream_t stream) {
  if (!dims_ok(nx, ny, nz) || !sdf) return LBM_E_ARG;
  const size_t N = size_t(nx) * ny * nz;
  hipStream_t s = (hipStream_t)stream;
#define LBM_SDF(C, O) hipLaunchKernelGGL((k_sphere_sdf<C, O>), grid_for(N), kBlock, 0, s, C(cx), C(cy), C(cz), \
                                         C(radius), (O*)sdf, nx, ny, nz)
  if (coord_dtype == LBM_F32 && out_dtype == LBM_F32) { LBM_SDF(float, float); }
  else if (coord_dtype == LBM_F32 && out_dtype == LBM_F64) { LBM_SDF(float, double); }
  else if (coord_dtype == LBM_F64 && out_dtype == LBM_F32) { LBM_SDF(double, float); }
  else if (coord_dtype == LBM_F64 && out_dtype == LBM_F64) { LBM_SDF(double, double); }
  else return LBM_E_DTYPE;
#undef LBM_SDF
  return launch_status();
}

int lbm_porosity_from_sdf(int dtype, const void* sdf, double sharpness, void* out, int nx, int ny, int nz,
                          lbm_stream_t stream) {
  if (!dims_ok(nx, ny, nz) || !sdf || !out) return LBM_E_ARG;
  const size_t N = size_t(nx) * ny * nz;
  LBM_DISPATCH_RF(dtype, hipLaunchKernelGGL(k_sigmoid<R>, grid_for(N), kBlock, 0, (hipStream_t)stream,
                                            (const R*)sdf, R(sharpness), (R*)out, N));
  return launch_status();
}

int lbm_q_from_solid(int dtype, const int32_t* solid, double default_q, void* q, int nx, int ny, int nz,
                     lbm_stream_t stream) {
  if (!dims_ok(nx, ny, nz) || !solid || !q) return LBM_E_ARG;
  const size_t N = size_t(nx) * ny * nz;
  LBM_DISPATCH_RF(dtype, hipLaunchKernelGGL(k_q_from_solid<R>, grid_for(N), kBlock, 0, (hipStream_t)stream,
                                            solid, R(default_q), (R*)q, nx, ny, nz));
  return launch_status();
}

int lbm_links_count(int dtype, const void* q, int32_t* block_counts, int32_t* block_offsets, int32_t* total,
                    int nx, int ny, int nz, lbm_stream_t stream) {
  if (!dims_ok(nx, ny, nz) || !q || !block_counts || !block_offsets || !total) return LBM_E_ARG;
  const size_t N = size_t(nx) * ny * nz;
  const unsigned nb = grid_for(N);
  LBM_DISPATCH_RF(dtype, hipLaunchKernelGGL(k_links_count<R>, nb, kBlock, 0, (hipStream_t)stream,
                                            (const R*)q, block_counts, N));
  hipLaunchKernelGGL(k_exclusive_scan, 1, kScanThreads, 0, (hipStream_t)stream, block_counts, block_offsets, int(nb),
                     total);
  return launch_status();
}

int lbm_links_build(int dtype, const void* q, const void* dqdr, const int32_t* solid, const int32_t* block_offsets,
                    int n_links, int32_t* link_cell, double* link_q, double* link_dq, uint8_t* flags, int nx, int ny,
                    int nz, lbm_stream_t stream) {
  if (!dims_ok(nx, ny, nz) || !q || !block_offsets || !flags || n_links < 0) return LBM_E_ARG;
  if (n_links > 0 && (!link_cell || !link_q)) return LBM_E_ARG;
  const size_t N = size_t(nx) * ny * nz;
  LBM_DISPATCH_RF(dtype, hipLaunchKernelGGL(k_links_build<R>, grid_for(N), kBlock, 0, (hipStream_t)stream,
                                            (const R*)q, (const R*)dqdr, solid, block_offsets, n_links, link_cell,
                                            link_q, link_dq, flags, N));
  return launch_status();
}

}  // extern "C"
