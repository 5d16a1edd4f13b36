// Shared device-side definitions for the D3Q19 kernels (gfx950).
//
// Lattice storage ("x-fastest SoA"): population i of cell (x, y, z) lives at
//   f[i * N + (z * ny + y) * nx + x],   N = nx * ny * nz.
// Scalar fields use the same cell index; vector fields put the component first.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

#include <type_traits>
#include <utility>

#include "../../include/lbm_hip.h"

namespace lbm {

constexpr int Q = 19;

// Compile-time loop: sfor<N>(f) calls f(std::integral_constant<int, 0>{}) ...
// f(std::integral_constant<int, N-1>{}).  Inside f, `constexpr int i = I;`
// makes every lattice table lookup a constant expression, so the moment
// matrix, weights and opposite map fold into immediates with no run-time
// table code.
template <typename F, int... Is>
__host__ __device__ __forceinline__ void sfor_impl(F&& f, std::integer_sequence<int, Is...>) {
  (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, typename F>
__host__ __device__ __forceinline__ void sfor(F&& f) {
  sfor_impl(f, std::make_integer_sequence<int, N>{});
}

// D3Q19 velocity set in the order of jax_lbm/core/lattice.py:29-49: rest, the
// six faces (+x,-x,+y,-y,+z,-z), then the xy, xz and yz edges.
constexpr int EX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
constexpr int EY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
constexpr int EZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};

__host__ __device__ constexpr int opp(int i) {
  // the direction with the negated velocity
  for (int j = 0; j < Q; ++j)
    if (EX[j] == -EX[i] && EY[j] == -EY[i] && EZ[j] == -EZ[i]) return j;
  return -1;
}

__host__ __device__ constexpr int c2(int i) { return EX[i] * EX[i] + EY[i] * EY[i] + EZ[i] * EZ[i]; }

// Weights rounded to float32, as the reference keeps them (lattice.py:53-60)
// even in its float64 runs.
__host__ __device__ constexpr float weight(int i) {
  return i == 0 ? 1.0f / 3.0f : (c2(i) == 1 ? 1.0f / 18.0f : 1.0f / 36.0f);
}

// d'Humieres moment matrix, row k evaluated from its defining polynomial of
// the lattice velocity c_i (row order of lattice.py:140-145).  All entries are
// small integers, so fully unrolled products fold to adds.
__host__ __device__ constexpr int mrt_m(int k, int i) {
  const int cx = EX[i], cy = EY[i], cz = EZ[i];
  const int s = c2(i);
  switch (k) {
    case 0: return 1;
    case 1: return 19 * s - 30;
    case 2: return (21 * s * s - 53 * s + 24) / 2;
    case 3: return cx;
    case 4: return (5 * s - 9) * cx;
    case 5: return cy;
    case 6: return (5 * s - 9) * cy;
    case 7: return cz;
    case 8: return (5 * s - 9) * cz;
    case 9: return 3 * cx * cx - s;
    case 10: return (3 * s - 5) * (3 * cx * cx - s);
    case 11: return cy * cy - cz * cz;
    case 12: return (3 * s - 5) * (cy * cy - cz * cz);
    case 13: return cx * cy;
    case 14: return cy * cz;
    case 15: return cx * cz;
    case 16: return (cy * cy - cz * cz) * cx;
    case 17: return (cz * cz - cx * cx) * cy;
    case 18: return (cx * cx - cy * cy) * cz;
  }
  return 0;
}

// ||M_k||^2 = sum_i M_ki^2 (19, 2394, 252, 10, 40, ... 8).
__host__ __device__ constexpr int mrt_norm(int k) {
  int acc = 0;
  for (int i = 0; i < Q; ++i) acc += mrt_m(k, i) * mrt_m(k, i);
  return acc;
}

// Fixed relaxation rates (lattice.py:106-126), float32-rounded like the
// reference.  Slots 9, 11, 13, 14, 15 take omega = 1/tau; 0, 3, 5, 7 are
// conserved (rate 0).
__host__ __device__ constexpr float mrt_rate_fixed(int k) {
  return k == 1 ? 1.19f : (k == 2 || k == 10 || k == 12) ? 1.4f
         : (k == 4 || k == 6 || k == 8) ? 1.2f : (k >= 16) ? 1.98f : 0.0f;
}
__host__ __device__ constexpr bool mrt_is_stress(int k) {
  return k == 9 || k == 11 || k == 13 || k == 14 || k == 15;
}
__host__ __device__ constexpr bool mrt_is_conserved(int k) {
  return k == 0 || k == 3 || k == 5 || k == 7;
}

// ---------------------------------------------------------------------------
// Storage formats.  Compute is float for the float32 / half / bfloat16
// formats and double for the float64 format.  Half formats hold the deviation
// f_i - off_i from a per-direction offset (the weights w_i by default).
//   ld / ld_c / st : one element (st is a non-temporal store)
//   ldv<V> / stv<V>: V consecutive elements at an index that is a multiple of V
// ---------------------------------------------------------------------------
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// Raw buffer access (gfx9 buffer resource, untyped, 32-bit byte offsets): a
// wave-uniform resource and scalar byte offset (soffset) plus a per-lane byte
// offset, so the address needs no 64-bit vector arithmetic.  `bytes` bounds
// the resource; every caller keeps its offsets below it.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, int(bytes), 0x00020000);
}
constexpr int kBufNT = 2;  // cache policy of the non-temporal stores

__device__ __forceinline__ float bf16_bits_to_float(unsigned short raw) {
  return __uint_as_float(static_cast<unsigned>(raw) << 16);
}
__device__ __forceinline__ unsigned short float_to_bf16_bits(float d) {
  // round to nearest even, NaN stays NaN: one v_cvt_pk_bf16_f32 on gfx950
  return __builtin_bit_cast(unsigned short, static_cast<__bf16>(d));
}
__device__ __forceinline__ float f16_bits_to_float(unsigned short raw) { return __half2float(__ushort_as_half(raw)); }
__device__ __forceinline__ unsigned short float_to_f16_bits(float d) { return __half_as_ushort(__float2half_rn(d)); }
// float(h) + off for binary16 bits h, in one v_fma_mix_f32 (op_sel_hi marks
// the half operand).  For a wave-uniform `off` (every caller but the link
// paths' per-lane offsets, which use ld_cv) the operand is a scalar register.
__device__ __forceinline__ float f16_bits_add(unsigned short raw, float off) {
  float r;
  asm("v_fma_mix_f32 %0, %1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(uint32_t(raw)), "s"(off));
  return r;
}
// The binary16 bits of v - off, rounded once to nearest even, in one
// v_fma_mixlo_f16.  `off` is wave-uniform (a scalar register) at every
// caller.
__device__ __forceinline__ unsigned short f16_bits_of_diff(float v, float off) {
  uint32_t r;
  asm("v_fma_mixlo_f16 %0, %1, 1.0, -%2 op_sel_hi:[0,0,0]" : "=v"(r) : "v"(v), "s"(off));
  return static_cast<unsigned short>(r);
}

// 32-bit integer hash ("lowbias32", C. Wellons, 2018: xor-shift-multiply
// rounds with the published constants).  Supplies the uniform variates of
// stochastic rounding.
__host__ __device__ constexpr uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

struct StoreF32 {
  using T = float;
  using R = float;
  static constexpr int tag = LBM_F32;
  __device__ static R ld(const T* p, size_t k, R) { return __builtin_nontemporal_load(p + k); }
  __device__ static R ld_c(const T* p, size_t k, R) { return p[k]; }
  __device__ static R ld_cv(const T* p, size_t k, R) { return p[k]; }
  __device__ static void st(T* p, size_t k, R v, R) { __builtin_nontemporal_store(v, p + k); }
  __device__ static R bld(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so, R) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, int(vo), int(so), 0));
  }
  __device__ static void bst(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so, R v, R) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, int(vo), int(so), kBufNT);
  }
  template <int V>
  __device__ static void ldv(const T* p, size_t k, R, R* o) {
    if constexpr (V == 4) {
      const f32x4 v = *reinterpret_cast<const f32x4*>(p + k);
      o[0] = v[0]; o[1] = v[1]; o[2] = v[2]; o[3] = v[3];
    } else if constexpr (V == 2) {
      const f32x2 v = *reinterpret_cast<const f32x2*>(p + k);
      o[0] = v[0]; o[1] = v[1];
    } else {
      o[0] = p[k];
    }
  }
  template <int V>
  __device__ static void stv(T* p, size_t k, const R* o, R) {
    if constexpr (V == 4) {
      const f32x4 v = {o[0], o[1], o[2], o[3]};
      __builtin_nontemporal_store(v, reinterpret_cast<f32x4*>(p + k));
    } else if constexpr (V == 2) {
      const f32x2 v = {o[0], o[1]};
      __builtin_nontemporal_store(v, reinterpret_cast<f32x2*>(p + k));
    } else {
      __builtin_nontemporal_store(o[0], p + k);
    }
  }
};
struct StoreF64 {
  using T = double;
  using R = double;
  static constexpr int tag = LBM_F64;
  __device__ static R ld(const T* p, size_t k, R) { return __builtin_nontemporal_load(p + k); }
  __device__ static R ld_c(const T* p, size_t k, R) { return p[k]; }
  __device__ static R ld_cv(const T* p, size_t k, R) { return p[k]; }
  __device__ static void st(T* p, size_t k, R v, R) { __builtin_nontemporal_store(v, p + k); }
  __device__ static R bld(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so, R) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, int(vo), int(so), 0));
  }
  __device__ static void bst(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so, R v, R) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, int(vo), int(so), kBufNT);
  }
  template <int V>
  __device__ static void ldv(const T* p, size_t k, R, R* o) {
#pragma unroll
    for (int h = 0; h < V; h += 2) {
      if constexpr (V >= 2) {
        const f64x2 v = *reinterpret_cast<const f64x2*>(p + k + h);
        o[h] = v[0];
        o[h + 1] = v[1];
      }
    }
    if constexpr (V == 1) o[0] = p[k];
  }
  template <int V>
  __device__ static void stv(T* p, size_t k, const R* o, R) {
#pragma unroll
    for (int h = 0; h < V; h += 2) {
      if constexpr (V >= 2) {
        const f64x2 v = {o[h], o[h + 1]};
        __builtin_nontemporal_store(v, reinterpret_cast<f64x2*>(p + k + h));
      }
    }
    if constexpr (V == 1) __builtin_nontemporal_store(o[0], p + k);
  }
};
// Every access of a half format but ld_cv takes a wave-uniform offset (the
// offset of a compile-time population, or of a population uniform over the
// workgroup).
// For binary16 the conversion and the offset fold into one mixed-precision
// FMA each way: v_fma_mix_f32 (half operand * 1.0 + off, rounded once to
// float) on load and v_fma_mixlo_f16 (v * 1.0 - off, rounded once to half,
// nearest even) on store.  bfloat16 shifts the bits on load and rounds with
// v_cvt_pk_bf16_f32 on store.
template <bool BF>
struct StoreHalfDev {
  using T = typename std::conditional<BF, __hip_bfloat16, __half>::type;
  using R = float;
  static constexpr int tag = BF ? LBM_BF16DEV : LBM_F16DEV;
  __device__ static float dec(unsigned short raw, float off) {
    if constexpr (BF) return bf16_bits_to_float(raw) + off;
    else return f16_bits_add(raw, off);
  }
  __device__ static unsigned short enc(float v, float off) {
    if constexpr (BF) return float_to_bf16_bits(v - off);
    else return f16_bits_of_diff(v, off);
  }
  __device__ static const unsigned short* bits(const T* p) { return reinterpret_cast<const unsigned short*>(p); }
  __device__ static unsigned short* bits(T* p) { return reinterpret_cast<unsigned short*>(p); }
  __device__ static R ld(const T* p, size_t k, R off) { return dec(__builtin_nontemporal_load(bits(p) + k), off); }
  __device__ static R ld_c(const T* p, size_t k, R off) { return dec(bits(p)[k], off); }
  // a load whose offset differs between lanes (the Bouzidi pulls of link
  // cells pick the offset of one of two populations per lane)
  __device__ static R ld_cv(const T* p, size_t k, R off) {
    const unsigned short raw = bits(p)[k];
    return (BF ? bf16_bits_to_float(raw) : f16_bits_to_float(raw)) + off;
  }
  __device__ static void st(T* p, size_t k, R v, R off) { __builtin_nontemporal_store(enc(v, off), bits(p) + k); }
  // Stochastic rounding of the deviation d = v - off with gfx950's
  // v_cvt_sr_f16_f32 / v_cvt_sr_bf16_f32: the 16-bit value above or below d,
  // the upper one with probability (d - lo) / (hi - lo), decided by the
  // random bits rbits (binary16 subnormals included).
  // The conversion writes the low half of its destination and keeps the
  // high half, which is never stored: d's own register serves as the old
  // value, so no register has to be zeroed first.
  __device__ static unsigned short enc_sr(float v, float off, uint32_t rbits) {
    const float d = v - off;
    if constexpr (BF) {
      const bf16x2 h = __builtin_amdgcn_cvt_sr_bf16_f32(__builtin_bit_cast(bf16x2, d), d, rbits, false);
      return __builtin_bit_cast(u16x2, h)[0];
    } else {
      const f16x2 h = __builtin_amdgcn_cvt_sr_f16_f32(__builtin_bit_cast(f16x2, d), d, rbits, false);
      return __builtin_bit_cast(u16x2, h)[0];
    }
  }
  __device__ static void st_sr(T* p, size_t k, R v, R off, uint32_t rbits) {
    __builtin_nontemporal_store(enc_sr(v, off, rbits), bits(p) + k);
  }
  __device__ static void bst_sr(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so, R v, R off, uint32_t rbits) {
    __builtin_amdgcn_raw_buffer_store_b16(enc_sr(v, off, rbits), r, int(vo), int(so), kBufNT);
  }
  __device__ static R bld(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so, R off) {
    return dec(__builtin_amdgcn_raw_buffer_load_b16(r, int(vo), int(so), 0), off);
  }
  __device__ static void bst(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so, R v, R off) {
    __builtin_amdgcn_raw_buffer_store_b16(enc(v, off), r, int(vo), int(so), kBufNT);
  }
  template <int V>
  __device__ static void ldv(const T* p, size_t k, R off, R* o) {
    if constexpr (V == 4) {
      const u16x4 v = *reinterpret_cast<const u16x4*>(bits(p) + k);
      o[0] = dec(v[0], off); o[1] = dec(v[1], off); o[2] = dec(v[2], off); o[3] = dec(v[3], off);
    } else if constexpr (V == 2) {
      const u16x2 v = *reinterpret_cast<const u16x2*>(bits(p) + k);
      o[0] = dec(v[0], off); o[1] = dec(v[1], off);
    } else {
      o[0] = ld_c(p, k, off);
    }
  }
  template <int V>
  __device__ static void stv(T* p, size_t k, const R* o, R off) {
    if constexpr (V == 4) {
      const u16x4 v = {enc(o[0], off), enc(o[1], off), enc(o[2], off), enc(o[3], off)};
      __builtin_nontemporal_store(v, reinterpret_cast<u16x4*>(bits(p) + k));
    } else if constexpr (V == 2) {
      const u16x2 v = {enc(o[0], off), enc(o[1], off)};
      __builtin_nontemporal_store(v, reinterpret_cast<u16x2*>(bits(p) + k));
    } else {
      st(p, k, o[0], off);
    }
  }
};
template <class S>
constexpr bool is_half_storage = sizeof(typename S::T) == 2;

// The random word of population I of a cell, from the cell's hash h: h
// rotated right by 5 I mod 32 (19 distinct rotations).
template <int I>
__device__ __forceinline__ uint32_t pop_bits(uint32_t h) {
  constexpr int sh = (5 * I) & 31;
  if constexpr (sh == 0) return h;
  else return __builtin_amdgcn_alignbit(h, h, sh);
}

using StoreF16 = StoreHalfDev<false>;
using StoreBF16 = StoreHalfDev<true>;

// Order of the dense tiles.  The hardware deals a launch's workgroups to the
// 8 XCDs in turn, and a tile's pulls shifted by one cell in x share cache
// lines with the neighbouring tile of the same x-row.  When a tile holds
// part of a row (nx = 384 with 256-cell tiles), those lines were fetched into
// two XCDs' L2 caches.  xcd_group_tile gives workgroups d, d + 8, ...,
// d + 8 (G - 1), which share an XCD, the G consecutive tiles G m, ...,
// G m + G - 1, a whole number of rows; the last incomplete group keeps its
// order, and G = 1 is the identity.  At 384 x 192 x 192 the FP16 adjoint's
// access pattern alone took 335 us per sweep in tile order and 281 us so.
__host__ __device__ constexpr uint32_t tile_group_size(uint32_t nx, uint32_t cells) {
  uint32_t a = nx, b = cells;
  while (b) {
    const uint32_t t = a % b;
    a = b;
    b = t;
  }
  const uint32_t g = nx / a;  // lcm(nx, cells) / cells
  return g <= 16u ? g : 1u;
}
__device__ __forceinline__ uint32_t xcd_group_tile(uint32_t d, uint32_t nd, uint32_t G) {
  const uint32_t sg = 8u * G;
  if (G <= 1u || d >= nd / sg * sg) return d;
  const uint32_t r = d % sg;
  return d / sg * sg + (r % 8u) * G + r / 8u;
}

// Fast unsigned division by a run-time constant (round-up multiplier method
// of Granlund & Montgomery, 1994): q = (mulhi(n, m) + n) >> s with
// s = ceil(log2 d), m = floor(2^32 (2^s - d) / d) + 1.  Exact for n < 2^31.
struct FastDiv {
  uint32_t d, m, s;
};
inline FastDiv make_fastdiv(uint32_t d) {
  FastDiv r{d, 0, 0};
  uint32_t s = 0;
  while ((1ull << s) < d) ++s;
  r.s = s;
  r.m = static_cast<uint32_t>(((1ull << 32) * ((1ull << s) - d)) / d + 1);
  return r;
}
__device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv& f) {
  return (__umulhi(n, f.m) + n) >> f.s;
}

// Strides of a fused-path state buffer (lbm_step_desc.pop_stride/row_stride,
// 0 = dense): population i of cell (x, y, z) is element
// i * pstride + (z * ny + y) * rstride + x.  Accepted: whole populations
// apart (rstride = nx, pstride >= N) or the populations of a row interleaved
// (pstride >= nx, rstride >= 19 pstride); every element offset below 2^31.
inline bool state_layout(const lbm_step_desc* d, long long& pstride, long long& rstride) {
  const long long nx = d->nx, rows = 1LL * d->ny * d->nz, n = nx * rows;
  pstride = d->pop_stride ? d->pop_stride : n;
  rstride = d->row_stride ? d->row_stride : nx;
  const bool apart = rstride == nx && pstride >= n;
  const bool rowwise = pstride >= nx && rstride >= Q * pstride;
  if (!apart && !rowwise) return false;
  const long long extent = (Q - 1) * pstride + (rows - 1) * rstride + nx;
  return extent < (1LL << 31);
}

// Link cells per link workgroup of the fused step and of its adjoint, one
// per lane of a wave.  A link cell gathers its populations from scattered
// rows, and on small grids the step waits for the link workgroups: 64 cells
// per workgroup (rather than one per thread of a 256- or 512-thread
// workgroup) spreads them over 4x or 8x as many CUs.  Measured at
// 96 x 48 x 48 against 256 per workgroup: forward 14.4 -> 11.4 us per step,
// 2000-step store-all dCd/dr 87 -> 73.5 ms; 64 x 32 x 32 forward
// 9.7 -> 8.6 us; 256 x 128 x 128 unchanged within 1%.
constexpr int kLinkCellsPerBlock = 64;

// Sum of a float64 over the 64 lanes of a wave (every lane active), in every
// lane: DPP exchanges inside each row of 16 lanes (quad_perm [1,0,3,2] and
// [2,3,0,1], row_half_mirror, row_mirror) leave each lane with its row's
// sum, then the four row sums are read from lanes 0, 16, 32 and 48.  Only
// VALU operations: the shuffle form (ds_bpermute through LDS) puts six LDS
// round trips on the dependent path.
__device__ __forceinline__ double wave_sum_dpp(double v) {
  auto step = [](double x, auto ctrl) {
    const u32x2 b = __builtin_bit_cast(u32x2, x);
    u32x2 o;
    o[0] = __builtin_amdgcn_update_dpp(0u, b[0], decltype(ctrl)::value, 0xf, 0xf, false);
    o[1] = __builtin_amdgcn_update_dpp(0u, b[1], decltype(ctrl)::value, 0xf, 0xf, false);
    return x + __builtin_bit_cast(double, o);
  };
  v = step(v, std::integral_constant<int, 0xB1>{});   // quad_perm [1,0,3,2]
  v = step(v, std::integral_constant<int, 0x4E>{});   // quad_perm [2,3,0,1]
  v = step(v, std::integral_constant<int, 0x141>{});  // row_half_mirror
  v = step(v, std::integral_constant<int, 0x140>{});  // row_mirror
  auto lane_value = [&](auto l) {
    const u32x2 b = __builtin_bit_cast(u32x2, v);
    u32x2 o;
    o[0] = __builtin_amdgcn_readlane(b[0], decltype(l)::value);
    o[1] = __builtin_amdgcn_readlane(b[1], decltype(l)::value);
    return __builtin_bit_cast(double, o);
  };
  const double r0 = lane_value(std::integral_constant<int, 0>{});
  const double r1 = lane_value(std::integral_constant<int, 16>{});
  const double r2 = lane_value(std::integral_constant<int, 32>{});
  const double r3 = lane_value(std::integral_constant<int, 48>{});
  return (r0 + r1) + (r2 + r3);
}

template <typename R>
__device__ __forceinline__ R wave_sum(R v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

}  // namespace lbm
