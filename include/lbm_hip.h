/*
 * lbm_hip.h -- C ABI of liblbm_hip.so, the gfx950 (MI355X) implementation of
 * the jax_lbm per-timestep hot path (MarcosAsh/LBM_JAX_Autodiff).
 *
 * The reference has no FFI: its boundary is the Python surface re-exported at
 * jax_lbm/__init__.py:23-60.  Each entry point below replaces the reference
 * function named in its comment; the Python package lbm_jax_autodiff_amd binds
 * them with ctypes and keeps the reference's names and signatures.
 *
 * Conventions
 *  - All pointers are device pointers unless marked "host".
 *  - Fields use "lattice storage": x fastest.  Population i of cell (x,y,z) is
 *    element i*N + (z*ny + y)*nx + x with N = nx*ny*nz; scalar fields drop i;
 *    vector fields (velocity) put the component first.
 *  - dtype selects the element type: LBM_F32 float, LBM_F64 double, and for the
 *    fused step also LBM_F16DEV / LBM_BF16DEV (half-precision deviations
 *    f_i - off_i, see lbm_step_desc.offset).
 *  - Every call is asynchronous on `stream` (a hipStream_t; 0 = null stream),
 *    allocates nothing, and returns 0 or an LBM_E_* / hipError_t code.
 *  - Calls are reentrant; there is no global mutable state.
 */
#ifndef LBM_HIP_H_
#define LBM_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* lbm_stream_t; /* hipStream_t */

enum {
  LBM_F32 = 0,
  LBM_F64 = 1,
  LBM_F16DEV = 2,
  LBM_BF16DEV = 3
};

enum {
  LBM_OK = 0,
  LBM_E_ARG = 1001,     /* bad size, null pointer or unsupported flag */
  LBM_E_DTYPE = 1002,   /* dtype not supported by this entry point */
  LBM_E_LAUNCH = 1003   /* kernel launch failed (hipGetLastError != 0) */
};

/* Library version and a probe that the gfx950 code object is loadable. */
int lbm_version(void);
int lbm_device_check(void);

/* ---- Equilibrium and moments (jax_lbm/core/equilibrium.py) ---------------- */

/* f_eq = w rho (1 + 3 e.u + 4.5 (e.u)^2 - 1.5 u^2).  Replaces equilibrium()
 * at jax_lbm/core/equilibrium.py:90-151.  rho: N, u: 3N, f_out: 19N. */
int lbm_equilibrium(int dtype, const void* rho, const void* u, void* f_out,
                    int nx, int ny, int nz, lbm_stream_t stream);

/* rho = sum_i f_i.  Replaces compute_density() at equilibrium.py:22-48. */
int lbm_density(int dtype, const void* f, void* rho_out, int nx, int ny, int nz,
                lbm_stream_t stream);

/* u = sum_i f_i e_i / rho.  Replaces compute_velocity() at equilibrium.py:51-87. */
int lbm_velocity(int dtype, const void* f, const void* rho, void* u_out,
                 int nx, int ny, int nz, lbm_stream_t stream);

/* ---- Collision (jax_lbm/core/collision.py) -------------------------------- */

/* BGK.  tau_field (N values) overrides tau when non-null.  Replaces bgk() at
 * jax_lbm/core/collision.py:46-100. */
int lbm_bgk(int dtype, const void* f, double tau, const void* tau_field, void* f_out,
            int nx, int ny, int nz, lbm_stream_t stream);

/* MRT (d'Humieres basis, omega in the stress slots).  Replaces mrt() at
 * collision.py:103-184 with mrt_forward/mrt_inverse/mrt_equilibrium_moments
 * at jax_lbm/core/lattice.py:132-391. */
int lbm_mrt(int dtype, const void* f, double tau, const void* tau_field, void* f_out,
            int nx, int ny, int nz, lbm_stream_t stream);

/* tau_eff = (tau + sqrt(tau^2 + 18 cs^2 |Pi_neq| / (rho + 1e-10))) / 2.
 * Replaces smagorinsky_tau() at collision.py:187-270. */
int lbm_smagorinsky_tau(int dtype, const void* f, double tau, double cs, void* tau_out,
                        int nx, int ny, int nz, lbm_stream_t stream);

/* (1 - phi) f + phi f[opp].  Replaces soft_bounce_back() at
 * jax_lbm/geometry/smooth.py:118-155. */
int lbm_soft_bounce_back(int dtype, const void* f, const void* porosity, void* f_out,
                         int nx, int ny, int nz, lbm_stream_t stream);

/* f_out = f[opp] where solid >= 1, else f (solid: int32 N).  Replaces the hard
 * solid bounce of jax_lbm/state.py:146-149. */
int lbm_hard_bounce(int dtype, const void* f, const int32_t* solid, void* f_out,
                    int nx, int ny, int nz, lbm_stream_t stream);

/* Post-step stability guard: cells with rho < 0.3, rho > 2, |u| > 0.4 or
 * rho = NaN are reset to f_eq(1, 0), the others copied.  Replaces the inline
 * guard of step() at jax_lbm/state.py:158-169. */
int lbm_stability_guard(int dtype, const void* f, void* f_out, int nx, int ny, int nz,
                        lbm_stream_t stream);

/* ---- Streaming (jax_lbm/core/streaming.py) -------------------------------- */

/* Pull streaming with Bouzidi links from a dense q (19N, same dtype; q<0 = no
 * wall).  Replaces stream() at jax_lbm/core/streaming.py:40-185. */
int lbm_stream(int dtype, const void* f, const void* q, int periodic_yz, void* f_out,
               int nx, int ny, int nz, lbm_stream_t stream);

/* ---- Zou-He boundaries (jax_lbm/core/boundary.py), in place on f ---------- */

/* Velocity inlet on the x = 0 plane.  has_rho_target != 0 uses rho_target.
 * Replaces zou_he_inlet() at jax_lbm/core/boundary.py:35-124. */
int lbm_zou_he_inlet(int dtype, void* f, double ux, double uy, double uz,
                     int has_rho_target, double rho_target,
                     int nx, int ny, int nz, lbm_stream_t stream);

/* Pressure outlet on the x = nx-1 plane.  Replaces zou_he_outlet() at
 * boundary.py:127-200. */
int lbm_zou_he_outlet(int dtype, void* f, double rho_out, int nx, int ny, int nz,
                      lbm_stream_t stream);

/* ---- Forces (jax_lbm/core/forces.py) -------------------------------------- */

/* Bytes of device workspace lbm_momentum_exchange needs. */
size_t lbm_reduce_workspace_bytes(int nx, int ny, int nz);

/* out[0:3] = total force, out[3:6] = pressure part (when want_pressure), both
 * accumulated in float64 with a fixed reduction order.  Replaces
 * momentum_exchange() (forces.py:37-120) and momentum_exchange_decomposed()
 * (forces.py:123-196).  q is dense (19N, same dtype as f). */
int lbm_momentum_exchange(int dtype, const void* f_post, const void* f_streamed,
                          const void* q, int want_pressure, double* out, void* workspace,
                          int nx, int ny, int nz, lbm_stream_t stream);

/* out[0] = sum_k a_k b_k over n elements in float64 with a fixed reduction
 * order; workspace as for lbm_momentum_exchange.  Chains per-element adjoints
 * with geometry derivatives (e.g. dL/dq . dq/dr). */
int lbm_dot(int a_dtype, const void* a, int b_dtype, const void* b, size_t n,
            double* out, void* workspace, lbm_stream_t stream);

/* Number of columns along `axis` holding a body cell (solid == 1), written as
 * one float to out.  `counter` is one device uint32 of scratch.  Replaces
 * projected_area() at forces.py:199-234. */
int lbm_projected_area(const int32_t* solid, int axis, float* out, uint32_t* counter,
                       int nx, int ny, int nz, lbm_stream_t stream);

/* cd[t] = |F_t,x| / (0.5 rho U^2 A + 1e-10) for T force rows (float64).
 * area is a device float64 scalar.  Replaces drag_coefficient() at
 * forces.py:237-276. */
int lbm_drag_coefficient(const double* force, int T, double inlet_velocity,
                         const double* area, double rho, double* cd,
                         lbm_stream_t stream);

/* ---- Geometry (jax_lbm/geometry) ----------------------------------------- */

/* Sphere geometry on the world box [-4,4]x[-2,2]x[-2,2].
 *  mode 0: hard mask dist^2 <= r^2 with ray q  (create_sphere,
 *          jax_lbm/geometry/primitives.py:56-174; disc floor 0)
 *  mode 1: porosity = sigmoid(sharpness (r - dist)), solid = porosity > 1/2,
 *          ray q with sqrt(max(disc, disc_floor))  (smooth_sphere_geometry,
 *          jax_lbm/geometry/smooth.py:158-248)
 * coord_dtype (LBM_F32/LBM_F64) is the precision of the coordinate and ray
 * arithmetic; out_dtype that of porosity/q/dqdr/dphidr.  porosity, dqdr and
 * dphidr may be null.  dqdr = dq/dr in world units (0 where q is clipped or
 * invalid); dphidr = d porosity / dr = sharpness phi (1 - phi) (0 in mode 0). */
int lbm_sphere_geometry(int mode, int coord_dtype, int out_dtype,
                        double cx, double cy, double cz, double radius, double sharpness,
                        double disc_floor, double denom_eps,
                        int32_t* solid, void* porosity, void* q, void* dqdr, void* dphidr,
                        int nx, int ny, int nz, lbm_stream_t stream);

/* r - |x_w - c| (positive inside), world units.  Replaces
 * signed_distance_sphere() at jax_lbm/geometry/smooth.py:34-78. */
int lbm_sphere_sdf(int coord_dtype, int out_dtype, double cx, double cy, double cz, double radius,
                   void* sdf, int nx, int ny, int nz, lbm_stream_t stream);

/* sigmoid(sharpness * sdf).  Replaces porosity_from_sdf() at
 * jax_lbm/geometry/smooth.py:81-115. */
int lbm_porosity_from_sdf(int dtype, const void* sdf, double sharpness, void* out,
                          int nx, int ny, int nz, lbm_stream_t stream);

/* q = default_q on every fluid cell whose in-bounds neighbour along e_i is a
 * body cell, else -1.  Replaces compute_q_from_solid() at
 * jax_lbm/geometry/bouzidi.py:32-106. */
int lbm_q_from_solid(int dtype, const int32_t* solid, double default_q, void* q,
                     int nx, int ny, int nz, lbm_stream_t stream);

/* ---- Link list (compiled form of a dense q for the fused step) ------------ */

/* Pass 1: per-256-cell-block count of link cells (cells with any q_j >= 0,
 * j = 1..18) into block_counts[ceil(N/256)], and their exclusive prefix into
 * block_offsets (same size) plus the total into total[0].  The host reads
 * total[0] to size pass 2's outputs. */
int lbm_links_count(int dtype, const void* q, int32_t* block_counts, int32_t* block_offsets,
                    int32_t* total, int nx, int ny, int nz, lbm_stream_t stream);

/* Pass 2: link_cell[n] (cell index, ascending), link_q[18][n] (q_j for j=1..18,
 * -1 = none), link_dq[18][n] (dq/dr, optional), and flags[N] (bit0 solid >= 1,
 * bit1 link cell; solid may be null). */
int lbm_links_build(int dtype, const void* q, const void* dqdr, const int32_t* solid,
                    const int32_t* block_offsets, int n_links, int32_t* link_cell,
                    double* link_q, double* link_dq, uint8_t* flags,
                    int nx, int ny, int nz, lbm_stream_t stream);

/* ---- Fused timestep (jax_lbm/state.py:87-171 and the examples' scan bodies) */

enum { LBM_COLL_BGK = 0, LBM_COLL_MRT = 1 };
enum { LBM_BOUNCE_NONE = 0, LBM_BOUNCE_HARD = 1, LBM_BOUNCE_SOFT = 2 };
enum { LBM_GUARD_NONE = 0, LBM_GUARD_POST = 1, LBM_GUARD_PRE = 2 };
enum {
  LBM_MODE_FULL = 0,     /* P_t -> pull/Bouzidi/Zou-He/[guard] -> S_{t+1} -> collide/bounce -> P_{t+1} */
  LBM_MODE_COLLIDE = 1,  /* S_t -> [pre guard]/collide/bounce -> P_t   (rollout prologue) */
  LBM_MODE_STREAM = 2    /* P_t -> pull/Bouzidi/Zou-He/[post guard] -> S_{t+1} (epilogue) */
};

typedef struct lbm_step_desc {
  int32_t nx, ny, nz;
  int32_t storage;       /* LBM_F32 / LBM_F64 / LBM_F16DEV / LBM_BF16DEV */
  int32_t collision;     /* LBM_COLL_* */
  int32_t smagorinsky;   /* 0/1: tau_eff from smagorinsky_tau before collision */
  int32_t bounce;        /* LBM_BOUNCE_* */
  int32_t guard;         /* LBM_GUARD_* */
  int32_t periodic_yz;   /* 1 periodic y/z, 0 clamp (streaming.py:134-142) */
  int32_t n_links;       /* number of link cells in link_cell / link_q */
  double tau;
  double smag_cs;
  double u_in[3];        /* Zou-He inlet velocity */
  double rho_out;        /* Zou-He outlet density */
  double u_ref[3];       /* pre-guard reset target eq(1, u_ref) */
  double offset[19];     /* half-storage offsets off_i (deviation f_i - off_i) */
  const uint8_t* flags;     /* N: bit0 solid (hard bounce), bit1 link cell */
  const void* porosity;     /* N, float (double for LBM_F64): soft bounce */
  const int32_t* link_cell; /* n_links */
  const double* link_q;     /* 18 x n_links */
  int32_t cells_per_thread; /* dense sweep: 0 = automatic (one cell per
                               thread; where nx % 64 == 0 each wave's row
                               addresses are scalar -- the adjoint, and the
                               forward with 16-bit populations), 1 = one cell
                               per thread with per-lane addresses, 2 or 4 =
                               x-consecutive cells per thread when they divide
                               nx and pop_stride (else 1) */
  int64_t pop_stride;       /* elements between population i and i+1 of every
                               state and cotangent buffer the fused entry
                               points read or write; 0 = N */
  int64_t row_stride;       /* elements between consecutive (y, z) rows; 0 = nx.
                               Population i of cell (x, y, z) is element
                                 i * pop_stride + (z * ny + y) * row_stride + x.
                               Two layouts are accepted:
                               - populations apart: row_stride = nx and
                                 pop_stride >= N ([19][pop_stride], cells
                                 x-fastest; lbm_state_stride gives a padded
                                 pop_stride);
                               - rows interleaved: pop_stride >= nx and
                                 row_stride >= 19 * pop_stride ([nz][ny][19][nx]
                                 for pop_stride = nx, row_stride = 19 nx: the
                                 19 populations of an x-row are adjacent) */
  int32_t rounding;         /* 16-bit storage formats: LBM_ROUND_NEAREST
                               (default) or LBM_ROUND_STOCHASTIC; ignored by
                               the float32/float64 formats */
  uint32_t seed;            /* stochastic rounding: seed of the dither */
  int64_t step;             /* stochastic rounding: index of the first step
                               of the call (lbm_step_chain's step k is
                               step + k); the dither of a stored value is a
                               hash of (seed, step, mode, cell, population),
                               so recomputing a step reproduces it exactly */
} lbm_step_desc;

/* Rounding of the 16-bit storage formats.  NEAREST: round to nearest even.
 * STOCHASTIC: the stored deviation is rounded up or down with probabilities
 * that make its expectation the exact value (a deterministic hash supplies
 * the uniform variate).  No counterpart in the reference (float32 only). */
enum { LBM_ROUND_NEAREST = 0, LBM_ROUND_STOCHASTIC = 1 };

/* Recommended pop_stride (>= N) for the grid of d: a padded stride that keeps
 * the 19 population streams of a cell out of the same HBM channels.  No
 * counterpart in the reference (its state is a dense jnp array); 0 on bad
 * dimensions. */
int64_t lbm_state_stride(const lbm_step_desc* d);

/* Number of float64 force partials (3 each) one lbm_step launch writes. */
int lbm_step_force_slots(const lbm_step_desc* d);

/* One fused timestep (one kernel launch).  src and dst are distinct state
 * buffers in d->storage format.  force_partial (optional, 3 x
 * lbm_step_force_slots doubles) receives the momentum-exchange force
 * sum over links of e_j (P_t(x)_j + S_{t+1}(x)_opp(j)) as per-workgroup float64
 * partials (modes FULL and STREAM).  Replaces one body of the reference's
 * scan loops: jax_lbm/state.py:129-171, examples/10_cd_convergence.py:69-83,
 * examples/11_optimisation_v2.py:152-160, examples/01_channel_flow.py:51-57. */
int lbm_step(const lbm_step_desc* d, int mode, const void* src, void* dst,
             double* force_partial, lbm_stream_t stream);

/* n consecutive LBM_MODE_FULL steps in one call: states[k] -> states[k+1]
 * for k = 0..n-1 (n+1 device buffers, host array; ping-pong by repeating
 * two pointers).  force_partial (optional) receives step k's partials at
 * force_partial + k * partial_stride doubles.  Replaces the body of the
 * reference's jax.lax.scan time loops (examples/10_cd_convergence.py:69-83,
 * examples/11_optimisation_v2.py:152-160, jax_lbm/state.py:174-230) driven
 * one launch per host call. */
int lbm_step_chain(const lbm_step_desc* d, int n, void* const* states, double* force_partial,
                   int64_t partial_stride, lbm_stream_t stream);

/* force[t][0:3] = fixed-order float64 sum of the `slots` partials of row t,
 * for T rows.  Completes momentum_exchange (forces.py:37-120) for the fused
 * path. */
int lbm_force_reduce(const double* partials, int T, int slots, double* force,
                     lbm_stream_t stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): lbm_step_chain followed by lbm_force_reduce in one call, n full
 * steps states[k] -> states[k+1] (host array of n+1 device pointers) and the
 * per-step momentum-exchange force force[n][3] (float64), with the
 * per-workgroup force partials held in a library-owned device array of
 * LBM_NOALLOC_PARTIALS doubles per GPU.  Allocates nothing (the array is a
 * file-scope device variable; its address is looked up once per device).
 * Returns LBM_E_ARG when n * lbm_step_force_slots(d) * 3 exceeds
 * LBM_NOALLOC_PARTIALS.  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (different streams, or host threads) race. */
#define LBM_NOALLOC_PARTIALS (1u << 20)
int lbm_step_chain_forces_noalloc(const lbm_step_desc* d, int n, void* const* states,
                                  double* force, lbm_stream_t stream);

/* ---- Adjoint of the fused timestep --------------------------------------
 * The reference differentiates its rollout with jax.grad / value_and_grad
 * (examples/11_optimisation_v2.py:176, tests/test_gradients_f64.py:129-217);
 * these entry points are the hand-written transpose of lbm_step that
 * replaces that XLA transpose.  Cotangent buffers ("lambda") are float for
 * the float32/half storages and double for LBM_F64, in lattice storage. */

typedef struct lbm_adjoint_desc {
  double* dq_acc;          /* 18 x n_links: += dL/dq of each link (may be null) */
  void* mu_link;           /* 19 x n_links scratch: per-step link-cell cotangents
                              (with fold_links: their additions to the link
                              cells' own entries) */
  double* phi_acc;         /* N: += dL/dporosity (soft bounce; may be null) */
  double* u_acc;           /* 3 x (ny*nz): += dL/du_in per inlet cell (may be null) */
  int32_t n_targets;       /* size of the scatter plan */
  const int32_t* tgt_index;/* n_targets: lambda entry i*pop_stride + cell */
  const int32_t* tgt_start;/* n_targets + 1: CSR offsets into src/coef */
  const int32_t* src;      /* >= 0: mu_link index i*n_links + b; < 0: force term of direction -1-src */
  const double* coef;      /* multiplier of mu_link[src] */
  int32_t lam_storage;     /* cotangent buffers: 0 = the arithmetic type (float,
                              double for LBM_F64); LBM_F16DEV = IEEE binary16
                              holding lambda * lam_scale (float arithmetic only) */
  double lam_scale;        /* power of two for LBM_F16DEV (<= 0 means 1) */
  void* mu_link_alt;       /* second 19 x n_links buffer for fold_links (a chain
                              alternates mu_link and mu_link_alt) */
  int32_t fold_links;      /* 0: the link additions come from lbm_adjoint_scatter
                              and the plan above, one launch per step;
                              1: no plan -- lbm_step_adjoint writes the Bouzidi
                              second-neighbour terms with its owner writes and
                              lbm_adjoint_link_finish adds the link cells' own
                              terms; lbm_adjoint_chain folds the latter into the
                              next step's loads and finishes its last step */
} lbm_adjoint_desc;

/* One adjoint step.  mode LBM_MODE_FULL: state = P_t,
 * lam_in = dL/dP_{t+1} -> lam_out = owner part of dL/dP_t.  mode
 * LBM_MODE_STREAM: lam_in = dL/dS_{t+1} (the epilogue).  mode
 * LBM_MODE_COLLIDE: state = S_0, lam_in = dL/dP_0 -> lam_out = dL/dS_0 (in the
 * arithmetic type, whatever lam_storage is).
 * gF (3 doubles on the device, may be null) = dL/dF_t of this step's
 * momentum-exchange force.  Requires periodic_yz and nx >= 2.  Modes FULL and
 * STREAM must be followed by lbm_adjoint_scatter, or with a->fold_links by
 * lbm_adjoint_link_finish. */
int lbm_step_adjoint(const lbm_step_desc* d, const lbm_adjoint_desc* a, int mode,
                     const void* state, const void* lam_in, void* lam_out,
                     const double* gF, lbm_stream_t stream);

/* Adds the Bouzidi-link and force-term cotangents to lam (same step's gF). */
int lbm_adjoint_scatter(const lbm_step_desc* d, const lbm_adjoint_desc* a, void* lam,
                        const double* gF, lbm_stream_t stream);

/* a->fold_links = 1: completes a lbm_step_adjoint as lbm_adjoint_scatter does,
 * without a plan: adds to lam the Bouzidi and force cotangents of the link
 * cells' own populations (streaming.py:155-178, forces.py:95-118), which that
 * step stored per link cell in mu (19 x n_links; null = a->mu_link).  One
 * thread per link cell. */
int lbm_adjoint_link_finish(const lbm_step_desc* d, const lbm_adjoint_desc* a, const void* mu,
                            void* lam, lbm_stream_t stream);

/* Adjoint of n consecutive FULL steps: for k = n-1 down to 0,
 * lbm_step_adjoint(FULL, states[k]) followed by lbm_adjoint_scatter, with
 * gF + 3k the force cotangent of step k.  With a->fold_links, one launch per
 * step: each launch adds the previous one's link-cell additions to the
 * cotangents it loads (mu_link and mu_link_alt alternate), and the chain
 * ends with lbm_adjoint_link_finish.  On entry lam_a holds the
 * cotangent of the state after the last step; the two buffers alternate, so
 * the result (the cotangent of states[0]) is in lam_a for even n and in
 * lam_b for odd n.  The transpose of the reference's scan loop under
 * jax.grad (examples/11_optimisation_v2.py:176). */
int lbm_adjoint_chain(const lbm_step_desc* d, const lbm_adjoint_desc* a, int n, const void* const* states,
                      void* lam_a, void* lam_b, const double* gF, lbm_stream_t stream);

/* Convert between storage formats (e.g. float32 <-> half deviations) for a
 * full state; src and dst both use d->pop_stride and d->row_stride.  Encoding
 * into a 16-bit format follows d->rounding (stochastic: the dither of
 * d->seed and d->step, drawn as by a step of its own). */
int lbm_convert_state(const lbm_step_desc* d, int src_dtype, const void* src,
                      int dst_dtype, void* dst, lbm_stream_t stream);

/* ---- Reverse-mode products of the per-operator surface ---------------------
 * Each replaces what jax.grad does through the cited reference function
 * (JAX's transposition of its jnp graph; the reference has no hand-written
 * derivatives).  Dense (19, nz, ny, nx) / (nz, ny, nx) / (3, nz, ny, nx)
 * arrays as in the forward per-operator entry points; g_out is the
 * cotangent of the forward output, the g_* outputs are the cotangents of the
 * inputs (optional ones may be NULL).  Every output entry is written by one
 * thread: deterministic, no atomics. */

/* equilibrium (jax_lbm/core/equilibrium.py:90-151): g_rho, g_u from g_feq. */
int lbm_equilibrium_vjp(int dtype, const void* rho, const void* u, const void* g_feq,
                        void* g_rho, void* g_u, int nx, int ny, int nz, lbm_stream_t stream);

/* compute_density (equilibrium.py:22-48): g_f_i = g_rho. */
int lbm_density_vjp(int dtype, const void* g_rho, void* g_f, int nx, int ny, int nz,
                    lbm_stream_t stream);

/* compute_velocity with rho an input (equilibrium.py:51-87): g_f, and g_rho
 * (optional). */
int lbm_velocity_vjp(int dtype, const void* f, const void* rho, const void* g_u, void* g_f,
                     void* g_rho, int nx, int ny, int nz, lbm_stream_t stream);

/* bgk / mrt (collision.py:46-184), collision = LBM_COLL_BGK / LBM_COLL_MRT,
 * tau scalar or tau_field: g_f, and the per-cell dL/dtau (optional). */
int lbm_collide_vjp(int dtype, int collision, const void* f, double tau, const void* tau_field,
                    const void* g_out, void* g_f, void* g_tau, int nx, int ny, int nz,
                    lbm_stream_t stream);

/* smagorinsky_tau (collision.py:187-270): g_f from the cotangent of tau_eff,
 * and the per-cell dL/dtau (optional).  At |Pi_neq| = 0 the square root's
 * derivative is taken as 0 (JAX gives NaN there). */
int lbm_smagorinsky_tau_vjp(int dtype, const void* f, double tau, double cs, const void* g_tau_eff,
                            void* g_f, void* g_tau, int nx, int ny, int nz, lbm_stream_t stream);

/* soft_bounce_back (jax_lbm/geometry/smooth.py:118-155): g_f, g_porosity (optional). */
int lbm_soft_bounce_back_vjp(int dtype, const void* f, const void* porosity, const void* g_out,
                             void* g_f, void* g_porosity, int nx, int ny, int nz,
                             lbm_stream_t stream);

/* hard bounce where solid >= 1 (jax_lbm/state.py:146-149). */
int lbm_hard_bounce_vjp(int dtype, const int32_t* solid, const void* g_out, void* g_f, int nx,
                        int ny, int nz, lbm_stream_t stream);

/* stability guard (jax_lbm/state.py:158-169): g_f = g_out where the guard
 * input f is stable, 0 at reset cells (the reset value is a constant). */
int lbm_stability_guard_vjp(int dtype, const void* f, const void* g_out, void* g_f, int nx,
                            int ny, int nz, lbm_stream_t stream);

/* stream with Bouzidi links (jax_lbm/core/streaming.py:40-185): g_f, and
 * g_q (optional; needs f) on the taken branch inside the clip bounds. */
int lbm_stream_vjp(int dtype, const void* f, const void* q, int periodic_yz, const void* g_out,
                   void* g_f, void* g_q, int nx, int ny, int nz, lbm_stream_t stream);

/* zou_he_inlet (jax_lbm/core/boundary.py:35-124), in place on g (a copy of
 * the output cotangent, which becomes the input cotangent); per (y, z) row
 * dL/du (3 doubles, optional) and dL/drho_target (optional). */
int lbm_zou_he_inlet_vjp(int dtype, const void* f, double ux, double uy, double uz,
                         int has_rho_target, double rho_target, void* g, double* g_u_rows,
                         double* g_rho_target_rows, int nx, int ny, int nz, lbm_stream_t stream);

/* zou_he_outlet (boundary.py:127-200), in place on g; per-row dL/drho_out (optional). */
int lbm_zou_he_outlet_vjp(int dtype, double rho_out, void* g, double* g_rho_out_rows, int nx,
                          int ny, int nz, lbm_stream_t stream);

/* momentum_exchange(_decomposed) (jax_lbm/core/forces.py:37-196): g_post and
 * g_streamed from dL/dF (3 doubles) and, for the pressure part, dL/dF_p
 * (3 doubles, optional; needs f_post). */
int lbm_momentum_exchange_vjp(int dtype, const void* f_post, const void* q, const double* g_force,
                              const double* g_pressure, void* g_post, void* g_streamed, int nx,
                              int ny, int nz, lbm_stream_t stream);

/* Sphere geometry (jax_lbm/geometry/smooth.py:158-248 for mode 1,
 * primitives.py:56-174 for mode 0; same arguments as lbm_sphere_geometry,
 * solid = its output): out[4] = dL/d(radius, cx, cy, cz) in float64 from the
 * cotangents of the porosity (N, optional) and of q (19 x N, optional).
 * The solid mask and the link set carry no cotangent (comparisons).
 * workspace: lbm_sphere_geometry_vjp_workspace_bytes() bytes; fixed-order
 * reduction. */
size_t lbm_sphere_geometry_vjp_workspace_bytes(void);
int lbm_sphere_geometry_vjp(int mode, int coord_dtype, int out_dtype, double cx, double cy, double cz,
                            double radius, double sharpness, double disc_floor, double denom_eps,
                            const int32_t* solid, const void* g_porosity, const void* g_q, double* out,
                            double* workspace, int nx, int ny, int nz, lbm_stream_t stream);

/* drag_coefficient (forces.py:237-276) of a (T, 3) force history:
 * dL/dF from dL/dCd (T doubles). */
int lbm_drag_coefficient_vjp(const double* force, const double* g_cd, int T,
                             double inlet_velocity, const double* area, double rho,
                             double* g_force, lbm_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* LBM_HIP_H_ */
