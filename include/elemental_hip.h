/*
 * elemental_hip.h -- C ABI of libelemental_hip.so, the MI355X (gfx950)
 * elemental-generation engine.
 *
 * The reference has no FFI; its drop-in seam for this path is Python: the
 * backend module object (lattice/backend.py:7-33) and the ElementalGenerator
 * contract (lattice/generator/elemental.py:17-26 constructor, :102-105 load,
 * :290-338 calc).  Each entry point below replaces one piece of that
 * contract; the Python class easydistillation_amd.ElementalGenerator binds
 * them through ctypes.
 *
 * Conventions
 *   - every function returns int: ED_OK (0) or a negative ED_ERR_* code;
 *     ed_last_error(ctx) gives the message of the last failure on ctx.
 *   - no C++ exception crosses this boundary.
 *   - a context binds one HIP device and one stream; it is not re-entrant
 *     (one context per host thread).  Calls on one context are ordered on its
 *     stream.  Calls on two contexts that share a device run independently.
 *   - caller-owned pointers are only read during the call; the library copies
 *     what it keeps (gauge links, eigenvectors, weights).
 *   - complex arrays are interleaved (re, im); ED_C128 = complex<double>,
 *     ED_C64 = complex<float>.
 */
#ifndef ELEMENTAL_HIP_H
#define ELEMENTAL_HIP_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ed_context ed_context;

enum ed_status {
  ED_OK = 0,
  ED_ERR_ARG = -1,    /* invalid argument (shape, dtype, range, null pointer)      */
  ED_ERR_STATE = -2,  /* call order: lattice / gauge / eigvecs / momenta not set   */
  ED_ERR_HIP = -3,    /* a HIP runtime call failed                                 */
  ED_ERR_ALLOC = -4,  /* device or host allocation failed                          */
  ED_ERR_UNSUPPORTED = -5
};

enum ed_dtype { ED_C128 = 0, ED_C64 = 1 };

enum ed_mode {
  ED_MODE_FP64 = 0,  /* complex128 stencil and contraction (reference numerics)          */
  ED_MODE_MIXED = 1  /* hops in FP64 arithmetic stored as complex64; contraction with   */
                     /* the split sum on FP32 MFMA and the momentum phase on FP16 MFMA   */
                     /* (three-term hi/lo split, FP32 partial sums per 32 sites) from    */
                     /* V3 = 4096 sites on, FP64 kernels below; FP64 phase table, zero-  */
                     /* momentum sum and cross-chunk spatial reduction                   */
};

/* ed_create / ed_destroy: one context per device and host thread.
 * Replaces the per-instance state of ElementalGenerator.__init__
 * (lattice/generator/elemental.py:17-59).  stream: the context creates its
 * own stream (default priority) unless ed_set_stream is called. */
int ed_create(int device, ed_context** ctx);
int ed_destroy(ed_context* ctx);
const char* ed_last_error(const ed_context* ctx);

/* Use an external stream (e.g. torch.cuda.current_stream().cuda_stream);
 * NULL restores the context's own stream. */
int ed_set_stream(ed_context* ctx, void* hip_stream);

/* latt_size = [Lx, Ly, Lz, Lt] and Ne = eigenvector.Ne
 * (elemental.py:30,53; lattice/preset.py:16-19). */
int ed_set_lattice(ed_context* ctx, int Lx, int Ly, int Lz, int Lt, int Ne);

/* Spatial links U[3][Lt][Lz][Ly][Lx][3][3] (mu = x, y, z), row-major,
 * U[..][a][b] acting on colour b -- the array elemental.py:103 builds from
 * gauge_field.load(key)[:].transpose(4,0,1,2,3,5,6)[:3].
 * dtype ED_C128 or ED_C64; on_device != 0 means U is a device pointer. */
int ed_set_gauge(ed_context* ctx, const void* U, int dtype, int on_device);

/* Eigenvectors V[Lt][Ne][Lz][Ly][Lx][3] (eigenvector.load(key)[t, e],
 * elemental.py:297-298).  Stored as complex64: complex128 input is rounded
 * to complex64 exactly as the reference's c8 staging buffer does
 * (elemental.py:55).  The call also measures, in FP64, the orthonormality
 * of what it stored, max over t, e, f of |(V_t^dag V_t)[e][f] - delta_ef|
 * (read by ed_orthonormality; gated by ED_OPT_ORTHO_GATE). */
int ed_set_eigvecs(ed_context* ctx, const void* V, int dtype, int on_device);

/* FP64 max|V^dag V - 1| of the eigenvectors the last ed_set_eigvecs stored
 * (no counterpart in the reference, which reads V without a check;
 * elemental.py:297-298).  ED_ERR_STATE before any ed_set_eigvecs. */
int ed_orthonormality(ed_context* ctx, double* max_dev);

/* Momentum list n[P][3] = (nx, ny, nz); output index p follows this order
 * and Phi_n(x) = exp(+2 pi i (nx x/Lx + ny y/Ly + nz z/Lz))
 * (lattice/insertion/phase.py:11-13,41-46; elemental.py:322). */
int ed_set_momenta(ed_context* ctx, const int32_t* n, int P);

/* num_nabla in [0, 3]: D = (3^(num_nabla+1)-1)/2 derivative combinations in
 * the index order of lattice/insertion/derivative.py:23-33
 * (elemental.py:48-49). */
int ed_set_derivatives(ed_context* ctx, int num_nabla);

/* Displacement elementals instead of derivative elementals: replaces
 * DisplacementElementalGenerator(..., distance, ...)
 * (lattice/generator/displacement_elemental.py:11-51, _D :53-71,
 * calc :78-96).  Output n = 0 .. distance of ed_compute (D = distance + 1) is
 *   E[n][p][e][f] = sum_{x,c} Phi_p(x) conj(V[e](x)_c) M_n[f](x)_c,
 * M_0 = V and M_n the mean of the six straight paths of length n,
 * U_mu(x) ... U_mu(x + (n-1) mu) V(x + n mu) and
 * U_mu(x - mu)^dag ... U_mu(x - n mu)^dag V(x - n mu), mu = x, y, z.
 * distance in [0, 64]; ed_set_derivatives returns to derivative elementals. */
int ed_set_displacement(ed_context* ctx, int distance);

/* Blending weights coeff[Ne][Ne] (float64, host), multiplied elementwise into
 * every E[d][p] (elemental.py:61-100,330-337); NULL disables blending. */
int ed_set_weights(ed_context* ctx, const double* coeff);

/* Elementals for timeslices t0 .. t0+nt-1:
 *   out[nt][D][P][Ne][Ne] complex128,
 *   E[d][p][e][f] = sum_splits (-1)^{#right} sum_{x,c} Phi_p(x)
 *                   conj(W_left[e](x)_c) W_right[f](x)_c
 * (elemental.py:290-338; calc(t) returns out[t - t0]).
 * mode: ED_MODE_FP64 or ED_MODE_MIXED.  out_on_device != 0: out is a device
 * pointer written on the context stream (the call returns after enqueueing);
 * otherwise out is host memory and the call returns when it is filled. */
int ed_compute(ed_context* ctx, int t0, int nt, int mode, void* out, int out_on_device);

/* No-allocation form of ed_compute.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.
 *   - out_device: a device array [nt][D][P][Ne][Ne] complex128; the call
 *     enqueues on the context stream and returns without waiting.
 *   - the displaced vectors go to a fixed workspace of
 *     ED_NOALLOC_SCRATCH_BYTES compiled into the library (one per device,
 *     held until the process ends) instead of the context's workspace;
 *     timeslices run in batches that fit it.  A lattice whose single
 *     timeslice needs more (slots x NeP x 3 x V3P x 8 bytes in the mixed
 *     form, x 16 in FP64, NeP = Ne rounded up to 16, V3P = V3 rounded up to
 *     32) fails with ED_ERR_UNSUPPORTED.
 *   - the plan tables must be built: ed_compute or ed_query once after the
 *     last ed_set_* / ed_set_option call; otherwise ED_ERR_STATE.
 *   - nothing is allocated or freed and no event is recorded; the used part of
 *     the workspace is cleared by a fill kernel on each call (the workspace
 *     keeps the previous call's contents).
 *   - every call on one device shares that device's workspace: calls that
 *     overlap on one device (two streams, two contexts, two host threads)
 *     race and give wrong results; order them on one stream.
 * Results equal ed_compute's bit for bit. */
#define ED_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
int ed_compute_noalloc(ed_context* ctx, int t0, int nt, int mode, void* out_device);

/* Block until all work enqueued on the context stream is done. */
int ed_synchronize(ed_context* ctx);

/* Integer tuning / testing knobs (no counterpart in the reference):
 *   ED_OPT_HERMITIAN      1 (default): derive E[rev d] from E[d] through
 *                         E_{rev d}[p] = (-1)^{|d|} E_d[-p]^dagger; 0: compute all d.
 *   ED_OPT_TBATCH         timeslices per batch (0 = automatic: a multiple of 8 when 8 fit the
 *                         budget, exactly 8 when a timeslice needs >= 1 GiB of workspace).
 *   ED_OPT_WS_MB          workspace budget in MiB for the automatic batch (default 65536; never
 *                         more than the context's workspace plus 3/4 of the free device memory).
 *   ED_OPT_PROFILE        1: record HIP events around every kernel launch.
 *   ED_OPT_MIXED_STAGE2   mixed form's phase contraction: 1 FP16 MFMA on a three-term
 *                         hi/lo split, 0 FP32 MFMA, -1 (default) FP16 when V3 >= 4096; with
 *                         the default, lattices with V3 < 4096 run the FP64 kernels (there
 *                         the FP32 stage-1 sum reaches ~1.2e-6 on (d, p) blocks that are
 *                         small against their derivative's largest block).
 *   ED_OPT_EXACT_V        1: the next ed_set_eigvecs also keeps V in complex128 and the FP64
 *                         form and ed_backward use it unrounded (for derivative checks by
 *                         finite differences); 0 (default): V is rounded to complex64 as
 *                         in the reference.
 *   ED_OPT_MIXED_WS       FP16 stage 2 kernel: 1 (default) warp-specialized (4 waves run the FP32
 *                         split sum, 4 the FP16 phase contraction, overlapping on each SIMD),
 *                         2 warp-specialized with the split sum on FP16 MFMA too (the
 *                         workspace then holds FP16 hi/lo pairs of W 2^k, k chosen from the
 *                         largest eigenvector component; up to 16 (split, colour) terms and
 *                         V3 >= 4096, else kind 1 runs),
 *                         0 every wave runs both stages; derivatives with more (split,
 *                         colour) terms than the chosen kernel takes use the latter.
 *                         Kinds 0 and 2 address one timeslice's workspace with 32-bit
 *                         offsets; from 2^31 bytes per timeslice (e.g. 32^3 with Nev 256 and
 *                         two derivatives) kinds 0 and 2 run as kind 1, which has a
 *                         64-bit-address form (derivatives of more than four split terms then
 *                         take the FP32 phase contraction).
 *   ED_OPT_WIDE_WS        1: kind 1 uses its 64-bit-address form below 2^31 bytes too (same
 *                         results bit for bit; for testing).
 *   ED_OPT_ORTHO_GATE     n in [1, 15]: ed_set_eigvecs fails with ED_ERR_ARG (and the context
 *                         holds no eigenvectors) when the FP64 max|V^dag V - 1| of the stored
 *                         eigenvectors exceeds 10^-n; 0 (default): measured and reported by
 *                         ed_orthonormality only.
 *   ED_OPT_STRIP          tile order of the contraction: its 16 x 16 (e, f) tiles run in strips
 *                         of this many e-tile rows, f-tile by f-tile inside a strip (0, default:
 *                         4 for the mixed kernels, 2 for the FP64 / FP32 kernel).  Chosen by
 *                         ed_tune.
 *   ED_OPT_ITEM_ORDER     contraction work items over the 8 XCDs: 0 (default) contiguous ranges
 *                         per XCD, 1 round robin.  Chosen by ed_tune.
 *   ED_OPT_H16_UNIT       1 (default): ED_OPT_MIXED_WS 2 runs the unit plan where it applies
 *                         (every split coefficient +1: the workspace stores each derivative
 *                         level k times i^k, splits of multiplicity m enter m times, results
 *                         are multiplied back by i^{|d|}); 0: coefficients scale the operands.
 *   ED_OPT_HOP_LDS        derivative hops: 1 LDS-staged (an 8 x 8 (x, y) tile with its halo and
 *                         the planes z - 1, z, z + 1 of 128 bytes of eigenvectors per site and
 *                         colour in LDS, marching over z) where Lx and Ly are multiples of 8 and
 *                         Lz >= 3; 0 one wave per site; -1 (default) LDS-staged for the FP64
 *                         form's 16-byte workspace elements, one wave per site for the 8-byte
 *                         ones of the mixed form.  Same results bit for bit. */
enum ed_option {
  ED_OPT_HERMITIAN = 1,
  ED_OPT_TBATCH = 2,
  ED_OPT_WS_MB = 3,
  ED_OPT_PROFILE = 4,
  ED_OPT_MIXED_STAGE2 = 5,
  ED_OPT_EXACT_V = 6,
  ED_OPT_MIXED_WS = 7,
  ED_OPT_WIDE_WS = 8,
  ED_OPT_ORTHO_GATE = 9,
  ED_OPT_STRIP = 10,
  ED_OPT_ITEM_ORDER = 11,
  ED_OPT_H16_UNIT = 12,
  ED_OPT_HOP_LDS = 13
};
int ed_set_option(ed_context* ctx, int option, int64_t value);
int ed_get_option(ed_context* ctx, int option, int64_t* value);

/* Autotuner of the contraction (no counterpart in the reference; the caller
 * is easydistillation_amd.tune, which keeps the choices per shape in a JSON
 * file): times ed_compute of timeslices t0 .. t0+nt-1 into a device buffer,
 * reps runs after one warm-up, for every candidate -- tile-order strip
 * height (ED_OPT_STRIP 1, 2, 4, 8) x item order over the XCDs
 * (ED_OPT_ITEM_ORDER 0, 1) x, in the mixed form, each distinct mixed kernel
 * that ED_OPT_MIXED_WS 1, 0 and 2 select for this shape -- leaves the fastest
 * set on the context and returns its time per run in *best_ms. */
int ed_tune(ed_context* ctx, int mode, int t0, int nt, int reps, double* best_ms);

/* Kernel timing collected while ED_OPT_PROFILE = 1, per kernel family:
 *   ED_KERNEL_STENCIL  covariant hops (one launch per derivative level),
 *   ED_KERNEL_CONTRACT the two-stage contraction,
 *   ED_KERNEL_STAGE    staging of V into the workspace,
 *   ED_KERNEL_ADJOINT  the contraction adjoint of ed_backward.
 * total_ms is the sum of HIP-event durations on the launch stream and
 * launches their number; both are reset by ed_profile_reset. */
enum ed_kernel { ED_KERNEL_STENCIL = 0, ED_KERNEL_CONTRACT = 1, ED_KERNEL_STAGE = 2, ED_KERNEL_ADJOINT = 3 };
int ed_profile_read(ed_context* ctx, int kernel, double* total_ms, int64_t* launches);
int ed_profile_reset(ed_context* ctx);

/* Plan figures for the current lattice / derivative / momentum settings
 * (used for roofline accounting; no counterpart in the reference):
 *   ED_Q_CONTRACT_FLOPS_PER_T  MFMA flops the contraction executes per timeslice
 *   ED_Q_ALGO_FLOPS_PER_T      8 Ne^2 (3 V3) S P, the direct split-by-split
 *                              formulation (S = number of split terms)
 *   ED_Q_PRIMARY_DERIVATIVES   derivatives contracted (the others follow by
 *                              Hermiticity)
 *   ED_Q_MOMENTUM_BLOCKS       blocks of 16 base momenta
 *   ED_Q_HOPS_PER_T            covariant hops per timeslice (D - 1)
 *   ED_Q_SPLITS                split terms over all derivatives (S)
 *   ED_Q_MIXED_F32_FLOPS_PER_T FP32 MFMA flops of the mixed FP16-stage-2 kernel per timeslice
 *   ED_Q_MIXED_F16_FLOPS_PER_T FP16 MFMA flops of the same kernel per timeslice
 *   ED_Q_MIXED_KERNEL          the contraction the mixed form runs: 0 the FP64 kernels
 *                              (V3 < 4096), 1 FP32 phase contraction, 2 FP16 phase contraction
 *                              with uniform waves, 3 warp-specialized (32-bit workspace
 *                              offsets), 4 warp-specialized with 64-bit addresses, 5 FP16
 *                              stage 1
 * (ED_Q_CONTRACT_FLOPS_PER_T counts the FP64 / FP32-stage-2 kernels.) */
enum ed_query {
  ED_Q_CONTRACT_FLOPS_PER_T = 1,
  ED_Q_ALGO_FLOPS_PER_T = 2,
  ED_Q_PRIMARY_DERIVATIVES = 3,
  ED_Q_MOMENTUM_BLOCKS = 4,
  ED_Q_HOPS_PER_T = 5,
  ED_Q_SPLITS = 6,
  ED_Q_MIXED_F32_FLOPS_PER_T = 7,
  ED_Q_MIXED_F16_FLOPS_PER_T = 8,
  ED_Q_MIXED_KERNEL = 9
};
int ed_query(ed_context* ctx, int what, int64_t* value);

/* Reverse pass of ed_compute's FP64 form (no counterpart in the reference;
 * the caller is easydistillation_amd.autograd.ElementalFunction.backward).
 * Given grad_E = dL/dRe(E) + i dL/dIm(E) for a real loss L, laid out as
 * ed_compute's out ([nt][D][P][Ne][Ne] complex128), writes
 *   grad_V [nt][Ne][Lz][Ly][Lx][3]        dL/dRe(V) + i dL/dIm(V)
 *   grad_U [3][nt][Lz][Ly][Lx][3][3]      dL/dRe(U) + i dL/dIm(U)
 * for timeslices t0 .. t0+nt-1 (either output may be NULL).  The displaced
 * vectors are recomputed in FP64, the blending weights of ed_set_weights
 * apply, and the gradient passes straight through the complex64 rounding of
 * V unless ED_OPT_EXACT_V is set.  grad_on_device / out_on_device: the
 * pointers are device memory (the call returns after enqueueing) or host
 * memory (the call returns when the outputs are filled). */
int ed_backward(ed_context* ctx, int t0, int nt, const void* grad_E, int grad_on_device, void* grad_V, void* grad_U,
                int out_on_device);

/* Link preprocessing on the context's resident spatial links (run after
 * ed_set_gauge, before ed_compute; the links are replaced in place).
 *
 * ed_project_su3 replaces ElementalGenerator.project_SU3
 * (lattice/generator/elemental.py:107-117): the polar iteration
 * U <- (U + (U^-1)^dagger)/2, applied to every link until
 * max|U - (U^-1)^dagger| <= 1e-15 and max|U U^dagger - 1| <= 1e-15 hold for all
 * links (the reference's whole-array test).  *iterations (may be NULL)
 * receives the number of updates; more than max_iter updates fail with
 * ED_ERR_UNSUPPORTED, leaving the last iterate in place.
 *
 * ed_stout_smear replaces ElementalGenerator.stout_smear(nstep, rho)
 * (elemental.py:175-242,264-277): nstep 3-D stout steps with staples over the two
 * other spatial directions in the link's own timeslice,
 * U <- exp(iQ) U, Q = (i/2)(Omega^dagger - Omega) made traceless,
 * Omega = rho C U^dagger.  Where the reference divides 0/0 (Q = 0) the
 * identity limit exp(i0) = 1 is used.
 *
 * ed_get_gauge copies the current links out as complex128
 * U[3][Lt][Lz][Ly][Lx][3][3] (host or device pointer). */
int ed_project_su3(ed_context* ctx, int max_iter, int* iterations);
/* Gauge-covariant 3-D Laplacian of timeslice t on the resident links, applied
 * to nvec vectors (the operator of EigenvectorGenerator,
 * lattice/generator/eigenvector.py:11-26 _Laplacian, whose low modes are
 * the eigenvectors V(t)):
 *   out[e](x) = 6 F[e](x) - sum_mu [U_mu(x) F[e](x + mu) + U_mu(x - mu)^dagger F[e](x - mu)]
 * F, out: [nvec][Lz][Ly][Lx][3] complex128, both host (on_device = 0) or both
 * device pointers (the call then returns after enqueueing on the context
 * stream). */
int ed_laplacian(ed_context* ctx, int t, const void* F, void* out, int nvec, int on_device);
int ed_stout_smear(ed_context* ctx, int nstep, double rho);
int ed_get_gauge(ed_context* ctx, void* U, int on_device);

/* Batched form (no counterpart in the reference, whose callers loop over
 * configurations and timeslices one calc(t) at a time): several contexts --
 * lattices of different size, Ne, momenta and derivative depth, each with its
 * inputs set -- computed together.  One run enqueues, for every context, the
 * staging and hop kernels over all its timeslices and then one persistent,
 * grouped contraction launch over the work items of every context; with
 * ED_BOPT_GRAPH = 1 (default) this sequence is captured once into a hipGraph
 * and each run is a single graph launch (recaptured automatically when a
 * context's inputs, plan or buffers change).  Each context holds the
 * workspace of all its timeslices at once.
 *   ed_batch_add: out_device is a device array [Lt][D][P][Ne][Ne] complex128
 *                 receiving every timeslice of ctx in mode (FP64 or MIXED).
 *   ed_batch_run: enqueue on the batch's own stream; ed_batch_synchronize waits.
 *                 The run first waits (stream-ordered, no host wait) for
 *                 work already enqueued on each member context's stream, and
 *                 each member's stream then waits for the run, so calls on a
 *                 context before and after a run are ordered with it.
 *                 The batch keeps the member context pointers: destroy the
 *                 batch before any of its contexts.
 *   ed_batch_time: mean duration (HIP events) of reps runs after one warm-up.
 *   ed_batch_tune: autotuner -- times every launch shape (persistent grid of
 *                 1, 2 or 4 workgroups per CU, or one workgroup per work item)
 *                 x item order (contiguous ranges per XCD, or round robin
 *                 over XCDs) x tile-order strip height (1, 2, 4, 8) x
 *                 ED_BOPT_H16 (0, and 1 when some mixed member fits it) with
 *                 reps runs each, keeps the fastest and returns its time. */
typedef struct ed_batch ed_batch;
enum ed_batch_option {
  ED_BOPT_GRID_PER_CU = 1, /* persistent workgroups per CU (0: one per work item; default 2) */
  ED_BOPT_ORDER = 2,       /* 0: contiguous item ranges per XCD (default), 1: round robin     */
  ED_BOPT_GRAPH = 3,       /* 1: capture and replay a hipGraph (default), 0: direct launches  */
  ED_BOPT_STRIP = 4,       /* tile-order strip height for every member (0: each context's own) */
  ED_BOPT_H16 = 5          /* mixed members' split sum on FP16 MFMA (the hi/lo FP16 workspace of
                              ED_OPT_MIXED_WS 2, one workgroup per work item): -1 (default) as each
                              context's ED_OPT_MIXED_WS, 1 wherever it fits (V3 >= 4096, at most two
                              k-groups, < 2^31 workspace bytes per timeslice), 0 never */
};
int ed_batch_create(int device, ed_batch** b);
int ed_batch_destroy(ed_batch* b);
const char* ed_batch_last_error(const ed_batch* b);
int ed_batch_add(ed_batch* b, ed_context* ctx, int mode, void* out_device);
int ed_batch_set_option(ed_batch* b, int option, int64_t value);
int ed_batch_get_option(ed_batch* b, int option, int64_t* value);
int ed_batch_run(ed_batch* b);
int ed_batch_synchronize(ed_batch* b);
int ed_batch_time(ed_batch* b, int reps, double* ms_per_run);
int ed_batch_tune(ed_batch* b, int reps, double* best_ms);

/* Library version string (static storage). */
const char* ed_version(void);

#ifdef __cplusplus
}
#endif
#endif /* ELEMENTAL_HIP_H */
