/*
 * cfft_hip.h — C ABI of the MI355X (gfx950) KRFFT / pseudo-spectral diffusion
 * library (libcfft_hip.so).
 *
 * Each entry point replaces one operation of the reference Julia package
 * liuyxpp/CrystallographicFFT.jl; the reference interface it stands in for is
 * cited beside it (path:line in that repository).  Plain pointers and sizes
 * only.  Every function returns a status code (CFFT_OK = 0); the message of
 * the last failure on the calling thread is available from cfft_last_error.
 *
 * Conventions
 *   - Grids are column-major (x fastest) with 0-based coordinates, as the
 *     reference's Julia arrays.  A subgrid field f0 has M0*M1*M2 entries,
 *     M = N / L (L from auto_L); batched fields are stored one after another.
 *   - Operations: int32 array of nops x 12 entries, R row-major (9) then t (3)
 *     in grid units: x -> R x + t (mod N).  2-D operations use the top-left
 *     2x2 block of R (R[2][2] = 1) and t[2] = 0.
 *   - Spectral coefficients: interleaved complex (re, im), one per spectral-ASU
 *     point, ordered lexicographically by (h1, h2, h3) as the reference's
 *     SpectralIndexing.points.
 *   - Host-pointer entry points are blocking and use FP64 at the boundary.
 *     *_dev entry points take device pointers of the plan's precision (double
 *     for CFFT_FP64, float for CFFT_MIXED) for fields, except the interaction
 *     factors E = exp(-w dt/2), which are double in both precisions, and a
 *     hipStream_t; the work runs
 *     on exactly that stream (NULL = the null stream, which is PyTorch's
 *     default stream), in order with the caller's other work on it, and the
 *     call returns without synchronising.
 *   - A plan is not thread-safe; use one plan per stream.
 */
#ifndef CFFT_HIP_H
#define CFFT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference's error() / @assert sites) ------------- */
#define CFFT_OK 0
#define CFFT_ERR_INVALID 1          /* null pointer / bad argument              */
#define CFFT_ERR_NOT_COMMENSURATE 2 /* src/symmetry_ops.jl:41-43                */
#define CFFT_ERR_NOT_DIVISIBLE 3    /* src/krfft.jl:309-311, 1109-1111          */
#define CFFT_ERR_UNREACHABLE 4      /* src/krfft.jl:983, q_fused_krfft.jl:127    */
#define CFFT_ERR_FILL_MAP 5         /* src/q_fused_krfft.jl:777-778             */
#define CFFT_ERR_CENTERING 6        /* src/q_fused_krfft.jl:833-836, prefold:56 */
#define CFFT_ERR_ODD_M 7            /* src/centering_fold.jl:61                 */
#define CFFT_ERR_SIZE 8             /* src/q_fused_krfft.jl:998-1001            */
#define CFFT_ERR_UNKNOWN_METHOD 9   /* src/diffusion_solver.jl:250              */
#define CFFT_ERR_BOUNDS 10          /* src/spectral_indexing.jl:196-198         */
#define CFFT_ERR_UNSUPPORTED 11     /* FFT length / table size outside support  */
#define CFFT_ERR_HIP 12             /* HIP allocation / launch failure          */

/* ---- pipeline paths ---------------------------------------------------- */
#define CFFT_PATH_AUTO 0 /* M7 when centred and foldable, else M2             */
#define CFFT_PATH_M2 1   /* fwd + spectral K + bwd, plan_m2_scft               */
#define CFFT_PATH_M7 2   /* centred fwd + K + bwd, plan_centered_scft          */
#define CFFT_PATH_M2Q 3  /* Q-fused multiply on the M-grid, plan_m2_q          */
#define CFFT_PATH_M7Q 4  /* centring fold + Q multiply, plan_m7_scft           */
#define CFFT_PATH_QR 5   /* fold + Q multiply at the M-grid orbit representatives
                          * under the subgrid-preserving ops, then the orbit
                          * expansion of src/centering_fold.jl:630-731; diffusion
                          * and steps only (no F_spec output)                   */
#define CFFT_PATH_QS 6   /* fold + one kernel per plane pair: FFT y, z, the
                          * Q-fused multiply in separable form over the sign
                          * images of each point, IFFT z, y (the separable
                          * multiply of src/q_fused_krfft.jl:606-667 on the
                          * channel array); needs cubic H, L = (2,2,2), the 8
                          * representative ops diagonal with distinct sign
                          * patterns and an orthogonal cell; diffusion, steps
                          * and the forward transform (no F_spec backward)     */

#define CFFT_FP64 0  /* all arithmetic FP64                                   */
#define CFFT_MIXED 1 /* FP32 data/FFT/propagator, FP64 tables and reductions  */

typedef struct cfft_plan cfft_plan;

/* Message of the last failure on this thread; returns its full length. */
int cfft_last_error(char* buf, size_t n);

/* ---- plan-time maps (host only; no GPU needed) ------------------------- */

/* find_optimal_shift, src/asu.jl:57-123.  shift_out[dim] = b/N;
 * ops_out (nops x 12) = shifted ops t' = t + (R b - b) N. */
int cfft_find_optimal_shift(int32_t dim, const int32_t* N, const int32_t* ops, int32_t nops,
                            double* shift_out, int32_t* ops_out);

/* _count_special_positions_fast, src/asu.jl:148-214. */
int cfft_count_special_positions(int32_t dim, const int32_t* N, const int32_t* ops,
                                 int32_t nops, int64_t* count_out);

/* calc_spectral_asu, src/spectral_indexing.jl:22-176.  Writes at most cap
 * points (idx_out: cap x 3, 0-based; mult_out: cap); *n_out = total count.
 * Call with cap = 0 to query the size. */
int cfft_calc_spectral_asu(int32_t dim, const int32_t* N, const int32_t* ops, int32_t nops,
                           int32_t* idx_out, int32_t* mult_out, int64_t cap, int64_t* n_out);

/* auto_L, src/krfft.jl:225-279. */
int cfft_auto_L(int32_t dim, const int32_t* ops, int32_t nops, int32_t* L_out);

/* _select_rep_ops, src/krfft.jl:958-987: rep_out[prod(L)] op indices (0-based). */
int cfft_select_rep_ops(int32_t dim, const int32_t* N, const int32_t* ops, int32_t nops,
                        const int32_t* L, int32_t* rep_out);

/* detect_centering_type, src/symmetry_ops.jl:133-179 (0 P, 1 C, 2 A, 3 I, 4 F). */
int cfft_detect_centering(int32_t dim, const int32_t* N, const int32_t* ops, int32_t nops,
                          int32_t* cent_out);

/* _build_fill_map, src/q_fused_krfft.jl:733-783: map_out[N0*N1*N2] = 0-based
 * linear M-grid index of the subgrid point each full-grid point copies. */
int cfft_fill_map(int32_t dim, const int32_t* N, const int32_t* ops, int32_t nops,
                  const int32_t* L, int32_t* map_out);

/* ---- plans ------------------------------------------------------------- */

/* plan_m2_scft (src/q_fused_krfft.jl:948-986), plan_centered_scft
 * (src/centering_fold.jl:868-894), plan_m2_q (src/q_fused_krfft.jl:83-169),
 * plan_m7_scft (:826-855), plan_krfft(spec, ops) (src/krfft.jl:301-389) and
 * plan_m2_backward (:1101-1170) in one object.  ops are the SHIFTED ops
 * (output of cfft_find_optimal_shift); lattice is 3x3 column-major (columns
 * a, b, c); batch = number of fields processed per call (>= 1). */
int cfft_plan_create(const int32_t* N, const int32_t* ops, int32_t nops, double ds,
                     const double* lattice, int32_t path, int32_t precision, int32_t batch,
                     cfft_plan** out);

/* Plan metadata: info[18] = {M0, M1, M2, L0, L1, L2, n_spec, n_ch, centering,
 * d, path, precision, batch, H0, H1, H2, n_rep (path QR orbit representatives,
 * else 0), n_alive (channel elements per field)}. */
int cfft_plan_info(const cfft_plan* plan, int32_t* info);

/* Spectral-ASU points of the plan: idx_out (n_spec x 3), mult_out (n_spec);
 * K_out (n_spec) = diffusion kernel; rep_out (d) = representative op indices.
 * Any pointer may be NULL. */
int cfft_plan_tables(const cfft_plan* plan, int32_t* idx_out, int32_t* mult_out, double* K_out,
                     int32_t* rep_out);

/* update_m2_kernel! / update_kernel! (src/q_fused_krfft.jl:1036-1052,
 * src/centering_fold.jl:939-953): recompute K for a new ds (and lattice;
 * NULL keeps the plan's).  A rejected update (a non-orthogonal cell on path
 * QS) leaves the plan unchanged. */
int cfft_update_kernel(cfft_plan* plan, double ds, const double* lattice);

int cfft_plan_destroy(cfft_plan* plan);

/* ---- execute (host pointers, FP64, blocking) --------------------------- */

/* fft_reconstruct! / fft_reconstruct_centered! (src/krfft.jl:554-562,
 * src/centering_fold.jl:421-431): f0 (batch x M^3 real) -> F_spec
 * (batch x n_spec complex). */
int cfft_forward(cfft_plan* plan, const double* f0, double* F_spec);

/* execute_m2_backward! / execute_centered_ikrfft! (src/krfft.jl:1269-1277,
 * src/centering_fold.jl:745-764): F_spec -> f0 (real part). */
int cfft_backward(cfft_plan* plan, const double* F_spec, double* f0);

/* Subgrid spectrum G0 = FFT_M(f0) as assembled by assemble_G0!
 * (src/centering_fold.jl:277-296): f0 -> G0 (batch x M^3 complex). */
int cfft_forward_g0(cfft_plan* plan, const double* f0, double* G0);

/* execute_m2_scft! / execute_centered_scft! / execute_m2_q! / execute_m7_scft!
 * (src/q_fused_krfft.jl:989-1026, src/centering_fold.jl:908-931,
 * src/q_fused_krfft.jl:553-576, :870-895): f0 <- D f0 in place. */
int cfft_diffuse(cfft_plan* plan, double* f0);

/* step_diffusion! (src/diffusion_solver.jl:254-267) repeated nsteps times:
 * q <- e^{-w dt/2} D e^{-w dt/2} q.  D uses the plan's ds. */
int cfft_step(cfft_plan* plan, double* q, const double* w, double dt, int32_t nsteps);

/* subgrid_to_fullgrid! / fullgrid_to_subgrid! (src/q_fused_krfft.jl:681-719),
 * one field (batch entry 0 layout). */
int cfft_sub_to_full(cfft_plan* plan, const double* f0, double* u);
int cfft_full_to_sub(cfft_plan* plan, const double* u, double* f0);

/* ---- execute (device pointers) ----------------------------------------- */
int cfft_forward_dev(cfft_plan* plan, const void* f0, void* F_spec, void* stream);
int cfft_backward_dev(cfft_plan* plan, const void* F_spec, void* f0, void* stream);
int cfft_diffuse_dev(cfft_plan* plan, void* f0, void* stream);
/* As cfft_diffuse_dev, allocating nothing: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under hipGraph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The work
 * buffers (channel spectra, Q-multiplied channels, spectral values) live in
 * one static device array of 16 MiB per device, held until the process
 * ends; a plan needing more (e.g. more than 8 fields of Im-3m 128^3 FP64)
 * gets CFFT_ERR_SIZE and nothing runs.  Paths M2, M7, M2Q, M7Q and QS (QR:
 * CFFT_ERR_UNSUPPORTED); the plan's device must be current.  Every call on a
 * device shares that device's array: calls that overlap in time on one
 * device (two streams, two threads) race and give wrong results.  The array
 * needs no clearing (each call writes its buffers before reading them). */
int cfft_diffuse_noalloc_dev(cfft_plan* plan, void* f0, void* stream);
/* E = exp(-w dt/2) per batch entry (batch x M^3, FP64 in both precisions),
 * applied on both sides of D: q <- E D (E q), each product formed in FP64. */
int cfft_step_dev(cfft_plan* plan, void* q, const void* E, int32_t nsteps, void* stream);
/* As cfft_step_dev with one factor array per batch entry: E is a host array
 * of `batch` device pointers (batch <= 64); the SCFT sweep uses it to switch
 * between w_A and w_B along the contour independently for q and q+. */
int cfft_step_ptrs_dev(cfft_plan* plan, void* q, const void* const* E, int32_t nsteps,
                       void* stream);
/* Contour sweep: nsteps Strang steps q <- E D (E q), batch entry b using
 * E1[b] at absolute step indices < sw[b] and E2[b] after (host arrays of
 * `batch` device pointers / ints, batch <= 64; steps are numbered from step0).
 * No counterpart in the reference (it has no SCFT driver); on cubic H-grids
 * the whole sweep is one cooperative launch with grid barriers between the
 * pipeline phases. */
int cfft_sweep_dev(cfft_plan* plan, void* q, const void* const* E1, const void* const* E2,
                   const int32_t* sw, int32_t step0, int32_t nsteps, void* stream);
/* Projection of n_fields FP64 subgrid fields (M^3 each, x fastest, one after
 * another) onto the fields symmetric under the plan's ops that map the
 * stride-L subgrid onto itself: out(x) = mean over those ops g of in(g x).
 * in and out must differ.  No counterpart in the reference (its symmetric
 * test fields are averaged over the ops on the full grid,
 * test/test_helpers.jl:11-29); the SCFT context uses it on the densities in
 * mixed precision on paths QS, M2Q and M7Q. */
int cfft_symmetrize_f64_dev(cfft_plan* plan, const void* in, void* out, int64_t n_fields, void* stream);
/* E <- exp(-w * factor), n entries: w of the plan precision, E FP64. */
int cfft_exp_dev(cfft_plan* plan, const void* w, void* E, double factor, int64_t n, void* stream);

/* ---- per-stage timer (hipEvents around each kernel launch) ------------- */
#define CFFT_STAGE_FOLD 0   /* P1: fold + E + FFT x                 */
#define CFFT_STAGE_FFT_X 1  /* (unused: FFT x is fused into P1)      */
#define CFFT_STAGE_FFT_Y 2  /* P2: FFT y + FFT z of a plane          */
#define CFFT_STAGE_FFT_Z 3  /* z pass of the non-cubic path          */
#define CFFT_STAGE_RECON 4  /* P3: reconstruction + K                */
#define CFFT_STAGE_IRECON 5 /* P4: inverse recon + IFFT z + IFFT y   */
#define CFFT_STAGE_IFFT_X 6 /* (unused)                              */
#define CFFT_STAGE_IFFT_Y 7 /* y pass of the non-cubic path          */
#define CFFT_STAGE_IFFT_Z 8 /* z pass (non-cubic) / Q-path IFFT z,y  */
#define CFFT_STAGE_UNFOLD 9 /* P5: IFFT x + unfold + E               */
#define CFFT_STAGE_QMUL 10
#define CFFT_STAGE_SWEEP 11 /* persistent sweep (all phases, all steps) */
#define CFFT_STAGE_UNFOLD_FOLD 12 /* P5 of step s fused with P1 of s + 1 */
#define CFFT_STAGE_PLANE_Q 13     /* path QS: FFT y, z + Q + IFFT z, y      */
#define CFFT_N_STAGES 14
/* No counterpart in the reference (its benches time stages with @elapsed,
 * test/profile_centering_fold.jl:48-107).  enable != 0 brackets every
 * subsequent kernel launch of the plan with a pair of hipEvents. */
int cfft_profile(cfft_plan* plan, int32_t enable);
/* Sum of the recorded kernel durations (ms_out[CFFT_N_STAGES]) and launch
 * counts per stage since the last read; clears the record. */
int cfft_profile_read(cfft_plan* plan, double* ms_out, int64_t* count_out);
/* Average duration of one launch of a chain kernel (no counterpart in the
 * reference): stage CFFT_STAGE_UNFOLD_FOLD (P5 fused with P1, as between two
 * steps of a plain sweep: q(s + 1) is not stored), CFFT_STAGE_FFT_Y (P2) or,
 * on path QS, CFFT_STAGE_PLANE_Q (the plane-pair
 * kernel), launched once untimed and then reps times back to
 * back on stream between two hipEvents, on the device field q (batch * M^3
 * entries, overwritten) with interaction factors E[b] (device pointers).
 * Plans on the fused-chain paths only (M2, M7, QR, QS); profiling must be off. */
int cfft_time_stage_dev(cfft_plan* plan, int32_t stage, void* q, const void* const* E, int32_t reps,
                        void* stream, double* ms_per_launch);

/* Propagator sweep keeping every contour point (SCFT densities; no
 * counterpart in the reference).  hist holds nsteps + 1 slices of batch
 * fields (batch * M0*M1*M2 entries each); step i reads slice i and writes
 * slice i + 1 with the factors of cfft_sweep_dev. */
int cfft_sweep_hist_dev(cfft_plan* plan, void* hist, const void* const* E1, const void* const* E2,
                        const int32_t* sw, int32_t step0, int32_t nsteps, void* stream);

/* ---- plan options -------------------------------------------------------- */
/* value != 0: run cfft_sweep_dev / cfft_step_dev on cubic H-grids as one
 * cooperative launch with grid barriers between phases (default 0: one
 * launch per phase and step). */
#define CFFT_OPT_PERSISTENT 1
/* value != 0 (default 1): in multi-step calls, fuse the unfold of a step with
 * the fold of the next (one tile kernel, q(s + 1) is not read back). */
#define CFFT_OPT_FUSE 2
/* value != 0 (default 1): multi-step calls (cfft_step*, cfft_sweep*) capture
 * their launch sequence into a hipGraph once per argument set (pointers,
 * factors, step range) and replay it on the caller's stream; at most 16
 * graphs are cached per plan.  Not used while profiling is enabled. */
#define CFFT_OPT_GRAPH 3
/* y rows per tile of the fold / unfold kernels (P1, P5, P51): 0 (default) =
 * the largest power of two <= 8 dividing H1; otherwise 1, 2, 4 or 8 dividing
 * H1 (CFFT_ERR_INVALID else).  Results do not depend on it; see
 * cfft_hip.tune for timing the choices on a workload. */
#define CFFT_OPT_TILE_ROWS 4
/* Mixed precision on path QS: 0 (default) forms the transforms' twiddle
 * products and the fold/unfold factors in FP64 (data and butterflies FP32),
 * 1 in FP32 (about 4 % faster per step; the roundings of those factors then
 * repeat on every contour step and bias F by about 1e-6).  The plane-pair Q
 * multiply and the interaction factors are FP64 either way. */
#define CFFT_OPT_FP32_TWIDDLES 5
/* Radix schedule of the path-QS plane-pair and tile transforms: the lane
 * split R1 x R2 of a length-H transform (in-lane R2-point passes, then one
 * cross-lane radix-R1 pass).  0 (default): R1 = 4 for H divisible by 16,
 * else 2; 2 selects R1 = 2 for H = 16 or 32 (results differ from the default
 * by rounding only).  Other values: CFFT_ERR_UNSUPPORTED. */
#define CFFT_OPT_FFT_LANES 6
/* Threads per plane-pair workgroup on path QS: 0 or 256 (default), or 512
 * for H = 16, 24, 32 (same results bit for bit). */
#define CFFT_OPT_PAIR_THREADS 7
/* Path QS, 0 (default) or 1: the per-axis chain.  With the separable multiply
 * (orthogonal cell, translations separable in the sign pattern) the Q multiply
 * factors into one 2 x 2 mixing per axis; 1 applies the x mixing with the x
 * transforms in the unfold/fold tile kernel and the y and z mixings in a
 * kernel over single x-planes (twice the work items of the plane pairs),
 * with 32-point lines over eight lanes (other lengths: the default lane
 * split).  H = 8, 12, 16, 24 or 32, FP64 twiddle products, default lane
 * split; results differ from the default chain by rounding only.  Ensembles
 * refuse such plans.  Otherwise CFFT_ERR_UNSUPPORTED. */
#define CFFT_OPT_AXIS_SPLIT 8
/* No counterpart in the reference. */
int cfft_plan_set_option(cfft_plan* plan, int32_t key, int64_t value);

/* Wait for the plan's own stream (host-pointer calls) and the null stream. */
int cfft_sync(cfft_plan* plan);

/* ---- real-space ASU and block CFFT ---------------------------------------- */
/* calc_asu(N, ops), src/asu.jl:216-325 (recursive parity filter; ops already
 * magic-shifted).  Writes at most cap points sorted by idx: idx_out and
 * depth_out cap x 3 (unused dimensions 0), mult_out cap; *n_out = total. */
int cfft_calc_asu(int32_t dim, const int32_t* N, const int32_t* ops, int32_t nops,
                  int32_t* idx_out, int32_t* depth_out, int32_t* mult_out, int64_t cap,
                  int64_t* n_out);

/* pack_asu, src/pack_asu.jl:6-102: greedy unit-stride boxes per depth group
 * (groups in sorted depth order) over n_points ASU points (idx and depth:
 * n_points x 3).  Writes at most cap blocks as 9 ints (start[3], shape[3],
 * depth[3]); *n_out = block count. */
int cfft_pack_asu(int32_t dim, const int32_t* idx, const int32_t* depth, int64_t n_points,
                  int32_t* blocks_out, int64_t cap, int64_t* n_out);

typedef struct cfft_bplan cfft_bplan;

/* plan_cfft, src/cfft_plan.jl:15-40: one DFT per packed block (blocks: 9 ints
 * each, as cfft_pack_asu writes them); every block's data stored one after
 * another in this order, column-major inside a block. */
int cfft_block_plan_create(int32_t dim, const int32_t* blocks, int64_t n_blocks,
                           cfft_bplan** out);
/* Block count, total element count and the block list (9 ints per block,
 * as cfft_pack_asu; blocks_out may be NULL). */
int cfft_block_plan_info(const cfft_bplan* plan, int64_t* n_blocks, int64_t* total,
                         int32_t* blocks_out);
/* mul!(y, p, x), src/cfft_plan.jl:43-53 / src/execution.jl:10-31: unnormalised
 * forward DFT of every block; in/out interleaved complex, total entries. */
int cfft_block_forward(cfft_bplan* plan, const double* in, double* out);
/* ldiv!(y, p, x), src/cfft_plan.jl:56-66 / src/execution.jl:37-64: inverse
 * DFT of every block with the 1/size normalisation. */
int cfft_block_inverse(cfft_bplan* plan, const double* in, double* out);
int cfft_block_plan_destroy(cfft_bplan* plan);

/* ---- dense real-ASU diffusion operator ------------------------------------ */
typedef struct cfft_dense cfft_dense;

/* MatrixDiffusionSolver (src/diffusion_solver.jl:20-71), plan_diffusion's
 * method=:matrix: the host builds A = Re(M diag(K) M^+) over the n real-ASU
 * points (M the real-ASU x spectral-ASU transform matrix, :29-60; M^+ its
 * pseudo-inverse, :175); A (n x n FP64, row-major) is held on the device. */
int cfft_dense_create(int64_t n, const double* A, cfft_dense** out);
/* apply_diffusion_operator!(solver, u) (:66-71) for a real u of n values, in
 * place (u <- A u). */
int cfft_dense_apply(cfft_dense* op, double* u);
/* step_diffusion!(solver, u, w, dt) (:254-282): u <- e^{-w dt/2} A e^{-w dt/2} u. */
int cfft_dense_step(cfft_dense* op, double* u, const double* w, double dt);
int cfft_dense_destroy(cfft_dense* op);

/* ---- full-grid centring pre-fold KRFFT ------------------------------------ */
typedef struct cfft_pplan cfft_pplan;

/* plan_krfft_centered(ops, N), src/centering_prefold.jl:175-205: I/F/C/A
 * pre-fold of the full N-grid into 4/2/2/2 channels on the halved grid, one
 * FFT per channel and the merge table onto the spectral ASU of ops
 * (src/centering_prefold.jl:207-230).  Errors: CFFT_ERR_CENTERING for CentP. */
int cfft_prefold_plan_create(const int32_t* N, const int32_t* ops, int32_t nops,
                             cfft_pplan** out);
/* info[6] = {n_spec, n_channels, centering, Nsub0, Nsub1, Nsub2};
 * idx_out (n_spec x 3, may be NULL) = the spectral-ASU points. */
int cfft_prefold_plan_info(const cfft_pplan* plan, int32_t* info, int32_t* idx_out);
/* execute_centered_krfft!, src/centering_prefold.jl:236-257: u real full grid
 * (N0*N1*N2) -> F_spec (n_spec interleaved complex; unmapped points 0). */
int cfft_prefold_forward(cfft_pplan* plan, const double* u, double* F_spec);
int cfft_prefold_plan_destroy(cfft_pplan* plan);

/* ---- SCFT driver (AB diblock melt on the subgrid) -------------------------- */
/* No counterpart in the reference, which stops at the diffusion step
 * (src/diffusion_solver.jl:254-282).  A context borrows a plan created with
 * batch = 2 (q and q+ advance together, ds = 1/ns) and owns the propagator
 * slices (ns + 1 of them), the FP64 fields w = [wA, wB], densities
 * phi = [phiA, phiB], deviations d = t - w and the Anderson history (m_hist
 * previous iterates).  One plan per context. */
typedef struct cfft_scft cfft_scft;

#define CFFT_SCFT_W 0   /* [wA, wB]   (2 x M0*M1*M2 doubles) */
#define CFFT_SCFT_PHI 1 /* [phiA, phiB] of the last iteration  */
#define CFFT_SCFT_DEV 2 /* [tA - wA, tB - wB] of the last iteration */

int cfft_scft_create(cfft_plan* plan, double f, double chiN, int32_t ns, int32_t m_hist,
                     cfft_scft** out);
/* New block fraction and chiN for the next iterations (the contour
 * discretisation ns is fixed). */
int cfft_scft_set_params(cfft_scft* ctx, double f, double chiN);
/* Anderson step length lam, simple-mixing step lam_simple, and the residual
 * above which simple mixing is used (defaults 1.0, 0.1, 1e-2). */
int cfft_scft_set_mixing(cfft_scft* ctx, double lam, double lam_simple, double simple_above);
/* Fields on the subgrid (M0*M1*M2 doubles each, x fastest); clears the
 * mixing history. */
int cfft_scft_set_fields(cfft_scft* ctx, const double* wA, const double* wB);
/* One iteration: sweep, densities (Simpson), F, field update and mixing.
 * out[8] = {Q, Q+, F, residual, <phiA + phiB>, <phiA>, 1 if Anderson else 0,
 * history entries used}. */
int cfft_scft_iterate(cfft_scft* ctx, double* out);
/* Iterate until residual < tol or max_iter; out[8] as cfft_scft_iterate. */
int cfft_scft_solve(cfft_scft* ctx, double tol, int32_t max_iter, int32_t* iters_out, double* out);
/* Copy a quantity (CFFT_SCFT_*) to the host: 2 x M0*M1*M2 doubles. */
int cfft_scft_get(cfft_scft* ctx, int32_t which, double* out);
/* Device pointer to the propagator slices [ns + 1][q, q+][M0*M1*M2] (plan
 * precision); with n instances [ns + 1][q_0, q+_0, ..., q_{n-1}, q+_{n-1}][...]. */
int cfft_scft_propagators(cfft_scft* ctx, void** hist_out);
int cfft_scft_destroy(cfft_scft* ctx);

/* Several melts in one context (no counterpart in the reference).  n_inst
 * (1 to 32) instances share the grid, space group, cell and ns of a plan created with
 * batch = 2 n_inst, and one sweep advances q, q+ of all of them (fields 2i,
 * 2i + 1); each has its own f[i], chiN[i], fields, densities and Anderson
 * history, with the results of the one-instance context bit for bit.  The
 * calls above that take no instance index reject such a context, except
 * set_mixing (shared), propagators and destroy. */
int cfft_scft_create_multi(cfft_plan* plan, int32_t n_inst, const double* f, const double* chiN,
                           int32_t ns, int32_t m_hist, cfft_scft** out);
int cfft_scft_set_params_inst(cfft_scft* ctx, int32_t inst, double f, double chiN);
int cfft_scft_set_fields_inst(cfft_scft* ctx, int32_t inst, const double* wA, const double* wB);
/* One iteration of every instance; out[8 n_inst], instance i at out + 8 i.
 * An instance whose partition function comes out non-positive or non-finite
 * is marked failed: its out[] reads {NaN, NaN, NaN, +inf, NaN, NaN, 0, 0},
 * the other instances go on, and it is skipped until its fields are set
 * again.  The call fails (CFFT_ERR_INVALID, message of the first failure)
 * only when every instance it iterated failed. */
int cfft_scft_iterate_multi(cfft_scft* ctx, double* out);
/* Iterate until each instance has residual < tol or max_iter iterations; an
 * instance that meets either, or fails as in cfft_scft_iterate_multi, stops
 * iterating while the others go on (its fields stay in the sweep).
 * iters_out[n_inst], out[8 n_inst]. */
int cfft_scft_solve_multi(cfft_scft* ctx, double tol, int32_t max_iter, int32_t* iters_out, double* out);
int cfft_scft_get_inst(cfft_scft* ctx, int32_t inst, int32_t which, double* out);
/* Hand a solve over between contexts of the same subgrid, instance count and
 * Anderson history length (typically a mixed-precision stage and its FP64
 * finish; no counterpart in the reference): each instance's fields, history
 * slots (w and deviations, FP64 in both precisions) and Gram matrix are
 * copied from src to dst, so dst's next iteration continues src's mixing
 * instead of restarting it from simple mixing. */
int cfft_scft_copy_state(cfft_scft* dst, cfft_scft* src);

/* Several contexts advanced together (no counterpart in the reference; the
 * reference's only parallelism is threads over blocks, src/execution.jl:20,
 * 50).  The contexts may differ in grid, space group, cell and instance
 * count; they share precision and ns, and each plan must be on path QS with
 * a cubic channel grid of length H in {8, 12, 16, 24, 32} (grids 32^3 to
 * 128^3 at L = 2) and FP64 twiddle arithmetic (else CFFT_ERR_UNSUPPORTED).
 * The contexts fall into classes by H; per iteration and class, one launch
 * per chain stage and contour step covers the plane pairs (or tiles) of
 * every context of the class with an active instance, from a work list.
 * Each class replays its sweep from one hipGraph on a stream of its own, so
 * the classes run concurrently, and each context's density, update and
 * mixing work follows on its class's stream.  Each instance's
 * results equal those of its own context bit for bit.  The ensemble borrows
 * the contexts: they must outlive it, and while an ensemble call runs no
 * other call may use them.  out and iters_out hold one entry block per
 * instance, contexts in the order given, instances in context order
 * (out[8] per instance as cfft_scft_iterate_multi). */
typedef struct cfft_scft_ensemble cfft_scft_ensemble;
int cfft_scft_ensemble_create(cfft_scft* const* ctx, int32_t n_ctx, cfft_scft_ensemble** out);
int cfft_scft_ensemble_iterate(cfft_scft_ensemble* e, double* out);
/* As cfft_scft_solve_multi over every instance of every context. */
int cfft_scft_ensemble_solve(cfft_scft_ensemble* e, double tol, int32_t max_iter, int32_t* iters_out,
                             double* out);
int cfft_scft_ensemble_destroy(cfft_scft_ensemble* e);

/* ---- P3c KRFFT families (src/recursive_blocks.jl:64-1323) ------------------
 * The recursive, sparse, selective-G0 and G0-ASU forward plans of the
 * reference for cubic groups: the spectral-ASU values from the FFTs of the
 * stride-4 sub-sub-grids u(4 k + 2 e) of one offset e per orbit of the 8
 * offsets under the body-diagonal 3-fold.  The host builds the family's
 * tables (src/recursive_blocks.jl:514-629 sparse, :64-144 recursive, :756-888
 * selective, :1034-1244 G0-ASU) as CSR matrices of complex (re, im) values:
 * stage 0 maps the n_rep M2^3 FFT outputs (representative r's block at
 * r M2^3, x fastest; M2 = N/4) to n_rows[0] values, stage 1 (when n_stage =
 * 2) maps those to n_rows[1]; the last stage's rows are the spectral points.
 * offsets: n_rep x 3 entries 0 or 1.  rowptr[s]: n_rows[s] + 1 int64
 * entries from 0; cols[s] (int32) and vals[s] (2 doubles each): rowptr[s]
 * [n_rows[s]] entries.  Columns out of range: CFFT_ERR_BOUNDS. */
typedef struct cfft_p3c cfft_p3c;
int cfft_p3c_create(int32_t N, int32_t n_rep, const int32_t* offsets, int32_t n_stage, const int64_t* n_rows,
                    const int64_t* const* rowptr, const int32_t* const* cols, const double* const* vals,
                    cfft_p3c** out);
/* Fractal KRFFT (plan_fractal_krfft, src/fractal_krfft.jl:678-722, and its
 * v2/v3 variants, src/fractal_krfft.jl:1021-1170, src/fractal_krfft_v3.jl:
 * 205-379): the leaves are the stride-`stride` sub-grids u(stride k + e) of
 * n_leaf offsets e (n_leaf x 3 entries in [0, stride)), leaf length
 * N / stride (2 to 64, a supported FFT length); 1 to 8 stages, stage 0
 * reading the n_leaf (N/stride)^3 leaf FFT outputs, each later stage the
 * previous stage's rows, the last stage's rows the spectral points.  The
 * handle is run and destroyed by the cfft_p3c_* calls below. */
int cfft_p3c_create_strided(int32_t N, int32_t stride, int32_t n_leaf, const int32_t* offsets, int32_t n_stage,
                            const int64_t* n_rows, const int64_t* const* rowptr, const int32_t* const* cols,
                            const double* const* vals, cfft_p3c** out);
/* u: N^3 real full grid (x fastest); F: 2 n_spec doubles.  Host pointers,
 * blocking (execute_*_krfft!, src/recursive_blocks.jl:323-330, 631-670,
 * 890-972, 1246-1323; execute_fractal_krfft!, src/fractal_krfft.jl:869-901). */
int cfft_p3c_forward(cfft_p3c* plan, const double* u, double* F);
/* Device pointers, on `stream`, without synchronising. */
int cfft_p3c_forward_dev(cfft_p3c* plan, const double* u, double* F, void* stream);
int cfft_p3c_destroy(cfft_p3c* plan);

/* ---- warm-start network (multilayer perceptron) ---------------------------- */
/* No counterpart in the reference.  dims[0..n_layers]: input width, hidden
 * widths, output width.  SiLU hidden activations, linear output.  Parameters:
 * FP32 master copy, flat as W_0 (dims[0] x dims[1], row-major), ..., W_{L-1},
 * then b_0, ..., b_{L-1}; the GEMMs read a BF16 copy and accumulate in FP32
 * on the matrix cores.  Host pointers, row-major FP32 rows. */
typedef struct cfft_mlp cfft_mlp;

int cfft_mlp_create(int32_t n_layers, const int32_t* dims, uint64_t seed, cfft_mlp** out);
int cfft_mlp_info(const cfft_mlp* mlp, int64_t* n_params);
int cfft_mlp_get_params(cfft_mlp* mlp, float* out);
/* Also resets the optimiser state. */
int cfft_mlp_set_params(cfft_mlp* mlp, const float* in);
int cfft_mlp_forward(cfft_mlp* mlp, const float* x, int32_t rows, float* y);
/* Mean-squared-error loss over rows x dims[L] and its gradient (n_params). */
int cfft_mlp_gradient(cfft_mlp* mlp, const float* x, const float* t, int32_t rows, double* loss,
                      float* grad);
/* Adam (0.9, 0.999, 1e-8) over `epochs` shuffled passes of mini-batches;
 * loss_per_epoch[epochs] (may be NULL) = mean training loss of each epoch. */
int cfft_mlp_train(cfft_mlp* mlp, const float* X, const float* T, int64_t n, int32_t epochs,
                   int32_t batch, float lr, uint64_t seed, double* loss_per_epoch);
int cfft_mlp_destroy(cfft_mlp* mlp);

#ifdef __cplusplus
}
#endif

#endif /* CFFT_HIP_H */
