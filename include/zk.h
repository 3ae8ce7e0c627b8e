/*
 * zk.h -- C ABI of libzk.so, the MI355X (gfx950) explicit 2D Euler update.
 *
 * One handle owns one whole-domain state on one GPU: the conserved variables
 * as structure-of-arrays planes in HBM, one HIP stream, and the scratch the
 * stage kernels need.  A handle is not thread-safe; different handles are
 * independent.  Every call returns ZK_OK (0) or a nonzero status, with the
 * message available from zk_last_error().  Host arrays passed in are
 * caller-owned and copied; nothing is retained after the call returns.
 *
 * Array convention: a field is a row-major ny*nx plane of doubles, row i is
 * the i-th row of the input .dat files (y downward) and column j is x -- the
 * order `file >> mat(i,j)` reads them in src_WENO/fluid.cpp:33-40.  A state
 * is 4*nm planes ordered (variable c, mode m) with nm = 1 for WENO and
 * (P+1)^2 for DG, mode m = px*(P+1)+py as src_DG/fluid.cpp:63-65.
 *
 * The reference is two executables (solver_WENO, solver_DG2/3/4) whose hot
 * call is `updateMesh(Mesh&, double dt, int rank, int num_procs)`; each entry
 * point below names the reference interface it replaces.
 */
#ifndef ZK_H
#define ZK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct zk_handle zk_handle;
typedef int zk_status;

#define ZK_OK 0
#define ZK_ERR_ARG 1      /* bad argument or call order */
#define ZK_ERR_HIP 2      /* HIP runtime error */
#define ZK_ERR_NUMERIC 3  /* non-finite state: the message names the variable, the cell
                             (i, j) and the step count of the check that found it */
#define ZK_ERR_CFL 4      /* DG CFL limit exceeded (src_DG/solver.cpp:245-260) */

/* Failure detection (no reference counterpart; the reference keeps stepping
 * and writing NaN files).  Every zk_step, zk_step_cfl, zk_load and
 * zk_set_state call ends by enqueueing a check of the state's cell means
 * (mode 0 of U0..U3, every cell).  The calls that wait for the stream --
 * zk_sync, zk_step_timed, zk_step_cfl, zk_get_state and zk_get_conserved (which
 * still fill their outputs) -- return ZK_ERR_NUMERIC when a check found a
 * value that is not finite, with the variable, the cell and the step count of
 * that check in zk_last_error.  zk_step and zk_step_cfl then refuse to step
 * (ZK_ERR_NUMERIC) until zk_load or zk_set_state replaces the state. */

/* method */
#define ZK_METHOD_WENO 0  /* src_WENO: WENO5 FV + HLLC, SSP-RK3 */
#define ZK_METHOD_DG 1    /* src_DG: modal DG P=1..3, SSP-RK3 */
/* precision of the WENO bulk arithmetic (state and RK accumulation stay FP64) */
#define ZK_PREC_FP64 0
#define ZK_PREC_FP32_BULK 1
/* numerical flux */
#define ZK_FLUX_HLLC 0    /* src_WENO/fluid.cpp:268-317 / src_DG/fluid.cpp:450-520 */
#define ZK_FLUX_HLL 1     /* hlleFlux, src_DG/fluid.cpp:525-580 */
/* DG limiter */
#define ZK_LIMITER_NONE 0
#define ZK_LIMITER_PERSSON_ZS 1 /* applyLimiter, src_DG/fluid.cpp:824-992 */
#define ZK_LIMITER_TVB 2        /* Cockburn-Shu TVB (M; at P >= 3 also the face-average
                                   test) + Zhang-Shu positivity */
#define ZK_LIMITER_LEARNED 3    /* MLP troubled-cell indicator gating the TVB limiter */
#define ZK_LIMITER_TVB_LINEAR 4 /* ZK_LIMITER_TVB with the linear-mode test alone at every P (the rule
                                   README_old.md:137-208 describes; at P >= 3 it lets the four-quadrant
                                   problem blow up, which the face-average test of ZK_LIMITER_TVB prevents) */

typedef struct zk_config {
    int method;      /* ZK_METHOD_* */
    int P;           /* DG degree 1..3 (DG_P_VAL, src_DG/fluid.h:35-39); ignored for WENO */
    int precision;   /* ZK_PREC_* */
    int flux;        /* ZK_FLUX_* */
    int limiter;     /* ZK_LIMITER_*; DG only */
    double tvb_M;    /* TVB constant M (50 in the benchmark configurations) */
    int device;      /* HIP device ordinal */
    int positivity;  /* 1 = positivity option (FP64), no reference counterpart; 0
                        reproduces the reference scheme.
                        WENO: face check: a face whose reconstructed states have rho or p
                        below 1e-12 uses the first-order states.  Cell fallback: an updated
                        cell whose new state has rho or p not above 1e-12 (or is not finite)
                        is bad, and every face touching a bad cell takes the first-order
                        flux for the stage (once; conservative, faces shared).
                        DG (Zhang-Shu scaling of every limiter): the minima of rho and p
                        are also taken at the face quadrature points the fluxes read and
                        at the cell corners the CFL rule reads, with the full basis,
                        after the reference's 4x4 points (src_DG/fluid.cpp:892-991). */
} zk_config;

/* Create a handle.  Replaces the `Mesh` object's lifetime
 * (src_WENO/fluid.h:26-80, src_DG/fluid.h:44-110). */
zk_status zk_create(const zk_config *cfg, zk_handle **out);

/* Release all device memory and the stream. */
zk_status zk_destroy(zk_handle *h);

/* Message of the last failed call on h (or of zk_create when h is NULL). */
const char *zk_last_error(const zk_handle *h);

/* Load a primitive-variable grid and convert to conserved variables; DG
 * modes >= 1 start at zero.  Replaces Mesh(folderPath) without the file
 * reading (src_WENO/fluid.cpp:68-94; src_DG/fluid.cpp:216-244 with
 * initDGModes :173-182).  bctype codes: 0 updated, -1 copy ghost, 1 slip
 * ghost (DG), -2 periodic ghost (extension). */
zk_status zk_load(zk_handle *h, int ny, int nx, double da, double gamma,
                  const double *rho, const double *u, const double *v,
                  const double *p, const int32_t *bctype);

/* Overwrite the whole state (4*nm planes, plane c*nm + m of variable c and
 * mode m, each ny*nx row-major), e.g. L2-projected DG modes or a restore.
 * No reference counterpart (the reference cannot resume).  A DG state lives
 * on the device with the planes of each grid row side by side; the call
 * converts on the device, through the first stage buffer. */
zk_status zk_set_state(zk_handle *h, const double *Q);

/* Copy the whole state (4*nm planes, as zk_set_state takes them) to the host. */
zk_status zk_get_state(zk_handle *h, double *Q);

/* Copy the cell averages U0..U3 (mode 0) into four caller-owned planes.
 * Replaces syncCellAverages + the field part of saveMeshData
 * (src_DG/fluid.cpp:190-196, src_WENO/fluid.cpp:558-576). */
zk_status zk_get_conserved(zk_handle *h, double *U0, double *U1, double *U2, double *U3);

/* Maximum wave speed.  WENO: computeMaxSpeed over every cell
 * (src_WENO/fluid.cpp:539-553).  DG: computeMaxSpeedDG_omp at the four cell
 * corners of bctype-0 cells (src_DG/solver.cpp:26-64). */
zk_status zk_max_speed(zk_handle *h, double *out);

/* Advance nsteps SSP-RK3 steps with a fixed dt.  Replaces nsteps calls of
 * updateMesh(mesh, dt, rank, nprocs) (src_WENO/fluid.cpp:468-493,
 * src_DG/fluid.cpp:1005-1054).  Asynchronous: returns after enqueueing. */
zk_status zk_step(zk_handle *h, double dt, int nsteps);

/* Advance nsteps steps with dt = cfl * da / lambda_max recomputed on the
 * device before every step (lambda from zk_max_speed's rule), clipped so that
 * time never passes t_end; stops early at t_end.  The time and step count
 * reached are returned.  The CFL-correct replacement of the `auto`/`time`
 * dt rule of src_WENO/solver.cpp:148-171, which sets dt = lambda * h.
 * DG: a cfl above the stable limit 0.9 / (2 (2P + 1)) returns ZK_ERR_CFL, as
 * the reference refuses to run (src_DG/solver.cpp:12-17, 245-260).  A wave
 * speed that is not finite stops the stepping and returns ZK_ERR_NUMERIC. */
zk_status zk_step_cfl(zk_handle *h, double cfl, double t_end, int nsteps,
                      double *t_out, int *steps_out);

/* Wait for all enqueued work on the handle's stream. */
zk_status zk_sync(zk_handle *h);

/* Time nsteps fixed-dt steps with HIP events recorded on the handle's own
 * stream; *ms receives the elapsed milliseconds (measurement helper). */
zk_status zk_step_timed(zk_handle *h, double dt, int nsteps, float *ms);

/* Per-cell limiter flags of the last stage (ny*nx bytes; TVB / learned
 * limiter: 1 = limited).  No reference counterpart. */
zk_status zk_get_flags(zk_handle *h, uint8_t *flags);

/* Kernel timing (measurement helper, no reference counterpart): while
 * enabled, every kernel launch is bracketed by a pair of HIP events on the
 * handle's stream.  zk_profile(h, 1) clears earlier records and starts;
 * zk_profile_report writes a JSON object {"kernel": [launches, total_ms]}. */
zk_status zk_profile(zk_handle *h, int enable);
zk_status zk_profile_report(zk_handle *h, char *buf, size_t len);

/* Count of cells updated per stage (bctype 0 inside [3,ny-3)x[3,nx-3),
 * the loop domain of src_WENO/fluid.cpp:398-402). */
zk_status zk_updated_cells(zk_handle *h, int64_t *out);

/* Streaming bandwidth of GPU `device` for the read/write mixes of a stage
 * kernel (measurement helper, no reference counterpart): a copy (1 read +
 * 1 write) and an add (2 reads + 1 write) over three arrays of array_bytes
 * each, 16-byte accesses, best of five timed runs after a first one, in GB/s.
 * The practical ceiling a stage's roofline fraction is read against.  It
 * allocates and frees the arrays; errors: zk_last_error(NULL). */
zk_status zk_stream_bandwidth(int device, size_t array_bytes, double *copy_gbs, double *add_gbs);

/* ---------------------------------------------------------------------------
 * Batched ensembles of independent WENO5 problems (no reference counterpart:
 * the reference runs one problem per MPI job).  Instances of different grid
 * sizes are packed and advanced together, each with its own
 * dt = cfl * da / lambda_max recomputed on the device every step (the
 * CFL-correct rule, lambda as computeMaxSpeed, src_WENO/fluid.cpp:539-553).
 * One step is a fixed sequence of launches captured once as a HIP graph.
 * ------------------------------------------------------------------------- */
typedef struct zk_ensemble zk_ensemble;

zk_status zk_ensemble_create(const zk_config *cfg, zk_ensemble **out);
/* Append one instance (primitive fields, as zk_load); copied.  Refused
 * (ZK_ERR_ARG) once the ensemble is finalized. */
zk_status zk_ensemble_add(zk_ensemble *e, int ny, int nx, double da, double gamma,
                          const double *rho, const double *u, const double *v,
                          const double *p, const int32_t *bctype, double cfl);
/* Upload every instance; no zk_ensemble_add afterwards.  A second call
 * returns ZK_ERR_ARG and changes nothing. */
zk_status zk_ensemble_finalize(zk_ensemble *e);
/* tile: 1 = 16x15, 2 = 32x15 cells per work item; use_graph: replay a captured step. */
zk_status zk_ensemble_set_tuning(zk_ensemble *e, int tile, int use_graph);
/* Advance every instance nsteps steps (asynchronous). */
zk_status zk_ensemble_step(zk_ensemble *e, int nsteps);
/* Wait for the ensemble's stream.  An instance whose wave-speed maximum was
 * not finite stops (dt = 0 from then on); sync then returns ZK_ERR_NUMERIC
 * naming the first such instances and their step counts. */
zk_status zk_ensemble_sync(zk_ensemble *e);
zk_status zk_ensemble_size(zk_ensemble *e, int *ninst, int64_t *updated_cells);
/* Conserved planes (4*ny*nx) and time reached of instance idx. */
zk_status zk_ensemble_get(zk_ensemble *e, int idx, double *U, double *t_out);
const char *zk_ensemble_last_error(const zk_ensemble *e);
zk_status zk_ensemble_destroy(zk_ensemble *e);

/* Face reconstruction of the WENO path: ZK_RECON_WENO5 (default, the
 * reference's weno5_reconstruct) or MUSCL with the minmod, van Leer or
 * superbee limiter (muscl_reconstruct, src_WENO/fluid.cpp:203-225, which the
 * reference defines but its solver never calls).  WENO handles only. */
enum { ZK_RECON_WENO5 = 0, ZK_RECON_MUSCL_MINMOD = 1, ZK_RECON_MUSCL_VANLEER = 2, ZK_RECON_MUSCL_SUPERBEE = 3 };
zk_status zk_set_reconstruction(zk_handle *h, int kind);

/* Slip walls on the WENO path (no reference counterpart: the reference's
 * finite-volume fill, src_WENO/fluid.cpp:338-370, leaves bctype-1 cells
 * untouched, which stays the default).  With enable = 1 a bctype-1 cell
 * within 3 cells of an updated cell along a row or column takes the updated
 * cell mirrored across the wall between them (2k-1 cells away when the
 * nearest updated cell is k away), with the momentum normal to the wall
 * negated; the column pass comes first and the row pass overrides it, as in
 * the reference's fills.  Before or after zk_load; WENO handles only. */
zk_status zk_set_slip_walls(zk_handle *h, int enable);

/* Kernel-variant selection for the autotuner (no reference counterpart):
 * fv_tile 0 = automatic, 1 = 16x15, 2 = 32x15 WENO tiles; dg_variant 0 =
 * automatic (fused row-marching stage kernel, workgroup form, without the
 * Persson limiter; split face/cell kernels otherwise), 100 + R =
 * fused workgroup form with R-row segments (R in 4..64; 100 = R chosen to
 * fill whole rounds of the device), 10 + R = fused single-wave form,
 * 2/3/4 = split kernels with the P=3 MFMA cell kernel at that minimum of
 * waves per SIMD, -1 = split kernels with the VALU cell kernel; use_graph = 1
 * (default) lets zk_step replay a captured step as a HIP graph (WENO, and DG
 * unless the limiter is learned or harvesting is on). */
zk_status zk_set_tuning(zk_handle *h, int fv_tile, int dg_variant, int use_graph);

/* ---------------------------------------------------------------------------
 * Learned troubled-cell indicator for DG P=2 (no reference counterpart).
 * A 192-64-64-1 MLP (BF16 weights on v_mfma_f32_16x16x32_bf16, FP32
 * accumulation, FP32 Adam master weights) scores a cell from its own and its
 * four neighbours' normalised modes; the limiter asks it only about the TVB
 * candidates (cells with a linear mode above M h^2).  Training data are (features, TVB
 * flag) pairs harvested from the solver's own TVB runs on the device.
 * ------------------------------------------------------------------------- */
typedef struct zk_mlp zk_mlp;

/* He-initialised network (seeded) with room for `capacity` harvested samples. */
zk_status zk_mlp_create(int device, uint64_t seed, int capacity, zk_mlp **out);
zk_status zk_mlp_destroy(zk_mlp *m);
const char *zk_mlp_last_error(const zk_mlp *m);
/* Number of FP32 parameters: W1[64][192], b1[64], W2[64][64], b2[64], w3[64], b3. */
int zk_mlp_param_count(void);
zk_status zk_mlp_get_params(zk_mlp *m, float *w);
zk_status zk_mlp_set_params(zk_mlp *m, const float *w);
/* Loss and parameter gradients of n host samples (n % 32 == 0); X holds BF16
 * bit patterns [n][192]; binary cross-entropy on the sigmoid output. */
zk_status zk_mlp_grad(zk_mlp *m, const uint16_t *X, const float *y, int n, float *grad, float *loss);
/* Scores of n host samples (n % 16 == 0). */
zk_status zk_mlp_infer(zk_mlp *m, const uint16_t *X, int n, float *scores);
/* Copy out up to max harvested samples; *count receives the number stored. */
zk_status zk_mlp_dataset(zk_mlp *m, int *count, uint16_t *X, float *y, int max);
zk_status zk_mlp_clear_dataset(zk_mlp *m);
/* Adam over the harvested set: epochs x (count / B) steps of batch B (B % 32 == 0). */
zk_status zk_mlp_train(zk_mlp *m, int epochs, int B, float lr, uint64_t seed, float *loss_per_epoch);
/* Harvest (features, TVB flag) pairs during the stages of a DG P=2 TVB handle,
 * from candidate cells (a linear mode above M h^2; the only cells the learned
 * limiter scores): positives kept with probability keep_pos, negatives with
 * keep_neg; samples past m's capacity are dropped; m = NULL stops. */
zk_status zk_harvest(zk_handle *h, zk_mlp *m, float keep_pos, float keep_neg);
/* Use m in a ZK_LIMITER_LEARNED handle: scores above hi mark a troubled cell,
 * below lo a smooth one; scores in [lo, hi] and a random `audit` fraction of
 * cells fall back to the TVB test, which also scores the network. */
zk_status zk_attach_mlp(zk_handle *h, zk_mlp *m, float lo, float hi, float audit);
/* Counters since the last reset, over candidate cells (those with a linear
 * mode above M h^2: every other cell passes the TVB test unchanged and is
 * never scored): audited, agreed, missed (TVB troubled, net not), false
 * flags (net troubled, TVB not), net-flagged, in-band, candidates, audited
 * cells the TVB test found troubled. */
zk_status zk_learned_stats(zk_handle *h, int64_t out[8], int reset);
/* Features (BF16 bits, [n][192]) and, when m is given, scores of n cells of the
 * current state (cells as row-major indices inside [3,ny-3)x[3,nx-3)). */
zk_status zk_mlp_features(zk_handle *h, zk_mlp *m, const int32_t *cells, int n, uint16_t *X, float *scores);

/* ---------------------------------------------------------------------------
 * WENO stepping that allocates nothing.  No counterpart in the reference: for
 * callers that cannot call hipMalloc (code under HIP graph capture, tight
 * serving loops, harnesses that forbid allocation).  It advances the caller's
 * device state Q (4 row-major ny*nx planes of conserved variables rho, rho*u,
 * rho*v, E in device memory on cfg->device) by nsteps fixed-dt SSP-RK3 steps
 * in place, reading the bctype codes from device memory (0 updated, -1 copy
 * ghost, -2 periodic ghost; other codes are never filled), and enqueues on
 * `stream` (a hipStream_t, or NULL).  Results equal zk_step on a WENO handle
 * loaded with the same state, bit for bit.  cfg: ZK_METHOD_WENO, FP64 or FP32
 * bulk, HLLC or HLL, positivity 0; the reconstruction is WENO5.
 *
 * Its stage buffers, ghost map, ring list and step scalars live in one device
 * array of ZK_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends), so
 * grids are limited to ZK_NOALLOC_MAX_CELLS cells (ny*nx); larger ones return
 * ZK_ERR_ARG.  Calls on one GPU share that array: calls that overlap on one
 * GPU (different streams or threads) race.  Errors: zk_last_error(NULL).  The
 * first call on each device looks up the array's address; make it outside
 * stream capture.  The call is asynchronous and does no failure check. */
#define ZK_NOALLOC_MAX_CELLS (1u << 20)
#define ZK_NOALLOC_SCRATCH_BYTES 76546304u
zk_status zk_weno_step_noalloc(const zk_config *cfg, int ny, int nx, double da, double gamma,
                               double *Q, const int32_t *bctype, double dt, int nsteps, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* ZK_H */
