/*
 * lym.h -- C ABI of the MI355X (gfx950) hot path of LatticeYangMills.
 *
 * The reference has no FFI: its boundary is the C++ virtual class API of
 * Action / Observable / App (include/Actions/action.h:54-68,
 * include/Observables/observable.h:52-66, include/Apps/app.h:59-77).
 * Each entry point below replaces one member function of that API; the C++
 * host classes in latticeyangmills_amd/host call these.
 *
 * Conventions
 *   - Link data crossing the boundary is the reference's GluonField layout:
 *     four arrays (one per direction mu), each V x 18 doubles, site index
 *     s = x + Lx*(y + Ly*(z + Lz*t)), element m[2*(3i+j)] = Re U_ij,
 *     m[2*(3i+j)+1] = Im U_ij (include/Math/field.h:52-76,
 *     include/Math/lattice.h:87-89, include/Math/su3.h:63).
 *   - Every function returns LYM_OK (0) or a negative LYM_ERR_* code and
 *     never throws; lym_last_error() gives the message of the last failure.
 *   - A context owns its device buffers and one HIP stream; calls on one
 *     context are not thread-safe, distinct contexts are independent.
 */
#ifndef LYM_H
#define LYM_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LYM_OK 0
#define LYM_ERR_ARG (-1)    /* invalid argument */
#define LYM_ERR_HIP (-2)    /* HIP runtime error */
#define LYM_ERR_NOMEM (-3)  /* device allocation failed */
#define LYM_ERR_NODEV (-4)  /* no GPU visible */
#define LYM_ERR_TOO_LARGE (-5) /* lattice beyond a fixed maximum (lym_observables_noalloc) */

#define LYM_PREC_FP64 0  /* FP64 links, FP64 arithmetic throughout */
#define LYM_PREC_MIXED 1 /* FP32 links/staples/clover; FP64 exp, reunitarisation, reductions */
/* An FP32 field keeps rows 0, 1 of every link rounded to FP32 and row 2 =
 * conj(row0 x row1) formed from them in FP32 (exact SU(3) completion, so the
 * upload of an SU(3) link moves row 2 by FP32 rounding only; a non-unitary
 * link keeps its rows 0, 1 and takes the derived row 2). */

#define LYM_CAND_SU2 0  /* X = R*S*T of SU(2) subgroups, Random::randSU3Transf (src/Math/random.cpp:71-156) */
#define LYM_CAND_EXPH 1 /* X = exp(i eps H), H = sum_a h_a lambda_a, h_a uniform in [-1, 1) */

typedef struct lym_ctx lym_ctx;

/* Creates a context for one Lx*Ly*Lz*Lt periodic lattice on GPU `device`;
 * links start cold (all 1).  Extents >= 2 and V < 44.7M sites in either
 * precision (64^4 and 64^3 x 128 fit; the limit keeps every buffer offset a
 * 31-bit integer).  Replaces WilsonFlow::createLattice /
 * App::createLattice (src/Apps/wilsonflow.cpp:87-91, src/Apps/app.cpp:41-44). */
int lym_create(lym_ctx **ctx, const int dims[4], int precision, int device);
int lym_destroy(lym_ctx *ctx);
/* The message of the last failure on ctx; with ctx == NULL, that of the last
 * failed lym_observables_noalloc call on this thread ("null context" if none). */
const char *lym_last_error(const lym_ctx *ctx);
/* V = number of sites; precision as given to lym_create. */
int lym_info(const lym_ctx *ctx, int dims[4], int64_t *volume, int *precision);

/* Host <-> device copies of the gauge field, reference layout (see top).
 * Upload replaces InputConf::readConf's fill of the GluonField
 * (src/InputOutput/inputconf.cpp:114-155); in LYM_PREC_MIXED it stores the
 * FP32 form above. */
int lym_upload_links(lym_ctx *ctx, const double *const by_mu[4]);
int lym_download_links(lym_ctx *ctx, double *const by_mu[4]);
/* Cold start, all U = 1 (setToIdentity, src/Math/lattice.cpp:33-41). */
int lym_set_cold(lym_ctx *ctx);
/* Hot start from Haar-random SU(3) (the reference's GaugeFieldFactory "Hot"
 * start, src/Apps/gaugefieldfactory.cpp:67-71, whose setToRandom is missing;
 * Philox-seeded here). */
int lym_set_hot(lym_ctx *ctx, uint64_t seed);

/* PureGauge::computeStaples(mu) (src/Actions/puregauge.cpp:36-54):
 * omega_out = V x 18 doubles. */
int lym_staples(lym_ctx *ctx, int mu, double *omega_out);
/* PureGauge::computeDerivative(mu) (src/Actions/puregauge.cpp:62-79):
 * z_out = V x 18 doubles, Z = (i/2)[U Omega - (U Omega)^+]_traceless. */
int lym_derivative(lym_ctx *ctx, int mu, double *z_out);

/* SuperObs::compute (src/Observables/superobs.cpp:31-128):
 * out[0] = plaq, out[1] = energy (raw sum of Re Tr C_{mu nu}^2, <= 0),
 * out[2] = topc, out[3] = sum over sites of the absolute values of the three
 * plane-pair terms of q(x), normalised like topc (the scale of the cancellations
 * in topc). */
int lym_observables(lym_ctx *ctx, double out[4]);

/* nsteps x WilsonFlow::flowStep (src/Apps/wilsonflow.cpp:141-172); enqueued
 * on the context's stream, returns before completion. */
int lym_flow_steps(lym_ctx *ctx, double eps, int nsteps);
/* WilsonFlow::applyWilsonFlow (src/Apps/wilsonflow.cpp:115-136): rows =
 * int(tau/eps) series rows, rows-1 flow steps; row t = [t, plaq, topc, energy]
 * with t accumulated by += eps.  Rows with t % measure_every != 0 hold NaN
 * observables (measure_every = 1 is the reference).  series holds max_rows*4
 * doubles; *rows_out receives rows. */
int lym_flow(lym_ctx *ctx, double eps, double tau, int measure_every, double *series, int max_rows,
             int *rows_out);
/* No counterpart in the reference (its flow loop, src/Apps/wilsonflow.cpp:115-136,
 * calls each step directly).  on != 0: lym_flow records its launches (steps
 * and measurements) once as a HIP graph and replays it while eps, rows,
 * measure_every and the start buffer repeat; one graph is kept per start
 * buffer, and a change of schedule drops both.  Results are those of the
 * launches it replays.  It does not make the flow faster on MI355X: even on
 * 8^4 each stage kernel runs ~10 us, which hides the asynchronous launches
 * (99 steps: 3.05 ms eager, 2.98 ms replayed).  It serves callers that want
 * one replayable unit per flow.  Default off. */
int lym_set_graph(lym_ctx *ctx, int on);

/* GaugeFieldFactory::MCUpdate / updateLink (src/Apps/gaugefieldfactory.cpp:192-220)
 * with the Delta S of PureGauge::compute (src/Actions/puregauge.cpp:28), as a
 * checkerboard: per sweep, sub-sweeps mu = 0..3, parity 0 then 1.  Sweeps are
 * numbered first_sweep .. first_sweep+nsweeps-1 (Philox counter).  predrawn ==
 * NULL: Philox-4x32-10 uniforms, counter {mu*V+s, sweep, hit, k/4}, key seed.
 * predrawn != NULL (test mode): uniform k of hit h of link l in sweep w is
 * predrawn[((w*4V + l)*nhit + h)*n + k], n = 13 (LYM_CAND_SU2) or 9 (LYM_CAND_EXPH).
 * Extents must be even.  *accepted = number of accepted hits; the call
 * synchronises the context's stream.  The sweeps run on the links reordered
 * into even-odd site order (a copy into the context's second link buffer),
 * and the context keeps them in that order until another call reads them,
 * which reorders them back first: consecutive calls, one sweep each as the
 * reference's MCUpdate, pay no reorder. */
int lym_metropolis(lym_ctx *ctx, double beta, double eps, int nhit, int nsweeps, int first_sweep,
                   int candidate, uint64_t seed, const double *predrawn, int64_t *accepted);

/* Health of the stored links, reported and never fatal.  The reference
 * (src/Math/su3.h, no checks anywhere) lets NaN propagate silently.
 * *max_dev = largest over all links of max_ij |(U U^+ - 1)_ij| and |det U - 1|
 * (finite entries only); *nonfinite = number of links with a NaN or infinite
 * entry.  Either pointer may be NULL. */
int lym_check_links(lym_ctx *ctx, double *max_dev, int64_t *nonfinite);

/* Blocks until all work on the context's stream is done. */
int lym_synchronize(lym_ctx *ctx);

/* Measurement helper: runs nsteps flow steps with a HIP event pair around
 * every kernel launch on the context's stream.  stage_ms[k] = mean duration
 * of RK stage k+1 (k = 0..2), stage_ms[3] = mean duration of the observables
 * kernel (measured every measure_every steps, 0 = never); *total_ms = wall
 * time of the whole sequence on the stream. */
int lym_profile_flow(lym_ctx *ctx, double eps, int nsteps, int measure_every, double *total_ms,
                     double stage_ms[4]);

/* Work schedule of the stencil kernels (no counterpart in the reference,
 * whose loops run in site order on one core).  kind 0 = LYM_SCHED_LINEAR
 * (workgroup b takes 64-site run b), 1 = LYM_SCHED_CONTIG (each of the 8 XCDs
 * takes a contiguous eighth), 2 = LYM_SCHED_COLUMN (each XCD marches columns of
 * about `target` sites of the (x, y, z) volume along t, so neighbouring time
 * slices stay in its L2).  Every schedule and staging mode computes each link
 * the same way, so the links agree bit for bit across them (the observables
 * are the same site terms summed in another order), the two-row form
 * included: it drops the row 2 that an FP32 field derives anyway.  Default:
 * COLUMN | LDS with 4096 sites, with TWO_ROW in LYM_PREC_MIXED; the fastest
 * choice depends on shape and precision, which lym_autotune measures. */
#define LYM_SCHED_LINEAR 0
#define LYM_SCHED_CONTIG 1
#define LYM_SCHED_COLUMN 2
#define LYM_SCHED_LDS 16 /* or-ed into a kind: stage each 64-site run's own links in LDS */
/* or-ed into a kind, LYM_PREC_MIXED only: the flow calls (lym_flow,
 * lym_flow_steps, lym_profile_flow) keep the links as their first two rows,
 * row 2 = conj(row0 x row1) rebuilt where a kernel reads a link: a third
 * fewer bytes per link, more arithmetic.  lym_upload_links stores that form
 * directly, a flow call converts any other field on entry, and the field
 * stays in two-row form, which lym_observables reads as it is;
 * the calls that read links as three rows (download, staples, derivative,
 * Metropolis, link check, autotune) convert it back first.  lym_autotune
 * tries it per shape. */
#define LYM_SCHED_TWO_ROW 32
int lym_set_schedule(lym_ctx *ctx, int kind, int target);
/* Times every candidate schedule on the context's own field (RK stage 1,
 * which leaves the field unchanged; three interleaved rounds, each candidate
 * keeping its fastest time) and keeps the default schedule unless another is
 * faster by more than 1.5 %; cache_path (may be NULL) holds one line per
 * lattice shape and precision; summary receives the choice. */
int lym_autotune(lym_ctx *ctx, const char *cache_path, char *summary, int summary_len);

/* No counterpart in the reference: SuperObs::compute (src/Observables/superobs.cpp:31-128)
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation).  It is
 * lym_observables on a link field the caller already holds on the device,
 * and it calls no hipMalloc, hipFree or allocating library routine: its one
 * scratch buffer, the per-workgroup partial sums, is a file-scope device
 * array of LYM_NOALLOC_MAX_SITES / 64 slots (16 MiB) that every GPU holds
 * until the process ends.
 *   links:  the field on the calling thread's current device, in the device
 *           layout of the library: direction planes mu = 0..3, each 9
 *           complex elements e = 3i + j by V sites,
 *           links[((mu*9 + e)*V + s)] = (Re U_ij, Im U_ij) of site s, as
 *           double2 (LYM_PREC_FP64) or float2 (LYM_PREC_MIXED), read as
 *           given (a field a context wrote holds the FP32 form above);
 *   out:    4 doubles in device memory, written as lym_observables' out;
 *   stream: the hipStream_t to run on (NULL: the null stream); the call is
 *           asynchronous on it.
 * Extents >= 2 and V <= LYM_NOALLOC_MAX_SITES, else LYM_ERR_ARG or
 * LYM_ERR_TOO_LARGE before any device work; the message of a failure is
 * lym_last_error(NULL) on the calling thread.  Every call on one GPU shares
 * that GPU's scratch array (each call writes every slot it reads, so nothing
 * needs clearing between calls), so calls that overlap on one GPU, on
 * different streams or threads, race. */
#define LYM_NOALLOC_MAX_SITES (1 << 24)
int lym_observables_noalloc(const void *links, const int dims[4], int precision, double *out, void *stream);

/* Measurement helper with no counterpart in the reference: a streaming
 * device-to-device copy of `bytes` bytes repeated `reps` times on GPU
 * `device`; *gbs = (read + written bytes) / mean kernel time.  Used to confirm
 * the HBM bandwidth of the box and to calibrate HBM byte counters. */
int lym_copy_bandwidth(int device, int64_t bytes, int reps, double *gbs);

#ifdef __cplusplus
}
#endif
#endif
