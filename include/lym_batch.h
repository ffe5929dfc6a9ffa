/*
 * lym_batch.h -- C ABI of the batched ensemble path: many independent
 * periodic lattices of uneven sizes, flowed together (one launch per RK stage
 * for all of them, or one persistent launch per measurement interval), with
 * the flow loop captured in a HIP graph.
 *
 * The reference processes its input configurations one after another
 * (WilsonFlow::flowConfigurations, src/Apps/wilsonflow.cpp:96-109, one
 * readConf + applyWilsonFlow per file); these entry points replace that loop
 * for a whole ensemble at once.  Layout and error conventions as in lym.h.
 */
#ifndef LYM_BATCH_H
#define LYM_BATCH_H
#include <stdint.h>

#include "lym.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lym_batch lym_batch;

/* nconf lattices, dims = nconf x 4 extents; all start cold.  The lattices'
 * list replaces WilsonFlow's m_inputConfList (src/Apps/wilsonflow.cpp:96-109). */
int lym_batch_create(lym_batch **b, int nconf, const int *dims, int precision, int device);
int lym_batch_destroy(lym_batch *b);
const char *lym_batch_last_error(const lym_batch *b);
int lym_batch_info(const lym_batch *b, int *nconf, int64_t *total_sites, int *grid, int *kind, int *target);

/* Per-lattice transfer in the reference's GluonField layout
 * (InputConf::readConf, src/InputOutput/inputconf.cpp:114-155). */
int lym_batch_upload_links(lym_batch *b, int conf, const double *const by_mu[4]);
int lym_batch_download_links(lym_batch *b, int conf, double *const by_mu[4]);
/* Haar-random hot start per lattice, seeds[c] as lym_set_hot. */
int lym_batch_set_hot(lym_batch *b, const uint64_t *seeds);
/* Checkerboard Metropolis on every lattice at once (GaugeFieldFactory::MCUpdate,
 * src/Apps/gaugefieldfactory.cpp:192-220), Philox keyed by seeds[c];
 * accepted[c] = accepted hits of lattice c (may be NULL).  Each lattice gets
 * the same updates as lym_metropolis on its own context, and like it the
 * batch keeps the links in even-odd order until another call reads them. */
int lym_batch_metropolis(lym_batch *b, double beta, double eps, int nhit, int nsweeps, int first_sweep,
                         int candidate, const uint64_t *seeds, int64_t *accepted);
/* SuperObs::compute for every lattice (src/Observables/superobs.cpp:31-128):
 * out[4c + k] as lym_observables. */
int lym_batch_observables(lym_batch *b, double *out);
/* WilsonFlow::applyWilsonFlow for every lattice (src/Apps/wilsonflow.cpp:115-136):
 * series[(c*max_rows + t)*4 + k] as lym_flow.  use_graph != 0 records the whole
 * flow loop (all steps and measurements) as one HIP graph and replays it. */
int lym_batch_flow(lym_batch *b, double eps, double tau, int measure_every, int use_graph, double *series,
                   int max_rows, int *rows_out);
/* Launch shape of the stage kernel: the schedule kind and target as
 * lym_set_schedule (lattices stay whole, in order; LYM_SCHED_TWO_ROW for a
 * mixed batch keeps every lattice in two-row form while lym_batch_flow runs);
 * grid = 0 means one workgroup per schedule slot, otherwise a persistent grid
 * of that many workgroups (rounded up to a multiple of 8).
 * With LYM_SCHED_WHOLE_FLOW or-ed into kind, lym_batch_flow runs the steps
 * between two measurements (all of them when measure_every exceeds the step
 * count) of every lattice as ONE persistent launch, with a grid-wide barrier
 * between RK stages; grid = 0 then takes as many workgroups as the device
 * holds at once (rounded down to a multiple of 8), and a larger grid is
 * refused, since every workgroup must be resident for the barriers.  The
 * barrier waits are bounded (lym_batch_set_barrier_timeout, 5 s): a launch
 * whose workgroups cannot all be resident ends, and lym_batch_flow returns
 * LYM_ERR_HIP.  The links equal
 * those of the per-stage launches bit for bit. */
#define LYM_SCHED_WHOLE_FLOW 64
int lym_batch_set_schedule(lym_batch *b, int grid, int kind, int target);
/* Bound of each grid-barrier wait of the whole-flow launch, in seconds
 * (default 5; no counterpart in the reference, which has no device).  A wait
 * past it ends the launch and lym_batch_flow returns LYM_ERR_HIP; 0 makes
 * every wait give up at once, which exercises that path from the host. */
int lym_batch_set_barrier_timeout(lym_batch *b, double seconds);
/* Times the candidate shapes on the batch itself and keeps the fastest, as
 * lym_autotune (the default schedule unless another is faster by more than
 * 1.5 %; then persistent grids of the chosen order, then the whole-flow
 * launch, under the same rule);
 * cache_path (may be NULL) holds one line per batch shape and precision.
 * summary receives a one-line description of the choice. */
int lym_batch_autotune(lym_batch *b, const char *cache_path, char *summary, int summary_len);

#ifdef __cplusplus
}
#endif
#endif
