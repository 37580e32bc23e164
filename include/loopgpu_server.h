/*
 * loopgpu server: continuous batching of loop-update simulation jobs on one
 * GPU.  No counterpart in the reference (parapack farms independent clones
 * over processes); this is the serving form of the same worker.
 *
 * Jobs (lattice, L, beta, XXZ anisotropy, thermalization and measurement
 * sweeps) are admitted while the batch has room for their device footprint
 * (upper bound of the operator count beta*Nb*(offset+|Jz|/4+|Jxy|/2) plus
 * 10 standard deviations, candidate slots, sites and bonds); admitted jobs run
 * together, every kernel launch covering all of them, each chain resident in
 * HBM for its whole life.  Jobs that arrive while others are running join at
 * the next chunk boundary; finished jobs leave without moving any other
 * chain's data.  A job's trajectory depends only on its parameters (seed and
 * chain_id key its random streams), so its result is the same alone and in
 * any mix, up to FP64 summation order in the estimators.
 *
 * All functions are thread-safe; a background thread drives the device.
 */
#ifndef LOOPGPU_SERVER_H
#define LOOPGPU_SERVER_H

#include <stddef.h>
#include <stdint.h>

#include "loopgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lg_server lg_server;

typedef struct lg_server_config {
  int32_t max_jobs;        /* chains resident at once */
  int64_t max_sites;       /* sum of sites of resident jobs */
  int64_t max_bonds;
  int64_t max_ops;         /* sum of operator footprints of resident jobs */
  int64_t max_cand;        /* sum of candidate slots */
  uint32_t max_chunk;      /* sweeps between admission points (default 64) */
} lg_server_config;

typedef struct lg_job {
  lg_params params;
  uint32_t thermalization;
  uint32_t sweeps;         /* measurement sweeps after thermalization */
} lg_job;

enum { LG_JOB_QUEUED = 0, LG_JOB_RUNNING = 1, LG_JOB_DONE = 2, LG_JOB_CANCELLED = 3, LG_JOB_UNKNOWN = -1 };

/* mean and binning error of the job's measurements (after thermalization):
 * index 0 Energy Density, 1 Susceptibility, 2 Staggered Structure Factor
 * (= Staggered Magnetization^2 / V), 3 Staggered Susceptibility,
 * 4 Number of Clusters, 5 Stiffness, 6 operators per sweep */
#define LG_JOB_NOBS 7
typedef struct lg_job_result {
  int32_t status;
  uint64_t sweeps_done;
  double latency_s;        /* submit -> done */
  double queue_s;          /* submit -> admitted */
  double mean[LG_JOB_NOBS];
  double error[LG_JOB_NOBS];
} lg_job_result;

typedef struct lg_server_stats {
  uint64_t submitted, done, running, queued;
  double p50_latency_s, p99_latency_s, mean_latency_s;
  double updates_per_s;    /* operator updates of all jobs / busy wall time */
  double busy_s;           /* wall time with at least one job resident */
  uint64_t total_updates;
  uint64_t sweep_launches; /* batched sweeps executed */
  uint64_t preemptions;    /* jobs saved and re-queued by lg_server_preempt */
  uint64_t graph_captures; /* sweep graphs captured (one per admission that changes the layout, two buffers each) */
  double graph_capture_s;  /* host time spent capturing and instantiating them (part of the job latencies) */
} lg_server_stats;

int lg_server_last_error(char* buf, size_t len);
int lg_server_start(const lg_server_config* cfg, lg_server** out);
/* LG_ERR_CAPACITY if the job can never fit the configured budget */
int lg_server_submit(lg_server* s, const lg_job* job, uint64_t* id);
int lg_server_poll(lg_server* s, uint64_t id, lg_job_result* res);
/* block until the job is done or cancelled, or timeout_ms elapses (0: poll) */
int lg_server_wait(lg_server* s, uint64_t id, int64_t timeout_ms, lg_job_result* res);
int lg_server_cancel(lg_server* s, uint64_t id);
/* Preemption: a running job leaves the device at the next chunk boundary
 * with its configuration saved (spins, operators and sweep count, the content
 * of the reference worker's checkpoint, path_integral.C:111-124) and its
 * measurements kept, and waits in the queue behind the jobs already there;
 * when admitted again it continues from that configuration, so its
 * trajectory and results are those of an uninterrupted run.  Its device
 * footprint is free meanwhile.  A queued job is left as it is; a done or
 * cancelled one gives LG_ERR_STATE. */
int lg_server_preempt(lg_server* s, uint64_t id);
int lg_server_stats_get(lg_server* s, lg_server_stats* st);
int lg_server_stop(lg_server* s);

#ifdef __cplusplus
}
#endif

#endif /* LOOPGPU_SERVER_H */
