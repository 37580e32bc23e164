/*
 * loopgpu worker: the reference worker's parameter and update/measure
 * interface over libloopgpu.so.
 *
 * Mirrors alps::parapack::mc_worker as implemented by loop_worker
 * (path_integral.C:57-351): construction from parameters
 * (loop_worker(alps::Parameters const&), path_integral.C:202-305),
 * run() = one Monte Carlo step (path_integral.C:323-351), is_thermalized() /
 * progress() (montecarlo.h:55-61), save()/load() (path_integral.C:111-124),
 * and the results of the evaluator (energy.h:89-102,
 * susceptibility.h:340-374) with binning errors -- the observable names of
 * init_observables (path_integral.C:307-321, energy.h:49-53,
 * susceptibility.h:59-94, stiffness.h:78-80).
 *
 * Parameters are text in the reference's .ip syntax ("KEY = value;" or one
 * per line; expressions over other parameters allowed), keys: ALGORITHM
 * ("loop; path integral" (default) | "loop; sse"), LATTICE
 * ("chain lattice" | "square lattice"), L, W, T | TEMPERATURE | BETA |
 * INVERSE_TEMPERATURE, the annealing schedule T_START_i / T_DURATION_i (or
 * TEMPERATURE_START_i / TEMPERATURE_DURATION_i, BETA_START_i /
 * BETA_DURATION_i, INVERSE_TEMPERATURE_START_i / INVERSE_TEMPERATURE_DURATION_i;
 * temperature.h:56-113, no longer than THERMALIZATION, path_integral.C:
 * 212-213), J, Jxy, Jz, MODEL (= spin), local_S (1/2, 1, 3/2, ... 4), SWEEPS
 * (default "[65536:]"), THERMALIZATION (default SWEEPS.min/8), SEED,
 * FORCE_SCATTER (default 0.1 where D > 0 and S > 1/2, as looper), CHAIN_ID, h,
 * Gamma (transverse field; sign-free with Jxy <= 0, with Jxy > 0 the worker
 * measures "Sign" and reports sign-weighted normal estimators), D (single-ion term),
 * DISABLE_IMPROVED_ESTIMATOR,
 * MEASURE[Correlations], MEASURE[Structure Factor], INITIAL_SITE
 * (correlation.h:59-109).  Invalid or unsupported parameters give
 * LG_ERR_ARG (the reference throws std::invalid_argument).
 */
#ifndef LOOPGPU_WORKER_H
#define LOOPGPU_WORKER_H

#include <stddef.h>
#include <stdint.h>

#include "loopgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lg_worker lg_worker;

int lg_worker_last_error(char* buf, size_t len);
/* reference: loop_worker(alps::Parameters const&) + init_observables */
int lg_worker_create(const char* params, lg_worker** out);
/* up to max_steps calls of run() while the schedule can_work()
 * (montecarlo.h:55-57); measurements start once is_thermalized() */
int lg_worker_run(lg_worker* h, uint64_t max_steps, uint64_t* done);
/* reference: mcs(), is_thermalized(), progress(), can_work() */
int lg_worker_status(lg_worker* h, uint64_t* mcs, int* thermalized, double* progress, int* can_work);
/* mean, binning error and integrated autocorrelation time of an observable
 * or evaluated quantity ("Specific Heat", "Binder Ratio of ...") */
int lg_worker_result(lg_worker* h, const char* name, double* mean, double* error, double* tau);
/* "[results]" block, one "Name: mean +/- error; tau = t" line per observable */
int lg_worker_report(lg_worker* h, char* buf, size_t len, size_t* needed);
/* checkpoint blob: "LGWORKR2", the Monte Carlo counters (mcs, SWEEPS min
 * and max, THERMALIZATION; mc_steps::save, montecarlo.h:66), the chain's
 * blob (lg_chain_save: mcs, spins, operators -- the content of
 * loop_worker::save, path_integral.C:111-124), then every observable's
 * binning state (what parapack checkpoints beside the worker), so a worker
 * loaded from it ends with the results of an uninterrupted run.  Valid at
 * any step count: lg_worker_run never advances the device beyond the steps
 * it was asked for.  *len receives the size (buf may be NULL to query it). */
int lg_worker_save(lg_worker* h, void* buf, size_t* len);
int lg_worker_load(lg_worker* h, const void* buf, size_t len);
/* sweeps fetched from the device per batch of run() calls (default 1024;
 * 1 while correlations are measured, which need every sweep's spins) */
int lg_worker_set_chunk(lg_worker* h, uint32_t sweeps);
int lg_worker_destroy(lg_worker* h);

/* The temperature schedule a parameter set defines (reference:
 * looper::temperature::init, temperature.h:43-81, as test/temperature.C
 * prints it): *n (start step, temperature) pairs, the last one (first step
 * at the final temperature, final temperature); n = 1 without annealing.
 * Initial temperature = temperatures[0], annealing steps = steps[n-1] - 1.
 * The worker runs Monte Carlo step m at 1 / T(m) (lg_chain_set_schedule).
 * Host computation, no device needed.  LG_ERR_ARG when no temperature is
 * defined or more than max_n pairs result (*n still receives the count). */
int lg_temperature_schedule(const char* params, int32_t max_n, int32_t* n, int32_t* steps, double* temperatures);

/* The correlation estimators of one tau = 0 configuration (reference:
 * looper::correlation::normal_estimator::collector::commit,
 * correlation.h:230-296), as the worker measures them with
 * MEASURE[Correlations] / MEASURE[Structure Factor] (and INITIAL_SITE):
 * spins[N] (0 = up) of an even-L chain (N = L) or square (N = L^2, site
 * x*L + y) lattice; origin < 0 for all pairs.  Outputs N values each:
 * ucorr / scorr ("Spin Correlations" / "Staggered Spin Correlations") by
 * displacement index (dx*L + dy), or by site with an origin; sfac ("Spin
 * Structure Factor") by momentum index (n1*L + n2, k = 2 pi n / L folded into
 * (-pi, pi]).  Host computation, no device needed. */
int lg_spin_correlations(int32_t lattice, int32_t L, const uint8_t* spins, int32_t origin, double* ucorr,
                         double* scorr, double* sfac);

#ifdef __cplusplus
}
#endif

#endif /* LOOPGPU_WORKER_H */
