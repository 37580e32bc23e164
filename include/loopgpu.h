/*
 * loopgpu: gfx950 (MI355X) continuous-imaginary-time loop-cluster update for
 * S=1/2 Heisenberg / XXZ models -- the C-ABI of libloopgpu.so.
 *
 * It replaces the hot path of ALPS/looper's "loop; path integral" worker
 * (path_integral.C:323-864): one call to lg_chain_sweep(h, n, rec) performs n
 * calls of loop_worker::run() (path_integral.C:323-351) and returns, per
 * sweep, the raw collector sums that loop_worker::dispatch hands to
 * coll.commit() (path_integral.C:846-862; energy.h:76-80, susceptibility.h:
 * 200-256, stiffness.h:123-128).  Everything below the worker -- diagonal
 * update and world-line linking (path_integral.C:380-588), cluster
 * identification (looper/union_find.h:242-350), cluster flip
 * (path_integral.C:781-825) and improved estimators (path_integral.C:651-778)
 * -- runs on the device.
 *
 * Conventions: plain C types only; every function returns LG_OK (0) or a
 * negative status and never throws; lg_last_error() gives the message of the
 * last failure on the calling thread.  Device memory is owned by the library;
 * caller buffers are host memory.  A handle is used from one thread at a time.
 */
#ifndef LOOPGPU_H
#define LOOPGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  LG_OK = 0,
  LG_ERR_ARG = -1,       /* invalid argument (reference: std::invalid_argument) */
  LG_ERR_DEVICE = -2,    /* HIP runtime failure or no device */
  LG_ERR_CAPACITY = -3,  /* device buffer capacity exceeded */
  LG_ERR_STATE = -4,     /* call not valid in the handle's current state */
  LG_ERR_UNSUPPORTED = -5
};

enum { LG_LATTICE_CHAIN = 0, LG_LATTICE_SQUARE = 1 };

/* largest 2S taken (S = 4): nv = 2S virtual sites per real site */
enum { LG_MAX_TWO_S = 8 };

/* Parameters of one Markov chain.  Mirrors the worker's parameters:
 * LATTICE ("chain lattice" / "square lattice"), L, T or BETA, J or Jxy/Jz,
 * SEED, FORCE_SCATTER, local_S (loop_config.h, lattice.h:304-307,
 * temperature.h:43-79, weight_impl.h:165-188); the spin model is the XXZ
 * model H = sum_b [Jz Sz Sz + Jxy/2 (S+S- + S-S+)] with spin S = two_s / 2.
 * S > 1/2 (two_s = 2 .. LG_MAX_TWO_S, both lattices) is swept as looper
 * sweeps it: nv = 2S virtual S=1/2 sites per real site
 * (virtual site r*nv + a), nv^2 virtual bonds per real bond (real bond rb =
 * (rs, rt) carries rb*nv^2 + a*nv + c = (rs*nv + a, rt*nv + c)), and a random
 * spin-conserving permutation joining the tau = 1 and tau = 0 ends of each
 * real site's virtual world lines (path_integral.C:589-615); sites, bonds,
 * spins and operators at the ABI are then the virtual ones.  A transverse
 * field (hx != 0) adds one bond per (virtual) site after the lattice's bonds:
 * bond Nb_lattice + s holds site s's operators (graph 0, one leg below and
 * one above on site s; an off-diagonal one flips that spin), so the bond
 * count at the ABI (bond_offsets of lg_chain_get_config / set_config) is
 * Nb_lattice + N.  A single-ion term (d != 0, S > 1/2) adds, after the
 * lattice's bonds and before those site bonds, the in-site bonds of each real
 * site r: for its virtual sites a < c (a outer), bond (r*nv + a, r*nv + c),
 * nv (nv - 1) / 2 per real site (looper's in-real-site bonds).
 * chain_id selects an independent random stream for the same seed. */
typedef struct lg_params {
  int32_t lattice;       /* LG_LATTICE_CHAIN or LG_LATTICE_SQUARE */
  int32_t L;             /* linear size (even, >= 2); chain: L sites, square: L*L */
  double beta;           /* inverse temperature 1/T */
  double jxy;            /* transverse coupling Jxy (J for the Heisenberg model) */
  double jz;             /* longitudinal coupling Jz */
  double force_scatter;  /* FORCE_SCATTER in [0,1] (0 = standard weights) */
  uint32_t seed;         /* SEED */
  uint32_t chain_id;     /* independent stream index */
  double h;              /* longitudinal field h (site term -h S^z, looper's FIELD path) */
  int32_t two_s;         /* 2S: 1 (S=1/2; 0 is taken as 1) .. LG_MAX_TWO_S */
  double hx;             /* transverse field hx (site term -hx S^x, looper/model_parameter.h:67);
                            sign-free only with jxy <= 0: with jxy > 0 every sweep records the
                            configuration's sign (lg_sweep_record.sign) */
  double d;              /* single-ion term D (S^z)^2 (model_parameter.h:69); for S > 1/2 it adds
                            in-site bonds; looper's worker scatters (FORCE_SCATTER 0.1) when D > 0 */
  int32_t algorithm;     /* LG_ALGORITHM_PATH_INTEGRAL ("loop; path integral") or LG_ALGORITHM_SSE
                            ("loop; sse", sse.C): the stochastic series expansion, without a
                            longitudinal field (h = 0), at most 4096 (virtual) sites and 4096 bonds
                            (in-site and site bonds included); a batch holds chains of one
                            algorithm */
} lg_params;

enum { LG_ALGORITHM_PATH_INTEGRAL = 0, LG_ALGORITHM_SSE = 1 };

/* Per-sweep collector sums (all improved estimators) -- see the collector of
 * susceptibility.h:177-198 and stiffness.h:117-121. */
typedef struct lg_sweep_record {
  double n_op;          /* operators after the diagonal update ("nop") */
  double n_clusters;    /* Number of Clusters */
  double energy;        /* energy_offset - nop/beta (path_integral.C:851) */
  double umag0;         /* sum_c umag0_c            (tau = 0 estimates) */
  double usize2, umag2, usize4, umag4;   /* sum_c usize0^2, umag0^2, usize0^4, umag0^4 */
  double usize, umag;   /* sum_c usize_c^2, umag_c^2 (tau-integrated) */
  double smag0;
  double ssize2, smag2, ssize4, smag4;
  double ssize, smag;
  double w2;            /* sum_c sum_d (w_{c,d}/2)^2 */
  double beta;
  double sweep;         /* sweep index (mcs) of this record */
  /* normal estimator (susceptibility.h:258-339), on the configuration before
   * the cluster flip: tau-integrated magnetization sum_c umag_c and its
   * staggered form (umag0 / smag0 above are the tau = 0 ones) */
  double umag_a, smag_a;
  double e_field;       /* Zeeman energy after the flip, sum_c (flip ? -w_c : w_c), included in energy */
  double sign;          /* sign of the configuration's weight, (-1)^(off-diagonal operators of negative
                           weight) (path_integral.C:551-576); 1 for sign-free models */
} lg_sweep_record;

typedef struct lg_chain lg_chain;   /* one chain, plain one-chain-per-launch path */
typedef struct lg_batch lg_batch;   /* many chains packed into each launch */

/* ---- library ------------------------------------------------------------ */
int lg_last_error(char* buf, size_t len);
int lg_device_count(int* n);

/* ---- plain path: one chain per launch ------------------------------------ */
/* reference: loop_worker::loop_worker(alps::Parameters const&)
 * (path_integral.C:202-305); all-up spins, empty operator string. */
int lg_chain_create(const lg_params* p, lg_chain** out);
/* reference: n x loop_worker::run(obs) (path_integral.C:323-351); rec may be
 * NULL, else it receives n records. */
int lg_chain_sweep(lg_chain* h, uint32_t n, lg_sweep_record* rec);
/* sizes of the current configuration */
int lg_chain_info(lg_chain* h, int32_t* n_sites, int32_t* n_bonds, int64_t* n_ops, uint64_t* sweep);
/* configuration (checkpoint content of path_integral.C:111-124): spins at
 * tau=0 [n_sites] (0 = up), operators in bond-major order with bond offsets
 * [n_bonds+1], times [n_ops] and type (offdiagonal | graph << 2, operator.h
 * layout) [n_ops]. */
int lg_chain_get_config(lg_chain* h, int32_t* spins, int32_t* bond_offsets, double* times,
                        int32_t* types);
int lg_chain_set_config(lg_chain* h, const int32_t* spins, const int32_t* bond_offsets,
                        const double* times, const int32_t* types, uint64_t sweep);
int lg_chain_destroy(lg_chain* h);
/* Annealing (reference: looper::temperature, looper/temperature.h:56-113, as
 * loop_worker::run applies it, path_integral.C:325): n pairs (steps[k],
 * temperatures[k]) with steps ascending from 1, the last pair giving the
 * first step at the chain's own final temperature 1/p->beta.  Sweep (Monte
 * Carlo step) m runs at 1/T(m), T piecewise linear between the pairs
 * (temperatures[0] up to steps[0], the final temperature from the last step
 * on; step 0 takes temperatures[0]).  The inverse temperature advances on the
 * device, so captured sweeps stay valid; records carry each sweep's beta.
 * Capacity grows to the schedule's largest beta (the configuration is kept).
 * n = 0 returns to the fixed temperature.  Single chains only. */
int lg_chain_set_schedule(lg_chain* h, int32_t n, const int32_t* steps, const double* temperatures);
/* Checkpoint blob of a chain, in the content order of the worker's save/load
 * (path_integral.C:111-124): "LGCHAIN2", mcs (uint64 sweep count), spins
 * (uint64 count, int32 each, 0 = up), operators (uint64 count, then in
 * imaginary-time order {int32 type = offdiagonal | graph << 2, int32
 * location = bond << 1 | 1, double time}, operator.h:108-109,138-139,
 * location_impl.h:43), then the parameters (lg_params), whose seed and chain
 * id with the sweep count are the Philox key and counter, so a loaded chain
 * continues the same trajectory.  load rejects a blob of another chain.
 * Little-endian host byte order. */
int lg_chain_state_size(lg_chain* h, size_t* bytes);
int lg_chain_save(lg_chain* h, void* buf, size_t len);
int lg_chain_load(lg_chain* h, const void* buf, size_t len);

/* Cluster stage only, on a given configuration (stages 7-15 of
 * path_integral.C without the diagonal update): links world lines, identifies
 * clusters, draws flip bits for sweep index `sweep` and flips.  Outputs:
 * labels[n_sites + 4*n_ops] canonical cluster label of every element
 * (element s = tau=0 boundary of site s, n_sites + 4k + {0,1,2,3} = legs
 * l0,l1,u0,u1 of operator k), label = minimum element index of the cluster;
 * spins_out / types_out the flipped configuration; n_clusters. */
int lg_cluster_config(const lg_params* p, const int32_t* spins, const int32_t* bond_offsets,
                      const double* times, const int32_t* types, uint64_t sweep, int32_t* labels,
                      int32_t* spins_out, int32_t* types_out, int64_t* n_clusters);

/* ---- batched form: many chains per launch --------------------------------- */
int lg_batch_create(const lg_params* p, int32_t n_chains, lg_batch** out);
/* n sweeps of every chain; rec receives n * n_chains records, sweep-major */
int lg_batch_run(lg_batch* h, uint32_t n, lg_sweep_record* rec);
/* choose lane-group and grid sizes by timing candidate settings for a few
 * sweeps (the configuration advances by those sweeps) */
int lg_batch_tune(lg_batch* h, uint32_t sweeps_per_trial);
/* use a captured HIP graph for the sweep (1) or plain launches (0) */
int lg_batch_set_graph(lg_batch* h, int enable);
int lg_batch_chain_info(lg_batch* h, int32_t chain, int64_t* n_ops, uint64_t* sweep);
/* append n chains (all spins up, no operators, path_integral.C:225); the
 * batch's chains keep their configurations and sweep counts.  Rebuilds the
 * batch's device memory (launch settings return to the automatic choice). */
int lg_batch_add(lg_batch* h, const lg_params* p, int32_t n);
int lg_batch_destroy(lg_batch* h);

/* ---- replica exchange ------------------------------------------------------
 * "loop; path integral; exchange" (path_integral.C:874-875: parapack's
 * exchange worker over loop_worker's set_beta / weight_parameter /
 * log_weight, path_integral.C:97-109).  n replicas of model p at inverse
 * temperatures betas[0..n) (neighbours in this order exchange; replica t is
 * chain t with chain_id p->chain_id + t).  After every sweep the chains at
 * temperatures t and t+1 (t even on even steps, odd on odd steps) swap
 * temperatures with probability min(1, (beta[t+1]/beta[t])^(n_a - n_b)),
 * n = operator count (weight_parameter), decided on the device from Philox
 * (key seed, 0xffffffff; counter step, 4, t, 0).  The handle is a batch:
 * lg_batch_run returns each chain's records, whose `beta` field is the
 * temperature the chain had during that sweep. */
int lg_exchange_create(const lg_params* p, const double* betas, int32_t n, uint32_t seed, lg_batch** out);
/* chain_at[n]: chain holding each temperature; accepted[n-1]: swaps accepted
 * between temperatures t and t+1; steps: exchange steps done (any may be NULL) */
int lg_exchange_state(lg_batch* h, int32_t* chain_at, uint64_t* accepted, uint64_t* steps);

/* ---- launch settings ----------------------------------------------------------
 * group: lanes per bond in the diagonal update (1, 2, 4, 8, 16, 32 or 64); blocks:
 * grid of the per-operator and per-node kernels; link: linking kernel (0 one
 * thread per operator; 1 one thread per site merging its bond lists; 128, 512
 * or 1536 a group of threads per site staging that many events in LDS: 16
 * threads, one wave, four waves); compress: 1 a
 * union-find compression pass before the estimator, 0 (the default) the
 * estimator finds the roots itself; flush: the estimator flushes its LDS
 * cluster table to global memory every 256 * flush segments of its chunk
 * (4, 8 or 16).  By default they follow the expected operator counts, and
 * lg_batch_tune times the candidates.  set: fields < 0 keep their value; the settings then
 * stay as given (no automatic choice) until the next tune.  Results do not
 * depend on them.  No counterpart in the reference (launch configuration of
 * the device kernels). */
typedef struct lg_launch {
  int32_t group;
  int32_t blocks;
  int32_t link;
  int32_t compress;
  int32_t flush;
} lg_launch;
int lg_chain_get_launch(lg_chain* h, lg_launch* out);
int lg_chain_set_launch(lg_chain* h, const lg_launch* l);
int lg_batch_get_launch(lg_batch* h, lg_launch* out);
int lg_batch_set_launch(lg_batch* h, const lg_launch* l);

/* ---- measurement ------------------------------------------------------------
 * n sweeps with HIP events recorded between the kernel launches on the
 * handle's stream; avg_ms[i] = mean duration per sweep of kernel i, in the
 * order of lg_profile_kernel_name(i), i < lg_profile_num_kernels().  The
 * configuration advances by n sweeps (records are not returned).  These are
 * the reference's timer regions (path_integral.C:284-299) as device kernels. */
int lg_profile_num_kernels(void);
const char* lg_profile_kernel_name(int i);
int lg_chain_profile(lg_chain* h, uint32_t n, double* avg_ms);
int lg_batch_profile(lg_batch* h, uint32_t n, double* avg_ms);
/* Plain device copy of `bytes` bytes repeated `iters` times (kernel
 * k_copy_stream); gbs = (read + write bytes) / time.  No counterpart in the
 * reference: it calibrates the HBM byte counters and measures the achievable
 * bandwidth against which the kernels are judged. */
int lg_copy_bandwidth(uint64_t bytes, int iters, double* gbs);

#ifdef __cplusplus
}
#endif

#endif /* LOOPGPU_H */
