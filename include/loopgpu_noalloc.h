/*
 * loopgpu: an entry point that allocates nothing.
 *
 * lg_cluster_config_noalloc is lg_cluster_config (include/loopgpu.h) for
 * callers that cannot allocate device memory: code that runs while a HIP
 * graph is being captured, tight serving loops, harnesses that forbid
 * hipMalloc.  It has no counterpart in the reference.  It calls no hipMalloc,
 * hipFree or other allocating routine: every buffer it works in lies in one
 * static device array of lg_noalloc_scratch_bytes() bytes per GPU (about
 * 16 MiB), which each GPU the library runs on holds until the process ends.  Its
 * address is looked up once per device, at the first call on that device.
 *
 * Inputs and outputs are the caller's DEVICE pointers, with the meaning of
 * lg_cluster_config's host arrays: spins[n_sites] (0 = up), bond_offsets
 * [n_bonds + 1], times[n_ops], types[n_ops] (offdiagonal | graph << 2) in;
 * labels[n_sites + 4 * n_ops] (canonical cluster label of every element),
 * spins_out[n_sites], types_out[n_ops] out (any output may be NULL).
 * n_clusters is a host pointer (may be NULL).  n_ops must equal
 * bond_offsets[n_bonds].  The configuration is validated on the device
 * (LG_ERR_ARG as lg_cluster_config).  At most LG_NOALLOC_MAX_SITES sites and
 * LG_NOALLOC_MAX_OPS operators: larger inputs give LG_ERR_CAPACITY before any
 * device work.  The call is synchronous on the device's default stream.
 *
 * Every call on a GPU shares that GPU's array, which keeps the last call's
 * contents (each call clears what must start at zero).  Calls that overlapped
 * on one GPU would race on the array: the library serialises the calls of
 * one process's threads (each process holds its own array).
 */
#ifndef LOOPGPU_NOALLOC_H
#define LOOPGPU_NOALLOC_H

#include <stddef.h>
#include <stdint.h>

#include "loopgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

#define LG_NOALLOC_MAX_SITES 16384
#define LG_NOALLOC_MAX_OPS 65536

/* size of the static device array (bytes) */
size_t lg_noalloc_scratch_bytes(void);

int lg_cluster_config_noalloc(const lg_params* p, const int32_t* spins, const int32_t* bond_offsets,
                              const double* times, const int32_t* types, int64_t n_ops, uint64_t sweep,
                              int32_t* labels, int32_t* spins_out, int32_t* types_out, int64_t* n_clusters);

#ifdef __cplusplus
}
#endif

#endif /* LOOPGPU_NOALLOC_H */
