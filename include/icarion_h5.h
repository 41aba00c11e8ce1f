/*
 * icarion_h5.h - C ABI of the trajectory file writer (libicarion_h5.so, built
 * against the HDF5 C library): the /ions and /trajectory groups of the
 * reference's simulation output file, fed by the snapshots of
 * icarion_gpu_step_snapshots.
 *
 * Layout (src/core/io/hdf5Writer.cpp; docs/HDF5_OUTPUT_STRUCTURE.md):
 *   /ions/initial_species_id   [N]        variable-length string
 *   /ions/initial_pos_{x,y,z}  [N]        float64
 *   /ions/initial_vel_{x,y,z}  [N]        float64
 *   /ions/birth_time_s         [N]        float64 (0 when not given)
 *   /ions/death_time_s         [N]        float64 (-1 while alive; rewritten by
 *                                          icarion_h5_update_death_times)
 *   /ions/charge_C             [N]        float64
 *   /trajectory/time           [T]        float64, chunk min(T0, 1000)
 *   /trajectory/positions      [T, N, 3]  float64, chunk (min(T0, 100), N, 3)
 *   /trajectory/velocities     [T, N, 3]  float64, chunk (min(T0, 100), N, 3)
 *   /trajectory/domain_indices [T, N]     int32,   chunk (min(T0, 100), N)
 *   /trajectory/species_id_indices [T, N] uint32,  chunk (min(T0, 100), N)
 *   /trajectory/time_per_ion   [T, N]     float64, only when given
 *   /trajectory/species_ids    [T, N]     variable-length string, only with a
 *                                          species pool
 * Every trajectory dataset is extendible along T and compressed with deflate
 * at the level given at creation (the reference's is 2); T0 is the size of the
 * first batch written.  Snapshots are held in memory and appended in batches of
 * up to `batch` rows, as OutputManager buffers them before
 * HDF5Writer::append_trajectory_batch_flat.
 *
 * Conventions as icarion_gpu.h: int status (0 = ok), no exception crosses the
 * ABI, the message of the last failure on the calling thread from
 * icarion_h5_last_error().  One writer is driven by one thread.
 */
#ifndef ICARION_H5_H
#define ICARION_H5_H

#include <stddef.h>
#include <stdint.h>

#include "icarion_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct icarion_h5 icarion_h5;

/* Create (truncate) the file, write /ions from the initial state and an empty
 * /trajectory group (HDF5Writer::create_output_file's ion part,
 * hdf5Writer.cpp:137-147, write_ion_metadata :1656-1730).  species_pool maps
 * species indices to names (NULL: /ions/initial_species_id holds the index as
 * text and no /trajectory/species_ids is written).  deflate_level 0-9;
 * batch = snapshots per append (0: 100). */
int icarion_h5_create(const char *path, const icarion_ion_soa *initial, const char *const *species_pool,
                      uint32_t n_species_pool, int deflate_level, uint32_t batch, icarion_h5 **out);

/* Add one snapshot (n ions of the initial state's count).  time_per_ion may be
 * NULL (the dataset is then never written); domain_index may be NULL (0). */
int icarion_h5_append(icarion_h5 *h, double t_s, const double *positions, const double *velocities,
                      const int32_t *domain_index, const uint32_t *species, const double *time_per_ion);

/* The callback form of icarion_h5_append for icarion_gpu_step_snapshots
 * (user = the icarion_h5 handle); returns non-zero (stop) after a failure. */
int icarion_h5_on_snapshot(void *user, uint64_t index, uint64_t step, double t_s, size_t n,
                           const double *positions, const double *velocities, const uint32_t *species,
                           const int32_t *domain_index, const uint8_t *flags);

/* Append the buffered snapshots now. */
int icarion_h5_flush(icarion_h5 *h);

/* Rewrite /ions/death_time_s from the final state
 * (HDF5Writer::update_death_times, hdf5Writer.cpp:711-729). */
int icarion_h5_update_death_times(icarion_h5 *h, const double *death_time_s, size_t n);

/* Flush and close; the handle is freed also when the flush fails. */
int icarion_h5_close(icarion_h5 *h);

/* Snapshots written so far (after the last flush) and buffered. */
uint64_t icarion_h5_rows(const icarion_h5 *h);

size_t icarion_h5_last_error(char *buf, size_t len);

#ifdef __cplusplus
}
#endif

#endif /* ICARION_H5_H */
