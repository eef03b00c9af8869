/*
 * sphgpu_hdf5.h -- HDF5 snapshot files in the reference's layout.
 *
 * A separate library (libsphgpu_hdf5.so) holds these entry points, so that
 * libsphgpu.so does not depend on the HDF5 runtime; it links libsphgpu.so for
 * the context calls and the metadata JSON.  Conventions as in sphgpu.h: 0 on
 * success or a negative SPHGPU_E* code, host field blocks are
 * [SPHGPU_NFIELDS][n] FP64 with index == id.
 *
 * Layout (HDF5Writer, src/writers/hdf5_writer.cpp:20-271):
 *   /metadata   attributes format_version, step, particle_count (int),
 *               time_code, time_physical, gamma (double), and the strings
 *               simulation_name, timestamp, sph_type, kernel_type,
 *               unit_system (UnitSystem JSON) and metadata_json (the whole
 *               OutputMetadata as JSON)
 *   /energy     datasets kinetic, thermal, potential, total ([1] double)
 *   /particles  datasets id (int), pos_x .. acc_z, mass, dens, pres, ene,
 *               sml, sound, alpha, balsara, gradh, phi (double), neighbor
 *               (int), each [n]; with compression > 0 chunked by
 *               min(n, 1024) and deflated at that level (clamped to 0..9)
 */
#ifndef SPHGPU_HDF5_H
#define SPHGPU_HDF5_H

#include "sphgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* HDF5Writer::open + write_particles + close (hdf5_writer.cpp:20-59,
 * 128-230, 237-271).  Axes beyond dim are written as 0.  An empty set leaves
 * the file with /metadata and /energy and returns SPHGPU_E_ARG, as the
 * reference's write_particles fails there. */
int sphgpu_write_hdf5_fields(const char *path, const sphgpu_output_meta *meta, int32_t dim, int32_t n,
                             const double *fields, int32_t compression);
/* The same file for the device state of a context. */
int sphgpu_write_snapshot_hdf5(sphgpu_ctx *ctx, const char *path, const sphgpu_output_meta *meta,
                               int32_t compression);
/* Number of particles (the extent of /particles/id). */
int sphgpu_hdf5_row_count(const char *path, int32_t *rows);
/* HDF5Writer::read_particles + read_metadata (hdf5_writer.cpp:316-454):
 * every dataset into the host block (rows must number n and ids be
 * 0..n-1), and metadata_json through OutputMetadata::from_json into meta
 * (may be NULL).  A missing /particles group or id dataset returns
 * SPHGPU_E_ARG; another missing dataset leaves its field at 0, and a missing
 * metadata_json leaves the metadata at OutputMetadata()'s defaults, as the
 * reference's reader does. */
int sphgpu_read_hdf5_fields(const char *path, sphgpu_output_meta *meta, int32_t n, double *fields);
/* OutputManager::load_for_resume for .h5 (src/output_manager.cpp:171-183):
 * the particles into the context and the context time to time_code. */
int sphgpu_read_snapshot_hdf5(sphgpu_ctx *ctx, const char *path, sphgpu_output_meta *meta);

#ifdef __cplusplus
}
#endif
#endif /* SPHGPU_HDF5_H */
