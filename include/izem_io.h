/*
 * izem_io.h -- restart files of the IZEM solver state in the reference's
 * HDF5 layout (libizem_io.so; host code, HDF5 C library).
 *
 * Replaces save_restart_hdf5 and the file part of read_restart_hdf5
 * (devel-izem/src/module_io.f90:51-104, 107-172): group /consvar with one
 * FP64 dataset per conserved variable ("rho", "rhoU", "rhoV", "rhoW",
 * "rhoE", then each species under its name) of C shape [ntz][nty][ntx],
 * and /simulation/time.  v is a HOST array in the solver layout of
 * include/izem_hip.h (ghosts included, field-major); only the interior is
 * written or read.  One process holds the whole grid (the reference writes
 * each rank's hyperslab of the same datasets).
 */
#ifndef IZEM_IO_H
#define IZEM_IO_H

#include "izem_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define IZEM_IO_OK 0
#define IZEM_IO_E_ARG -1
#define IZEM_IO_E_HDF5 -2     /* the file could not be created, opened, read or written */
#define IZEM_IO_E_NOMEM -3
#define IZEM_IO_E_SHAPE -4    /* a dataset's extent differs from the grid */
#define IZEM_IO_E_MISSING -5  /* a variable (species) is not in the file */

/* save_restart_hdf5 (module_io.f90:51-104): nrv species names in order */
int izem_restart_write(const char *path, const izem_grid *grid, int nrv,
                       const char *const *species, const double *v, double time);
/* the interior extent (nx, ny, nz) of a restart file and its time (may be
 * NULL) */
int izem_restart_dims(const char *path, int dims[3], double *time);
/* the datasets of read_restart_hdf5 (module_io.f90:146-169) into the
 * interior of v; *time (may be NULL) from /simulation/time */
int izem_restart_read(const char *path, const izem_grid *grid, int nrv,
                      const char *const *species, double *v, double *time);

#ifdef __cplusplus
}
#endif
#endif
