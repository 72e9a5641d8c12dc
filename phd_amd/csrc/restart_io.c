/*
 * restart_io.c -- restart files in the reference's HDF5 layout.
 *
 * Reference: save_restart_hdf5 / read_restart_hdf5
 * (devel-izem/src/module_io.f90:51-195).  A restart holds the conserved
 * variables of the interior, one 3-D FP64 dataset per variable in group
 * /consvar ("rho", "rhoU", "rhoV", "rhoW", "rhoE", then one per species
 * under its name), each of C shape [ntz][nty][ntx] (x fastest: the
 * reference's Fortran (ntx, nty, ntz)), and the time as a 1-element FP64
 * dataset /simulation/time.  Values are the reference's nondimensional ones.
 *
 * The arrays handed in and out are host arrays in the solver layout
 * v(sx-ng:ex+ng, sy-ng:ey+ng, sz-ng:ez+ng, 1:nv) (include/izem_hip.h); only
 * the interior is written or read.  Host-side I/O: built with the host
 * compiler against the HDF5 C library, no GPU code.
 */
#include <hdf5.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/izem_io.h"

static const char *k_flow_names[5] = {"rho", "rhoU", "rhoV", "rhoW", "rhoE"};

static long cell_index(const izem_grid *g, int i, int j, int k) {
  const long NX = g->nx + 2L * g->ng, NY = g->ny + 2L * g->ng;
  return (long)(i + g->ng) + NX * ((long)(j + g->ng) + NY * (long)(k + g->ng));
}

static long field_size(const izem_grid *g) {
  return (long)(g->nx + 2 * g->ng) * (g->ny + 2 * g->ng) * (g->nz + 2 * g->ng);
}

static int grid_ok(const izem_grid *g) {
  return g && g->nx > 0 && g->ny > 0 && g->nz > 0 && g->ng >= 0;
}

/* interior of field l of v <-> a packed [nz][ny][nx] buffer */
static void pack(const izem_grid *g, const double *v, int l, double *buf) {
  const long nc = field_size(g);
  long q = 0;
  for (int k = 0; k < g->nz; k++)
    for (int j = 0; j < g->ny; j++)
      for (int i = 0; i < g->nx; i++) buf[q++] = v[l * nc + cell_index(g, i, j, k)];
}

static void unpack(const izem_grid *g, const double *buf, int l, double *v) {
  const long nc = field_size(g);
  long q = 0;
  for (int k = 0; k < g->nz; k++)
    for (int j = 0; j < g->ny; j++)
      for (int i = 0; i < g->nx; i++) v[l * nc + cell_index(g, i, j, k)] = buf[q++];
}

int izem_restart_write(const char *path, const izem_grid *g, int nrv,
                       const char *const *species, const double *v, double time) {
  if (!path || !grid_ok(g) || nrv < 0 || (nrv > 0 && !species) || !v) return IZEM_IO_E_ARG;
  const long n = (long)g->nx * g->ny * g->nz;
  double *buf = malloc(sizeof(double) * n);
  if (!buf) return IZEM_IO_E_NOMEM;
  int rc = IZEM_IO_E_HDF5;
  hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  hid_t grp = -1, gs = -1, sp = -1, sp1 = -1;
  if (f < 0) goto done;
  grp = H5Gcreate2(f, "/consvar", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  if (grp < 0) goto done;
  hsize_t dims[3] = {(hsize_t)g->nz, (hsize_t)g->ny, (hsize_t)g->nx};
  sp = H5Screate_simple(3, dims, NULL);
  if (sp < 0) goto done;
  for (int l = 0; l < 5 + nrv; l++) {
    const char *name = l < 5 ? k_flow_names[l] : species[l - 5];
    if (!name || !*name) {
      rc = IZEM_IO_E_ARG;
      goto done;
    }
    pack(g, v, l, buf);
    hid_t ds = H5Dcreate2(grp, name, H5T_IEEE_F64LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (ds < 0) goto done;
    herr_t e = H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf);
    H5Dclose(ds);
    if (e < 0) goto done;
  }
  gs = H5Gcreate2(f, "/simulation", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  if (gs < 0) goto done;
  hsize_t one = 1;
  sp1 = H5Screate_simple(1, &one, NULL);
  if (sp1 < 0) goto done;
  {
    hid_t ds = H5Dcreate2(gs, "time", H5T_IEEE_F64LE, sp1, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (ds < 0) goto done;
    herr_t e = H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, &time);
    H5Dclose(ds);
    if (e < 0) goto done;
  }
  rc = IZEM_IO_OK;
done:
  if (sp1 >= 0) H5Sclose(sp1);
  if (sp >= 0) H5Sclose(sp);
  if (gs >= 0) H5Gclose(gs);
  if (grp >= 0) H5Gclose(grp);
  if (f >= 0 && H5Fclose(f) < 0 && rc == IZEM_IO_OK) rc = IZEM_IO_E_HDF5;
  free(buf);
  return rc;
}

int izem_restart_dims(const char *path, int dims[3], double *time) {
  if (!path || !dims) return IZEM_IO_E_ARG;
  int rc = IZEM_IO_E_HDF5;
  hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) return IZEM_IO_E_HDF5;
  hid_t ds = H5Dopen2(f, "/consvar/rho", H5P_DEFAULT);
  if (ds >= 0) {
    hid_t sp = H5Dget_space(ds);
    hsize_t d[3] = {0, 0, 0};
    if (sp >= 0 && H5Sget_simple_extent_ndims(sp) == 3 && H5Sget_simple_extent_dims(sp, d, NULL) == 3) {
      dims[0] = (int)d[2];  /* nx */
      dims[1] = (int)d[1];
      dims[2] = (int)d[0];
      rc = IZEM_IO_OK;
    }
    if (sp >= 0) H5Sclose(sp);
    H5Dclose(ds);
  }
  if (rc == IZEM_IO_OK && time) {
    hid_t dt = H5Dopen2(f, "/simulation/time", H5P_DEFAULT);
    if (dt < 0 || H5Dread(dt, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, time) < 0)
      rc = IZEM_IO_E_HDF5;
    if (dt >= 0) H5Dclose(dt);
  }
  H5Fclose(f);
  return rc;
}

int izem_restart_read(const char *path, const izem_grid *g, int nrv,
                      const char *const *species, double *v, double *time) {
  if (!path || !grid_ok(g) || nrv < 0 || (nrv > 0 && !species) || !v) return IZEM_IO_E_ARG;
  int dims[3];
  int rc = izem_restart_dims(path, dims, time);
  if (rc != IZEM_IO_OK) return rc;
  if (dims[0] != g->nx || dims[1] != g->ny || dims[2] != g->nz) return IZEM_IO_E_SHAPE;
  const long n = (long)g->nx * g->ny * g->nz;
  double *buf = malloc(sizeof(double) * n);
  if (!buf) return IZEM_IO_E_NOMEM;
  rc = IZEM_IO_E_HDF5;
  hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  hid_t grp = -1;
  if (f < 0) goto done;
  grp = H5Gopen2(f, "/consvar", H5P_DEFAULT);
  if (grp < 0) goto done;
  for (int l = 0; l < 5 + nrv; l++) {
    const char *name = l < 5 ? k_flow_names[l] : species[l - 5];
    if (!name || !*name) {
      rc = IZEM_IO_E_ARG;
      goto done;
    }
    if (H5Lexists(grp, name, H5P_DEFAULT) <= 0) {
      rc = IZEM_IO_E_MISSING;
      goto done;
    }
    hid_t ds = H5Dopen2(grp, name, H5P_DEFAULT);
    if (ds < 0) goto done;
    hid_t sp = H5Dget_space(ds);
    hsize_t d[3] = {0, 0, 0};
    const int ok = sp >= 0 && H5Sget_simple_extent_ndims(sp) == 3 &&
                   H5Sget_simple_extent_dims(sp, d, NULL) == 3 && d[0] == (hsize_t)g->nz &&
                   d[1] == (hsize_t)g->ny && d[2] == (hsize_t)g->nx;
    if (sp >= 0) H5Sclose(sp);
    herr_t e = ok ? H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) : -1;
    H5Dclose(ds);
    if (!ok) {
      rc = IZEM_IO_E_SHAPE;
      goto done;
    }
    if (e < 0) goto done;
    unpack(g, buf, l, v);
  }
  rc = IZEM_IO_OK;
done:
  if (grp >= 0) H5Gclose(grp);
  if (f >= 0) H5Fclose(f);
  free(buf);
  return rc;
}
