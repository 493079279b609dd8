/*
 * chy_output.c -- per-step field output, the counterpart of OutputHDF5
 * (reference fileio.c:73-126): six f64 datasets of shape (Nz, Ny, Nx) named
 * as the DMDA fields of setup.c:53-56 plus "rhs" and "potential"; and the
 * solver's input, the counterpart of LoadVlasovQuantities (fileio.c:6-70).
 *
 * HDF5's C library is opened with dlopen so libchybrid.so has no link-time
 * dependency on it; the handful of entry points used are resolved by name.
 * CHY_NO_HDF5=1 in the environment skips the library (raw-file output), as on
 * a host without HDF5.
 */
#include "chybrid.h"

#include <dlfcn.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef int64_t hid_t;          /* HDF5 >= 1.10 */
typedef int herr_t;
typedef unsigned long long hsize_t;
typedef long long hssize_t;

static struct {
  int tried, ok;
  herr_t (*open)(void);
  hid_t (*fcreate)(const char *, unsigned, hid_t, hid_t);
  hid_t (*screate)(int, const hsize_t *, const hsize_t *);
  hid_t (*dcreate)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t, hid_t);
  herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void *);
  herr_t (*dclose)(hid_t), (*sclose)(hid_t), (*fclose)(hid_t);
  hid_t *native_double;
  /* reading */
  int rok;
  hid_t (*fopen)(const char *, unsigned, hid_t);
  hid_t (*dopen)(hid_t, const char *, hid_t);
  hid_t (*dspace)(hid_t);
  hssize_t (*npoints)(hid_t);
  herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void *);
} H5;

static void h5_load_once(void);
static pthread_once_t h5_once = PTHREAD_ONCE_INIT;
/* resolved once per process, whichever thread writes first */
static void h5_load(void) { pthread_once(&h5_once, h5_load_once); }

static void h5_load_once(void)
{
  if (H5.tried) return;
  H5.tried = 1;
  const char *off = getenv("CHY_NO_HDF5");
  if (off && off[0] && strcmp(off, "0") != 0) return;
  const char *names[] = {"libhdf5.so", "libhdf5.so.103", "libhdf5_serial.so", "/opt/conda/lib/libhdf5.so", NULL};
  void *h = NULL;
  for (int i = 0; names[i] && !h; i++) h = dlopen(names[i], RTLD_NOW | RTLD_LOCAL);
  if (!h) return;
  *(void **)&H5.open = dlsym(h, "H5open");
  *(void **)&H5.fcreate = dlsym(h, "H5Fcreate");
  *(void **)&H5.screate = dlsym(h, "H5Screate_simple");
  *(void **)&H5.dcreate = dlsym(h, "H5Dcreate2");
  *(void **)&H5.dwrite = dlsym(h, "H5Dwrite");
  *(void **)&H5.dclose = dlsym(h, "H5Dclose");
  *(void **)&H5.sclose = dlsym(h, "H5Sclose");
  *(void **)&H5.fclose = dlsym(h, "H5Fclose");
  H5.native_double = (hid_t *)dlsym(h, "H5T_NATIVE_DOUBLE_g");
  *(void **)&H5.fopen = dlsym(h, "H5Fopen");
  *(void **)&H5.dopen = dlsym(h, "H5Dopen2");
  *(void **)&H5.dspace = dlsym(h, "H5Dget_space");
  *(void **)&H5.npoints = dlsym(h, "H5Sget_simple_extent_npoints");
  *(void **)&H5.dread = dlsym(h, "H5Dread");
  H5.ok = H5.open && H5.fcreate && H5.screate && H5.dcreate && H5.dwrite && H5.dclose && H5.sclose &&
          H5.fclose && H5.native_double && H5.open() >= 0;
  H5.rok = H5.ok && H5.fopen && H5.dopen && H5.dspace && H5.npoints && H5.dread;
}

static int fail(char *msg, size_t len, int rc, const char *what, const char *path)
{
  if (msg && len) snprintf(msg, len, "%s: %s", what, path);
  return rc;
}

int chy_read_vlasov(const chy_params *p, const char *path, double *vl, char *msg, size_t len)
{
  if (msg && len) msg[0] = 0;
  if (!p || !path || !vl) return CHY_ERR_ARG_WRONG;
  const int64_t nc = p->N[0] * p->N[1] * p->N[2];
  if (nc <= 0) return fail(msg, len, CHY_ERR_ARG_WRONG, "empty grid", path);
  double *n = malloc(sizeof(double) * (size_t)nc);
  if (!n) return CHY_ERR_MEM;
  int rc = CHY_OK;
  const size_t pl = strlen(path);
  if (pl > 4 && strcmp(path + pl - 4, ".f64") == 0) {
    FILE *fp = fopen(path, "rb");
    if (!fp) {
      rc = fail(msg, len, CHY_ERR_ARG_WRONG, "cannot open", path);
    } else {
      size_t got = fread(n, sizeof(double), (size_t)nc, fp);
      char extra;
      if (got != (size_t)nc || fread(&extra, 1, 1, fp) != 0)
        rc = fail(msg, len, CHY_ERR_ARG_WRONG, "density size differs from the grid in", path);
      fclose(fp);
    }
  } else if (h5_load(), !H5.rok) {
    rc = fail(msg, len, CHY_ERR_SUP, "no HDF5 library to read", path);
  } else {
    const hid_t file = H5.fopen(path, 0x0000u /* H5F_ACC_RDONLY */, 0 /* H5P_DEFAULT */);
    if (file < 0) {
      rc = fail(msg, len, CHY_ERR_ARG_WRONG, "cannot open HDF5 file", path);
    } else {
      const hid_t ds = H5.dopen(file, "density-kji", 0);
      const hid_t sp = ds >= 0 ? H5.dspace(ds) : -1;
      const hssize_t np = sp >= 0 ? H5.npoints(sp) : -1;
      if (ds < 0) rc = fail(msg, len, CHY_ERR_ARG_WRONG, "no dataset \"density-kji\" in", path);
      else if (np != (hssize_t)nc) rc = fail(msg, len, CHY_ERR_ARG_WRONG, "density size differs from the grid in", path);
      else if (H5.dread(ds, *H5.native_double, 0 /* H5S_ALL */, 0, 0, n) < 0)
        rc = fail(msg, len, CHY_ERR_LIB, "HDF5 read failed on", path);
      if (sp >= 0) H5.sclose(sp);
      if (ds >= 0) H5.dclose(ds);
      H5.fclose(file);
    }
  }
  for (int64_t c = 0; c < nc && !rc; c++) {
    vl[4 * c] = n[c];
    for (int a = 0; a < 3; a++) vl[4 * c + 1 + a] = p->flux_scale[a] * n[c];
  }
  free(n);
  return rc;
}

/* The six fields (f[0..3] moments, f[4] rhs, f[5] potential, each nc f64,
 * x fastest) to <stem><insert>.hdf, or to raw files without HDF5 (then
 * CHY_ERR_SUP after writing them).  Host memory only: the synchronous and the
 * asynchronous output share it. */
int chyo_write_fields(const chy_params *p, const char *stem, const char *insert, const double *const f[6])
{
  const int64_t nc = p->N[0] * p->N[1] * p->N[2];
  const char *names[6] = {"density", "x flux", "y flux", "z flux", "rhs", "potential"};
  char path[8300];
  int rc = CHY_OK;
  h5_load();
  if (H5.ok) {
    snprintf(path, sizeof path, "%s%s.hdf", stem, insert);
    const hid_t file = H5.fcreate(path, 0x0002u /* H5F_ACC_TRUNC */, 0, 0);
    if (file < 0) rc = CHY_ERR_LIB;
    const hsize_t dims[3] = {(hsize_t)p->N[2], (hsize_t)p->N[1], (hsize_t)p->N[0]};
    for (int a = 0; a < 6 && !rc; a++) {
      const hid_t sp = H5.screate(3, dims, NULL);
      const hid_t ds = H5.dcreate(file, names[a], *H5.native_double, sp, 0, 0, 0);
      if (sp < 0 || ds < 0 || H5.dwrite(ds, *H5.native_double, 0, 0, 0, f[a]) < 0) rc = CHY_ERR_LIB;
      if (ds >= 0) H5.dclose(ds);
      if (sp >= 0) H5.sclose(sp);
    }
    if (file >= 0) H5.fclose(file);
  } else {
    for (int a = 0; a < 6 && !rc; a++) {
      snprintf(path, sizeof path, "%s%s.%s.f64", stem, insert, names[a]);
      for (char *q = path + strlen(stem); *q; q++) if (*q == ' ') *q = '_';
      FILE *fp = fopen(path, "wb");
      if (!fp || fwrite(f[a], sizeof(double), (size_t)nc, fp) != (size_t)nc) rc = CHY_ERR_LIB;
      if (fp) fclose(fp);
    }
    if (!rc) rc = CHY_ERR_SUP;
  }
  return rc;
}

int chy_output_write(chy_ctx *ctx, const char *stem, const char *insert)
{
  chy_params p;
  /* one writer at a time: finish an asynchronous output first */
  int rc = chy_output_wait(ctx);
  if (rc && rc != CHY_ERR_SUP) return rc;
  rc = chy_get_params(ctx, &p);
  if (rc) return rc;
  const int64_t nc = p.N[0] * p.N[1] * p.N[2];
  double *vl = malloc(sizeof(double) * 4 * nc), *f[6];
  for (int a = 0; a < 6; a++) f[a] = malloc(sizeof(double) * nc);
  rc = chy_get_moments(ctx, vl);
  if (!rc) rc = chy_get_rhs(ctx, f[4]);
  if (!rc) rc = chy_get_potential(ctx, f[5]);
  for (int64_t c = 0; c < nc && !rc; c++)
    for (int a = 0; a < 4; a++) f[a][c] = vl[4 * c + a];
  if (!rc) rc = chyo_write_fields(&p, stem, insert, (const double *const *)f);
  free(vl);
  for (int a = 0; a < 6; a++) free(f[a]);
  return rc;
}
