/*
 * rbn_driver.c — drop-in host program over librbn (include/rbn.h).
 *
 * Replays the call sequence of the reference's two programs with their parameter blocks
 * and file names:
 *   truth run  (Trajectories/code/temppart_dump.c:1573-1769): init -> particles (writes
 *     init_prtcl{x,y,T,u,v}<nstart>.in and init_prtcl_id<nstart>.in, Trajectories/code/mympi2.h:256-289)
 *     -> loop with trajectory records Prtcl_Veldata1/Particles<istep>.h5 every noutconfig
 *     steps (/id /x /y /u /v /xo /yo /uo /vo /temp, plus /dtu../fv with Prtcl_Writefull)
 *     and evolution.dat ("%d %e %e") every nout steps;
 *   nudged run (Nudging/code/tempvel_nudge.c:2254-2419): observers from the truth run's
 *     init_prtcl*0.in, records read from <obs>/Prtcl_Veldata1/Particles<step>.h5,
 *     Tempfolder/snaptemp<istep>.h5 (/T /u /v, [nx][ny]) every noutconfig steps,
 *     profiles.<istep>.dat every nout steps, evolution.dat every 5000 steps.
 * HDF5 is loaded at run time; without it records are raw files Particles<step>.bin:
 *   int32 n, int32 id[n], then double arrays in the dataset order above.
 *
 * usage: rbn_driver [options] params-file [params-file ...]
 *   --out DIR        output directory (default .)
 *   --obs DIR        nudged run: directory of the truth run's outputs (default: --out)
 *   --device N       GPU index (default 0)
 *   --init ref | conduction:AMP:SEED
 *                    scratch temperature: the reference tanh front (default) or a linear
 *                    conduction profile with seeded uniform noise of amplitude AMP
 *   --particles FILE raw float64 x[n] then y[n] (default: the reference's 40x40 grid)
 *   --raw            write raw records even when HDF5 is available
 *   --steps N        override nsteps
 *   --verbose        print "Loop <istep> particles propagated" every step, as the reference
 *   --restart-dir DIR  scratch 0: read the checkpoint of step nstart from DIR (default --out)
 *                    instead of the reference's fixed absolute paths
 *   --dump-every N   checkpoint every N steps (default 2000000, as the reference) and after
 *                    the last step: confnew/congnew/confpre/congpre<istep>.h5 (/f, /g
 *                    [nx][ny][9]) and prtcl_dump<istep>.in (+ prtcl_dump_id<istep>.in)
 */
#define _POSIX_C_SOURCE 200809L
#include <dlfcn.h>
#include <errno.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include "rbn.h"

/* ---------------------------------------------------------------- HDF5 at run time */

typedef int64_t h5id;
typedef int h5err;
typedef unsigned long long h5size;

static struct {
  int ok;
  h5err (*open)(void);
  h5id (*fcreate)(const char *, unsigned, h5id, h5id);
  h5id (*fopen)(const char *, unsigned, h5id);
  h5err (*fclose)(h5id);
  h5id (*screate_simple)(int, const h5size *, const h5size *);
  h5err (*sclose)(h5id);
  h5id (*dcreate2)(h5id, const char *, h5id, h5id, h5id, h5id, h5id);
  h5id (*dopen2)(h5id, const char *, h5id);
  h5err (*dwrite)(h5id, h5id, h5id, h5id, h5id, const void *);
  h5err (*dread)(h5id, h5id, h5id, h5id, h5id, void *);
  h5err (*dclose)(h5id);
  h5id (*dget_space)(h5id);
  long long (*npoints)(h5id);
  h5id *t_double, *t_int;
} H5;

#define H5F_ACC_RDONLY_ 0u
#define H5F_ACC_TRUNC_ 2u
#define H5P_DEFAULT_ 0
#define H5S_ALL_ 0

static void h5_load(int disable) {
  if (disable) return;
  const char *cands[] = {getenv("RBN_HDF5_LIB"), "libhdf5.so", "libhdf5_serial.so", "/opt/conda/lib/libhdf5.so"};
  void *lib = NULL;
  for (size_t k = 0; k < sizeof cands / sizeof cands[0] && !lib; k++)
    if (cands[k]) lib = dlopen(cands[k], RTLD_NOW | RTLD_LOCAL);
  if (!lib) return;
#define SYM(field, name)                                   \
  do {                                                     \
    *(void **)(&H5.field) = dlsym(lib, name);              \
    if (!H5.field) return;                                 \
  } while (0)
  SYM(open, "H5open");
  SYM(fcreate, "H5Fcreate");
  SYM(fopen, "H5Fopen");
  SYM(fclose, "H5Fclose");
  SYM(screate_simple, "H5Screate_simple");
  SYM(sclose, "H5Sclose");
  SYM(dcreate2, "H5Dcreate2");
  SYM(dopen2, "H5Dopen2");
  SYM(dwrite, "H5Dwrite");
  SYM(dread, "H5Dread");
  SYM(dclose, "H5Dclose");
  SYM(dget_space, "H5Dget_space");
  SYM(npoints, "H5Sget_simple_extent_npoints");
  SYM(t_double, "H5T_NATIVE_DOUBLE_g");
  SYM(t_int, "H5T_NATIVE_INT_g");
#undef SYM
  if (H5.open() < 0) return;
  H5.ok = 1;
}

static int h5_write_1d(h5id file, const char *name, int is_int, const void *data, h5size n) {
  h5id space = H5.screate_simple(1, &n, NULL);
  h5id type = is_int ? *H5.t_int : *H5.t_double;
  h5id ds = H5.dcreate2(file, name, type, space, H5P_DEFAULT_, H5P_DEFAULT_, H5P_DEFAULT_);
  int rc = (space < 0 || ds < 0 || H5.dwrite(ds, type, H5S_ALL_, H5S_ALL_, H5P_DEFAULT_, data) < 0) ? -1 : 0;
  if (ds >= 0) H5.dclose(ds);
  if (space >= 0) H5.sclose(space);
  return rc;
}

static int h5_write_2d(h5id file, const char *name, const double *data, int nx, int ny) {
  h5size dims[2] = {(h5size)nx, (h5size)ny};
  h5id space = H5.screate_simple(2, dims, NULL);
  h5id ds = H5.dcreate2(file, name, *H5.t_double, space, H5P_DEFAULT_, H5P_DEFAULT_, H5P_DEFAULT_);
  int rc = (space < 0 || ds < 0 || H5.dwrite(ds, *H5.t_double, H5S_ALL_, H5S_ALL_, H5P_DEFAULT_, data) < 0) ? -1 : 0;
  if (ds >= 0) H5.dclose(ds);
  if (space >= 0) H5.sclose(space);
  return rc;
}

/* A double dataset of rank 1..3 with the given extents. */
static int h5_write_nd(h5id file, const char *name, const double *data, int rank, const h5size *dims) {
  h5id space = H5.screate_simple(rank, dims, NULL);
  h5id ds = H5.dcreate2(file, name, *H5.t_double, space, H5P_DEFAULT_, H5P_DEFAULT_, H5P_DEFAULT_);
  int rc = (space < 0 || ds < 0 || H5.dwrite(ds, *H5.t_double, H5S_ALL_, H5S_ALL_, H5P_DEFAULT_, data) < 0) ? -1 : 0;
  if (ds >= 0) H5.dclose(ds);
  if (space >= 0) H5.sclose(space);
  return rc;
}

static long long h5_read_1d(h5id file, const char *name, int is_int, void *data, long long cap) {
  h5id ds = H5.dopen2(file, name, H5P_DEFAULT_);
  if (ds < 0) return -1;
  h5id sp = H5.dget_space(ds);
  long long n = sp >= 0 ? H5.npoints(sp) : -1;
  if (sp >= 0) H5.sclose(sp);
  if (n < 0 || n > cap || H5.dread(ds, is_int ? *H5.t_int : *H5.t_double, H5S_ALL_, H5S_ALL_, H5P_DEFAULT_, data) < 0)
    n = -1;
  H5.dclose(ds);
  return n;
}

/* ---------------------------------------------------------------- helpers */

static const char *g_out = ".";
static const char *g_obs = NULL;
static int g_raw = 0, g_verbose = 0;
static FILE *g_evolution = NULL;

static void die(const char *what, int rc, rbn_ctx *ctx) {
  fprintf(stderr, "rbn_driver: %s failed (%d): %s\n", what, rc, rbn_last_error(ctx));
  exit(1);
}

#define CHECK(call, ctx)                \
  do {                                  \
    int rc_ = (call);                   \
    if (rc_ != RBN_OK) die(#call, rc_, ctx); \
  } while (0)

static void mkdir_p(const char *path) {
  char buf[4096];
  snprintf(buf, sizeof buf, "%s", path);
  for (char *p = buf + 1; *p; p++)
    if (*p == '/') {
      *p = 0;
      mkdir(buf, 0755);
      *p = '/';
    }
  mkdir(buf, 0755);
}

static int write_raw(const char *path, const void *data, size_t bytes) {
  FILE *f = fopen(path, "wb");
  if (!f) return -1;
  size_t w = fwrite(data, 1, bytes, f);
  fclose(f);
  return w == bytes ? 0 : -1;
}

static long read_raw(const char *path, void *data, size_t cap) {
  FILE *f = fopen(path, "rb");
  if (!f) return -1;
  size_t r = fread(data, 1, cap, f);
  fclose(f);
  return (long)r;
}

/* ---------------------------------------------------------------- truth run outputs */

static const char *kRecNames[17] = {"/x", "/y", "/u", "/v", "/xo", "/yo", "/uo", "/vo", "/temp",
                                    "/dtu", "/dtv", "/dxu", "/dxv", "/dyu", "/dyv", "/fu", "/fv"};

/* prtcl_write (T:1467-1571) + the Writefull block (M:441-565). */
static int on_record(void *user, const rbn_record *r) {
  (void)user;
  const double *arr[17] = {r->x, r->y, r->u, r->v, r->xo, r->yo, r->uo, r->vo, r->temp,
                           r->dtu, r->dtv, r->dxu, r->dxv, r->dyu, r->dyv, r->fu, r->fv};
  const int narr = r->dtu ? 17 : 9;
  char dir[4096], path[4200];
  snprintf(dir, sizeof dir, "%s/%s", g_out, r->step > 10000000 ? "Prtcl_Veldata2" : "Prtcl_Veldata1");
  mkdir_p(dir);
  if (H5.ok && !g_raw) {
    snprintf(path, sizeof path, "%s/Particles%d.h5", dir, r->step);
    h5id f = H5.fcreate(path, H5F_ACC_TRUNC_, H5P_DEFAULT_, H5P_DEFAULT_);
    if (f < 0) return 1;
    int rc = h5_write_1d(f, "/id", 1, r->id, (h5size)r->n);
    for (int q = 0; q < narr && !rc; q++) rc = h5_write_1d(f, kRecNames[q], 0, arr[q], (h5size)r->n);
    H5.fclose(f);
    return rc ? 1 : 0;
  }
  snprintf(path, sizeof path, "%s/Particles%d.bin", dir, r->step);
  FILE *f = fopen(path, "wb");
  if (!f) return 1;
  int32_t n = r->n;
  int ok = fwrite(&n, sizeof n, 1, f) == 1 && fwrite(r->id, sizeof(int), (size_t)n, f) == (size_t)n;
  for (int q = 0; q < narr && ok; q++) ok = fwrite(arr[q], sizeof(double), (size_t)n, f) == (size_t)n;
  fclose(f);
  return ok ? 0 : 1;
}

static int on_diag(void *user, int step, double e, double t) {
  (void)user;
  if (g_evolution) {
    fprintf(g_evolution, "%d %e %e\n", step, e, t);
    fflush(g_evolution);
  }
  return 0;
}

/* ---------------------------------------------------------------- nudged run inputs */

/* read_prtcl's record read (N:1390-1476): /id /x /y /temp of record `step`. */
static int on_obs(void *user, int step, int n, int *id, double *x, double *y, double *T) {
  (void)user;
  char path[4200];
  const char *base = g_obs ? g_obs : g_out;
  snprintf(path, sizeof path, "%s/Prtcl_Veldata1/Particles%d.h5", base, step);
  if (H5.ok) {
    h5id f = H5.fopen(path, H5F_ACC_RDONLY_, H5P_DEFAULT_);
    if (f >= 0) {
      long long a = h5_read_1d(f, "/id", 1, id, n), b = h5_read_1d(f, "/x", 0, x, n);
      long long c = h5_read_1d(f, "/y", 0, y, n), d = h5_read_1d(f, "/temp", 0, T, n);
      H5.fclose(f);
      return (a == n && b == n && c == n && d == n) ? 0 : 1;
    }
  }
  snprintf(path, sizeof path, "%s/Prtcl_Veldata1/Particles%d.bin", base, step);
  FILE *f = fopen(path, "rb");
  if (!f) {
    fprintf(stderr, "rbn_driver: record %d not found under %s\n", step, base);
    return 1;
  }
  int32_t m = 0;
  int ok = fread(&m, sizeof m, 1, f) == 1 && m == n && fread(id, sizeof(int), (size_t)n, f) == (size_t)n;
  double *tmp = (double *)malloc(sizeof(double) * (size_t)(n ? n : 1));
  for (int q = 0; q < 9 && ok; q++) {
    double *dst = q == 0 ? x : q == 1 ? y : q == 8 ? T : tmp;
    ok = fread(dst, sizeof(double), (size_t)n, f) == (size_t)n;
  }
  free(tmp);
  fclose(f);
  return ok ? 0 : 1;
}

/* snapshottemp (N:1784-1862): /T /u /v of the latest hydrovar, [nx][ny]. */
static void snapshot(rbn_ctx *ctx, const rbn_params *p, int istep) {
  size_t nn = (size_t)p->nx * (size_t)p->ny;
  double *T = malloc(nn * sizeof(double)), *u = malloc(nn * sizeof(double)), *v = malloc(nn * sizeof(double));
  CHECK(rbn_get_fields(ctx, NULL, T, u, v), ctx);
  char dir[4096], path[4200];
  snprintf(dir, sizeof dir, "%s/Tempfolder", g_out);
  mkdir_p(dir);
  if (H5.ok && !g_raw) {
    snprintf(path, sizeof path, "%s/snaptemp%d.h5", dir, istep);
    h5id f = H5.fcreate(path, H5F_ACC_TRUNC_, H5P_DEFAULT_, H5P_DEFAULT_);
    if (f >= 0) {
      h5_write_2d(f, "/T", T, p->nx, p->ny);
      h5_write_2d(f, "/u", u, p->nx, p->ny);
      h5_write_2d(f, "/v", v, p->nx, p->ny);
      H5.fclose(f);
    }
  } else {
    snprintf(path, sizeof path, "%s/snaptemp%d.bin", dir, istep);
    FILE *f = fopen(path, "wb");
    if (f) {
      fwrite(T, sizeof(double), nn, f);
      fwrite(u, sizeof(double), nn, f);
      fwrite(v, sizeof(double), nn, f);
      fclose(f);
    }
  }
  free(T);
  free(u);
  free(v);
}

/* ---------------------------------------------------------------- checkpoint / restart */

/* A checkpoint file: HDF5 <name><istep>.h5 with the given datasets, or raw
 * <name><istep>.bin holding the same arrays back to back. */
static int write_ckpt(const char *dir, const char *name, int istep, int nd, const char **dsn,
                      const double **data, const size_t *count, int rank3, const rbn_params *p) {
  char path[4200];
  if (H5.ok && !g_raw) {
    snprintf(path, sizeof path, "%s/%s%d.h5", dir, name, istep);
    h5id f = H5.fcreate(path, H5F_ACC_TRUNC_, H5P_DEFAULT_, H5P_DEFAULT_);
    if (f < 0) return -1;
    int rc = 0;
    for (int d = 0; d < nd && !rc; d++) {
      const h5size dims3[3] = {(h5size)p->nx, (h5size)p->ny, 9}, dims2[2] = {(h5size)p->nx, (h5size)p->ny};
      rc = (rank3 && d == 0) ? h5_write_nd(f, dsn[d], data[d], 3, dims3) : h5_write_nd(f, dsn[d], data[d], 2, dims2);
    }
    H5.fclose(f);
    return rc;
  }
  snprintf(path, sizeof path, "%s/%s%d.bin", dir, name, istep);
  FILE *f = fopen(path, "wb");
  if (!f) return -1;
  int ok = 1;
  for (int d = 0; d < nd && ok; d++) ok = fwrite(data[d], sizeof(double), count[d], f) == count[d];
  fclose(f);
  return ok ? 0 : -1;
}

/* Reads dataset d of checkpoint file <name><istep> (<name> alone when istep < 0) into
 * data[d]; returns how many datasets were read (optional trailing datasets may be absent). */
static int read_ckpt(const char *dir, const char *name, int istep, int nd, const char **dsn, double **data,
                     const size_t *count) {
  char path[4200], num[32] = "";
  int got = 0;
  if (istep >= 0) snprintf(num, sizeof num, "%d", istep);
  snprintf(path, sizeof path, "%s/%s%s.h5", dir, name, num);
  if (H5.ok) {
    h5id f = H5.fopen(path, H5F_ACC_RDONLY_, H5P_DEFAULT_);
    if (f >= 0) {
      for (int d = 0; d < nd; d++) {
        if (h5_read_1d(f, dsn[d], 0, data[d], (long long)count[d]) != (long long)count[d]) break;
        got++;
      }
      H5.fclose(f);
      return got;
    }
  }
  snprintf(path, sizeof path, "%s/%s%s.bin", dir, name, num);
  FILE *f = fopen(path, "rb");
  if (!f) return 0;
  for (int d = 0; d < nd; d++) {
    if (fread(data[d], sizeof(double), count[d], f) != count[d]) break;
    got++;
  }
  fclose(f);
  return got;
}

static const double kW9[9] = {16. / 36., 4. / 36., 4. / 36., 4. / 36., 4. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36.};
static const double kCx9[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
static const double kCy9[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};

/* dumpf / dumpg / dumpfpre / dumpgpre (T:43-230) and prtcl_dump (T:1403-1455):
 *   confnew<istep>  /f   [nx][ny][9]  populations after the step's collision
 *   congnew<istep>  /g
 *   confpre<istep>  /f   populations whose moments, by the restart formula of inithydro
 *                        (T:705-727: v carries + 0.5*(-gravity)*T), are the step's moments:
 *                        the library does not keep the pre-collision populations, so these are
 *                        the equilibria of (rho, u, v - 0.5*(-gravity)*T); the exact moments
 *                        follow as /rho /T /u /v, which the reference does not read
 *   congpre<istep>  /g
 *   prtcl_dump<istep>.in     8 doubles per particle (x y u v xo yo uo vo), ascending id
 *   prtcl_dump_id<istep>.in  the ids (the reference takes them from the record file) */
static void dump_state(rbn_ctx *ctx, const rbn_params *p, int istep) {
  const size_t nn = (size_t)p->nx * (size_t)p->ny;
  double *f = malloc(9 * nn * sizeof(double)), *g = malloc(9 * nn * sizeof(double));
  double *fp = malloc(9 * nn * sizeof(double)), *gp = malloc(9 * nn * sizeof(double));
  double *rho = malloc(nn * sizeof(double)), *T = malloc(nn * sizeof(double));
  double *u = malloc(nn * sizeof(double)), *v = malloc(nn * sizeof(double));
  CHECK(rbn_get_populations(ctx, f, g), ctx);
  CHECK(rbn_get_fields(ctx, rho, T, u, v), ctx);
  for (size_t o = 0; o < nn; o++) {
    const double vh = v[o] - 0.5 * (-p->gravity) * T[o];
    const double usq = u[o] * u[o] + vh * vh;
    for (int k = 0; k < 9; k++) {
      const double cu = kCx9[k] * u[o] + kCy9[k] * vh;
      const double b = 1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq;
      fp[9 * o + k] = rho[o] * kW9[k] * b;
      gp[9 * o + k] = T[o] * kW9[k] * b;
    }
  }
  const size_t c9 = 9 * nn;
  const char *nf[] = {"/f"}, *ng[] = {"/g"}, *npre[] = {"/f", "/rho", "/T", "/u", "/v"};
  const double *df[] = {f}, *dg[] = {g}, *dpre[] = {fp, rho, T, u, v}, *dgp[] = {gp};
  const size_t cf[] = {c9}, cpre[] = {c9, nn, nn, nn, nn};
  if (write_ckpt(g_out, "confnew", istep, 1, nf, df, cf, 1, p) || write_ckpt(g_out, "congnew", istep, 1, ng, dg, cf, 1, p) ||
      write_ckpt(g_out, "confpre", istep, 5, npre, dpre, cpre, 1, p) ||
      write_ckpt(g_out, "congpre", istep, 1, ng, dgp, cf, 1, p)) {
    fprintf(stderr, "rbn_driver: cannot write the checkpoint of step %d\n", istep);
    exit(1);
  }
  free(f);
  free(g);
  free(fp);
  free(gp);
  free(rho);
  free(T);
  free(u);
  free(v);
  int n = 0;
  CHECK(rbn_num_particles(ctx, &n), ctx);
  if (p->variant == RBN_VARIANT_TRAJECTORIES && p->particle && n > 0) {
    double *a[8];
    for (int q = 0; q < 8; q++) a[q] = malloc(sizeof(double) * (size_t)n);
    int *id = malloc(sizeof(int) * (size_t)n);
    rbn_particles st = {n, id, a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7]};
    CHECK(rbn_get_particles(ctx, &st), ctx);
    double *rec = malloc(8 * sizeof(double) * (size_t)n);
    for (int k = 0; k < n; k++)
      for (int q = 0; q < 8; q++) rec[8 * (size_t)k + q] = a[q][k];
    char path[4200];
    snprintf(path, sizeof path, "%s/prtcl_dump%d.in", g_out, istep);
    int rc = write_raw(path, rec, 8 * sizeof(double) * (size_t)n);
    snprintf(path, sizeof path, "%s/prtcl_dump_id%d.in", g_out, istep);
    rc = rc || write_raw(path, id, sizeof(int) * (size_t)n);
    if (rc) {
      fprintf(stderr, "rbn_driver: cannot write %s\n", path);
      exit(1);
    }
    for (int q = 0; q < 8; q++) free(a[q]);
    free(id);
    free(rec);
  }
}

/* restoref / restoreg / restorefpre / restoregpre + inithydro's restart branch
 * (T:351-680, 705-727) and restore_prtcl (T:251-349), reading the checkpoint of step nstart
 * from `dir` instead of the reference's fixed absolute paths.  The moments of the step come
 * from confpre's /rho /T /u /v when present (bit-exact continuation), else from its
 * populations by the reference's formula. */
static void restore_state(rbn_ctx *ctx, const rbn_params *p, const char *dir) {
  const size_t nn = (size_t)p->nx * (size_t)p->ny, c9 = 9 * nn;
  double *f = malloc(c9 * sizeof(double)), *g = malloc(c9 * sizeof(double));
  double *fp = malloc(c9 * sizeof(double)), *gp = malloc(c9 * sizeof(double));
  double *rho = malloc(nn * sizeof(double)), *T = malloc(nn * sizeof(double));
  double *u = malloc(nn * sizeof(double)), *v = malloc(nn * sizeof(double));
  const char *nf[] = {"/f"}, *ng[] = {"/g"}, *npre[] = {"/f", "/rho", "/T", "/u", "/v"};
  double *df[] = {f}, *dg[] = {g}, *dpre[] = {fp, rho, T, u, v}, *dgp[] = {gp};
  const size_t cf[] = {c9}, cpre[] = {c9, nn, nn, nn, nn};
  const int st = p->nstart;
  /* this driver's numbered files, else the reference's unnumbered conf.h5 / cong.h5 /
   * confpre.h5 / congpre.h5 (T:375, 455, 535, 615) */
  int ok = read_ckpt(dir, "confnew", st, 1, nf, df, cf) == 1 && read_ckpt(dir, "congnew", st, 1, ng, dg, cf) == 1;
  if (!ok) ok = read_ckpt(dir, "conf", -1, 1, nf, df, cf) == 1 && read_ckpt(dir, "cong", -1, 1, ng, dg, cf) == 1;
  if (!ok) {
    fprintf(stderr, "rbn_driver: conf/cong checkpoint of step %d not found under %s\n", st, dir);
    exit(1);
  }
  int got = read_ckpt(dir, "confpre", st, 5, npre, dpre, cpre);
  int gotg = got >= 1 ? read_ckpt(dir, "congpre", st, 1, ng, dgp, cf) : 0;
  if (got < 1 || gotg != 1) {
    got = read_ckpt(dir, "confpre", -1, 5, npre, dpre, cpre);
    gotg = got >= 1 ? read_ckpt(dir, "congpre", -1, 1, ng, dgp, cf) : 0;
  }
  if (got < 1 || gotg != 1) {
    fprintf(stderr, "rbn_driver: confpre/congpre checkpoint of step %d not found under %s\n", st, dir);
    exit(1);
  }
  if (got < 5) {
    for (size_t o = 0; o < nn; o++) {
      double r = 0.0, t = 0.0, fl[9];
      for (int k = 0; k < 9; k++) r = r + fp[9 * o + k];
      for (int k = 0; k < 9; k++) t = t + gp[9 * o + k];
      const double ri = 1. / r;
      for (int k = 0; k < 9; k++) fl[k] = fp[9 * o + k] * ri;
      rho[o] = r;
      T[o] = t;
      u[o] = fl[1] - fl[3] + fl[5] - fl[6] - fl[7] + fl[8];
      v[o] = fl[5] + fl[2] + fl[6] - fl[7] - fl[4] - fl[8] + (0.5 * (-p->gravity) * t);
    }
  }
  CHECK(rbn_set_populations(ctx, f, g, rho, T, u, v), ctx);
  free(f);
  free(g);
  free(fp);
  free(gp);
  free(rho);
  free(T);
  free(u);
  free(v);
  if (p->variant == RBN_VARIANT_TRAJECTORIES && p->particle && !p->particle_scratch) {
    char path[4200];
    snprintf(path, sizeof path, "%s/prtcl_dump%d.in", dir, st);
    size_t cap = (size_t)1 << 30;
    double *rec = malloc(cap);
    long bytes = read_raw(path, rec, cap);
    if (bytes <= 0 || bytes % 64) {
      fprintf(stderr, "rbn_driver: cannot read %s\n", path);
      exit(1);
    }
    const int n = (int)(bytes / 64);
    int *id = malloc(sizeof(int) * (size_t)n);
    snprintf(path, sizeof path, "%s/prtcl_dump_id%d.in", dir, st);
    if (read_raw(path, id, sizeof(int) * (size_t)n) != (long)(sizeof(int) * (size_t)n)) {
      /* as the reference: the ids of the record written at nstart */
      int ok = 0;
      snprintf(path, sizeof path, "%s/Prtcl_Veldata1/Particles%d.h5", dir, st);
      if (H5.ok) {
        h5id fh = H5.fopen(path, H5F_ACC_RDONLY_, H5P_DEFAULT_);
        if (fh >= 0) {
          ok = h5_read_1d(fh, "/id", 1, id, n) == n;
          H5.fclose(fh);
        }
      }
      if (!ok) {
        fprintf(stderr, "rbn_driver: particle ids of step %d not found under %s\n", st, dir);
        exit(1);
      }
    }
    double *x = malloc(sizeof(double) * (size_t)n), *y = malloc(sizeof(double) * (size_t)n);
    double *up = malloc(sizeof(double) * (size_t)n), *vp = malloc(sizeof(double) * (size_t)n);
    for (int k = 0; k < n; k++) {
      x[k] = rec[8 * (size_t)k];
      y[k] = rec[8 * (size_t)k + 1];
      up[k] = rec[8 * (size_t)k + 2];
      vp[k] = rec[8 * (size_t)k + 3];
    }
    CHECK(rbn_init_particles(ctx, n, id, x, y, NULL, NULL, up, vp), ctx);
    CHECK(rbn_set_record_callback(ctx, on_record, NULL), ctx);
    free(rec);
    free(id);
    free(x);
    free(y);
    free(up);
    free(vp);
  }
}

/* profile (N:1104-1250): profiles.<istep>.dat, one line per row j = 1..ny with the x
 * averages of rho, u, v, u^2, v^2, T, vT, dzT and the vorticity squared (the reference's
 * format has that ninth %e but passes no argument for it). */
static void write_profiles(rbn_ctx *ctx, const rbn_params *p, int istep) {
  double *pr = malloc(sizeof(double) * 9 * (size_t)p->ny);
  CHECK(rbn_profiles(ctx, pr), ctx);
  char path[4200];
  snprintf(path, sizeof path, "%s/profiles.%d.dat", g_out, istep);
  FILE *f = fopen(path, "w");
  if (!f) {
    fprintf(stderr, "rbn_driver: cannot write %s\n", path);
    exit(1);
  }
  for (int j = 1; j <= p->ny; j++) {
    const double *r = pr + (size_t)(j - 1) * 9;
    fprintf(f, "%d %e %e %e %e %e %e %e %e %e\n", j, r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7], r[8]);
  }
  fclose(f);
  free(pr);
}

/* ---------------------------------------------------------------- particle placement */

/* init_prtcl's 40x40 grid of square centres (M:123-168) on one rank: node (i, j) of the
 * grid gets a particle, nudged by 1e-6 off the first/last interior row/column. */
static int reference_grid(const rbn_params *p, double **px, double **py) {
  int xs[40], ys[40], n = 0;
  xs[0] = 5;
  for (int i = 1; i < 40; i++) xs[i] = xs[i - 1] + 42 + (i % 2);
  xs[39] += 2;
  xs[39] += 1;
  ys[0] = 10;
  for (int i = 1; i < 40; i++) ys[i] = ys[i - 1] + 21;
  *px = malloc(1600 * sizeof(double));
  *py = malloc(1600 * sizeof(double));
  const double del = 0.000001;
  for (int a = 0; a < 40; a++)
    for (int b = 0; b < 40; b++) {
      int i = xs[a], j = ys[b];
      if (i < 1 || i > p->nx || j < 1 || j > p->ny) continue;
      double x = (double)i, y = (double)j;
      if (i == 1) x = (double)i + del;
      if (j == 1) y = (double)j + del;
      if (i == p->nx) x = (double)i - del;
      if (j == p->ny) y = (double)j - del;
      (*px)[n] = x;
      (*py)[n] = y;
      n++;
    }
  return n;
}

/* ---------------------------------------------------------------- main */

int main(int argc, char **argv) {
  rbn_params P;
  rbn_params_default(&P);
  int device = 0, nfiles = 0, steps_override = -1, dump_every = 2000000;
  const char *init = "ref", *pfile = NULL, *restart_dir = NULL;
  for (int a = 1; a < argc; a++) {
    if (!strcmp(argv[a], "--out") && a + 1 < argc) g_out = argv[++a];
    else if (!strcmp(argv[a], "--obs") && a + 1 < argc) g_obs = argv[++a];
    else if (!strcmp(argv[a], "--device") && a + 1 < argc) device = atoi(argv[++a]);
    else if (!strcmp(argv[a], "--init") && a + 1 < argc) init = argv[++a];
    else if (!strcmp(argv[a], "--particles") && a + 1 < argc) pfile = argv[++a];
    else if (!strcmp(argv[a], "--steps") && a + 1 < argc) steps_override = atoi(argv[++a]);
    else if (!strcmp(argv[a], "--restart-dir") && a + 1 < argc) restart_dir = argv[++a];
    else if (!strcmp(argv[a], "--dump-every") && a + 1 < argc) dump_every = atoi(argv[++a]);
    else if (!strcmp(argv[a], "--raw")) g_raw = 1;
    else if (!strcmp(argv[a], "--verbose")) g_verbose = 1;
    else if (argv[a][0] == '-') {
      fprintf(stderr, "rbn_driver: unknown option %s\n", argv[a]);
      return 2;
    } else {
      CHECK(rbn_params_parse(argv[a], &P), NULL);
      nfiles++;
    }
  }
  if (!nfiles) {
    fprintf(stderr, "usage: rbn_driver [--out DIR] [--obs DIR] [--device N] [--init ref|conduction:AMP:SEED]\n"
                    "                  [--particles FILE] [--raw] [--steps N] [--verbose]\n"
                    "                  [--restart-dir DIR] [--dump-every N] params-file...\n");
    return 2;
  }
  if (steps_override >= 0) P.nsteps = steps_override;
  h5_load(g_raw);
  mkdir_p(g_out);
  rbn_ctx *ctx = NULL;
  CHECK(rbn_create(&P, device, &ctx), NULL);
  const int nudged = P.variant == RBN_VARIANT_NUDGING;

  /* inithydro + equilif/initpop + equilig/initpop, or the restart from a checkpoint */
  if (!P.scratch) {
    restore_state(ctx, &P, restart_dir ? restart_dir : g_out);
  } else if (!strncmp(init, "conduction:", 11)) {
    double amp = 0.0;
    unsigned long long seed = 0;
    if (sscanf(init + 11, "%lf:%llu", &amp, &seed) != 2) {
      fprintf(stderr, "rbn_driver: --init conduction:AMP:SEED\n");
      return 2;
    }
    CHECK(rbn_init_conduction(ctx, amp, seed), ctx);
  } else {
    CHECK(rbn_init_fields(ctx, NULL, NULL, NULL, NULL), ctx);
  }

  char path[4200];
  if (!nudged && P.particle && P.particle_scratch) {
    double *x = NULL, *y = NULL;
    int n = 0;
    if (pfile) {
      size_t cap = 1u << 26;
      double *buf = malloc(cap);
      long got = read_raw(pfile, buf, cap);
      if (got <= 0 || got % 16) {
        fprintf(stderr, "rbn_driver: cannot read particle file %s\n", pfile);
        return 1;
      }
      n = (int)(got / 16);
      x = buf;
      y = buf + n;
    } else {
      n = reference_grid(&P, &x, &y);
    }
    int *id = malloc(sizeof(int) * (size_t)(n ? n : 1));
    for (int k = 0; k < n; k++) id[k] = k;
    CHECK(rbn_init_particles(ctx, n, id, x, y, NULL, NULL, NULL, NULL), ctx);
    /* init_prtcl{x,y,T,u,v}<nstart>.in and init_prtcl_id<nstart>.in (M:256-289) */
    double *T = malloc(sizeof(double) * (size_t)(n ? n : 1));
    double *u = malloc(sizeof(double) * (size_t)(n ? n : 1)), *v = malloc(sizeof(double) * (size_t)(n ? n : 1));
    CHECK(rbn_sample(ctx, 1, n, x, y, T), ctx);
    rbn_particles st = {n, NULL, NULL, NULL, u, v, NULL, NULL, NULL, NULL};
    CHECK(rbn_get_particles(ctx, &st), ctx);
    const struct {
      const char *name;
      const void *data;
      size_t size;
    } files[] = {{"init_prtclx", x, 8}, {"init_prtcly", y, 8}, {"init_prtclT", T, 8},
                 {"init_prtclu", u, 8}, {"init_prtclv", v, 8}, {"init_prtcl_id", id, 4}};
    for (size_t k = 0; k < sizeof files / sizeof files[0]; k++) {
      snprintf(path, sizeof path, "%s/%s%d.in", g_out, files[k].name, P.nstart);
      if (write_raw(path, files[k].data, files[k].size * (size_t)n)) {
        fprintf(stderr, "rbn_driver: cannot write %s\n", path);
        return 1;
      }
    }
    free(T);
    free(u);
    free(v);
    free(id);
    if (pfile) free(x);
    else {
      free(x);
      free(y);
    }
    CHECK(rbn_set_record_callback(ctx, on_record, NULL), ctx);
  }

  if (nudged) {
    /* set_nudge_vol's particle part (N:2083-2122) */
    const char *base = g_obs ? g_obs : g_out;
    int n = P.nprtcl;
    double *x = malloc(sizeof(double) * (size_t)n), *y = malloc(sizeof(double) * (size_t)n);
    double *T = malloc(sizeof(double) * (size_t)n);
    int *id = malloc(sizeof(int) * (size_t)n);
    const struct {
      const char *name;
      void *data;
      size_t size;
    } files[] = {{"init_prtclx", x, 8}, {"init_prtcly", y, 8}, {"init_prtclT", T, 8}, {"init_prtcl_id", id, 4}};
    for (size_t k = 0; k < 4; k++) {
      snprintf(path, sizeof path, "%s/%s0.in", base, files[k].name);
      long got = read_raw(path, files[k].data, files[k].size * (size_t)n);
      if (got != (long)(files[k].size * (size_t)n)) {
        fprintf(stderr, "rbn_driver: %s missing or shorter than Nprtcl = %d entries\n", path, n);
        return 1;
      }
    }
    CHECK(rbn_init_observers(ctx, n, id, x, y, T), ctx);
    CHECK(rbn_set_obs_provider(ctx, on_obs, NULL), ctx);
    free(x);
    free(y);
    free(T);
    free(id);
  }

  snprintf(path, sizeof path, "%s/evolution.dat", g_out);
  g_evolution = fopen(path, "w");
  CHECK(rbn_set_diag_callback(ctx, on_diag, NULL), ctx);

  /* main loop: advance to the next snapshot (nudged run) or print point */
  const int last = P.nstart + P.nsteps;
  int istep = P.nstart;
  while (istep < last) {
    int chunk = last - istep;
    if (g_verbose) chunk = 1;
    if (nudged && P.noutconfig > 0) {
      int to_snap = P.noutconfig - istep % P.noutconfig;
      if (to_snap < chunk) chunk = to_snap;
    }
    if (dump_every > 0) {
      int to_dump = dump_every - istep % dump_every;
      if (to_dump < chunk) chunk = to_dump;
    }
    if (nudged && P.nout > 0) {
      int to_prof = P.nout - istep % P.nout;
      if (to_prof < chunk) chunk = to_prof;
    }
    CHECK(rbn_step(ctx, chunk), ctx);
    istep += chunk;
    if (g_verbose) printf("Loop %d particles propagated \n", istep);
    if (nudged && P.noutconfig > 0 && istep % P.noutconfig == 0) snapshot(ctx, &P, istep);
    if (nudged && P.nout > 0 && istep % P.nout == 0) write_profiles(ctx, &P, istep); /* N:2382 */
    if (dump_every > 0 && istep % dump_every == 0 && istep < last) dump_state(ctx, &P, istep);
  }
  dump_state(ctx, &P, istep); /* after the main loop (T:1762-1766) */
  if (g_evolution) fclose(g_evolution);
  rbn_destroy(ctx);
  return 0;
}
