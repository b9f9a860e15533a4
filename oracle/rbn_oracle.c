/*
 * rbn_oracle.c — CPU FP64 restatement of the reference RB step (test infrastructure only;
 * see rbn_oracle.h for what may use it and what pins it).
 *
 * Citations are to the reference repository (Lokahith-Agasthya/LagrangianRBNudging):
 *   T: = Trajectories/code/temppart_dump.c   M: = Trajectories/code/mympi2.h
 *   N: = Nudging/code/tempvel_nudge.c
 * All arithmetic keeps the reference's operand order (C left-to-right association,
 * divisions by the cs2 macros rather than multiplications); build with
 * -std=c99 -ffp-contract=off.
 */
#include "rbn_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NPOP 9

/* Lattice constants, T:9-12 (macros) and dellar T:1363-1401.  The reverse directions are
 * negations of the forward ones, which makes cy[3], cx[4] negative zeros. */
#define CS2 (1. / 3.)
#define CS22 (2. * CS2)
#define CSSQ (2. / 9.)
static const double W[NPOP] = {16. / 36., 4. / 36., 4. / 36., 4. / 36., 4. / 36.,
                               1. / 36.,  1. / 36., 1. / 36., 1. / 36.};
static const double CXD[NPOP] = {0., 1., 0., -1., -0., 1., -1., -1., 1.};
static const double CYD[NPOP] = {0., 0., 1., -0., -1., 1., 1., -1., -1.};
static const int CXI[NPOP] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
static const int CYI[NPOP] = {0, 0, 1, 0, -1, 1, 1, -1, -1};

typedef struct {
  double p[NPOP];
} or_pop;

typedef struct {
  int step, n;
  int *id;
  double *a; /* 18 arrays of n: id, 9 record fields, 8 Writefull fields */
} or_rec;

typedef struct {
  int step, n;
  int *id;
  double *x, *y, *T;
} or_obsrec;

struct or_sim {
  rbn_params P;
  int NX, NY, H; /* H = NY+2, the column stride */
  size_t nn;     /* (NX+2)*(NY+2) */
  int faithful;
  char err[256];
  /* lattice */
  or_pop *f, *g, *feq, *geq, *memup1, *memdown1;
  double *u, *v, *rho, *T, *Fx, *Fy, *ubef, *vbef;
  double *dtu, *dtv, *dxu, *dxv, *dyu, *dyv;
  double *upost, *vpost, *rhopost, *Tpost, *utot, *vtot, *tau, *taug;
  int istep;
  int initialised;
  /* particles, id order */
  int np;
  int *pid;
  unsigned char *heavy;
  double *px, *py, *pu, *pv, *pxo, *pyo, *puo, *pvo, *pinvtau, *pbeta;
  or_rec *recs;
  int nrec, caprec;
  /* nudging */
  int *nudge_vol;
  double *tempreadint; /* [NX*NY], idx (j-1)+NY*(i-1) as N:1563 */
  double *nudge_w;     /* [NX*NY]: bilinear deposit weight per node (RBN_DEPOSIT_BILINEAR) */
  double *nud_term;
  int nobs;
  int *oid, *chk, *bcy, *xidx, *yidx;
  double *dx, *dy, *dT, *xint, *yint, *Tint, *prx, *pry, *prT;
  or_obsrec *orecs;
  int norec, caporec;
  /* diagnostics */
  int ndiag_out, capdiag;
  int *dstep;
  double *de, *dt;
};

static int fail(or_sim *s, int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(s->err, sizeof s->err, fmt, ap);
  va_end(ap);
  return code;
}

const char *or_last_error(const or_sim *s) { return s ? s->err : "null simulation"; }
void or_set_faithful(or_sim *s, int on) { s->faithful = on; }
int or_get_step(const or_sim *s) { return s->istep; }

static inline size_t IX(const or_sim *s, int i, int j) { return (size_t)j + (size_t)s->H * (size_t)i; }

/* ------------------------------------------------------------------ lifetime */

or_sim *or_create(const rbn_params *p) {
  if (!p || p->nx < 8 || p->ny < 4) return NULL;
  or_sim *s = (or_sim *)calloc(1, sizeof(or_sim));
  if (!s) return NULL;
  s->P = *p;
  s->NX = p->nx;
  s->NY = p->ny;
  s->H = p->ny + 2;
  s->nn = (size_t)(p->nx + 2) * (size_t)(p->ny + 2);
  size_t nn = s->nn;
  s->f = (or_pop *)calloc(nn, sizeof(or_pop));
  s->g = (or_pop *)calloc(nn, sizeof(or_pop));
  s->feq = (or_pop *)calloc(nn, sizeof(or_pop));
  s->geq = (or_pop *)calloc(nn, sizeof(or_pop));
  s->memup1 = (or_pop *)calloc((size_t)p->nx + 2, sizeof(or_pop));
  s->memdown1 = (or_pop *)calloc((size_t)p->nx + 2, sizeof(or_pop));
  double **fields[] = {&s->u,     &s->v,     &s->rho,     &s->T,     &s->Fx,   &s->Fy,
                       &s->ubef,  &s->vbef,  &s->dtu,     &s->dtv,   &s->dxu,  &s->dxv,
                       &s->dyu,   &s->dyv,   &s->upost,   &s->vpost, &s->rhopost,
                       &s->Tpost, &s->utot,  &s->vtot,    &s->tau,   &s->taug, &s->nud_term};
  for (size_t k = 0; k < sizeof fields / sizeof fields[0]; k++) *fields[k] = (double *)calloc(nn, sizeof(double));
  s->nudge_vol = (int *)calloc(nn, sizeof(int));
  s->tempreadint = (double *)calloc((size_t)p->nx * (size_t)p->ny, sizeof(double));
  s->nudge_w = (double *)calloc((size_t)p->nx * (size_t)p->ny, sizeof(double));
  if (!s->f || !s->g || !s->feq || !s->geq || !s->memup1 || !s->memdown1 || !s->nud_term ||
      !s->nudge_vol || !s->tempreadint || !s->nudge_w) {
    or_destroy(s);
    return NULL;
  }
  s->istep = p->nstart;
  return s;
}

static void free_particles(or_sim *s) {
  void *ptrs[] = {s->pid, s->heavy, s->px, s->py, s->pu, s->pv, s->pxo,
                  s->pyo, s->puo,   s->pvo, s->pinvtau, s->pbeta};
  for (size_t k = 0; k < sizeof ptrs / sizeof ptrs[0]; k++) free(ptrs[k]);
  s->pid = NULL; s->heavy = NULL; s->px = s->py = s->pu = s->pv = NULL;
  s->pxo = s->pyo = s->puo = s->pvo = s->pinvtau = s->pbeta = NULL;
  s->np = 0;
}

static void free_observers(or_sim *s) {
  void *ptrs[] = {s->oid, s->chk, s->bcy, s->xidx, s->yidx, s->dx,  s->dy,
                  s->dT,  s->xint, s->yint, s->Tint, s->prx,  s->pry, s->prT};
  for (size_t k = 0; k < sizeof ptrs / sizeof ptrs[0]; k++) free(ptrs[k]);
  s->oid = s->chk = s->bcy = s->xidx = s->yidx = NULL;
  s->dx = s->dy = s->dT = s->xint = s->yint = s->Tint = s->prx = s->pry = s->prT = NULL;
  s->nobs = 0;
}

void or_destroy(or_sim *s) {
  if (!s) return;
  void *ptrs[] = {s->f,     s->g,     s->feq,     s->geq,   s->memup1, s->memdown1,
                  s->u,     s->v,     s->rho,     s->T,     s->Fx,     s->Fy,
                  s->ubef,  s->vbef,  s->dtu,     s->dtv,   s->dxu,    s->dxv,
                  s->dyu,   s->dyv,   s->upost,   s->vpost, s->rhopost, s->Tpost,
                  s->utot,  s->vtot,  s->tau,     s->taug,  s->nud_term, s->nudge_vol,
                  s->tempreadint, s->nudge_w, s->dstep, s->de, s->dt};
  for (size_t k = 0; k < sizeof ptrs / sizeof ptrs[0]; k++) free(ptrs[k]);
  free_particles(s);
  free_observers(s);
  for (int r = 0; r < s->nrec; r++) { free(s->recs[r].id); free(s->recs[r].a); }
  free(s->recs);
  for (int r = 0; r < s->norec; r++) {
    free(s->orecs[r].id); free(s->orecs[r].x); free(s->orecs[r].y); free(s->orecs[r].T);
  }
  free(s->orecs);
  free(s);
}

/* ------------------------------------------------------------------ halos */

/* pbc (T:234-250) on one rank of a periodic Cartesian grid: whole ghost columns from the
 * opposite interior column, then ghost rows (including ghost columns) from the opposite
 * interior row. */
static void or_pbc(or_sim *s, or_pop *a) {
  const int NX = s->NX, NY = s->NY, H = s->H;
  memcpy(&a[0], &a[(size_t)NX * H], (size_t)H * sizeof(or_pop));
  memcpy(&a[(size_t)(NX + 1) * H], &a[(size_t)H], (size_t)H * sizeof(or_pop));
  for (int i = 0; i < NX + 2; i++) {
    a[IX(s, i, 0)] = a[IX(s, i, NY)];
    a[IX(s, i, NY + 1)] = a[IX(s, i, 1)];
  }
}

/* communication (M:313-348) on one rank: ghost columns for rows 1..NY, then ghost rows
 * for columns 0..NX+1 — periodic in x and in y. */
static void or_communication(or_sim *s, double *a) {
  const int NX = s->NX, NY = s->NY;
  for (int j = 1; j < NY + 1; j++) {
    a[IX(s, 0, j)] = a[IX(s, NX, j)];
    a[IX(s, NX + 1, j)] = a[IX(s, 1, j)];
  }
  for (int i = 0; i < NX + 2; i++) {
    double top = a[IX(s, i, NY)], bot = a[IX(s, i, 1)];
    a[IX(s, i, NY + 1)] = bot;
    a[IX(s, i, 0)] = top;
  }
}

/* ------------------------------------------------------------------ LBM pieces */

/* equilif / equilig (T:1097-1180): second-order D2Q9 equilibrium in the reference's
 * expression form; r1 is rho for f and T for g. */
void or_equilibrium(double r1, double u, double v, double q[9]) {
  double usq = u * u, vsq = v * v;
  double sumsq = (usq + vsq) / CS22;
  double sumsq2 = sumsq * (1. - CS2) / CS2;
  double u2 = usq / CSSQ, v2 = vsq / CSSQ;
  double ui = u / CS2, vi = v / CS2;
  double uv = ui * vi;
  q[0] = r1 * W[0] * (1. - sumsq);
  q[1] = r1 * W[1] * (1. - sumsq + u2 + ui);
  q[2] = r1 * W[2] * (1. - sumsq + v2 + vi);
  q[3] = r1 * W[3] * (1. - sumsq + u2 - ui);
  q[4] = r1 * W[4] * (1. - sumsq + v2 - vi);
  q[5] = r1 * W[5] * (1. + sumsq2 + ui + vi + uv);
  q[6] = r1 * W[6] * (1. + sumsq2 - ui + vi - uv);
  q[7] = r1 * W[7] * (1. + sumsq2 - ui - vi + uv);
  q[8] = r1 * W[8] * (1. + sumsq2 + ui - vi - uv);
}

static void or_equili_all(or_sim *s, const double *r, or_pop *eq) {
  for (int i = 0; i < s->NX + 2; i++)
    for (int j = 0; j < s->NY + 2; j++) {
      size_t id = IX(s, i, j);
      or_equilibrium(r[id], s->u[id], s->v[id], eq[id].p);
    }
}

/* One in-place pull sweep of direction k over rows jlo..jhi, columns 1..NX, visiting
 * nodes so that every source is read before it is overwritten (the loop orders of
 * T:781-854). */
static void or_stream_dir(or_sim *s, or_pop *a, int k, int jlo, int jhi) {
  const int cx = CXI[k], cy = CYI[k];
  const int i0 = cx > 0 ? s->NX : 1, i1 = cx > 0 ? 0 : s->NX + 1, di = cx > 0 ? -1 : 1;
  const int j0 = cy > 0 ? jhi : jlo, j1 = cy > 0 ? jlo - 1 : jhi + 1, dj = cy > 0 ? -1 : 1;
  for (int i = i0; i != i1; i += di)
    for (int j = j0; j != j1; j += dj) a[IX(s, i, j)].p[k] = a[IX(s, i - cx, j - cy)].p[k];
}

/* forcingconstructWW (T:744-755): the force of this step from the previous step's
 * moments, interior only. */
static void or_forcing(or_sim *s) {
  const double average = 0.0;
  for (int i = 1; i < s->NX + 1; i++)
    for (int j = 1; j < s->NY + 1; j++) {
      size_t id = IX(s, i, j);
      s->Fx[id] = 0.0;
      s->Fy[id] = -s->P.gravity * s->rho[id] * (s->T[id] - average);
    }
}

static void or_save_walls(or_sim *s) {
  for (int i = 0; i < s->NX + 2; i++) {
    s->memdown1[i] = s->f[IX(s, i, 1)];
    s->memup1[i] = s->f[IX(s, i, s->NY)];
  }
}

/* Half-way bounce-back of f from the saved pre-stream wall rows (T:859-894). */
static void or_f_bounce(or_sim *s) {
  const int NY = s->NY;
  for (int i = 1; i < s->NX + 1; i++) {
    or_pop *n = &s->f[IX(s, i, NY)];
    n->p[4] = s->memup1[i].p[2];
    n->p[7] = s->memup1[i].p[5];
    n->p[8] = s->memup1[i].p[6];
  }
  for (int i = 1; i < s->NX + 1; i++) {
    or_pop *n = &s->f[IX(s, i, 1)];
    n->p[2] = s->memdown1[i].p[4];
    n->p[5] = s->memdown1[i].p[7];
    n->p[6] = s->memdown1[i].p[8];
  }
}

/* moveplusforcingconstructWW, truth-run variant (T:757-922). */
static void or_move_traj(or_sim *s) {
  const int NY = s->NY;
  or_save_walls(s);
  for (int k = 1; k < NPOP; k++) {
    or_stream_dir(s, s->f, k, 1, NY);
    or_stream_dir(s, s->g, k, 1, NY);
  }
  or_forcing(s);
  or_f_bounce(s);
  /* g: fixed plate temperatures from the post-stream opposite populations (T:896-921). */
  const double Tup = s->P.Tup, Tdown = s->P.Tdown;
  for (int i = 1; i < s->NX + 1; i++) {
    or_pop *n = &s->g[IX(s, i, NY)];
    n->p[4] = -n->p[2] + 2.0 * W[2] * Tup;
    n->p[7] = -n->p[5] + 2.0 * W[5] * Tup;
    n->p[8] = -n->p[6] + 2.0 * W[6] * Tup;
  }
  for (int i = 1; i < s->NX + 1; i++) {
    or_pop *n = &s->g[IX(s, i, 1)];
    n->p[2] = -n->p[4] + 2.0 * W[4] * Tdown;
    n->p[5] = -n->p[7] + 2.0 * W[7] * Tdown;
    n->p[6] = -n->p[8] + 2.0 * W[8] * Tdown;
  }
}

/* moveplusforcingconstructWW, nudged-run variant (N:677-1032) with the semantics of a
 * decomposition with NPY >= 2 (the only one in which each wall block runs once): g rows 1
 * and ny take the pre-stream populations of rows 2 and ny-1 and are not streamed; g rows
 * 2..ny-1 stream from that modified array; f streams everywhere and bounces back. */
static void or_move_nudge(or_sim *s) {
  const int NY = s->NY;
  or_save_walls(s);
  for (int i = 0; i < s->NX + 2; i++) {
    s->g[IX(s, i, 1)] = s->g[IX(s, i, 2)];
    s->g[IX(s, i, NY)] = s->g[IX(s, i, NY - 1)];
  }
  for (int k = 1; k < NPOP; k++) {
    or_stream_dir(s, s->f, k, 1, NY);
    or_stream_dir(s, s->g, k, 2, NY - 1);
  }
  or_forcing(s);
  or_f_bounce(s);
}

/* hydrovar (T:957-987; nudged form N:1069-1102 adds half the persisted nudging term). */
static void or_hydrovar(or_sim *s, double *u, double *v, double *rho, double *T, int nudged) {
  for (int i = 0; i < s->NX + 2; i++)
    for (int j = 0; j < s->NY + 2; j++) {
      size_t id = IX(s, i, j);
      const double *fp = s->f[id].p, *gp = s->g[id].p;
      double r = 0.0, t = 0.0, fl[NPOP];
      for (int k = 0; k < NPOP; k++) r = r + fp[k];
      for (int k = 0; k < NPOP; k++) t = t + gp[k];
      if (nudged && s->nudge_vol[id] > 0) t = (t + 0.5 * s->nud_term[id]);
      rho[id] = r;
      T[id] = t;
      double rhoi = 1. / r;
      for (int k = 0; k < NPOP; k++) fl[k] = fp[k] * rhoi;
      u[id] = fl[1] - fl[3] + fl[5] - fl[6] - fl[7] + fl[8] + 0.5 * rhoi * s->Fx[id];
      v[id] = fl[5] + fl[2] + fl[6] - fl[7] - fl[4] - fl[8] + 0.5 * rhoi * s->Fy[id];
    }
}

/* collisf (T:1314-1328): BGK with Guo forcing. */
static void or_collide_f(or_sim *s) {
  const double tau = s->P.relax;
  for (int i = 1; i < s->NX + 1; i++)
    for (int j = 1; j < s->NY + 1; j++) {
      size_t id = IX(s, i, j);
      double om = 1. / tau;
      double uu = s->u[id], vv = s->v[id], fx = s->Fx[id], fy = s->Fy[id];
      for (int k = 0; k < NPOP; k++) {
        double source = (1.0 - 0.5 * om) * W[k];
        double factor = ((CXD[k] - uu) * fx + (CYD[k] - vv) * fy) / CS2;
        factor = factor + ((CXD[k] * uu + CYD[k] * vv) * (CXD[k] * fx + CYD[k] * fy)) / (CS2 * CS2);
        s->f[id].p[k] = s->f[id].p[k] * (1. - om) + om * s->feq[id].p[k] + source * factor;
      }
    }
}

/* collisg: plain (T:1330-1341) or nudged (N:1917-1945).  In the nudged form a
 * non-nudged node adds B*w*Q with B = 0, i.e. a signed zero, which is omitted here. */
static void or_collide_g(or_sim *s, int nudged) {
  const double taug = s->P.relaxg, alpha = s->P.alpha;
  for (int i = 1; i < s->NX + 1; i++)
    for (int j = 1; j < s->NY + 1; j++) {
      size_t id = IX(s, i, j);
      double om = 1. / taug;
      if (nudged && s->nudge_vol[id] > 0) {
        size_t ir = (size_t)(j - 1) + (size_t)s->NY * (size_t)(i - 1);
        double B = 1.0 - 0.5 * om;
        const double Wn = s->nudge_w[ir] < 1.0 ? s->nudge_w[ir] : 1.0;  /* min(W, 1) */
        double Q = s->P.deposit == RBN_DEPOSIT_BILINEAR ? alpha * (Wn * (s->T[id] - s->tempreadint[ir]))
                                                        : alpha * (s->T[id] - s->tempreadint[ir]);
        for (int k = 0; k < NPOP; k++)
          s->g[id].p[k] = s->g[id].p[k] * (1. - om) + om * s->geq[id].p[k] - B * W[k] * Q;
        s->nud_term[id] = -1. * Q;
      } else {
        for (int k = 0; k < NPOP; k++) s->g[id].p[k] = s->g[id].p[k] * (1. - om) + om * s->geq[id].p[k];
      }
    }
}

/* ------------------------------------------------------------------ particles */

/* bilinear (M:292-311) at a halo-coordinate position of one field with halos filled by
 * or_communication.  Out-of-array corners (the reference would read out of bounds) set
 * *err and return 0. */
static double or_bilinear(const or_sim *s, const double *a, double xx, double yy, int *err) {
  if (!(xx > -1.0 && xx < (double)s->NX + 1.0 && yy > -1.0 && yy < (double)s->NY + 1.0)) {
    if (err) *err = 1; /* also rejects NaN before the int conversion */
    return 0.0;
  }
  int x1 = (int)xx, x2 = x1 + 1, y1 = (int)yy, y2 = y1 + 1;
  double u11 = a[IX(s, x1, y1)], u12 = a[IX(s, x1, y2)];
  double u21 = a[IX(s, x2, y1)], u22 = a[IX(s, x2, y2)];
  double val1 = u11 * ((double)x2 - xx) * ((double)y2 - yy);
  double val2 = u21 * (xx - (double)x1) * ((double)y2 - yy);
  double val3 = u12 * ((double)x2 - xx) * (yy - (double)y1);
  double val4 = u22 * (xx - (double)x1) * (yy - (double)y1);
  return val1 + val2 + val3 + val4;
}

double or_bilinear_global(const or_sim *s, int field, double x, double y, int *err) {
  const double *a = field == 0 ? s->u : field == 1 ? s->v : field == 2 ? s->T : s->rho;
  return or_bilinear(s, a, x, y, err);
}

/* Node fields of the Maxey–Riley right-hand side (M:374-411): time and centred space
 * differences of u and v, one-sided at the walls, then periodic halos. */
static void or_derivatives(or_sim *s) {
  const int NX = s->NX, NY = s->NY;
  const double *u = s->u, *v = s->v;
  for (int i = 1; i < NX + 1; i++)
    for (int j = 1; j < NY + 1; j++) {
      size_t id = IX(s, i, j), ir = IX(s, i + 1, j), il = IX(s, i - 1, j);
      size_t it = IX(s, i, j + 1), ib = IX(s, i, j - 1);
      s->dtu[id] = u[id] - s->ubef[id];
      s->dtv[id] = v[id] - s->vbef[id];
      s->dxu[id] = (u[ir] - u[il]) * 0.5;
      s->dxv[id] = (v[ir] - v[il]) * 0.5;
      s->dyu[id] = (u[it] - u[ib]) * 0.5;
      s->dyv[id] = (v[it] - v[ib]) * 0.5;
    }
  for (int i = 1; i < NX + 1; i++) {
    size_t id = IX(s, i, 1), it = IX(s, i, 2);
    s->dyu[id] = (u[it] - u[id]);
    s->dyv[id] = (v[it] - v[id]);
  }
  for (int i = 1; i < NX + 1; i++) {
    size_t id = IX(s, i, NY), ib = IX(s, i, NY - 1);
    s->dyu[id] = (u[id] - u[ib]);
    s->dyv[id] = (v[id] - v[ib]);
  }
  double *d[6] = {s->dtu, s->dtv, s->dxu, s->dxv, s->dyu, s->dyv};
  for (int k = 0; k < 6; k++) or_communication(s, d[k]);
}

static int any_heavy(const or_sim *s) {
  for (int p = 0; p < s->np; p++)
    if (s->heavy[p]) return 1;
  return 0;
}

static int push_record(or_sim *s, int step) {
  if (s->nrec == s->caprec) {
    int cap = s->caprec ? 2 * s->caprec : 16;
    or_rec *r = (or_rec *)realloc(s->recs, (size_t)cap * sizeof(or_rec));
    if (!r) return -1;
    s->recs = r;
    s->caprec = cap;
  }
  or_rec *r = &s->recs[s->nrec];
  r->step = step;
  r->n = s->np;
  r->id = (int *)malloc((size_t)(s->np ? s->np : 1) * sizeof(int));
  r->a = (double *)calloc((size_t)18 * (size_t)(s->np ? s->np : 1), sizeof(double));
  if (!r->id || !r->a) return -1;
  s->nrec++;
  return 0;
}

/* propagate_prtcl (M:350-440) + prtcl_write (T:1467-1509, Writefull M:441-501) +
 * toggle_prtcl (M:568-760) on one rank, particles kept in id order. */
static int or_particles(or_sim *s) {
  const int NX = s->NX, NY = s->NY, np = s->np;
  const double delta_t = 1.0;
  const int record = (s->istep % s->P.noutconfig) == 0;
  const int full = record && s->P.prtcl_writefull;
  int err = 0;
  or_communication(s, s->u);
  or_communication(s, s->v);
  if (any_heavy(s) || full) or_derivatives(s);
  for (int p = 0; p < np; p++) {
    double xx = s->px[p], yy = s->py[p];
    s->puo[p] = s->pu[p];
    s->pvo[p] = s->pv[p];
    double up = or_bilinear(s, s->u, xx, yy, &err);
    double vp = or_bilinear(s, s->v, xx, yy, &err);
    if (s->heavy[p]) {
      double tdtu = or_bilinear(s, s->dtu, xx, yy, &err);
      double tdtv = or_bilinear(s, s->dtv, xx, yy, &err);
      double tdxu = or_bilinear(s, s->dxu, xx, yy, &err);
      double tdxv = or_bilinear(s, s->dxv, xx, yy, &err);
      double tdyu = or_bilinear(s, s->dyu, xx, yy, &err);
      double tdyv = or_bilinear(s, s->dyv, xx, yy, &err);
      const double beta = s->pbeta[p], onebytau = s->pinvtau[p];
      s->pxo[p] = xx;
      s->pyo[p] = yy;
      s->pu[p] = s->pu[p] + delta_t * (beta * (tdtu + up * tdxu + vp * tdyu) - onebytau * (s->pu[p] - up));
      s->pv[p] = s->pv[p] + delta_t * (beta * (tdtv + up * tdxv + vp * tdyv) - onebytau * (s->pv[p] - vp));
      s->px[p] = xx + delta_t * s->pu[p];
      s->py[p] = yy + delta_t * s->pv[p];
    } else {
      s->pxo[p] = xx;
      s->pyo[p] = yy;
      s->px[p] = xx + 1.5 * delta_t * up - 0.5 * delta_t * s->puo[p];
      s->py[p] = yy + 1.5 * delta_t * vp - 0.5 * delta_t * s->pvo[p];
      s->pu[p] = up;
      s->pv[p] = vp;
    }
  }
  if (record) {
    or_communication(s, s->T);
    if (push_record(s, s->istep)) return fail(s, RBN_ENOMEM, "record allocation failed");
    or_rec *r = &s->recs[s->nrec - 1];
    double *a = r->a;
    for (int p = 0; p < np; p++) {
      r->id[p] = s->pid[p];
      a[0 * (size_t)np + p] = (double)s->pid[p];
      a[1 * (size_t)np + p] = s->px[p];
      a[2 * (size_t)np + p] = s->py[p];
      a[3 * (size_t)np + p] = s->pu[p];
      a[4 * (size_t)np + p] = s->pv[p];
      a[5 * (size_t)np + p] = s->pxo[p];
      a[6 * (size_t)np + p] = s->pyo[p];
      a[7 * (size_t)np + p] = s->puo[p];
      a[8 * (size_t)np + p] = s->pvo[p];
      a[9 * (size_t)np + p] = or_bilinear(s, s->T, s->px[p], s->py[p], &err);
      if (full) {
        double xo = s->pxo[p], yo = s->pyo[p];
        const double *src[8] = {s->dtu, s->dtv, s->dxu, s->dxv, s->dyu, s->dyv, s->u, s->v};
        for (int q = 0; q < 8; q++) a[(size_t)(10 + q) * np + p] = or_bilinear(s, src[q], xo, yo, &err);
      }
    }
  }
  for (int p = 0; p < np; p++) {
    if (s->px[p] >= (double)NX) {
      s->px[p] = s->px[p] - (double)NX;
    } else if (s->px[p] <= 0.0) {
      s->px[p] = s->px[p] + (double)NX;
    }
    if (s->py[p] < 1.0) {
      s->py[p] = s->py[p] + 2.0 * (1.0 - s->py[p]);
      s->pv[p] = -s->pv[p];
    } else if (s->py[p] >= (double)NY) {
      s->py[p] = s->py[p] + 2.0 * ((double)NY - s->py[p]);
      s->pv[p] = -s->pv[p];
    }
  }
  if (err) return fail(s, RBN_ERANGE, "particle left the lattice at step %d", s->istep);
  return 0;
}

/* ------------------------------------------------------------------ nudging */

/* One running-mean update of the deposit (N:1578-1579). */
static inline void fold(or_sim *s, int i, int j, double Tk) {
  if ((i > 0) && (i <= s->NX) && (j > 0) && (j <= s->NY)) {
    size_t id = IX(s, i, j), ir = (size_t)(j - 1) + (size_t)s->NY * (size_t)(i - 1);
    s->tempreadint[ir] = (s->tempreadint[ir] * (double)s->nudge_vol[id] + Tk) / ((double)s->nudge_vol[id] + 1.0);
    s->nudge_vol[id] += 1;
  }
}

/* The 7x7 running-mean deposit of read_prtcl with the x-wrap windows, in the global form
 * of the istep%ntempread==1 branch (N:1554-1620). */
static void or_deposit(or_sim *s, int n, const int *xi, const int *yi, const double *Tk) {
  const int NX = s->NX;
  const int right = s->P.fix_q3 ? NX - 3 : NX - 2;
  memset(s->nudge_vol, 0, s->nn * sizeof(int));
  memset(s->tempreadint, 0, (size_t)s->NX * (size_t)s->NY * sizeof(double));
  for (int k = 0; k < n; k++) {
    const int x_now = xi[k], y_now = yi[k];
    for (int i1 = -3; i1 < 4; i1++)
      for (int j1 = -3; j1 < 4; j1++) fold(s, x_now + i1, y_now + j1, Tk[k]);
    if (xi[k] < 4) {
      int t = NX + xi[k];
      for (int i1 = -3; i1 < 1; i1++)
        for (int j1 = -3; j1 < 4; j1++) fold(s, t + i1, y_now + j1, Tk[k]);
    }
    if (xi[k] > right) {
      int t = xi[k] - NX;
      for (int i1 = 0; i1 < 4; i1++)
        for (int j1 = -3; j1 < 4; j1++) fold(s, t + i1, y_now + j1, Tk[k]);
    }
  }
}

int or_deposit_only(or_sim *s, int n, const int *xidx, const int *yidx, const double *T) {
  or_deposit(s, n, xidx, yidx, T);
  return 0;
}

/* Fixed point of the bilinear deposit's sums: v * 2^40 rounded to the nearest integer, ties
 * to even (the scaling is exact, rint rounds in the default mode). */
static long long or_dep_fix(double v) { return (long long)rint(v * 0x1p40); }

/* Bilinear deposit (RBN_DEPOSIT_BILINEAR; no reference counterpart): observer k at the
 * continuous position (x_k, y_k) gives each of the four corners of its interpolation cell the
 * bilinear weight that the record interpolation (M:292-311) gives that corner, with ghost
 * corners folded onto the interior node they copy (periodic in x and y, as communication).
 * It is the transpose of the interpolation operator.  Per node the sums are taken in fixed
 * point, each term rounded to a multiple of 2^-40 and added exactly, so they do not depend
 * on the order of the observers:
 *   Wi = sum fix(w_k),  Si = sum fix(w_k T_k),  W = Wi 2^-40,  tempreadint = Si / Wi
 * (each converted to double; 0 where Wi = 0), nudge_vol = number of observers with w_k != 0.
 * The nudged collision then uses Q = alpha * (min(W, 1) * (T - tempreadint)):
 * sum_k alpha w_k (T - T_k) where the observers' weights at the node sum to at most 1, and a
 * full relaxation toward the weighted mean where several observers reach the node (W > 1),
 * so that alpha <= 1 never over-relaxes. */
static void or_deposit_bilinear(or_sim *s, int n, const double *x, const double *y, const double *Tk) {
  const int NX = s->NX, NY = s->NY;
  const size_t nn = (size_t)NX * (size_t)NY;
  long long *Si = (long long *)calloc(nn ? nn : 1, sizeof(long long));
  long long *Wi = (long long *)calloc(nn ? nn : 1, sizeof(long long));
  memset(s->nudge_vol, 0, s->nn * sizeof(int));
  memset(s->tempreadint, 0, nn * sizeof(double));
  memset(s->nudge_w, 0, nn * sizeof(double));
  if (!Si || !Wi) {
    free(Si);
    free(Wi);
    return;
  }
  for (int k = 0; k < n; k++) {
    const double xx = x[k], yy = y[k];
    if (!(xx > -1.0 && xx < (double)NX + 1.0 && yy > -1.0 && yy < (double)NY + 1.0)) continue;
    const int x1 = (int)xx, y1 = (int)yy, x2 = x1 + 1, y2 = y1 + 1;
    const int ci[4] = {x1, x2, x1, x2}, cj[4] = {y1, y1, y2, y2};
    const double w[4] = {((double)x2 - xx) * ((double)y2 - yy), (xx - (double)x1) * ((double)y2 - yy),
                         ((double)x2 - xx) * (yy - (double)y1), (xx - (double)x1) * (yy - (double)y1)};
    for (int c = 0; c < 4; c++) {
      if (w[c] == 0.0) continue;
      const int i = ci[c] == 0 ? NX : (ci[c] == NX + 1 ? 1 : ci[c]);
      const int j = cj[c] == 0 ? NY : (cj[c] == NY + 1 ? 1 : cj[c]);
      const size_t id = IX(s, i, j), ir = (size_t)(j - 1) + (size_t)NY * (size_t)(i - 1);
      Wi[ir] += or_dep_fix(w[c]);
      Si[ir] += or_dep_fix(w[c] * Tk[k]);
      s->nudge_vol[id] += 1;
    }
  }
  for (size_t ir = 0; ir < nn; ir++) {
    s->nudge_w[ir] = (double)Wi[ir] * 0x1p-40;
    s->tempreadint[ir] = Wi[ir] != 0 ? (double)Si[ir] / (double)Wi[ir] : 0.0;
  }
  free(Si);
  free(Wi);
}

int or_deposit_bilinear_only(or_sim *s, int n, const double *x, const double *y, const double *T) {
  or_deposit_bilinear(s, n, x, y, T);
  return 0;
}

/* The deposit of the configured kind for the current observer state. */
static void or_deposit_current(or_sim *s) {
  if (s->P.deposit == RBN_DEPOSIT_BILINEAR)
    or_deposit_bilinear(s, s->nobs, s->xint, s->yint, s->Tint);
  else
    or_deposit(s, s->nobs, s->xidx, s->yidx, s->Tint);
}

static const or_obsrec *find_obs(const or_sim *s, int step) {
  for (int r = 0; r < s->norec; r++)
    if (s->orecs[r].step == step) return &s->orecs[r];
  return NULL;
}

/* Order of (id, record index) pairs: by id, then by index. */
static int cmp_id_then_index(const void *a, const void *b) {
  const int *x = (const int *)a, *y = (const int *)b;
  if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
  return x[1] < y[1] ? -1 : (x[1] > y[1]);
}

/* abs() of the reference applied to a double: the argument converts to int (N:1496,1515). */
static inline int iabs_trunc(double d) { return abs((int)d); }

/* Advance of the interpolated y position and temperature shared by the branches
 * (N:1514-1550 and N:1697-1718). */
static void advance_y_T(or_sim *s, int i, const double *recT) {
  const double nyd = (double)s->P.ny;
  double yposnow = s->yint[i] + s->dy[i];
  if (yposnow > nyd) {
    yposnow = yposnow - nyd + 1.0;
    s->bcy[i] = 2;
  }
  if (yposnow < 1.0) {
    yposnow = yposnow + nyd - 1.0;
    s->bcy[i] = 2;
  }
  s->yint[i] = yposnow;
  s->yidx[i] = (int)round(yposnow);
  double Tnow = s->Tint[i] + s->dT[i];
  if (s->bcy[i] == 1) Tnow = s->Tint[i];
  if (s->bcy[i] == 2) {
    Tnow = recT[s->chk[i]];
    s->bcy[i] = 1;
  }
  s->Tint[i] = Tnow;
}

static double advance_x(or_sim *s, int i) {
  const double nxd = (double)s->P.nx;
  double xposnow = s->xint[i] + s->dx[i];
  if (xposnow > nxd) xposnow = xposnow - nxd;
  if (xposnow < 0.0) xposnow = xposnow + nxd;
  s->xint[i] = xposnow;
  s->xidx[i] = (int)round(xposnow);
  return xposnow;
}

/* read_prtcl (N:1380-1781): observer time interpolation and deposit. */
static int or_read_obs(or_sim *s) {
  const int n = s->nobs, ntr = s->P.ntempread;
  const double nxd = (double)s->P.nx, nyd = (double)s->P.ny;
  const int r = s->istep % ntr;
  if (r == 1) {
    const or_obsrec *rec = find_obs(s, s->istep + ntr - 1);
    if (!rec) return fail(s, RBN_ENOTFOUND, "observation record %d missing", s->istep + ntr - 1);
    if (rec->n != n) return fail(s, RBN_EINVAL, "record %d has %d entries, expected %d", rec->step, rec->n, n);
    /* chk[i] = the first record entry j whose id is observer i's (the reference's search,
     * N:1480-1487); found by binary search over the entries sorted by (id, j), the same j. */
    int *order = (int *)malloc((size_t)(n > 0 ? n : 1) * 2 * sizeof(int));
    if (!order) return fail(s, RBN_ENOMEM, "observer map allocation failed");
    for (int j = 0; j < n; j++) {
      order[2 * j] = rec->id[j];
      order[2 * j + 1] = j;
    }
    qsort(order, (size_t)n, 2 * sizeof(int), cmp_id_then_index);
    for (int i = 0; i < n; i++) {
      int lo = 0, hi = n;
      while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (order[2 * mid] < s->oid[i]) lo = mid + 1; else hi = mid;
      }
      if (lo == n || order[2 * lo] != s->oid[i]) {
        free(order);
        return fail(s, RBN_ENOTFOUND, "id %d missing from record %d", s->oid[i], rec->step);
      }
      s->chk[i] = order[2 * lo + 1];
    }
    free(order);
    for (int i = 0; i < n; i++) {
      s->bcy[i] = 0;
      double xread = rec->x[s->chk[i]], yread = rec->y[s->chk[i]];
      s->dx[i] = (xread - s->xint[i]) / ((double)ntr);
      if (iabs_trunc(xread - s->xint[i]) > (nxd / 2.0)) {
        if (xread > s->xint[i]) {
          xread = xread - nxd;
          s->dx[i] = (xread - s->xint[i]) / ((double)ntr);
        } else {
          xread = xread + nxd;
          s->dx[i] = (xread - s->xint[i]) / ((double)ntr);
        }
      }
      advance_x(s, i);
      s->dy[i] = (yread - s->yint[i]) / ((double)ntr);
      if (iabs_trunc(yread - s->yint[i]) > (nyd / 2.0)) {
        s->bcy[i] = 1;
        if (yread > s->yint[i]) {
          yread = yread - nyd + 1.0;
          s->dy[i] = (yread - s->yint[i]) / ((double)ntr);
        } else {
          yread = yread + nyd - 1.0;
          s->dy[i] = (yread - s->yint[i]) / ((double)ntr);
        }
      }
      double Tread = rec->T[s->chk[i]];
      s->dT[i] = (Tread - s->Tint[i]) / ((double)ntr);
      advance_y_T(s, i, rec->T);
    }
    or_deposit_current(s);
    for (int k = 0; k < n; k++) {
      s->prx[k] = rec->x[k];
      s->pry[k] = rec->y[k];
      s->prT[k] = rec->T[k];
    }
  } else if (r == 0) {
    for (int i = 0; i < n; i++) {
      s->xint[i] = s->prx[s->chk[i]];
      s->yint[i] = s->pry[s->chk[i]];
      s->Tint[i] = s->prT[s->chk[i]];
      s->xidx[i] = (int)round(s->prx[s->chk[i]]);
      s->yidx[i] = (int)round(s->pry[s->chk[i]]);
    }
    or_deposit_current(s);
  } else {
    for (int i = 0; i < n; i++) {
      advance_x(s, i);
      advance_y_T(s, i, s->prT);
    }
    or_deposit_current(s);
  }
  return 0;
}

/* ------------------------------------------------------------------ initial state */

static void fill_halo_copy(or_sim *s, double *a) { or_communication(s, a); }

static void init_pops(or_sim *s) {
  or_equili_all(s, s->rho, s->feq);
  memcpy(s->f, s->feq, s->nn * sizeof(or_pop));
  or_equili_all(s, s->T, s->geq);
  memcpy(s->g, s->geq, s->nn * sizeof(or_pop));
}

int or_init_fields(or_sim *s, const double *rho, const double *T, const double *u, const double *v) {
  const int NX = s->NX, NY = s->NY;
  if (!T) return or_init_scratch_reference(s);
  for (int i = 1; i < NX + 1; i++)
    for (int j = 1; j < NY + 1; j++) {
      size_t id = IX(s, i, j), h = (size_t)(i - 1) * NY + (size_t)(j - 1);
      s->rho[id] = rho ? rho[h] : s->P.rhom1;
      s->T[id] = T[h];
      s->u[id] = u ? u[h] : 0.0;
      s->v[id] = v ? v[h] : 0.0;
    }
  fill_halo_copy(s, s->rho);
  fill_halo_copy(s, s->T);
  fill_halo_copy(s, s->u);
  fill_halo_copy(s, s->v);
  init_pops(s);
  s->istep = s->P.nstart;
  s->initialised = 1;
  return 0;
}

/* inithydro scratch branch (T:680-700) for rank 0: rho = rhom1, u = v = 0, T a tanh front
 * around ny/2 whose centre is jittered per node by rand() after srand(0). */
int or_init_scratch_reference(or_sim *s) {
  const double randnorm = 1.0 / RAND_MAX, epstan = 10.;
  srand(0);
  for (int i = 0; i < s->NX + 2; i++)
    for (int j = 0; j < s->NY + 2; j++) {
      size_t id = IX(s, i, j);
      s->rho[id] = s->P.rhom1;
      double yyytemp = ((double)(j - 1 + (s->NY) * 0));
      double y_c = (double)s->P.ny * (1) / 2. + (randnorm * rand() - 0.5) * 2. * epstan;
      s->T[id] = s->P.Tdown + (s->P.Tup - s->P.Tdown) * (tanh((yyytemp - y_c) / epstan) + 1.0) / 2.0;
      s->v[id] = 0;
      s->u[id] = 0;
    }
  or_communication(s, s->u);
  or_communication(s, s->v);
  init_pops(s);
  s->istep = s->P.nstart;
  s->initialised = 1;
  return 0;
}

int or_set_populations(or_sim *s, const double *f, const double *g, const double *rho_prev,
                       const double *T_prev, const double *u_prev, const double *v_prev) {
  const int NX = s->NX, NY = s->NY;
  for (int i = 1; i < NX + 1; i++)
    for (int j = 1; j < NY + 1; j++) {
      size_t id = IX(s, i, j), h = (size_t)(i - 1) * NY + (size_t)(j - 1);
      for (int k = 0; k < NPOP; k++) {
        s->f[id].p[k] = f[h * NPOP + k];
        s->g[id].p[k] = g[h * NPOP + k];
      }
      s->rho[id] = rho_prev[h];
      s->T[id] = T_prev[h];
      s->u[id] = u_prev ? u_prev[h] : 0.0;
      s->v[id] = v_prev ? v_prev[h] : 0.0;
    }
  or_pbc(s, s->f);
  or_pbc(s, s->g);
  fill_halo_copy(s, s->u);
  fill_halo_copy(s, s->v);
  s->istep = s->P.nstart;
  s->initialised = 1;
  return 0;
}

int or_get_populations(const or_sim *s, double *f, double *g) {
  for (int i = 1; i < s->NX + 1; i++)
    for (int j = 1; j < s->NY + 1; j++) {
      size_t id = IX(s, i, j), h = (size_t)(i - 1) * s->NY + (size_t)(j - 1);
      for (int k = 0; k < NPOP; k++) {
        if (f) f[h * NPOP + k] = s->f[id].p[k];
        if (g) g[h * NPOP + k] = s->g[id].p[k];
      }
    }
  return 0;
}

int or_get_fields(const or_sim *s, double *rho, double *T, double *u, double *v) {
  for (int i = 1; i < s->NX + 1; i++)
    for (int j = 1; j < s->NY + 1; j++) {
      size_t id = IX(s, i, j), h = (size_t)(i - 1) * s->NY + (size_t)(j - 1);
      if (rho) rho[h] = s->rho[id];
      if (T) T[h] = s->T[id];
      if (u) u[h] = s->u[id];
      if (v) v[h] = s->v[id];
    }
  return 0;
}

/* evolution.dat quantities (T:1739-1757). */
int or_energy(const or_sim *s, double *mean_energy, double *mean_T) {
  double e = 0.0, t = 0.0;
  for (int i = 1; i < s->NX + 1; i++)
    for (int j = 1; j < s->NY + 1; j++) {
      size_t id = IX(s, i, j);
      e += 0.5 * s->rho[id] * (s->u[id] * s->u[id] + s->v[id] * s->v[id]);
      t += s->T[id];
    }
  e /= (((double)s->P.nx) * ((double)(s->P.ny)));
  t /= (((double)s->P.nx) * ((double)(s->P.ny)));
  if (mean_energy) *mean_energy = e;
  if (mean_T) *mean_T = t;
  return 0;
}

/* profile (N:1104-1250; the same routine, unused, at T:990-1080): per row j = 1..ny the x
 * averages (sums times 1/nx) of rho, u, v, u^2, v^2, T, vT, dzT = T(j+1) - T(j) (the top row
 * reads the ghost row, filled by pbc + hydrovar), and the vorticity squared, which the
 * reference computes but does not print (its format has one more %e than arguments); here
 * with the periodic interior neighbours in x and one-sided y differences at the walls.
 * out[(j-1)*9 + q]. */
int or_profiles(const or_sim *s, double *out) {
  const int NX = s->NX, NY = s->NY;
  const double norma = 1 / ((double)(s->P.nx));
  for (int j = 1; j <= NY; j++) {
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 1; i < NX + 1; i++) {
      const size_t id = IX(s, i, j), up = IX(s, i, j + 1);
      acc[7] += (s->T[up] - s->T[id]);
    }
    for (int i = 1; i < NX + 1; i++) {
      const size_t id = IX(s, i, j);
      const int il = i == 1 ? NX : i - 1, ir = i == NX ? 1 : i + 1;
      const double dvbydx = 0.5 * (s->v[IX(s, ir, j)] - s->v[IX(s, il, j)]);
      double dubydy;
      if (j == 1) dubydy = (s->u[IX(s, i, 2)] - s->u[id]);
      else if (j == NY) dubydy = (s->u[id] - s->u[IX(s, i, NY - 1)]);
      else dubydy = 0.5 * (s->u[IX(s, i, j + 1)] - s->u[IX(s, i, j - 1)]);
      const double vor = dvbydx - dubydy;
      acc[1] += (s->u[id]);
      acc[2] += (s->v[id]);
      acc[3] += (s->u[id] * s->u[id]);
      acc[4] += (s->v[id] * s->v[id]);
      acc[5] += s->T[id];
      acc[0] += s->rho[id];
      acc[6] += (s->v[id] * s->T[id]);
      acc[8] += (vor * vor);
    }
    for (int q = 0; q < 9; q++) out[(size_t)(j - 1) * 9 + q] = acc[q] * norma;
  }
  return 0;
}

/* ------------------------------------------------------------------ particle set-up */

static const int *g_sort_ids;
static int cmp_by_id(const void *a, const void *b) {
  int ia = *(const int *)a, ib = *(const int *)b;
  return (g_sort_ids[ia] > g_sort_ids[ib]) - (g_sort_ids[ia] < g_sort_ids[ib]);
}

/* init_prtcl (M:116-290) with given positions; the initial velocity rule of M:181-208. */
int or_init_particles(or_sim *s, int n, const int *id, const double *x, const double *y,
                      const double *tau_p, const double *beta_p, const double *up,
                      const double *vp) {
  if (n < 0) return fail(s, RBN_EINVAL, "negative particle count");
  free_particles(s);
  size_t m = (size_t)(n ? n : 1);
  s->pid = (int *)malloc(m * sizeof(int));
  s->heavy = (unsigned char *)malloc(m);
  double **arr[] = {&s->px, &s->py, &s->pu, &s->pv, &s->pxo, &s->pyo, &s->puo, &s->pvo, &s->pinvtau, &s->pbeta};
  for (size_t k = 0; k < sizeof arr / sizeof arr[0]; k++) *arr[k] = (double *)malloc(m * sizeof(double));
  int *order = (int *)malloc(m * sizeof(int));
  for (int p = 0; p < n; p++) order[p] = p;
  g_sort_ids = id;
  qsort(order, (size_t)n, sizeof(int), cmp_by_id);
  or_communication(s, s->u);
  or_communication(s, s->v);
  int err = 0;
  s->np = n;
  for (int q = 0; q < n; q++) {
    int p = order[q];
    s->pid[q] = id[p];
    s->px[q] = x[p];
    s->py[q] = y[p];
    s->pxo[q] = x[p];
    s->pyo[q] = y[p];
    double tp = tau_p ? tau_p[p] : (s->P.heavyprtcl && !s->P.tracer ? s->P.taulag : 0.0);
    s->heavy[q] = tp > 0.0;
    s->pinvtau[q] = tp > 0.0 ? (1 / tp) : 0.0;
    s->pbeta[q] = beta_p ? beta_p[p] : s->P.beta;
    if (up && vp) {
      s->pu[q] = up[p];
      s->pv[q] = vp[p];
    } else {
      double tu = or_bilinear(s, s->u, x[p], y[p], &err);
      double tv = or_bilinear(s, s->v, x[p], y[p], &err);
      if (s->heavy[q]) {
        s->pu[q] = s->P.heavy_init_factor * tu;
        s->pv[q] = s->P.heavy_init_factor * tv;
      } else {
        s->pu[q] = tu;
        s->pv[q] = tv;
      }
    }
    s->puo[q] = s->pu[q];
    s->pvo[q] = s->pv[q];
  }
  free(order);
  if (err) return fail(s, RBN_ERANGE, "particle outside the lattice at init");
  return 0;
}

int or_get_particles(const or_sim *s, int *id, double *x, double *y, double *u, double *v,
                     double *xo, double *yo, double *uo, double *vo) {
  for (int p = 0; p < s->np; p++) {
    if (id) id[p] = s->pid[p];
    if (x) x[p] = s->px[p];
    if (y) y[p] = s->py[p];
    if (u) u[p] = s->pu[p];
    if (v) v[p] = s->pv[p];
    if (xo) xo[p] = s->pxo[p];
    if (yo) yo[p] = s->pyo[p];
    if (uo) uo[p] = s->puo[p];
    if (vo) vo[p] = s->pvo[p];
  }
  return s->np;
}

int or_num_records(const or_sim *s) { return s->nrec; }

int or_get_record(const or_sim *s, int index, int *step, int *n, int *id, double *out18) {
  if (index < 0 || index >= s->nrec) return RBN_EINVAL;
  const or_rec *r = &s->recs[index];
  if (step) *step = r->step;
  if (n) *n = r->n;
  if (id) memcpy(id, r->id, (size_t)r->n * sizeof(int));
  if (out18) memcpy(out18, r->a, (size_t)18 * (size_t)r->n * sizeof(double));
  return 0;
}

/* set_nudge_vol's particle part (N:2083-2122). */
int or_init_observers(or_sim *s, int n, const int *id, const double *x, const double *y, const double *T) {
  free_observers(s);
  size_t m = (size_t)(n ? n : 1);
  int **ia[] = {&s->oid, &s->chk, &s->bcy, &s->xidx, &s->yidx};
  for (size_t k = 0; k < sizeof ia / sizeof ia[0]; k++) *ia[k] = (int *)calloc(m, sizeof(int));
  double **da[] = {&s->dx, &s->dy, &s->dT, &s->xint, &s->yint, &s->Tint, &s->prx, &s->pry, &s->prT};
  for (size_t k = 0; k < sizeof da / sizeof da[0]; k++) *da[k] = (double *)calloc(m, sizeof(double));
  s->nobs = n;
  for (int i = 0; i < n; i++) {
    s->oid[i] = id[i];
    s->prx[i] = s->xint[i] = x[i];
    s->pry[i] = s->yint[i] = y[i];
    s->prT[i] = s->Tint[i] = T[i];
  }
  memset(s->nud_term, 0, s->nn * sizeof(double));
  memset(s->nudge_vol, 0, s->nn * sizeof(int));
  return 0;
}

int or_add_obs_record(or_sim *s, int step, int n, const int *id, const double *x, const double *y, const double *T) {
  if (s->norec == s->caporec) {
    int cap = s->caporec ? 2 * s->caporec : 16;
    or_obsrec *r = (or_obsrec *)realloc(s->orecs, (size_t)cap * sizeof(or_obsrec));
    if (!r) return RBN_ENOMEM;
    s->orecs = r;
    s->caporec = cap;
  }
  or_obsrec *r = &s->orecs[s->norec++];
  size_t m = (size_t)(n ? n : 1);
  r->step = step;
  r->n = n;
  r->id = (int *)malloc(m * sizeof(int));
  r->x = (double *)malloc(m * sizeof(double));
  r->y = (double *)malloc(m * sizeof(double));
  r->T = (double *)malloc(m * sizeof(double));
  memcpy(r->id, id, (size_t)n * sizeof(int));
  memcpy(r->x, x, (size_t)n * sizeof(double));
  memcpy(r->y, y, (size_t)n * sizeof(double));
  memcpy(r->T, T, (size_t)n * sizeof(double));
  return 0;
}

int or_get_nudging_weights(const or_sim *s, double *w) {
  for (int i = 1; i < s->NX + 1; i++)
    for (int j = 1; j < s->NY + 1; j++)
      w[(size_t)(i - 1) * s->NY + (size_t)(j - 1)] = s->nudge_w[(size_t)(j - 1) + (size_t)s->NY * (size_t)(i - 1)];
  return 0;
}

int or_get_nudging(const or_sim *s, int *nudge_vol, double *tempreadint, double *nud_term) {
  for (int i = 1; i < s->NX + 1; i++)
    for (int j = 1; j < s->NY + 1; j++) {
      size_t id = IX(s, i, j), h = (size_t)(i - 1) * s->NY + (size_t)(j - 1);
      size_t ir = (size_t)(j - 1) + (size_t)s->NY * (size_t)(i - 1);
      if (nudge_vol) nudge_vol[h] = s->nudge_vol[id];
      if (tempreadint) tempreadint[h] = s->tempreadint[ir];
      if (nud_term) nud_term[h] = s->nud_term[id];
    }
  return 0;
}

int or_get_observer_state(const or_sim *s, double *xint, double *yint, double *Tint, int *xidx, int *yidx) {
  for (int i = 0; i < s->nobs; i++) {
    if (xint) xint[i] = s->xint[i];
    if (yint) yint[i] = s->yint[i];
    if (Tint) Tint[i] = s->Tint[i];
    if (xidx) xidx[i] = s->xidx[i];
    if (yidx) yidx[i] = s->yidx[i];
  }
  return s->nobs;
}

int or_num_diag(const or_sim *s) { return s->ndiag_out; }
int or_get_diag(const or_sim *s, int index, int *step, double *e, double *t) {
  if (index < 0 || index >= s->ndiag_out) return RBN_EINVAL;
  *step = s->dstep[index];
  *e = s->de[index];
  *t = s->dt[index];
  return 0;
}

static void push_diag(or_sim *s) {
  if (s->ndiag_out == s->capdiag) {
    int cap = s->capdiag ? 2 * s->capdiag : 64;
    s->dstep = (int *)realloc(s->dstep, (size_t)cap * sizeof(int));
    s->de = (double *)realloc(s->de, (size_t)cap * sizeof(double));
    s->dt = (double *)realloc(s->dt, (size_t)cap * sizeof(double));
    s->capdiag = cap;
  }
  double e, t;
  or_energy(s, &e, &t);
  s->dstep[s->ndiag_out] = s->istep;
  s->de[s->ndiag_out] = e;
  s->dt[s->ndiag_out] = t;
  s->ndiag_out++;
}

/* ------------------------------------------------------------------ steps */

/* AVERAGE (T:1343-1360), run only in faithful-cost mode: its outputs feed nothing. */
static void or_average(or_sim *s) {
  for (size_t id = 0; id < s->nn; id++) {
    double mp = (s->rho[id] * s->u[id]), mq = (s->rhopost[id] * s->upost[id]);
    s->utot[id] = 0.5 * (mp + mq) / (s->rho[id]);
    mp = (s->rho[id] * s->v[id]);
    mq = (s->rhopost[id] * s->vpost[id]);
    s->vtot[id] = 0.5 * (mp + mq) / (s->rho[id]);
  }
}

static void or_computetau(or_sim *s) {
  for (int i = 1; i < s->NX + 1; i++)
    for (int j = 1; j < s->NY + 1; j++) {
      s->tau[IX(s, i, j)] = s->P.relax;
      s->taug[IX(s, i, j)] = s->P.relaxg;
    }
}

/* One iteration of the truth-run main loop (T:1674-1759). */
static int or_step_traj(or_sim *s) {
  or_pbc(s, s->f);
  or_pbc(s, s->g);
  or_move_traj(s);
  or_pbc(s, s->f);
  or_pbc(s, s->g);
  memcpy(s->ubef, s->u, s->nn * sizeof(double));
  memcpy(s->vbef, s->v, s->nn * sizeof(double));
  or_hydrovar(s, s->u, s->v, s->rho, s->T, 0);
  if (s->P.particle) {
    int rc = or_particles(s);
    if (rc) return rc;
  }
  if (s->faithful) or_computetau(s);
  or_equili_all(s, s->rho, s->feq);
  or_collide_f(s);
  or_equili_all(s, s->T, s->geq);
  or_collide_g(s, 0);
  or_pbc(s, s->f);
  or_pbc(s, s->g);
  if (s->faithful) {
    or_hydrovar(s, s->upost, s->vpost, s->rhopost, s->Tpost, 0);
    or_average(s);
  }
  int nd = s->P.ndiag > 0 ? s->P.ndiag : s->P.nout;
  if (nd > 0 && s->istep % nd == 0) push_diag(s);
  return 0;
}

/* One iteration of the nudged-run main loop (N:2346-2411). */
static int or_step_nudge(or_sim *s) {
  or_pbc(s, s->f);
  or_pbc(s, s->g);
  int rc = or_read_obs(s);
  if (rc) return rc;
  or_move_nudge(s);
  or_pbc(s, s->f);
  or_pbc(s, s->g);
  or_hydrovar(s, s->u, s->v, s->rho, s->T, 1);
  if (s->faithful) or_computetau(s);
  or_equili_all(s, s->rho, s->feq);
  or_collide_f(s);
  or_equili_all(s, s->T, s->geq);
  or_collide_g(s, 1);
  or_pbc(s, s->f);
  or_pbc(s, s->g);
  int nd = s->P.ndiag > 0 ? s->P.ndiag : 5000;
  if (nd > 0 && s->istep % nd == 0) push_diag(s);
  return 0;
}

int or_step(or_sim *s, int nsteps) {
  if (!s->initialised) return fail(s, RBN_ESTATE, "fields not initialised");
  if (s->P.variant == RBN_VARIANT_NUDGING && s->P.ntempread < 2)
    return fail(s, RBN_EINVAL, "ntempread must be >= 2");
  for (int n = 0; n < nsteps; n++) {
    s->istep++;
    int rc = s->P.variant == RBN_VARIANT_NUDGING ? or_step_nudge(s) : or_step_traj(s);
    if (rc) return rc;
  }
  return 0;
}
