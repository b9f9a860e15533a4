/*
 * rbn_oracle.h — CPU FP64 restatement of the reference Rayleigh–Bénard step.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline may load this library, and only as the checker / the timed CPU restatement of
 * the reference algorithm.  librbn (the product) never links or calls it.
 *
 * It restates, on one domain (the reference at NPX = NPY = 1), the two main loops of the
 * reference: Trajectories/code/temppart_dump.c:1674-1759 (truth run with particles and
 * trajectory records) and Nudging/code/tempvel_nudge.c:2346-2411 (nudged run).  Layout and
 * sweep structure follow the reference (array of 9-double structs with a one-node halo,
 * idx = j + (ny+2)*i, per-direction in-place streaming sweeps, separate moment,
 * equilibrium and collision sweeps) so that its cost is representative of the reference's
 * CPU path; every arithmetic expression keeps the reference's operand order so that the
 * FP64 results are reproducible bit for bit.  Compile with -ffp-contract=off.
 *
 * Parity pins (tests/test_oracle_pins.py): closed-form equilibrium moments, Guo-source
 * invariants, rest-state and conduction fixed points, bilinear exactness, AB2/Maxey–Riley
 * closed forms, the wrap/reflect map, deposit known answers, the nudging recurrence, the
 * observer state machine's known answers across record boundaries and walls, and
 * independent numpy restatements of a whole truth-run step, a whole nudged step and a whole
 * particle step, and the reference-scratch initialisation against a numpy restatement of
 * the jittered tanh front over an independent restatement of glibc's rand() (its published
 * TYPE_3 additive feedback algorithm, checked against rand()'s known first outputs).
 */
#ifndef RBN_ORACLE_H
#define RBN_ORACLE_H

#include "../include/rbn.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_sim or_sim;

or_sim *or_create(const rbn_params *p);
void or_destroy(or_sim *s);
const char *or_last_error(const or_sim *s);
/* 1: also run the reference's per-step work that has no observable output (computetau,
 * the post-collision hydrovar + AVERAGE of the truth run), for cost-faithful timing. */
void or_set_faithful(or_sim *s, int on);

int or_init_fields(or_sim *s, const double *rho, const double *T, const double *u,
                   const double *v);
int or_init_scratch_reference(or_sim *s);
int or_set_populations(or_sim *s, const double *f, const double *g, const double *rho_prev,
                       const double *T_prev, const double *u_prev, const double *v_prev);
int or_get_populations(const or_sim *s, double *f, double *g);
int or_get_fields(const or_sim *s, double *rho, double *T, double *u, double *v);
int or_energy(const or_sim *s, double *mean_energy, double *mean_T);
/* Row profiles of the latest step: out[(j-1)*9 + q], q = rho, u, v, u^2, v^2, T, vT, dzT,
 * vorticity^2 (x averages). */
int or_profiles(const or_sim *s, double *out);

int or_init_particles(or_sim *s, int n, const int *id, const double *x, const double *y,
                      const double *tau_p, const double *beta_p, const double *up,
                      const double *vp);
int or_get_particles(const or_sim *s, int *id, double *x, double *y, double *u, double *v,
                     double *xo, double *yo, double *uo, double *vo);

/* Records produced so far (truth run).  Arrays of length n in ascending id order; fields
 * 0..9 are id,x,y,u,v,xo,yo,uo,vo,temp and 10..17 the Writefull set
 * dtu,dtv,dxu,dxv,dyu,dyv,fu,fv (zero when prtcl_writefull is 0). */
int or_num_records(const or_sim *s);
int or_get_record(const or_sim *s, int index, int *step, int *n, int *id, double *out18);

int or_init_observers(or_sim *s, int n, const int *id, const double *x, const double *y,
                      const double *T);
int or_add_obs_record(or_sim *s, int step, int n, const int *id, const double *x,
                      const double *y, const double *T);
int or_get_nudging(const or_sim *s, int *nudge_vol, double *tempreadint, double *nud_term);
/* Observer interpolation state: xint, yint, Tint, xidx, yidx (length n, observer order). */
int or_get_observer_state(const or_sim *s, double *xint, double *yint, double *Tint,
                          int *xidx, int *yidx);

/* Stand-alone pieces, exposed for closed-form tests. */
void or_equilibrium(double r1, double u, double v, double out9[9]);
double or_bilinear_global(const or_sim *s, int field, double x, double y, int *err);
/* Run only the deposit for given rounded positions and temperatures (observer order). */
int or_deposit_only(or_sim *s, int n, const int *xidx, const int *yidx, const double *T);
/* Run only the bilinear deposit (RBN_DEPOSIT_BILINEAR) for given continuous positions. */
int or_deposit_bilinear_only(or_sim *s, int n, const double *x, const double *y, const double *T);
/* Per-node bilinear deposit weights W ([nx][ny]; zero for the 7x7 deposit). */
int or_get_nudging_weights(const or_sim *s, double *w);

int or_step(or_sim *s, int nsteps);
int or_get_step(const or_sim *s);
/* Diagnostics recorded during stepping: step, <rho|u|^2/2>, <T>. */
int or_num_diag(const or_sim *s);
int or_get_diag(const or_sim *s, int index, int *step, double *e, double *t);

#ifdef __cplusplus
}
#endif
#endif
