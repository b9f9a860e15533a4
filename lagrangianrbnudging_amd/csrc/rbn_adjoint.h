// rbn_adjoint.h — device interface of the adjoint of the nudged step (rbn_adjoint.hip).
#pragma once

#include "rbn_internal.h"

namespace rbn {

// Arguments of one adjoint step s of the nudged run.
struct AdjStepArgs {
  Geom g;
  // Forward state entering step s (recomputed from the nearest checkpoint).
  const void *fA, *gA;      // 9 planes each; FP32 populations are stored shifted
  const void *Fin, *nudin;  // lagged force and nud_term entering the step
  const int *nvol;          // deposit counts of the step
  const double *Tbar;       // deposit means of the step
  const double *wq;         // bilinear deposit weights of the step (nullptr: weight 1)
  // Adjoint state.
  const void *lam_next;     // pulled-population adjoints of step s+1 (f: planes 0-8, g: 9-17),
                            // nullptr after the last step of the window
  void *lam;                // pulled-population adjoints of step s (output, 18 planes)
  void *lamF, *lamN;        // adjoints of the lagged force and of nud_term: step s+1's on
                            // entry, step s's on exit (updated in place, one node per thread)
  unsigned long long *lamT_fx;      // fixed-point misfit source of step s (read and cleared),
                                    // nullptr on steps without observations
  const unsigned long long *maxbits;// max |residual| of the observation (fixed-point scale)
  int nobs;
  double *dalpha_part;      // per-block partial sums of dJ/dalpha
  // Constants (FP64, the values of the forward step).
  double sf[9], sg[9];      // population shifts w_k*rho0, w_k*T0 (zero in FP64 mode)
  double om, one_m_om, omg, one_m_omg, g0, alpha;
  double src[9], bw[9];
};

int adj_blocks(const Geom &g);
hipError_t launch_adjoint_step(hipStream_t s, int prec, const AdjStepArgs &a);
// acc += sum of n partials, in a fixed order.
hipError_t launch_sum_partials(hipStream_t s, const double *part, int n, double *acc);
// dJ/dT0 (FP64, device layout) from the pulled adjoints of the first step of the window and
// the adjoint of the first lagged force, for the equilibrium initial state of rbn_init_fields.
hipError_t launch_adj_init(hipStream_t s, int prec, const Geom &g, const void *lam, const void *lamF,
                           const void *rho0, const void *uv0, double g0, double *dT0);
// Misfit of one observation: residuals r_k = T(x_k, y_k) - T_k (bilinear, observer order),
// per-block sums of r_k^2 and the max |r_k| (as ordered bits).
hipError_t launch_misfit(hipStream_t s, int prec, const Geom &g, const void *T, int n, const double *x,
                         const double *y, const double *Tobs, double *res, double *part,
                         unsigned long long *maxbits, int *err);
int misfit_blocks(int n);
// Scatter 2 r_k times the bilinear weights into the fixed-point source plane.
hipError_t launch_misfit_adjoint(hipStream_t s, const Geom &g, int n, const double *x, const double *y,
                                 const double *res, const unsigned long long *maxbits,
                                 unsigned long long *lamT_fx);

}  // namespace rbn
