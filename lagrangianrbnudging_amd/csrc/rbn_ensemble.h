// rbn_ensemble.h — device-side description of one ensemble member (see rbn_ensemble.hip).
#pragma once

#include "rbn_internal.h"

namespace rbn {

// Steps of a replayed cycle: after kCycle steps the A/B population planes and the velocity
// ring of every member are back where they started.  An ensemble keeps each member's ring in
// its first kCycle velocity slots (rbn_step uses all rbn_ctx::kVel).
constexpr int kCycle = 6;
static_assert(kCycle <= rbn_ctx::kVel, "the ensemble's ring lies inside the member's");
static_assert(kCycle % 2 == 0, "a cycle must return the A/B planes");

// Rows of a two-step tile (64 columns) in the ensemble: the single-lattice defaults.
constexpr int kEnsTH32 = 8, kEnsTH64 = 4;

// Pointers and constants of one member as of the descriptors' base state: pf[0]/pg[0] are the
// f/g planes read by the first launch, F[0] its force plane (F[1] the second plane a two-step
// launch writes), uv[cur0] the latest velocity ((u, v) interleaved).
template <typename R>
struct EnsMember {
  Geom g;
  StepConst<R> c;
  R *pf[2], *pg[2];
  R *F[2];
  R *uv[kCycle];
  R *rho, *T;
  int cur0;
  int tiles_x, tiles_y;  // one-step tiles
  int t2x, t2y;          // two-step tiles
  double *part;  // 10 arrays of np (x, y, u, v, xo, yo, uo, vo, 1/tau, beta)
  int np;
  int *err;
  int *bad;  // non-finite moments flag (the member's d_err[1])
  // nudged members (variant RBN_VARIANT_NUDGING): the deposits the steps read (packed,
  // DepPtrs::tw), [k & 1] for step k (1: the member's own set, 0: the ensemble's second set,
  // so that the deposit of step k + 1 can be built while step k runs, or both deposits of a
  // two-step launch exist), and the persistent nudging term, read and written in place by
  // one-step launches and swapped between nud[0] and nud[1] by two-step launches (with the
  // force planes)
  int variant;
  const void *tw[2];
  R *nud[2];
  // bilinear deposit: each set's bins (DepPtrs), from which a two-step launch may sum its
  // second step's deposit in the tiles (EnsStep::bins2); null for the 7x7 deposit
  const long long *bin[2];
  int *tcnt[2];
  long long *S[2], *Wi[2];
  int *ocnt[2];
  int bin_cap;
};

// items: the n_lbm lattice items, then the n_part particle items.  phase 0: the lattice items
// of step k; 1: the particle items of step k; 2: the particle items of step k-1 followed by
// the lattice items of step k in one launch.
hipError_t launch_ensemble_phase(hipStream_t s, int prec, int phase, const void *members,
                                 const int2 *items, int n_lbm, int n_part, int *counters, int k,
                                 int tile_w, int tile_h, int write_moments, int grid, int reverse);

// One launch of the two-step mode: the particle updates of steps k-ncarry .. k-1 (items
// n_lbm .. n_lbm+n_part), then the lattice of step k as one-step tiles (lat 1, items 0 ..
// n_lbm) or of steps k, k+1 as two-step tiles (lat 2, items n_lbm+n_part ..).  lpar: which
// of pf/pg is the input; fpar: which of F (and of a nudged member's nud) is the input;
// cset: the counter set.  bins2: the bilinear members' second-step deposits are summed in the
// two-step tiles from their bins (their tile passes were left out).  nudged (of the launch):
// 0 the ensemble holds truth runs only, 1 nudged runs only, 2 both.
struct EnsStep {
  int k, lpar, fpar, ncarry, lat, moments, reverse, cset, bins2;
};
// The observer update (with the bilinear deposit's binning) of a later step carried by a
// two-step launch as extra work items, one per block of the batched observer pass: the nudged
// members' descriptors of that step's deposit set, the pass's block map, its n blocks and the
// step t in the call.
struct EnsHead {
  const NudgeMember *M;
  const int4 *map;
  int n, t;
};
hipError_t launch_ensemble_pairs(hipStream_t s, int prec, const void *members, const int2 *items, int n_lbm,
                                 int n_part, int n_lbm2, int *counters, const EnsStep &st, int tile_w,
                                 int tile_h, int grid, int nudged, const EnsHead &head);

}  // namespace rbn
