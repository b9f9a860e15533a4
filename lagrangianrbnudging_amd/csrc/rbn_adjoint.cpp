// rbn_adjoint.cpp — C-ABI forward/backward pair of the adjoint window of the nudged run
// (rbn_adj_forward / rbn_adj_backward).  No reference counterpart: the reference has no
// gradient.
//
// Cost function of a window of W steps started from rbn_init_fields:
//   J = sum over observation steps s (istep % ntempread == 0) of
//       sum over observers k of (T_s(x_k, y_k) - T_k)^2,
// where T_s is the temperature of the nudged step s, (x_k, y_k, T_k) the record observer k
// snaps to at that step, and T_s(x, y) the bilinear interpolation of the trajectory
// records (Nudging/code/tempvel_nudge.c bilinear, as in the particle code).
//
// Forward: the nudged steps of rbn_step (same kernels, same results bit for bit) plus the
// misfit after the observation steps; every K steps the state entering the step (f, g,
// lagged force, nud_term and the observer state) is copied into a checkpoint, and every
// record fetched from the observation provider is kept on the device.
// The deposit (counts, means, bilinear weights) depends only on the records.  When all W
// steps of it fit in a quarter of the free device memory (12-20 B per node and step) the
// forward writes each step's deposit into its own slot and keeps it; otherwise the backward
// replays the observer state machine and the deposit from the checkpointed observer state.
// Backward: segments in reverse order.  A segment is recomputed from its checkpoint into a
// K-slot buffer (populations, force and nud_term entering each step; the lattice step writes
// its output straight into the next slot), then swept backwards with one adjoint launch per
// step (rbn_adjoint.hip).  Finally the pulled adjoints of the first step are projected on
// the initial temperature.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "rbn_adjoint.h"
#include "rbn_host.h"
#include "rbn_internal.h"

using rbn::Geom;
using namespace rbn::host;

namespace rbn {

struct AdjState {
  int W = 0, K = 0, nseg = 0;
  int start = 0;       // istep before the first step of the window
  int nrec = 0, nobs_steps = 0, nobs = 0;
  int prec = RBN_FP64;
  bool ready = false;  // forward complete, backward possible
  size_t plane_bytes = 0;
  // forward products
  unsigned char *ckpt = nullptr;      // nseg x 20 planes: f (9), g (9), F, nud
  int *ckpt_oi = nullptr;             // nseg x ONI*nobs
  double *ckpt_od = nullptr;          // nseg x OND*nobs
  double *rec = nullptr;              // nrec x (rx, ry, rT)
  int *rec_chk = nullptr;             // nrec x nobs
  double *obs = nullptr;              // nobs_steps x (x, y, r)
  unsigned long long *maxbits = nullptr;  // nobs_steps
  double *part = nullptr;             // block partials
  double *acc = nullptr;              // [0] J, [1] dJ/dalpha
  unsigned char *init = nullptr;      // rho0 plane, then (u0, v0) interleaved (2 planes)
  // backward workspace
  unsigned char *seg = nullptr;       // K x 18 planes (populations entering each step)
  unsigned char *segFN = nullptr;     // K x 2 planes (F_in, nud_in)
  int *seg_nvol = nullptr;            // K x plane ints
  double *seg_tbar = nullptr;         // K x plane doubles
  double *seg_w = nullptr;            // K x plane doubles (bilinear deposit weights)
  // every step's deposit, kept by the forward when it fits (the backward then does not
  // replay the observers)
  bool store_dep = false;
  int *all_nvol = nullptr;            // W x plane
  double *all_tbar = nullptr;         // W x plane
  double *all_w = nullptr;            // W x plane (bilinear deposit)
  unsigned char *work = nullptr;      // F, nud planes, then a velocity slot (2 planes)
  int *rep_oi = nullptr;              // replayed observer state
  double *rep_od = nullptr;
  unsigned char *lam = nullptr;       // 2 x 18 planes
  unsigned char *lamFN = nullptr;     // lamF, lamN planes
  unsigned long long *lamT = nullptr; // plane
  double *dT0 = nullptr;              // plane (FP64)
};

void adj_free(rbn_ctx *c) {
  AdjState *a = c->adj;
  if (!a) return;
  void *ptrs[] = {a->ckpt, a->ckpt_oi, a->ckpt_od, a->rec, a->rec_chk, a->obs, a->maxbits, a->part,
                  a->acc, a->init, a->seg, a->segFN, a->seg_nvol, a->seg_tbar, a->seg_w, a->work, a->rep_oi,
                  a->rep_od, a->lam, a->lamFN, a->lamT, a->dT0, a->all_nvol, a->all_tbar, a->all_w};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  delete a;
  c->adj = nullptr;
}

}  // namespace rbn

namespace {

using rbn::AdjState;

template <typename T>
int alloc_n(rbn_ctx *c, T **p, size_t n) {
  void *q = nullptr;
  int rc = dalloc(c, &q, n * sizeof(T));
  *p = (T *)q;
  return rc;
}

unsigned char *plane_at(unsigned char *base, size_t plane_bytes, size_t k) { return base + k * plane_bytes; }

// Step constants of the adjoint (FP64 values of the forward's coefficients).
void fill_adj_consts(const rbn_ctx *c, rbn::AdjStepArgs &a) {
  const rbn_params &P = c->P;
  const double om = 1. / P.relax, omg = 1. / P.relaxg;
  a.om = om;
  a.one_m_om = 1. - om;
  a.omg = omg;
  a.one_m_omg = 1. - omg;
  a.g0 = -P.gravity;
  a.alpha = P.alpha;
  const bool shifted = c->prec == RBN_FP32;
  for (int k = 0; k < 9; k++) {
    a.src[k] = (1.0 - 0.5 * om) * rbn::kW[k];
    a.bw[k] = (1.0 - 0.5 * omg) * rbn::kW[k];
    a.sf[k] = shifted ? rbn::kW[k] * P.rho0_shift : 0.0;
    a.sg[k] = shifted ? rbn::kW[k] * P.T0_shift : 0.0;
  }
}

int launch_step_kernels(rbn_ctx *c, int istep, const rbn::ObsPtrs &o, const rbn::DepPtrs &d) {
  const rbn_params &P = c->P;
  const int r = istep % P.ntempread;
  const int branch = r == 1 ? 1 : (r == 0 ? 0 : 2);
  const int right = P.fix_q3 ? P.nx - 3 : P.nx - 2;
  return timed(c, KT_DEPOSIT, c->stream,
               [&] { return rbn::launch_obs_deposit(c->stream, branch, P.ntempread, c->geom, right, o, d); });
}

int ensure_forward_storage(rbn_ctx *c, int W, int K) {
  const Geom &g = c->geom;
  const size_t pb = g.plane * c->rs;
  const int nseg = (W + K - 1) / K;
  const int ntr = c->P.ntempread;
  const int nrec = (W + ntr - 1) / ntr, nos = W / ntr;
  const bool bil = c->dep.mode == RBN_DEPOSIT_BILINEAR;
  const size_t dep_bytes = (size_t)W * g.plane * (sizeof(int) + sizeof(double) + (bil ? sizeof(double) : 0));
  bool want_store = false;
  if (c->adj_store_deposit) {
    size_t free_b = 0, total_b = 0;
    want_store = hipMemGetInfo(&free_b, &total_b) == hipSuccess && dep_bytes <= free_b / 4;
  }
  AdjState *a = c->adj;
  if (a && a->W == W && a->K == K && a->nobs == c->nobs && a->prec == c->prec && a->plane_bytes == pb &&
      (a->all_nvol != nullptr) == want_store && (a->all_w != nullptr) == (want_store && bil)) {
    a->ready = false;
    return RBN_OK;
  }
  rbn::adj_free(c);
  a = c->adj = new AdjState();
  a->W = W;
  a->K = K;
  a->nseg = nseg;
  a->nrec = nrec;
  a->nobs_steps = nos;
  a->nobs = c->nobs;
  a->prec = c->prec;
  a->plane_bytes = pb;
  const size_t n = (size_t)c->nobs;
  const size_t nparts = (size_t)std::max(rbn::adj_blocks(g), rbn::misfit_blocks(c->nobs));
  int rc;
  if ((rc = alloc_n(c, &a->ckpt, (size_t)nseg * 20 * pb)) || (rc = alloc_n(c, &a->ckpt_oi, (size_t)nseg * ONI * n)) ||
      (rc = alloc_n(c, &a->ckpt_od, (size_t)nseg * OND * n)) || (rc = alloc_n(c, &a->rec, (size_t)nrec * 3 * n)) ||
      (rc = alloc_n(c, &a->rec_chk, (size_t)nrec * n)) || (rc = alloc_n(c, &a->obs, (size_t)std::max(nos, 1) * 3 * n)) ||
      (rc = alloc_n(c, &a->maxbits, (size_t)std::max(nos, 1))) || (rc = alloc_n(c, &a->part, nparts)) ||
      (rc = alloc_n(c, &a->acc, 2)) || (rc = alloc_n(c, &a->init, 3 * pb))) {
    rbn::adj_free(c);
    return rc;
  }
  if (want_store) {
    void *p1 = nullptr, *p2 = nullptr, *p3 = nullptr;
    const size_t np = (size_t)W * g.plane;
    bool ok = hipMalloc(&p1, np * sizeof(int)) == hipSuccess && hipMalloc(&p2, np * sizeof(double)) == hipSuccess &&
              (!bil || hipMalloc(&p3, np * sizeof(double)) == hipSuccess);
    if (ok) {
      a->all_nvol = (int *)p1;
      a->all_tbar = (double *)p2;
      a->all_w = (double *)p3;
      a->store_dep = true;
    } else {  // not enough memory after all: replay instead
      for (void *p : {p1, p2, p3})
        if (p) (void)hipFree(p);
      (void)hipGetLastError();
    }
  }
  return RBN_OK;
}

int ensure_backward_storage(rbn_ctx *c) {
  AdjState *a = c->adj;
  if (a->seg) return RBN_OK;
  const Geom &g = c->geom;
  const size_t pb = a->plane_bytes, K = (size_t)a->K, n = (size_t)a->nobs;
  int rc;
  const bool replay = !a->store_dep;
  if ((rc = alloc_n(c, &a->seg, K * 18 * pb)) || (rc = alloc_n(c, &a->segFN, K * 2 * pb)) ||
      (replay && (rc = alloc_n(c, &a->seg_nvol, K * g.plane))) ||
      (replay && (rc = alloc_n(c, &a->seg_tbar, K * g.plane))) ||
      (rc = alloc_n(c, &a->work, 4 * pb)) || (replay && (rc = alloc_n(c, &a->rep_oi, ONI * n))) ||
      (replay && (rc = alloc_n(c, &a->rep_od, OND * n))) || (rc = alloc_n(c, &a->lam, 2 * 18 * pb)) ||
      (rc = alloc_n(c, &a->lamFN, 2 * pb)) || (rc = alloc_n(c, &a->lamT, g.plane)) ||
      (rc = alloc_n(c, &a->dT0, g.plane)) ||
      (replay && c->dep.mode == RBN_DEPOSIT_BILINEAR && (rc = alloc_n(c, &a->seg_w, K * g.plane)))) {
    // keep the forward products; free the partial workspace
    void **ws[] = {(void **)&a->seg, (void **)&a->segFN, (void **)&a->seg_nvol, (void **)&a->seg_tbar, (void **)&a->seg_w,
                   (void **)&a->work, (void **)&a->rep_oi, (void **)&a->rep_od, (void **)&a->lam,
                   (void **)&a->lamFN, (void **)&a->lamT, (void **)&a->dT0};
    for (void **p : ws) dfree(*p);
    return rc;
  }
  return RBN_OK;
}

}  // namespace

extern "C" {

int rbn_adj_forward(rbn_ctx *c, int nsteps, int ckpt_every, double *J) {
  int rc = check_ctx(c);
  if (rc) return rc;
  const rbn_params &P = c->P;
  if (P.variant != RBN_VARIANT_NUDGING) return fail(c, RBN_ESTATE, "the adjoint window belongs to the nudged variant");
  if (nsteps < 1 || ckpt_every < 1 || !J) return fail(c, RBN_EINVAL, "rbn_adj_forward: bad arguments");
  if (!c->at_init) return fail(c, RBN_ESTATE, "rbn_adj_forward needs the state set by rbn_init_fields (no step since)");
  if (c->nobs <= 0 || !c->obs_fn) return fail(c, RBN_ESTATE, "rbn_adj_forward needs observers and an observation provider");
  if (P.ny < 4) return fail(c, RBN_EINVAL, "rbn_adj_forward needs ny >= 4");
  if (c->istep % P.ntempread != 0) return fail(c, RBN_ESTATE, "the window must start at a multiple of ntempread");
  const Geom &g = c->geom;
  const int K = std::min(ckpt_every, nsteps);
  if ((rc = ensure_forward_storage(c, nsteps, K))) return rc;
  AdjState *a = c->adj;
  const size_t pb = a->plane_bytes, n = (size_t)c->nobs;
  a->start = c->istep;
  hipStream_t s = c->stream;
  // initial moments (rbn_init_fields left rho0 in rho and (u0, v0) in the first ring slot)
  HIPCHK(c, hipMemcpyAsync(plane_at(a->init, pb, 0), c->rho, pb, hipMemcpyDeviceToDevice, s));
  HIPCHK(c, hipMemcpyAsync(plane_at(a->init, pb, 1), c->uv[c->cur], 2 * pb, hipMemcpyDeviceToDevice, s));
  HIPCHK(c, hipMemsetAsync(a->acc, 0, 2 * sizeof(double), s));
  HIPCHK(c, hipMemsetAsync(a->maxbits, 0, (size_t)std::max(a->nobs_steps, 1) * sizeof(unsigned long long), s));
  c->at_init = false;
  const int ntr = P.ntempread;
  const rbn::ObsPtrs o = obs_ptrs(c);
  for (int t = 0; t < nsteps; t++) {
    const int istep = ++c->istep;
    if (t % K == 0) {
      unsigned char *ck = plane_at(a->ckpt, pb, (size_t)(t / K) * 20);
      HIPCHK(c, hipMemcpyAsync(ck, c->fA, 9 * pb, hipMemcpyDeviceToDevice, s));
      HIPCHK(c, hipMemcpyAsync(ck + 9 * pb, c->gA, 9 * pb, hipMemcpyDeviceToDevice, s));
      HIPCHK(c, hipMemcpyAsync(ck + 18 * pb, c->F, pb, hipMemcpyDeviceToDevice, s));
      HIPCHK(c, hipMemcpyAsync(ck + 19 * pb, c->nud, pb, hipMemcpyDeviceToDevice, s));
      HIPCHK(c, hipMemcpyAsync(a->ckpt_oi + (size_t)(t / K) * ONI * n, c->d_obs_i, ONI * n * sizeof(int), hipMemcpyDeviceToDevice, s));
      HIPCHK(c, hipMemcpyAsync(a->ckpt_od + (size_t)(t / K) * OND * n, c->d_obs_d, OND * n * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    const int r = istep % ntr;
    if (r == 1) {
      if ((rc = load_observation(c, istep + ntr - 1))) return rc;
      const size_t e = (size_t)(t / ntr);
      HIPCHK(c, hipMemcpyAsync(a->rec + e * 3 * n, (const void *)o.rx, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, s));
      HIPCHK(c, hipMemcpyAsync(a->rec_chk + e * n, o.chk, n * sizeof(int), hipMemcpyDeviceToDevice, s));
    }
    rbn::DepPtrs d = c->dep;
    if (a->store_dep) {
      d.cnt = a->all_nvol + (size_t)t * g.plane;
      d.Tbar = a->all_tbar + (size_t)t * g.plane;
      d.wq = a->all_w ? a->all_w + (size_t)t * g.plane : nullptr;
    }
    if ((rc = launch_step_kernels(c, istep, o, d))) return rc;
    const bool obs_step = r == 0;
    const bool moments = obs_step || t == nsteps - 1;
    const int nxt = (c->cur + 1) % rbn_ctx::kVel;
    rbn::LbmPtrs lp{};
    lp.fA = c->fA;
    lp.gA = c->gA;
    lp.fB = c->fB;
    lp.gB = c->gB;
    lp.F = c->F;
    lp.uv = c->uv[nxt];
    lp.rho = moments ? c->rho : nullptr;
    lp.T = moments ? c->T : nullptr;
    lp.nvol = d.cnt;
    lp.Tbar = d.Tbar;
    lp.nud = c->nud;
    lp.wq = c->dep.mode == RBN_DEPOSIT_BILINEAR ? d.wq : nullptr;
    lp.bad = c->d_err + 1;
    if ((rc = timed(c, KT_LBM, s, [&] { return rbn::launch_lbm(s, c->prec, P.variant, g, lp, c->consts.data(), c->tune); })))
      return rc;
    std::swap(c->fA, c->fB);
    std::swap(c->gA, c->gB);
    c->cur = nxt;
    c->moments_valid = moments;
    if (obs_step) {
      const size_t m = (size_t)((t + 1) / ntr - 1);
      double *ob = a->obs + m * 3 * n;
      HIPCHK(c, hipMemcpyAsync(ob, o.xint, 2 * n * sizeof(double), hipMemcpyDeviceToDevice, s));
      HIPCHK(c, rbn::launch_misfit(s, c->prec, g, c->T, c->nobs, o.xint, o.yint, o.Tint, ob + 2 * n, a->part,
                                   a->maxbits + m, c->d_err));
      HIPCHK(c, rbn::launch_sum_partials(s, a->part, rbn::misfit_blocks(c->nobs), a->acc));
    }
  }
  if (a->store_dep) {  // the context's nudging state is the last step's
    const size_t last = (size_t)(nsteps - 1) * g.plane;
    HIPCHK(c, hipMemcpyAsync(c->dep.cnt, a->all_nvol + last, g.plane * sizeof(int), hipMemcpyDeviceToDevice, s));
    HIPCHK(c, hipMemcpyAsync(c->dep.Tbar, a->all_tbar + last, g.plane * sizeof(double), hipMemcpyDeviceToDevice, s));
    if (a->all_w)
      HIPCHK(c, hipMemcpyAsync(c->dep.wq, a->all_w + last, g.plane * sizeof(double), hipMemcpyDeviceToDevice, s));
  }
  if (c->timing && (rc = drain_timing(c))) return rc;
  if ((rc = check_err_flag(c))) return rc;
  HIPCHK(c, hipMemcpyAsync(J, a->acc, sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  a->ready = true;
  return RBN_OK;
}

int rbn_adj_backward(rbn_ctx *c, double *dJ_dT0, double *dJ_dalpha) {
  int rc = check_ctx(c);
  if (rc) return rc;
  AdjState *a = c->adj;
  if (!a || !a->ready) return fail(c, RBN_ESTATE, "rbn_adj_backward needs a completed rbn_adj_forward");
  if (a->prec != c->prec) return fail(c, RBN_ESTATE, "adjoint state does not match the context");
  if ((rc = ensure_backward_storage(c))) return rc;
  const rbn_params &P = c->P;
  const Geom &g = c->geom;
  const size_t pb = a->plane_bytes, n = (size_t)a->nobs;
  const int ntr = P.ntempread;
  hipStream_t s = c->stream;
  unsigned char *lamF = a->lamFN, *lamN = a->lamFN + pb;
  unsigned char *wF = plane_at(a->work, pb, 0), *wN = plane_at(a->work, pb, 1);
  unsigned char *wuv = plane_at(a->work, pb, 2);
  HIPCHK(c, hipMemsetAsync(a->lamFN, 0, 2 * pb, s));
  HIPCHK(c, hipMemsetAsync(a->lamT, 0, g.plane * sizeof(unsigned long long), s));
  HIPCHK(c, hipMemsetAsync(a->acc + 1, 0, sizeof(double), s));
  const rbn::ObsPtrs ro = a->store_dep ? rbn::ObsPtrs{} : obs_view(a->nobs, a->rep_oi, a->rep_od);
  rbn::AdjStepArgs args{};
  args.g = g;
  args.nobs = a->nobs;
  args.dalpha_part = a->part;
  fill_adj_consts(c, args);
  const int nparts = rbn::adj_blocks(g);
  int cur = 0;          // lam[cur] holds the pulled adjoints of the step after the current one
  bool have_next = false;
  for (int sg = a->nseg - 1; sg >= 0; sg--) {
    const int t0 = sg * a->K, L = std::min(a->K, a->W - t0);
    const unsigned char *ck = plane_at(a->ckpt, pb, (size_t)sg * 20);
    HIPCHK(c, hipMemcpyAsync(a->seg, ck, 18 * pb, hipMemcpyDeviceToDevice, s));
    HIPCHK(c, hipMemcpyAsync(wF, ck + 18 * pb, pb, hipMemcpyDeviceToDevice, s));
    HIPCHK(c, hipMemcpyAsync(wN, ck + 19 * pb, pb, hipMemcpyDeviceToDevice, s));
    if (!a->store_dep) {
      HIPCHK(c, hipMemcpyAsync(a->rep_oi, a->ckpt_oi + (size_t)sg * ONI * n, ONI * n * sizeof(int), hipMemcpyDeviceToDevice, s));
      HIPCHK(c, hipMemcpyAsync(a->rep_od, a->ckpt_od + (size_t)sg * OND * n, OND * n * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    // the deposit of window step w: kept by the forward, or replayed into segment slot t
    auto dep_of = [&](int w, int t) {
      rbn::DepPtrs d = c->dep;
      if (a->store_dep) {
        d.cnt = a->all_nvol + (size_t)w * g.plane;
        d.Tbar = a->all_tbar + (size_t)w * g.plane;
        d.wq = a->all_w ? a->all_w + (size_t)w * g.plane : nullptr;
      } else {
        d.cnt = a->seg_nvol + (size_t)t * g.plane;
        d.Tbar = a->seg_tbar + (size_t)t * g.plane;
        d.wq = a->seg_w ? a->seg_w + (size_t)t * g.plane : nullptr;
      }
      return d;
    };
    // recompute the segment
    for (int t = 0; t < L; t++) {
      const int w = t0 + t, istep = a->start + w + 1;
      const rbn::DepPtrs d = dep_of(w, t);
      if (!a->store_dep) {
        if (istep % ntr == 1) {
          const size_t e = (size_t)(w / ntr);
          HIPCHK(c, hipMemcpyAsync((void *)ro.rx, a->rec + e * 3 * n, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, s));
          HIPCHK(c, hipMemcpyAsync((void *)ro.chk, a->rec_chk + e * n, n * sizeof(int), hipMemcpyDeviceToDevice, s));
        }
        if ((rc = launch_step_kernels(c, istep, ro, d))) return rc;
      }
      unsigned char *fn = plane_at(a->segFN, pb, 2 * (size_t)t);
      HIPCHK(c, hipMemcpyAsync(fn, wF, pb, hipMemcpyDeviceToDevice, s));
      HIPCHK(c, hipMemcpyAsync(fn + pb, wN, pb, hipMemcpyDeviceToDevice, s));
      if (t == L - 1) break;  // the segment's last output is not needed
      unsigned char *in = plane_at(a->seg, pb, 18 * (size_t)t), *out = plane_at(a->seg, pb, 18 * (size_t)(t + 1));
      rbn::LbmPtrs lp{};
      lp.fA = in;
      lp.gA = in + 9 * pb;
      lp.fB = out;
      lp.gB = out + 9 * pb;
      lp.F = wF;
      lp.uv = wuv;
      lp.nvol = d.cnt;
      lp.Tbar = d.Tbar;
      lp.nud = wN;
      lp.wq = d.wq;
      if ((rc = timed(c, KT_LBM, s, [&] { return rbn::launch_lbm(s, c->prec, P.variant, g, lp, c->consts.data(), c->tune); })))
        return rc;
    }
    // sweep the segment backwards
    for (int t = L - 1; t >= 0; t--) {
      const int w = t0 + t, istep = a->start + w + 1;
      const bool obs_step = istep % ntr == 0;
      const size_t m = obs_step ? (size_t)((w + 1) / ntr - 1) : 0;
      if (obs_step) {
        const double *ob = a->obs + m * 3 * n;
        HIPCHK(c, rbn::launch_misfit_adjoint(s, g, a->nobs, ob, ob + n, ob + 2 * n, a->maxbits + m, a->lamT));
      }
      const unsigned char *in = plane_at(a->seg, pb, 18 * (size_t)t);
      const unsigned char *fn = plane_at(a->segFN, pb, 2 * (size_t)t);
      args.fA = in;
      args.gA = in + 9 * pb;
      args.Fin = fn;
      args.nudin = fn + pb;
      const rbn::DepPtrs d = dep_of(w, t);
      args.nvol = d.cnt;
      args.Tbar = d.Tbar;
      args.wq = d.wq;
      args.lam_next = have_next ? plane_at(a->lam, pb, 18 * (size_t)cur) : nullptr;
      args.lam = plane_at(a->lam, pb, 18 * (size_t)(1 - cur));
      args.lamF = lamF;
      args.lamN = lamN;
      args.lamT_fx = obs_step ? a->lamT : nullptr;
      args.maxbits = a->maxbits + m;
      if ((rc = timed(c, KT_ADJOINT, s, [&] { return rbn::launch_adjoint_step(s, c->prec, args); }))) return rc;
      HIPCHK(c, rbn::launch_sum_partials(s, a->part, nparts, a->acc + 1));
      cur = 1 - cur;
      have_next = true;
    }
  }
  HIPCHK(c, rbn::launch_adj_init(s, c->prec, g, plane_at(a->lam, pb, 18 * (size_t)cur), lamF, plane_at(a->init, pb, 0),
                                 plane_at(a->init, pb, 1), -P.gravity, a->dT0));
  if (c->timing && (rc = drain_timing(c))) return rc;
  double h[2];
  std::vector<double> buf(dJ_dT0 ? g.plane : 0);
  HIPCHK(c, hipMemcpyAsync(h, a->acc, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  if (dJ_dT0) HIPCHK(c, hipMemcpyAsync(buf.data(), a->dT0, g.plane * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  if (dJ_dalpha) *dJ_dalpha = h[1];
  if (dJ_dT0) {
    for (int i = 0; i < g.nx; i++)
      for (int j = 0; j < g.ny; j++) dJ_dT0[(size_t)i * g.ny + j] = buf[(size_t)j * g.pitch + i];
  }
  return RBN_OK;
}

}  // extern "C"
