// rbn_particle_node.h — device code of K2, the Lagrangian particles of the truth run, for
// one particle; shared by the plain launch (rbn_particles.hip) and the persistent ensemble
// launch (rbn_ensemble.hip).
//
// One thread advances one particle through the reference's per-step sequence
// (Trajectories/code/mympi2.h):
//   bilinear interpolation of the fluid velocity (bilinear, :292-311), with the periodic
//   ghost values of communication (:313-348) resolved as an index wrap in x and in y;
//   tracer: second-order Adams–Bashforth position update (:358-372);
//   inertial: the Maxey–Riley velocity relaxation without gravity (:373-433), whose node
//   fields (time difference, centred x/y differences, one-sided at the walls, :384-405)
//   are evaluated on the fly from u, v and the previous step's u, v instead of being
//   stored as six full fields;
//   on record steps the trajectory-record quantities (prtcl_write,
//   Trajectories/code/temppart_dump.c:1467-1509, and the Writefull block, :441-501), taken
//   after the update and before the wrap;
//   the periodic-x wrap and reflecting-y rule of toggle_prtcl (:568-760).
// Particle state is FP64 in both precision modes.  With FP64 fields the arithmetic is
// operand-for-operand the reference's (compiled with -ffp-contract=off).
#pragma once

#include "rbn_internal.h"

namespace rbn {
namespace part {

// A node field: node o's value is a[o * cs] (cs = 1 for scalar planes; 2 for the u or v
// component of an interleaved velocity slot, a pointing at the component).
template <typename R>
struct Field {
  const R *__restrict__ a;
  int nx, ny, pitch, cs;
  // Value at halo coordinates ic in [0, nx+1], jc in [0, ny+1]: ghosts are periodic copies
  // (communication on one rank, periodic in both directions).
  __device__ __forceinline__ double at(int ic, int jc) const {
    const int i = ic == 0 ? nx : (ic == nx + 1 ? 1 : ic);
    const int j = jc == 0 ? ny : (jc == ny + 1 ? 1 : jc);
    return (double)a[((size_t)(j - 1) * pitch + (i - 1)) * cs];
  }
};

// An interleaved velocity slot: node o holds (u, v) at uv[2o], uv[2o + 1], fetched with one
// vector load per node.
template <typename R>
struct VelField {
  const R *__restrict__ uv;
  int nx, ny, pitch;
  template <typename D>
  __device__ __forceinline__ void at(int ic, int jc, D &u, D &v) const {
    typedef R R2 __attribute__((ext_vector_type(2)));
    const int i = ic == 0 ? nx : (ic == nx + 1 ? 1 : ic);
    const int j = jc == 0 ? ny : (jc == ny + 1 ? 1 : jc);
    const R2 w = ((const R2 *)uv)[(size_t)(j - 1) * pitch + (i - 1)];
    u = (D)w.x;
    v = (D)w.y;
  }
};

template <typename R>
__device__ __host__ __forceinline__ VelField<R> vel_of(const void *uv, int nx, int ny, int pitch) {
  return VelField<R>{(const R *)uv, nx, ny, pitch};
}
template <typename R>
__device__ __host__ __forceinline__ Field<R> scalar_of(const void *a, int nx, int ny, int pitch) {
  return Field<R>{(const R *)a, nx, ny, pitch, 1};
}

// Lower corner of the interpolation cell (C truncation, M:293-296) and whether all four
// corners lie on the halo-extended lattice, i.e. x in (-1, nx+1), y in (-1, ny+1).  The
// position is range-checked before the int conversion, so NaN or infinite positions are
// rejected instead of saturating into range.
__device__ __forceinline__ bool corners(double xx, double yy, int nx, int ny, int &x1, int &y1) {
  if (!(xx > -1.0 && xx < (double)nx + 1.0 && yy > -1.0 && yy < (double)ny + 1.0)) {
    x1 = y1 = 0;
    return false;
  }
  x1 = (int)xx;
  y1 = (int)yy;
  return true;
}

// bilinear (M:292-311) from four corner values.
__device__ __forceinline__ double blend(double u11, double u12, double u21, double u22, int x1,
                                        int y1, double xx, double yy) {
  const int x2 = x1 + 1, y2 = y1 + 1;
  double val1 = u11 * ((double)x2 - xx) * ((double)y2 - yy);
  double val2 = u21 * (xx - (double)x1) * ((double)y2 - yy);
  double val3 = u12 * ((double)x2 - xx) * (yy - (double)y1);
  double val4 = u22 * (xx - (double)x1) * (yy - (double)y1);
  return val1 + val2 + val3 + val4;
}

template <typename R>
__device__ __forceinline__ double interp(const Field<R> &F, double xx, double yy, int &err) {
  int x1, y1;
  if (!corners(xx, yy, F.nx, F.ny, x1, y1)) {
    err = 1;
    return 0.0;
  }
  return blend(F.at(x1, y1), F.at(x1, y1 + 1), F.at(x1 + 1, y1), F.at(x1 + 1, y1 + 1), x1, y1,
               xx, yy);
}

// bilinear of u and v from the same four corner loads.
template <typename R>
__device__ __forceinline__ void interp_uv(const VelField<R> &F, double xx, double yy, double &up,
                                          double &vp, int &err) {
  int x1, y1;
  if (!corners(xx, yy, F.nx, F.ny, x1, y1)) {
    err = 1;
    up = vp = 0.0;
    return;
  }
  double u11, v11, u12, v12, u21, v21, u22, v22;
  F.at(x1, y1, u11, v11);
  F.at(x1, y1 + 1, u12, v12);
  F.at(x1 + 1, y1, u21, v21);
  F.at(x1 + 1, y1 + 1, u22, v22);
  up = blend(u11, u12, u21, u22, x1, y1, xx, yy);
  vp = blend(v11, v12, v21, v22, x1, y1, xx, yy);
}

// The Maxey–Riley node fields (M:374-411) at the four corners of an interpolation cell
// need the velocity at the two corner rows on four columns (centred x differences) and at
// the rows below/above each corner row on the two corner columns (y differences, one-sided
// at the walls).  Ghost corners take the values of the periodic interior node they copy,
// so all indices are wrapped to interior nodes.  The neighbourhood is loaded once, u and v
// of a node together, and shared by the four corners.  It is held in the field's own
// precision (every value is widened to FP64 exactly where it is used).
template <typename R>
struct CellPatch {
  int I[2], J[2];  // wrapped corner columns / rows (1-based interior indices)
  R c[2][4];       // rows J[b], columns I0-1, I0, I1, I1+1
  R lo[2][2];      // row J[b]-1 (unused at J[b] == 1), columns I0, I1
  R hi[2][2];      // row J[b]+1 (unused at J[b] == ny), columns I0, I1
};

template <typename R>
__device__ __forceinline__ void load_patch(const VelField<R> &F, int x1, int y1, CellPatch<R> &Pu,
                                           CellPatch<R> &Pv) {
  const int nx = F.nx, ny = F.ny;
  Pu.I[0] = x1 == 0 ? nx : x1;
  Pu.I[1] = x1 + 1 == nx + 1 ? 1 : x1 + 1;
  Pu.J[0] = y1 == 0 ? ny : y1;
  Pu.J[1] = y1 + 1 == ny + 1 ? 1 : y1 + 1;
  for (int q = 0; q < 2; q++) {
    Pv.I[q] = Pu.I[q];
    Pv.J[q] = Pu.J[q];
  }
  for (int b = 0; b < 2; b++) {
    const int J = Pu.J[b];
    F.at(Pu.I[0] - 1, J, Pu.c[b][0], Pv.c[b][0]);
    F.at(Pu.I[0], J, Pu.c[b][1], Pv.c[b][1]);
    F.at(Pu.I[1], J, Pu.c[b][2], Pv.c[b][2]);
    F.at(Pu.I[1] + 1, J, Pu.c[b][3], Pv.c[b][3]);
  }
  for (int b = 0; b < 2; b++) {
    const int J = Pu.J[b];
    const int other = 1 - b;
    for (int a = 0; a < 2; a++) {
      // reuse the other corner row when it is the neighbour row
      if (J > 1) {
        if (Pu.J[other] == J - 1) {
          Pu.lo[b][a] = Pu.c[other][1 + a];
          Pv.lo[b][a] = Pv.c[other][1 + a];
        } else {
          F.at(Pu.I[a], J - 1, Pu.lo[b][a], Pv.lo[b][a]);
        }
      } else {
        Pu.lo[b][a] = Pv.lo[b][a] = R(0);
      }
      if (J < ny) {
        if (Pu.J[other] == J + 1) {
          Pu.hi[b][a] = Pu.c[other][1 + a];
          Pv.hi[b][a] = Pv.c[other][1 + a];
        } else {
          F.at(Pu.I[a], J + 1, Pu.hi[b][a], Pv.hi[b][a]);
        }
      } else {
        Pu.hi[b][a] = Pv.hi[b][a] = R(0);
      }
    }
  }
}

// d/dx and d/dy of the field at corner (a, b) of the patch (M:384-405).
template <typename R>
__device__ __forceinline__ void patch_dxdy(const CellPatch<R> &P, int a, int b, int ny, double &dx,
                                           double &dy) {
  const double u0 = (double)P.c[b][1 + a];
  dx = ((double)P.c[b][2 + a] - (double)P.c[b][a]) * 0.5;
  const int J = P.J[b];
  if (J == 1) dy = ((double)P.hi[b][a] - u0);
  else if (J == ny) dy = (u0 - (double)P.lo[b][a]);
  else dy = ((double)P.hi[b][a] - (double)P.lo[b][a]) * 0.5;
}

// U, V and the six Maxey–Riley interpolants at (xx, yy) from one patch load per field.
template <typename R>
__device__ __forceinline__ bool interp_all(const VelField<R> &UV, const VelField<R> &UVp, double xx,
                                           double yy, double &up, double &vp, double D[6]) {
  int x1, y1;
  if (!corners(xx, yy, UV.nx, UV.ny, x1, y1)) return false;
  CellPatch<R> pu, pv;
  load_patch(UV, x1, y1, pu, pv);
  // blend() of all eight node fields, corner by corner in its summation order (x1,y1),
  // (x1+1,y1), (x1,y1+1), (x1+1,y1+1), each term (value * x weight) * y weight: the same
  // operations as eight blend() calls, with eight running sums live instead of 32 corner
  // values.
  const double wx[2] = {(double)(x1 + 1) - xx, xx - (double)x1};
  const double wy[2] = {(double)(y1 + 1) - yy, yy - (double)y1};
  const int ca[4] = {0, 1, 0, 1}, cb[4] = {0, 0, 1, 1};
  double acc[8];  // u, v, dtu, dtv, dxu, dxv, dyu, dyv
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int a = ca[k], b = cb[k];
    const double u0 = (double)pu.c[b][1 + a], v0 = (double)pv.c[b][1 + a];
    double ub, vb;
    UVp.at(pu.I[a], pu.J[b], ub, vb);
    double n[8];
    n[0] = u0;
    n[1] = v0;
    n[2] = u0 - ub;
    n[3] = v0 - vb;
    patch_dxdy(pu, a, b, UV.ny, n[4], n[6]);
    patch_dxdy(pv, a, b, UV.ny, n[5], n[7]);
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const double t = n[q] * wx[a] * wy[b];
      acc[q] = k == 0 ? t : acc[q] + t;
    }
  }
  up = acc[0];
  vp = acc[1];
#pragma unroll
  for (int q = 0; q < 6; q++) D[q] = acc[2 + q];
  return true;
}

// One particle's step (index p < pp.n).
template <typename R>
__device__ __forceinline__ void particle_update(const VelField<R> &UV, const VelField<R> &UVp,
                                                const Field<R> &T, const PartPtrs &pp, int p, int record,
                                                int full) {
  const double delta_t = 1.0;
  int err = 0;
  const double xx = pp.x[p], yy = pp.y[p];
  const double uo = pp.u[p], vo = pp.v[p];
  double pu = uo, pv = vo;
  const double onebytau = pp.inv_tau[p];
  const bool heavy = onebytau > 0.0;
  double D[6] = {0, 0, 0, 0, 0, 0};
  double up = 0.0, vp = 0.0;
  if (heavy || (record && full)) {
    if (!interp_all(UV, UVp, xx, yy, up, vp, D)) err = 1;
  } else {
    interp_uv(UV, xx, yy, up, vp, err);
  }
  double x, y;
  if (heavy) {
    const double beta = pp.beta[p];
    pu = pu + delta_t * (beta * (D[0] + up * D[2] + vp * D[4]) - onebytau * (pu - up));
    pv = pv + delta_t * (beta * (D[1] + up * D[3] + vp * D[5]) - onebytau * (pv - vp));
    x = xx + delta_t * pu;
    y = yy + delta_t * pv;
  } else {
    x = xx + 1.5 * delta_t * up - 0.5 * delta_t * uo;
    y = yy + 1.5 * delta_t * vp - 0.5 * delta_t * vo;
    pu = up;
    pv = vp;
  }
  if (record && pp.rec) {
    const size_t n = (size_t)pp.n;
    double *r = pp.rec;
    r[0 * n + p] = x;
    r[1 * n + p] = y;
    r[2 * n + p] = pu;
    r[3 * n + p] = pv;
    r[4 * n + p] = xx;
    r[5 * n + p] = yy;
    r[6 * n + p] = uo;
    r[7 * n + p] = vo;
    r[8 * n + p] = interp(T, x, y, err);
    if (full) {
      // Writefull quantities at (xo, yo) = (xx, yy): the same interpolants as the update.
#pragma unroll
      for (int q = 0; q < 6; q++) r[(9 + q) * n + p] = D[q];
      r[15 * n + p] = up;
      r[16 * n + p] = vp;
    }
  }
  // toggle_prtcl on one rank (M:598-635, 694-717).
  const double NXd = (double)UV.nx, NYd = (double)UV.ny;
  if (x >= NXd) {
    x = x - NXd;
  } else if (x <= 0.0) {
    x = x + NXd;
  }
  if (y < 1.0) {
    y = y + 2.0 * (1.0 - y);
    pv = -pv;
  } else if (y >= NYd) {
    y = y + 2.0 * (NYd - y);
    pv = -pv;
  }
  pp.x[p] = x;
  pp.y[p] = y;
  pp.u[p] = pu;
  pp.v[p] = pv;
  pp.xo[p] = xx;
  pp.yo[p] = yy;
  pp.uo[p] = uo;
  pp.vo[p] = vo;
  if (err) atomicOr(pp.err, 1);
}

// The particle updates a lattice launch carries (PartJob), as kernel arguments: nblocks
// workgroups of the launch, one thread per particle, run them in step order.
template <typename R>
struct PartSteps {
  VelField<R> uv[kMaxCarried], uvp[kMaxCarried];
  Field<R> T;
  PartPtrs pp;
  double *rec[kMaxCarried];
  int record[kMaxCarried], full[kMaxCarried];
  int n, nblocks;
};

template <typename R>
inline PartSteps<R> part_steps(const Geom &g, const PartJob *job, int bw) {
  PartSteps<R> a{};
  if (!job || job->n <= 0 || job->pp.n <= 0) return a;
  for (int u = 0; u < job->n; u++) {
    a.uv[u] = vel_of<R>(job->uv[u], g.nx, g.ny, g.pitch);
    a.uvp[u] = vel_of<R>(job->uvp[u], g.nx, g.ny, g.pitch);
    a.rec[u] = job->rec[u];
    a.record[u] = job->record[u];
    a.full[u] = job->full[u];
  }
  a.T = scalar_of<R>(job->T, g.nx, g.ny, g.pitch);
  a.pp = job->pp;
  a.n = job->n;
  a.nblocks = (job->pp.n + bw - 1) / bw;
  return a;
}

template <typename R>
__device__ __forceinline__ void run_part_steps(const PartSteps<R> &a, int q) {
  if (q >= a.pp.n) return;
  PartPtrs pp = a.pp;
  for (int u = 0; u < a.n; u++) {
    pp.rec = a.rec[u];
    particle_update(a.uv[u], a.uvp[u], a.T, pp, q, a.record[u], a.full[u]);
  }
}

}  // namespace part
}  // namespace rbn
