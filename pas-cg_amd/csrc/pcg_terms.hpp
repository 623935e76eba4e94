// pcg_terms.hpp -- per-item physics of every force term, as device functions.
//
// Each function evaluates one list item of one term: it reads positions
// through a geometry accessor G (global arrays for the per-term kernels, LDS
// for the persistent batched kernel), pushes forces through an accumulator A
// (global or LDS atomics) and returns the item's energy in FP64. R is the
// force arithmetic type (float in the mixed mode, double in FP64 mode);
// coordinate differences are formed in FP64 and then rounded to R.
// Decisions that must agree with the CPU path bit for bit (cutoffs of the
// contact-state transitions, QA criteria, saturation) use the exact FP64
// helpers of pcg_core.hpp.
#pragma once
#include "pcg_core.hpp"

namespace pcg {

template <typename R> __device__ __forceinline__ v3<R> cast3(v3<double> d) {
  return v3<R>{(R)d.x, (R)d.y, (R)d.z};
}

// Without a periodic cell (all sides 0) both wraps are the identity, bit for
// bit (v - 0 * round(v * 0) = v); the test is uniform, so a system without
// one skips the FP64 rounding work.
__device__ __forceinline__ bool has_cell(SysDescC &s) {
  return s.cell[0] != 0.0 || s.cell[1] != 0.0 || s.cell[2] != 0.0;
}
__device__ __forceinline__ v3<double> wrapd(v3<double> d, SysDescC &s) {
  if (!has_cell(s)) return d;
  d.x -= s.cell[0] * round(d.x * s.cell_inv[0]);
  d.y -= s.cell[1] * round(d.y * s.cell_inv[1]);
  d.z -= s.cell[2] * round(d.z * s.cell_inv[2]);
  return d;
}
__device__ __forceinline__ v3<double> xwrapd(v3<double> d, SysDescC &s) {
  if (!has_cell(s)) return d;
  return v3<double>{xwrap1(d.x, s.cell[0], s.cell_inv[0]), xwrap1(d.y, s.cell[1], s.cell_inv[1]),
                    xwrap1(d.z, s.cell[2], s.cell_inv[2])};
}
__device__ __forceinline__ v3<double> xsub(v3<double> a, v3<double> b) {
  return v3<double>{__dsub_rn(a.x, b.x), __dsub_rn(a.y, b.y), __dsub_rn(a.z, b.z)};
}
__device__ __forceinline__ double xnorm3(v3<double> d) { return xnorm(d.x, d.y, d.z); }

// Base potentials (cg/include/cg/base_forces/{harmonic,lj,shifted_lj,sink_lj}.h)
template <typename R> __device__ __forceinline__ void harmonicR(R H1, R H2, R nat, R r, R &V, R &dV) {
  R dr = r - nat, dr2 = dr * dr;
  V = (R)0.5 * dr2 * (H1 + H2 * dr2);
  dV = clampR(dr * (H1 + (R)2 * H2 * dr2), (R)-1e3, (R)1e3);
}
template <typename R> __device__ __forceinline__ void ljR(R depth, R r_min, R r_inv, R &V, R &dV) {
  R s = r_inv * r_min, s6 = ipow6(s), s12 = s6 * s6;
  V = depth * (s12 - (R)2 * s6);
  dV = (R)12 * depth * r_inv * (s6 - s12);
}
template <typename R> __device__ __forceinline__ void shifted_ljR(R depth, R r_min, R r_inv, R &V, R &dV) {
  R s = r_inv * r_min, s6 = ipow6(s), s12 = s6 * s6;
  V = depth * (s12 - (R)2 * s6 + (R)1);
  dV = clampR((R)12 * depth * r_inv * (s6 - s12), (R)-1e3, (R)1e3);
}
template <typename R> __device__ __forceinline__ void sink_ljR(R depth, R r_low, R r_high, R r, R r_inv, R &V, R &dV) {
  R r_eff = r < r_low ? r_low : (r < r_high ? r : r_high);
  R s = r_inv * r_eff, s6 = ipow6(s), s12 = s6 * s6;
  V = depth * (s12 - (R)2 * s6);
  dV = (R)12 * depth * r_inv * (s6 - s12);
}

// ---------------------------------------------------------------- terms ----

// tether::eval_forces::iter (cg/src/cg/tether/eval_forces.cpp:18-33)
// (each *_core takes the item's residues and parameters; *_item reads them
// from the term's list)
template <typename R, typename G, typename A>
__device__ __forceinline__ double tether_core(ParamsC &p, const G &g, A &acc, int i1, int i2,
                                              double nat) {
  v3<R> r12 = cast3<R>(g.r(i2) - g.r(i1));
  const R r2 = nsq(r12), rinv = rsq(r2);
  R rn = r2 * rinv, V, dV;
  harmonicR<R>((R)p.tether_H1, (R)p.tether_H2, (R)nat, rn, V, dV);
  v3<R> f = (dV * rinv) * r12;
  acc.add(i1, f);
  acc.add(i2, -f);
  return (double)V;
}
template <typename R, typename G, typename A>
__device__ __forceinline__ double tether_item(const DevBatch &b, const G &g, A &acc, int k) {
  return tether_core<R>(b.P->p, g, acc, b.teth_i[0][k], b.teth_i[1][k], b.teth_nat[k]);
}

// local_rep::eval_forces::iter (cg/src/cg/local_rep/eval_forces.cpp:11-28)
template <typename R, typename G, typename A>
__device__ __forceinline__ double lrep_core(ParamsC &p, const G &g, A &acc, int i1, int i2) {
  v3<R> r12 = cast3<R>(g.r(i2) - g.r(i1));
  R rn = rsq(nsq(r12)), cut = (R)p.rep_cutoff;
  if ((R)1 > rn * cut) return 0.0;
  R V, dV;
  shifted_ljR<R>((R)p.lrep_depth, cut, rn, V, dV);
  v3<R> f = (dV * rn) * r12;
  acc.add(i1, f);
  acc.add(i2, -f);
  return (double)V;
}

template <typename R, typename G, typename A>
__device__ __forceinline__ double lrep_item(const DevBatch &b, const G &g, A &acc, int k) {
  return lrep_core<R>(b.P->p, g, acc, b.lrep_i[0][k], b.lrep_i[1][k]);
}

// pauli::eval_forces::iter (cg/src/cg/pauli/eval_forces.cpp:12-30)
template <typename R, typename G, typename A>
__device__ double pauli_item(const DevBatch &b, SysDescC &s, const G &g, A &acc, int k) {
  ParamsC &p = b.P->p;
  int i1 = b.pauli_i[0][k], i2 = b.pauli_i[1][k];
  v3<R> r12 = cast3<R>(wrapd(g.r(i2) - g.r(i1), s));
  R rn = rsq(nsq(r12)), cut = (R)p.rep_cutoff;
  if ((R)1 > rn * cut) return 0.0;
  R V, dV;
  shifted_ljR<R>((R)p.pauli_depth, cut, rn, V, dV);
  v3<R> f = (dV * rn) * r12;
  acc.add(i1, f);
  acc.add(i2, -f);
  return (double)V;
}

// nat_cont::eval_forces::iter (cg/src/cg/nat_cont/eval_forces.cpp:11-50).
// c indexes the list of all native contacts (where active/change_t live);
// the state transition is made only when update_state is set, so that a
// contact evaluated once per residue changes state once.
struct NcFetch {
  int k, c, i1, i2;
  double nat;
  int type;  // contact type | active << 7
  float th2;
};
template <typename R, typename G, typename A>
__device__ __forceinline__ double nat_cont_core(const DevBatch &b, SysDescC &s, SysState &st,
                                                const G &g, A &acc, const NcFetch &q, double t,
                                                bool update_state) {
#pragma clang fp contract(off)
  ParamsC &p = b.P->p;
  const int c = q.c, i1 = q.i1, i2 = q.i2;
  const double nat = q.nat;
  v3<double> d = xwrapd(xsub(g.r(i2), g.r(i1)), s);
  v3<R> r12 = cast3<R>(d);
  R rn, V, dV;
  double rnx = 0.0;
  int now = -1;  // transition decision, -1 while undecided
  if constexpr (sizeof(R) == 4) {
    // both decisions (cutoff, transition) from the FP32 square when it is
    // clear of the thresholds by 2e-5; the exact FP64 norm otherwise
    const float d2 = nsq(r12), cut2 = (float)(p.nl_cutoff * p.nl_cutoff);
    const float lo = 1.0f - 2e-5f, hi = 1.0f + 2e-5f;
    if (d2 > cut2 * hi) return 0.0;
    if (d2 < cut2 * lo && (d2 < q.th2 * lo || d2 > q.th2 * hi)) {
      now = d2 < q.th2 * lo ? 1 : 0;
      rn = sq(d2);
    }
  }
  if (now < 0) {
    rnx = xnorm3(d);
    if (rnx >= p.nl_cutoff) return 0.0;
    rn = (R)rnx;
  }
  if (p.nc_ss_kind != 0 && (q.type & 0x7F) == 5) {
    // disulfide_force::operator() receives the distance as its inverse-distance
    // argument (nat_cont/eval_forces.cpp:27, disulfide.h:12-20)
    if (p.nc_ss_kind == 1)
      harmonicR<R>((R)p.nc_ss_H1, (R)p.nc_ss_H2, (R)p.nc_ss_nat_r, rcpR(rn), V, dV);
    else
      ljR<R>((R)p.nc_ss_depth, (R)p.nc_ss_r_min, rn, V, dV);
  } else {
    ljR<R>((R)p.nat_cont_depth, (R)nat, rcpR(rn), V, dV);
  }
  dV = clampR(dV, (R)-1e3, (R)1e3);
  v3<R> f = (dV * rcpR(rn)) * r12;
  acc.add(i1, f);
  acc.add(i2, -f);
  if (!update_state) return (double)V;
  // state transition, exact FP64 unless decided above
  if (now < 0) now = rnx <= __dmul_rn(__dmul_rn(nat, C216_INV), p.counting_factor) ? 1 : 0;
  if (now != (q.type >> 7)) {
    b.nc_active[c] = (uint8_t)now;
    b.ncur_t[q.k] = (uint8_t)((q.type & 0x7F) | (now << 7));
    b.ncur_ijt[q.k] = (b.ncur_ijt[q.k] & 0x7FFFFFFFu) | ((uint32_t)now << 31);
    if (b.nc_change_t[c] < 0) {
      b.nc_change_t[c] = t;
      atomicAdd(&st.num_changed, 1);
    }
  }
  return (double)V;
}

// Item k of the current list: it carries the contact's index into the list
// of all contacts and a copy of its residues, length and type (independent
// loads, issued together).
__device__ __forceinline__ NcFetch nat_cont_fetch(const DevBatch &b, int k) {
  return NcFetch{k, b.ncur[k], b.ncur_i[0][k], b.ncur_i[1][k], b.ncur_nat[k], b.ncur_t[k],
                 b.ncur_th2[k]};
}
// every field in registers before the item's first decision (see keep)
__device__ __forceinline__ void keep_fetch(NcFetch &q) {
  keep(q.c);
  keep(q.i1);
  keep(q.i2);
  keep(q.nat);
  keep(q.type);
  keep(q.th2);
}
template <typename R, typename G, typename A>
__device__ __forceinline__ double nat_cont_item(const DevBatch &b, SysDescC &s, SysState &st,
                                                const G &g, A &acc, int k, double t) {
  NcFetch q = nat_cont_fetch(b, k);
  keep_fetch(q);
  return nat_cont_core<R>(b, s, st, g, acc, q, t, true);
}

// const_dh / rel_dh ::iter (cg/src/cg/dh/const.cpp:18-40, rel.cpp:17-40).
// With want_v the FP32 form returns the pair's energy in FP64, in the
// reference's order of operations: the charges' signs make the term's sum
// cancel, so FP32 pair energies would not hold its relative error to 1e-6.
template <typename R, typename G, typename A>
__device__ double dh_item(const DevBatch &b, SysDescC &s, const G &g, A &acc, int k, int variant,
                          bool want_v = true) {
  ParamsC &p = b.P->p;
  int i1 = b.dh_i[0][k], i2 = b.dh_i[1][k];
  double q1q2 = b.dh_q[k];
  keep(i1);
  keep(i2);
  keep(q1q2);
  v3<double> d = wrapd(g.r(i2) - g.r(i1), s);
  double rnd = sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
  if (rnd > p.dh_cutoff) return 0.0;
  v3<R> r12 = cast3<R>(d);
  R rn = (R)rnd, rinv = rcpR(rn), sinv = (R)b.P->screen_inv;
  R Vij, dV;
  if (variant == 0) {
    Vij = (R)b.P->V_factor * (R)q1q2 * expR(-rn * sinv) * rinv;
    dV = -Vij * (sinv + rinv);
  } else {
    Vij = (R)b.P->V_factor * (R)q1q2 * expR(-rn * sinv) * rinv * rinv;
    dV = -Vij * (sinv + (R)2 * rinv);
  }
  v3<R> f = (dV * rinv) * r12;
  acc.add(i1, f);
  acc.add(i2, -f);
  if constexpr (sizeof(R) == 4) {
    if (want_v) {
      const double rn64 = 1.0 / rnd;
      double V64 = b.P->V_factor * q1q2 * exp(-rnd * b.P->screen_inv) * rn64;
      if (variant != 0) V64 = V64 * rn64;
      return V64;
    }
  }
  return (double)Vij;
}

// chir::eval_forces::iter (cg/src/cg/chir/eval_forces.cpp:10-31); f0/f1 are
// the stored nat_factor/nat_chir fields of the quad.
template <typename R, typename G, typename A>
__device__ __forceinline__ double chir_core(ParamsC &p, const G &g, A &acc, int i1, int i2,
                                            int i3, int i4, double f0d, double f1d) {
  v3<double> r1 = g.r(i1), r2 = g.r(i2), r3 = g.r(i3), r4 = g.r(i4);
  v3<R> r12 = cast3<R>(r2 - r1), r23 = cast3<R>(r3 - r2), r34 = cast3<R>(r4 - r3);
  v3<R> x12_23 = cross(r12, r23), x12_34 = cross(r12, r34), x23_34 = cross(r23, r34);
  R f0 = (R)f0d, f1 = (R)f1d, e = (R)p.e_chi;
  R cd = dot(r12, x23_34) * f0 - f1;
  R f = e * cd * f0;
  acc.add(i1, f * x12_23);
  acc.add(i2, -(f * (x12_34 + x23_34)));
  acc.add(i3, f * (x12_23 + x12_34));
  acc.add(i4, -(f * x12_23));
  return (double)((R)0.5 * e * cd * cd);
}
template <typename R, typename G, typename A>
__device__ __forceinline__ double chir_item(const DevBatch &b, const G &g, A &acc, int k) {
  return chir_core<R>(b.P->p, g, acc, b.chir_i[0][k], b.chir_i[1][k], b.chir_i[2][k],
                      b.chir_i[3][k], b.chir_f0[k], b.chir_f1[k]);
}

// angle geometry of nat_ang / heur_ang (angles/*/eval_forces.cpp:10-38)
template <typename R, typename G>
__device__ __forceinline__ R angle_geom(const G &g, int i1, int i2, int i3, v3<R> &d1, v3<R> &d2, v3<R> &d3) {
  v3<R> r12 = cast3<R>(g.r(i2) - g.r(i1)), r23 = cast3<R>(g.r(i3) - g.r(i2));
  v3<R> x = cross(r12, r23);
  R r12_rn = rsq(nsq(r12)), r23_rn = rsq(nsq(r23)), x_rn = rsq(nsq(x));
  v3<R> c1 = cross(r12, x), c3 = cross(r23, x);
  // |r12 x x| = |r12| |x| and |r23 x x| = |r23| |x|, since x is normal to both
  d1 = (x_rn * r12_rn * r12_rn) * c1;
  d3 = (x_rn * r23_rn * r23_rn) * c3;
  d2 = -d1 - d3;
  R cos_t = clampR(-dot(r12, r23) * r12_rn * r23_rn, (R)-1, (R)1);
  if (cos_t == (R)0) d1 = d2 = d3 = v3<R>{0, 0, 0};
  return acosc(cos_t);
}

template <typename R, typename G, typename A>
__device__ __forceinline__ double nat_ang_core(ParamsC &p, const G &g, A &acc, int i1, int i2,
                                               int i3, double th0) {
  v3<R> d1, d2, d3;
  R th = angle_geom<R>(g, i1, i2, i3, d1, d2, d3);
  R CBA = (R)p.CBA, dth = th - (R)th0, dV = (R)2 * CBA * dth;
  acc.add(i1, -(dV * d1));
  acc.add(i2, -(dV * d2));
  acc.add(i3, -(dV * d3));
  return (double)(CBA * dth * dth);
}
template <typename R, typename G, typename A>
__device__ __forceinline__ double nat_ang_item(const DevBatch &b, const G &g, A &acc, int k) {
  return nat_ang_core<R>(b.P->p, g, acc, b.nang_i[0][k], b.nang_i[1][k], b.nang_i[2][k], b.nang_th[k]);
}

template <typename R, typename G, typename A>
__device__ __forceinline__ double heur_ang_core(ParamsC &p, const G &g, A &acc, int i1, int i2,
                                                int i3, int ty) {
  v3<R> d1, d2, d3;
  R th = angle_geom<R>(g, i1, i2, i3, d1, d2, d3);
  const double(*c)[9] = p.heur_ang_poly;
  // the coefficients (~1e5) cancel to O(1): evaluated as the reference writes
  // it, in the precision R (callers use FP64)
  R t2 = th * th, t3 = t2 * th, t4 = t3 * th, t5 = t4 * th, t6 = t3 * t3;
  R V = (R)c[0][ty] + (R)c[1][ty] * th + (R)c[2][ty] * t2 + (R)c[3][ty] * t3 +
        (R)c[4][ty] * t4 + (R)c[5][ty] * t5 + (R)c[6][ty] * t6;
  R dV = (R)c[1][ty] + (R)(2.0 * c[2][ty]) * th + (R)(3.0 * c[3][ty]) * t2 +
         (R)(4.0 * c[4][ty]) * t3 + (R)(5.0 * c[5][ty]) * t4 + (R)(6.0 * c[6][ty]) * t5;
  acc.add(i1, -(dV * d1));
  acc.add(i2, -(dV * d2));
  acc.add(i3, -(dV * d3));
  return (double)V;
}
template <typename R, typename G, typename A>
__device__ __forceinline__ double heur_ang_item(const DevBatch &b, const G &g, A &acc, int k) {
  return heur_ang_core<R>(b.P->p, g, acc, b.hang_i[0][k], b.hang_i[1][k], b.hang_i[2][k], b.hang_t[k]);
}

// Angles of the batched mixed form: FP32 geometry from the FP64 differences,
// and theta = atan2(|r12 x r23|, -r12.r23), which is well conditioned at
// every angle (an acos of a rounded cosine is not, near 0 and pi): theta
// carries about 1e-7 rad of absolute error. Near-collinear triples
// (sin theta < 0.05), where the FP32 directions of the derivatives lose
// accuracy, return false and are evaluated in FP64 by the caller.
template <typename G>
__device__ __forceinline__ bool angle_geom_f32(const G &g, int i1, int i2, int i3, v3<float> &d1,
                                               v3<float> &d2, v3<float> &d3, float &th) {
  v3<float> r12 = cast3<float>(g.r(i2) - g.r(i1)), r23 = cast3<float>(g.r(i3) - g.r(i2));
  v3<float> x = cross(r12, r23);
  const float a2 = nsq(r12), b2 = nsq(r23), x2 = nsq(x);
  if (!(x2 > 0.0025f * a2 * b2)) return false;
  const float ar = rsq(a2), br = rsq(b2), xr = rsq(x2);
  v3<float> c1 = cross(r12, x), c3 = cross(r23, x);
  d1 = (xr * ar * ar) * c1;
  d3 = (xr * br * br) * c3;
  d2 = -d1 - d3;
  const float dt = dot(r12, r23);
  if (dt == 0.0f) d1 = d2 = d3 = v3<float>{0.f, 0.f, 0.f};  // the cos == 0 case of angle_geom
  th = atan2f(x2 * xr, -dt);
  return true;
}
template <typename G, typename A>
__device__ __forceinline__ double nat_ang_mixed(ParamsC &p, const G &g, A &acc, int i1, int i2,
                                                int i3, double th0) {
  v3<float> d1, d2, d3;
  float th;
  if (!angle_geom_f32(g, i1, i2, i3, d1, d2, d3, th)) return nat_ang_core<double>(p, g, acc, i1, i2, i3, th0);
  const double dth = (double)th - th0;
  const float dV = (float)(2.0 * p.CBA * dth);
  acc.add(i1, -(dV * d1));
  acc.add(i2, -(dV * d2));
  acc.add(i3, -(dV * d3));
  return p.CBA * dth * dth;
}
// the polynomial cancels heavily: FP64 in theta
template <typename G, typename A>
__device__ __forceinline__ double heur_ang_mixed(ParamsC &p, const G &g, A &acc, int i1, int i2,
                                                 int i3, int ty) {
  v3<float> d1, d2, d3;
  float thf;
  if (!angle_geom_f32(g, i1, i2, i3, d1, d2, d3, thf)) return heur_ang_core<double>(p, g, acc, i1, i2, i3, ty);
  const double(*c)[9] = p.heur_ang_poly, th = thf;
  double t2 = th * th, t3 = t2 * th, t4 = t3 * th, t5 = t4 * th, t6 = t3 * t3;
  double V = c[0][ty] + c[1][ty] * th + c[2][ty] * t2 + c[3][ty] * t3 + c[4][ty] * t4 +
             c[5][ty] * t5 + c[6][ty] * t6;
  float dV = (float)(c[1][ty] + (2.0 * c[2][ty]) * th + (3.0 * c[3][ty]) * t2 + (4.0 * c[4][ty]) * t3 +
                     (5.0 * c[5][ty]) * t4 + (6.0 * c[6][ty]) * t5);
  acc.add(i1, -(dV * d1));
  acc.add(i2, -(dV * d2));
  acc.add(i3, -(dV * d3));
  return V;
}

// cnd / snd ::iter (angles/nat_dih/complex.cpp:10-47, simple.cpp:10-42)
//
// The complex variant needs only cos and sin of phi - phi0. They come from
// cos phi = x1u.x2u and sin phi = |r23| (x1.r34) / (|x1| |x2|), since
// x1 x x2 = r23 (x1.r34); the sign of x1.r34 is the sign the reference gives
// phi after its acos. cos phi0 and sin phi0 are the caller's (precomputed per
// item), so the term takes no acos and no sincos.
template <typename R, typename G, typename A>
__device__ __forceinline__ double nat_dih_core(ParamsC &p, const G &g, A &acc, int i1, int i2,
                                               int i3, int i4, double phi0, double cphi0,
                                               double sphi0, int variant) {
  v3<double> q1 = g.r(i1), q2 = g.r(i2), q3 = g.r(i3), q4 = g.r(i4);
  v3<R> r12 = cast3<R>(q2 - q1), r23 = cast3<R>(q3 - q2), r34 = cast3<R>(q4 - q3);
  v3<R> x1 = cross(r12, r23), x2 = cross(r23, r34);
  R x1rn = rsq(nsq(x1)), x2rn = rsq(nsq(x2));
  v3<R> x1u = x1 * x1rn, x2u = x2 * x2rn;
  const R r23sq = nsq(r23), r23rn = rsq(r23sq);
  R r23n = r23sq * r23rn, V, dVdphi;
  if (variant == 0) {
    // cos 3x = c (4c^2 - 3), sin 3x = s (3 - 4s^2)
    R cphi = clampR(dot(x1u, x2u), (R)-1, (R)1), sphi = r23n * dot(x1, r34) * x1rn * x2rn;
    R c1 = cphi * (R)cphi0 + sphi * (R)sphi0, s1 = sphi * (R)cphi0 - cphi * (R)sphi0;
    R CDA = (R)p.CDA, CDB = (R)p.CDB;
    R c3 = c1 * ((R)4 * c1 * c1 - (R)3), s3 = s1 * ((R)3 - (R)4 * s1 * s1);
    V = CDA * ((R)1 - c1) + CDB * ((R)1 - c3);
    dVdphi = CDA * s1 + CDB * (R)3 * s3;
  } else {
    R phi = acosc(dot(x1u, x2u));
    if (dot(x1, r34) < (R)0) phi = -phi;
    R diff = phi - (R)phi0;
    V = (R)0.5 * (R)p.CDH * diff * diff;
    dVdphi = (R)p.CDH * diff;
  }
  v3<R> dr1 = -(x1u * (r23n * x1rn)), dr4 = x2u * (r23n * x2rn);
  v3<R> df = (-dr1 * dot(r12, r23) + dr4 * dot(r23, r34)) * (r23rn * r23rn);
  v3<R> dr2 = -dr1 + df, dr3 = -dr4 - df;
  acc.add(i1, -(dVdphi * dr1));
  acc.add(i2, -(dVdphi * dr2));
  acc.add(i3, -(dVdphi * dr3));
  acc.add(i4, -(dVdphi * dr4));
  return (double)V;
}
template <typename R, typename G, typename A>
__device__ __forceinline__ double nat_dih_item(const DevBatch &b, const G &g, A &acc, int k, int variant) {
  const double phi0 = b.ndih_ph[k];
  double s0 = 0.0, c0 = 1.0;
  if (variant == 0) sincos(phi0, &s0, &c0);
  return nat_dih_core<R>(b.P->p, g, acc, b.ndih_i[0][k], b.ndih_i[1][k], b.ndih_i[2][k],
                         b.ndih_i[3][k], phi0, c0, s0, variant);
}

// heur_dih::eval_forces::iter (angles/heur_dih/eval_forces.cpp:10-55)
template <typename R, typename G, typename A>
__device__ __forceinline__ double heur_dih_core(ParamsC &p, const G &g, A &acc, int i1, int i2,
                                                int i3, int i4, int ty) {
  const double(*c)[9] = p.heur_dih_coeffs;
  v3<double> q1 = g.r(i1), q2 = g.r(i2), q3 = g.r(i3), q4 = g.r(i4);
  v3<R> r12 = cast3<R>(q2 - q1), r23 = cast3<R>(q3 - q2), r34 = cast3<R>(q4 - q3);
  v3<R> x1 = cross(r23, r12), x2 = cross(r34, r23);
  R x1rn = rsq(nsq(x1)), x2rn = rsq(nsq(x2));
  v3<R> x1u = x1 * x1rn, x2u = x2 * x2rn;
  // sin phi without acos: x1 x x2 = r23 ((r12 x r23).r34), and the sign the
  // reference gives phi (that of -r12.x2u) is the sign of (r12 x r23).r34
  R cphi = dot(x1u, x2u);
  const R r23sq = nsq(r23), r23rn = rsq(r23sq);
  R sphi = -(r23sq * r23rn * dot(x1, r34) * x1rn * x2rn);
  R s2 = sphi * sphi, c2 = cphi * cphi, sc = sphi * cphi;
  R V = (R)c[0][ty] + (R)c[1][ty] * sphi + (R)c[2][ty] * cphi + (R)c[3][ty] * s2 +
        (R)c[4][ty] * c2 + (R)c[5][ty] * sc;
  R dV = (R)c[1][ty] * cphi - (R)c[2][ty] * sphi + (R)2 * ((R)c[3][ty] - (R)c[4][ty]) * sc +
         (R)c[5][ty] * (c2 - s2);
  R r23n = r23sq * r23rn;
  v3<R> dr1 = x1u * (r23n * x1rn), dr4 = -(x2u * (r23n * x2rn));
  v3<R> df = (dr1 * (-dot(r12, r23)) - dr4 * (-dot(r34, r23))) * (r23rn * r23rn);
  v3<R> dr2 = -dr1 + df, dr3 = -dr4 - df;
  acc.add(i1, -(dV * dr1));
  acc.add(i2, -(dV * dr2));
  acc.add(i3, -(dV * dr3));
  acc.add(i4, -(dV * dr4));
  return (double)V;
}
template <typename R, typename G, typename A>
__device__ __forceinline__ double heur_dih_item(const DevBatch &b, const G &g, A &acc, int k) {
  return heur_dih_core<R>(b.P->p, g, acc, b.hdih_i[0][k], b.hdih_i[1][k], b.hdih_i[2][k],
                          b.hdih_i[3][k], b.hdih_t[k]);
}

// ---------------------------------------------------------------- PID ----
// pid::eval_forces, vectorised form (cg/src/cg/pid/eval_forces.cpp:337-538),
// lambda of cg/include/cg/pid/lambda.h:58-85. The i+-1 neighbours are raw
// index neighbours (eval_forces.cpp:45) bounded by the system's own residues.
template <typename R> __device__ __forceinline__ bool lam_supp(R alpha, R psi0, R psi) {
  return fabs((double)(alpha * (psi - psi0))) < M_PI;
}
template <typename R> __device__ __forceinline__ void lam_eval(int variant, R alpha, R psi0, R psi, R &val, R &der) {
  if (variant == 0) {
    R s = alpha * (psi - psi0);
    if (fabs((double)s) >= M_PI) { val = der = 0; return; }
    val = (R)0.5 * cosR(s) + (R)0.5;
    der = -(R)0.5 * alpha * sinR(s);
  } else {
    R t = alpha * (psi - psi0) * (R)M_1_PI, a = t < 0 ? -t : t;
    if ((double)a >= M_PI) { val = der = 0; return; }
    R x = (R)2 * a * (a - (R)1);
    val = (R)1 - (a * a) / (x + (R)1);
    der = alpha * (R)M_1_PI * (x / ((x + (R)1) * (x + (R)1)));
    if (t < (R)0) der = -der;
  }
}
template <typename R, typename G>
__device__ __forceinline__ bool pid_psi(const G &g, int i, int j, bool vp, bool vn, R &psi,
                                        v3<R> &d_self, v3<R> &d_next, v3<R> &d_prev, v3<R> &d_other) {
  v3<double> z{0, 0, 0};
  v3<double> rn_ = vn ? g.r(i + 1) : z, rp_ = vp ? g.r(i - 1) : z;
  v3<R> rij = cast3<R>(g.r(i) - rn_);
  v3<R> rkj = vp ? cast3<R>(rp_ - rn_) : v3<R>{0, 0, 0};
  v3<R> rkl = cast3<R>(rp_ - g.r(j));
  v3<R> rm = cross(rij, rkj), rn = cross(rkj, rkl);
  R rm_n = sq(nsq(rm)), rn_n = sq(nsq(rn));
  bool ok = !(rm_n < (R)0.1 || rn_n < (R)0.1);
  R rm_i = rcpR(rm_n), rn_i = rcpR(rn_n), rkj_i = rsq(nsq(rkj)), rkj_n = rcpR(rkj_i);
  v3<R> fi = rm * (rkj_n * rm_i * rm_i);
  v3<R> fl = -(rn * (rkj_n * rn_i * rn_i));
  v3<R> df = (fi * dot(rij, rkj) - fl * dot(rkl, rkj)) * (rkj_i * rkj_i);
  d_self = fi;
  d_next = -fi + df;
  d_prev = -fl - df;
  d_other = fl;
  R c = clampR(dot(rm, rn) * rm_i * rn_i, (R)-1, (R)1);
  psi = acosc(c);
  if (dot(rij, rn) < (R)0) psi = -psi;
  return ok;
}

template <typename R, typename G, typename A>
__device__ double pid_item(const DevBatch &b, SysDescC &s, const G &g, A &acc, int k) {
  ParamsC &p = b.P->p;
  int i1 = b.pid_i[0][k], i2 = b.pid_i[1][k];
  if (b.atype[i1] == 14 || b.atype[i2] == 14) return 0.0; // PRO
  v3<R> r12 = cast3<R>(wrapd(g.r(i2) - g.r(i1), s));
  R r12n = sq(nsq(r12)), r12rn = rcpR(r12n);
  if (!(r12n < (R)p.nl_cutoff)) return 0.0;
  int lo = s.res0, hi = s.res0 + s.n;
  bool v1p = i1 - 1 >= lo, v1n = i1 + 1 < hi, v2p = i2 - 1 >= lo, v2n = i2 + 1 < hi;
  R psi1, psi2;
  v3<R> a1, a1n, a1p, a2o, c2, c2n, c2p, c1o;
  if (!pid_psi<R>(g, i1, i2, v1p, v1n, psi1, a1, a1n, a1p, a2o)) return 0.0;
  if (!pid_psi<R>(g, i2, i1, v2p, v2n, psi2, c2, c2n, c2p, c1o)) return 0.0;
  bool sep3 = (i2 - i1 == 3);
  R bpa = (R)p.pid_bbp_alpha, bpp = (R)p.pid_bbp_psi0, bma = (R)p.pid_bbm_alpha, bmp = (R)p.pid_bbm_psi0;
  bool opt1 = lam_supp(bpa, bpp, psi1) || sep3;
  bool opt2 = !sep3 && lam_supp(bpa, bpp, psi2);
  R l1a = opt1 ? bpa : bma, l1p = opt1 ? bpp : bmp, l2a = opt2 ? bpa : bma, l2p = opt2 ? bpp : bmp;
  const double *bblj = opt2 ? p.pid_bbp_lj : p.pid_bbm_lj;
  int var = p.pid_lambda_variant;
  R dV1 = 0, dV2 = 0, dVr = 0, V = 0;
  if (lam_supp(l1a, l1p, psi1) && lam_supp(l2a, l2p, psi2)) {
    R lam1, der1, lam2, der2;
    lam_eval<R>(var, l1a, l1p, psi1, lam1, der1);
    lam_eval<R>(var, l2a, l2p, psi2, lam2, der2);
    if (lam1 * lam2 > (R)5e-5) {
      R lV, ldV;
      sink_ljR<R>((R)bblj[0], (R)bblj[1], (R)bblj[2], r12n, r12rn, lV, ldV);
      V += lam1 * lam2 * lV;
      dV1 += der1 * lam2 * lV;
      dV2 += der2 * lam1 * lV;
      dVr += lam1 * lam2 * ldV;
    }
  }
  int ty = b.pid_t[k];
  R sa = (R)p.pid_ss_alpha, sp = (R)p.pid_ss_psi0;
  if (lam_supp(sa, sp, psi1) && lam_supp(sa, sp, psi2)) {
    R lam1, der1, lam2, der2;
    lam_eval<R>(var, sa, sp, psi1, lam1, der1);
    lam_eval<R>(var, sa, sp, psi2, lam2, der2);
    R depth = (R)p.pid_ss_depth[ty];
    if (lam1 * lam2 > (R)5e-5 && depth > (R)5e-5) {
      R lV, ldV;
      sink_ljR<R>(depth, (R)p.pid_ss_r_low[ty], (R)p.pid_ss_r_high[ty], r12n, r12rn, lV, ldV);
      V += lam1 * lam2 * lV;
      dV1 += der1 * lam2 * lV;
      dV2 += der2 * lam1 * lV;
      dVr += lam1 * lam2 * ldV;
    }
  }
  v3<R> u = r12 * r12rn;
  if (v1p) acc.add(i1 - 1, -(dV1 * a1p));
  acc.add(i1, -(dV1 * a1 + dV2 * c1o - dVr * u));
  if (v1n) acc.add(i1 + 1, -(dV1 * a1n));
  if (v2p) acc.add(i2 - 1, -(dV2 * c2p));
  acc.add(i2, -(dV1 * a2o + dV2 * c2 + dVr * u));
  if (v2n) acc.add(i2 + 1, -(dV2 * c2n));
  return (double)V;
}

// ----------------------------------------------------------------- AFM ----
// afm::vel::eval_forces::iter (cg/src/cg/afm/vel/eval_forces.cpp:10-31); t is
// the time since the tips were placed (t - ref_t), now the current time.
// One thread takes a tip, so its moving averages have one writer.
template <typename R, typename G, typename A>
__device__ double vel_afm_item(const DevBatch &b, const G &g, A &acc, int k, double t,
                               double now) {
#pragma clang fp contract(off)
  ParamsC &p = b.P->p;
  int res = b.tip_res[k];
  v3<double> cur{b.tip_orig[0][k] + t * b.tip_vel[0][k], b.tip_orig[1][k] + t * b.tip_vel[1][k],
                 b.tip_orig[2][k] + t * b.tip_vel[2][k]};
  v3<R> ra = cast3<R>(cur - g.r(res));
  R n = sq(nsq(ra));
  if (!(n > (R)1e-10)) return 0.0;
  R V, dV;
  harmonicR<R>((R)p.afm_H1, (R)p.afm_H2, (R)0, n, V, dV);
  const v3<R> f = (dV / n) * ra;
  acc.add(res, f);
  if (b.tip_stat) {
    // the force on the tip, and its part perpendicular to the pulling
    // direction in the x-z plane (eval_forces.cpp:23-29); a tip at rest has
    // no direction and its perpendicular part is NaN, as in the reference
    const double ww = 1.0 / sqrt(b.tip_vel[0][k] * b.tip_vel[0][k] +
                                 b.tip_vel[1][k] * b.tip_vel[1][k] +
                                 b.tip_vel[2][k] * b.tip_vel[2][k]);
    const double wx = ww * b.tip_vel[0][k], wz = ww * b.tip_vel[2][k];
    const double fx = (double)f.x, fz = (double)f.z;
    const double smp[4] = {-fx, -(double)f.y, -fz,
                           fx * wz / (wx + wz * wz / wx) - fz / (1.0 + (wz / wx) * (wz / wx))};
    mavg_add(b.tip_stat[k], now, smp, 4);
  }
  return (double)V;
}

// afm::force::eval_forces::iter (cg/src/cg/afm/force/eval_forces.cpp:10-13):
// the tip's constant pulling force on its residue, no energy. Force tips sit
// after the velocity tips, their force in tip_vel. The tip's average
// velocity reads the state's velocities (y1 / dt as the last Langevin step
// left them, zero before the first) at the current time.
template <typename A>
__device__ __forceinline__ double force_afm_item(const DevBatch &b, A &acc, int k, double now) {
  const int res = b.tip_res[k];
  acc.add(res, v3<double>{b.tip_vel[0][k], b.tip_vel[1][k], b.tip_vel[2][k]});
  if (b.tip_stat) {
    const double smp[3] = {b.vel[0][res], b.vel[1][res], b.vel[2][res]};
    mavg_add(b.tip_stat[k], now, smp, 3);
  }
  return 0.0;
}

// ---------------------------------------------------------------- walls ----
// Walls are evaluated in FP64 in every precision (a few items per step).
__device__ __forceinline__ double sgnd(double x) { return (double)((0.0 < x) - (x < 0.0)); }
// dist^-10 by products of the reciprocal (a few ulp from pow(dist, -10),
// without pow's registers in the persistent kernel)
__device__ __forceinline__ double inv10(double d) {
  const double i1 = 1.0 / d, i2 = i1 * i1, i4 = i2 * i2;
  return i4 * i4 * i2;
}
__device__ __forceinline__ void wall_react(const DevBatch &b, int w, v3<double> f) {
  atomicAdd(&b.wall->wall_F[w][0], f.x);
  atomicAdd(&b.wall->wall_F[w][1], f.y);
  atomicAdd(&b.wall->wall_F[w][2], f.z);
}

// wall::solid::eval_forces::iter (cg/src/cg/wall/solid/eval_forces.cpp:4-17):
// residue i against every solid plane, d^-10 within min_dist, no energy; the
// planes' reaction forces are kept in a plain context (b.wall).
template <typename G, typename A>
__device__ __forceinline__ double solid_wall_item(const DevBatch &b, const G &g, A &acc, int i) {
#pragma clang fp contract(off)
  ParamsC &p = b.P->p;
  const v3<double> r = g.r(i);
  v3<double> f{0.0, 0.0, 0.0};
  bool hit = false;
  for (int w = 0; w < p.n_walls; ++w) {
    if (p.wall_kind[w] != 0) continue;
    const double nx = p.wall_normal[w][0], ny = p.wall_normal[w][1], nz = p.wall_normal[w][2];
    const double sd = nx * (r.x - p.wall_origin[w][0]) + ny * (r.y - p.wall_origin[w][1]) +
                      nz * (r.z - p.wall_origin[w][2]);
    const double dist = fabs(sd);
    if (dist >= p.wall_min_dist) continue;
    const double dV = p.wall_depth * inv10(dist) * sgnd(sd);
    const v3<double> fw{dV * nx, dV * ny, dV * nz};
    f = f + fw;
    hit = true;
    if (b.wall) wall_react(b, w, -fw);
  }
  if (hit) acc.add(i, f);
  return 0.0;
}

// wall::harmonic::eval_free::iter (cg/src/cg/wall/harmonic/eval_free.cpp:15-31):
// a residue not connected to a wall against every harmonic plane, the solid
// wall's d^-10 with the reference's opposite sign (F -= dV n), no energy.
template <typename G, typename A>
__device__ __forceinline__ double hw_free_item(const DevBatch &b, const G &g, A &acc, int i) {
#pragma clang fp contract(off)
  ParamsC &p = b.P->p;
  if (b.wall->is_conn[i]) return 0.0;
  const v3<double> r = g.r(i);
  for (int w = 0; w < p.n_walls; ++w) {
    if (p.wall_kind[w] != 1) continue;
    const v3<double> n{p.wall_normal[w][0], p.wall_normal[w][1], p.wall_normal[w][2]};
    const double sd = n.x * (r.x - p.wall_origin[w][0]) + n.y * (r.y - p.wall_origin[w][1]) +
                      n.z * (r.z - p.wall_origin[w][2]);
    const double dist = fabs(sd);
    if (dist >= p.wall_min_dist) continue;
    const double dV = p.wall_depth * inv10(dist) * sgnd(sd);
    const v3<double> force{dV * n.x, dV * n.y, dV * n.z};
    acc.add(i, -force);
    wall_react(b, w, force);
  }
  return 0.0;
}

// wall::harmonic::eval_connected::iter
// (cg/src/cg/wall/harmonic/eval_connected.cpp:17-29): connection k pulls its
// residue by harmonic(HH1, 0, nat) towards its bead (offset + plane origin),
// with the reference's sign F[res] -= dV_dr u, u from the residue to the bead.
template <typename G, typename A>
__device__ __forceinline__ double hw_conn_item(const DevBatch &b, const G &g, A &acc, int k) {
#pragma clang fp contract(off)
  ParamsC &p = b.P->p;
  const WallDev &W = *b.wall;
  const int res = W.hw_res[k], w = W.hw_wall[k];
  const v3<double> bead{W.hw_off[0][k] + p.wall_origin[w][0], W.hw_off[1][k] + p.wall_origin[w][1],
                        W.hw_off[2][k] + p.wall_origin[w][2]};
  const v3<double> r12 = bead - g.r(res);
  const double r12_n = sqrt(nsq(r12)), r12_rn = 1.0 / r12_n;
  const v3<double> u = r12 * r12_rn;
  const double dr = r12_n - W.hw_nat[k], dr2 = dr * dr;
  const double V = 0.5 * dr2 * (p.wall_hh1 + 0.0 * dr2);
  const double dV = clampR(dr * (p.wall_hh1 + 2.0 * 0.0 * dr2), -1.0e3, 1.0e3);
  const v3<double> f = dV * u;
  acc.add(res, -f);
  wall_react(b, w, f);
  return V;
}

// wall::lj::sift_free::iter (cg/src/cg/wall/lj/sift_free.cpp:15-37): a
// residue not connected to a wall against every LJ plane within min_dist:
// shifted LJ with r_min = min_dist, and a connection candidate while the
// plane is below its limit.
template <typename G, typename A>
__device__ __forceinline__ double ljw_free_item(const DevBatch &b, const G &g, A &acc, int i) {
#pragma clang fp contract(off)
  ParamsC &p = b.P->p;
  WallDev &W = *b.wall;
  if (W.is_conn[i]) return 0.0;
  const v3<double> r = g.r(i);
  double V = 0.0;
  for (int w = 0; w < p.n_walls; ++w) {
    if (p.wall_kind[w] != 2) continue;
    const v3<double> n{p.wall_normal[w][0], p.wall_normal[w][1], p.wall_normal[w][2]};
    const double sd = n.x * (r.x - p.wall_origin[w][0]) + n.y * (r.y - p.wall_origin[w][1]) +
                      n.z * (r.z - p.wall_origin[w][2]);
    const double dist = fabs(sd);
    if (dist > p.wall_min_dist) continue;
    const double rinv = 1.0 / dist, s = rinv * p.wall_min_dist, s6 = ipow6(s), s12 = s6 * s6;
    V += p.wall_lj_depth * (s12 - 2.0 * s6 + 1.0);
    const double dV = clampR(12.0 * p.wall_lj_depth * rinv * (s6 - s12), -1.0e3, 1.0e3);
    const double sg = sgnd(sd);
    const v3<double> F_{(dV * n.x) * sg, (dV * n.y) * sg, (dV * n.z) * sg};
    acc.add(i, -F_);
    wall_react(b, w, F_);
    if (W.num_conns[w] < p.wall_limit[w]) {
      const int slot = atomicAdd(&W.n_cand, 1);
      if (slot < W.cand_cap) {
        W.cand_res[slot] = i;
        W.cand_wall[slot] = w;
        W.cand_d[slot] = dist;
      } else {
        W.overflow = 1;
      }
    }
  }
  return V;
}

// wall::lj::eval_connected::iter (cg/src/cg/wall/lj/eval_connected.cpp:15-41):
// connection k, saturation-weighted LJ (r_min = min_dist) to its bead with the
// reference's sign (F[res] -= F_), then its saturation moves by
// saturation_diff towards 1 within the breaking distance, else towards 0;
// at 0 the connection is released (processed by k_ljw_process).
template <typename G, typename A>
__device__ __forceinline__ double ljw_conn_item(const DevBatch &b, const G &g, A &acc, int k) {
#pragma clang fp contract(off)
  ParamsC &p = b.P->p;
  WallDev &W = *b.wall;
  const int res = W.lj_res[k], w = W.lj_wall[k];
  const v3<double> bead{p.wall_origin[w][0] + W.lj_off[0][k], p.wall_origin[w][1] + W.lj_off[1][k],
                        p.wall_origin[w][2] + W.lj_off[2][k]};
  const v3<double> r12 = bead - g.r(res);
  const double r12_n = sqrt(nsq(r12)), r12_rn = 1.0 / r12_n;
  const v3<double> u = r12 * r12_rn;
  const double s = r12_rn * p.wall_min_dist, s6 = ipow6(s), s12 = s6 * s6;
  const double V_ = p.wall_lj_depth * (s12 - 2.0 * s6);
  const double dV = 12.0 * p.wall_lj_depth * r12_rn * (s6 - s12);
  double sat = W.lj_sat[k];
  const v3<double> F_ = (sat * dV) * u;
  acc.add(res, -F_);
  wall_react(b, w, F_);
  const double V = sat * V_;
  if (r12_n < p.ljw_breaking_dist) {
    sat = fmin(sat + p.ljw_saturation_diff, 1.0);
  } else {
    sat = fmax(sat - p.ljw_saturation_diff, 0.0);
    if (sat == 0.0 && !W.lj_gone[k]) {
      W.lj_gone[k] = 1;
      atomicAdd(&W.n_gone, 1);
    }
  }
  W.lj_sat[k] = sat;
  return V;
}

// ------------------------------------------------------------------ QA ----
// qa::prepare_nh::iter (cg/src/cg/qa/prepare_nh.cpp:10-19), exact FP64:
// n and h of residue i from its own and its chain neighbours' positions.
__device__ __forceinline__ void nh_from(v3<double> ri, v3<double> rp, v3<double> rn, double nh[6]) {
#pragma clang fp contract(off)
  v3<double> v1 = xsub(ri, rp), v2 = xsub(rn, ri);
  v3<double> nn = xsub(v2, v1);
  double ni = __ddiv_rn(1.0, xnorm3(nn));
  v3<double> h{__dsub_rn(__dmul_rn(v2.y, v1.z), __dmul_rn(v2.z, v1.y)),
               __dsub_rn(__dmul_rn(v2.z, v1.x), __dmul_rn(v2.x, v1.z)),
               __dsub_rn(__dmul_rn(v2.x, v1.y), __dmul_rn(v2.y, v1.x))};
  double hi = __ddiv_rn(1.0, xnorm3(h));
  nh[0] = __dmul_rn(ni, nn.x);
  nh[1] = __dmul_rn(ni, nn.y);
  nh[2] = __dmul_rn(ni, nn.z);
  nh[3] = __dmul_rn(hi, h.x);
  nh[4] = __dmul_rn(hi, h.y);
  nh[5] = __dmul_rn(hi, h.z);
}
template <typename G>
__device__ void prepare_nh_item(const DevBatch &b, const G &g, int i) {
  int ip = b.prev[i], in = b.next[i];
  if (ip < 0 || in < 0) return;
  double nh[6];
  nh_from(g.r(i), g.r(ip), g.r(in), nh);
  for (int c = 0; c < 6; ++c) b.nh[c][i] = nh[c];
}

__device__ __forceinline__ double xdot(double ax, double ay, double az, v3<double> u) {
#pragma clang fp contract(off)
  return __dadd_rn(__dadd_rn(__dmul_rn(ax, u.x), __dmul_rn(ay, u.y)), __dmul_rn(az, u.z));
}

// qa::loop_over_candidates::iter (cg/src/cg/qa/loop_over_candidates.cpp:12-99):
// criteria in exact FP64, repulsion force in R; a candidate is appended to
// the system's candidate buffer (its order is restored by the finish sort).
// Item hk of the near list (every entry passed orig_dist <= max_req_dist):
// its fields are loaded together, ahead of the geometry.
struct QhFetch {
  int hk, i1, i2;
  float thr2;
  int vac;
  double req;  // the fields below only from qa_cand_fetch_full
  int slot;    // -1: load it when the pair becomes a candidate
};
__device__ __forceinline__ QhFetch qa_cand_fetch(const DevBatch &b, SysDescC &s, int hk) {
  int h = s.qf0 + hk;
  return QhFetch{hk, b.qh_i[0][h], b.qh_i[1][h], b.qh_thr2[h], b.qh_vac[h], 0.0, 0};
}
__device__ __forceinline__ void keep_fetch(QhFetch &q) {
  keep(q.i1);
  keep(q.i2);
  keep(q.thr2);
  keep(q.vac);
}
// The first test's view of near pair hk, in two loads: its residues within
// the system and its FP32 radius, negative once the pair is vacant (the test
// then fails by itself).
struct QhLite {
  uint32_t ij;
  float thr2;
};
__device__ __forceinline__ QhLite qa_cand_fetch_lite(const DevBatch &b, SysDescC &s, int hk) {
  const int h = s.qf0 + hk;
  return QhLite{b.qh_ij[h], b.qh_thr2[h]};
}
// with what the exact path needs, in the same round trip
__device__ __forceinline__ QhFetch qa_cand_fetch_full(const DevBatch &b, SysDescC &s, int hk) {
  int h = s.qf0 + hk;
  return QhFetch{hk, b.qh_i[0][h], b.qh_i[1][h], b.qh_thr2[h], b.qh_vac[h], b.qh_req[h], b.qh_slot[h]};
}
// The cheap first test: a vacant pair, or one beyond its radius, has no
// effect at all.
template <typename G>
__device__ __forceinline__ bool qa_cand_may_act(SysDescC &s, const G &g, const QhFetch &q) {
  if (q.vac) return false;
  v3<double> d = xwrapd(xsub(g.r(q.i2), g.r(q.i1)), s);
  float dx = (float)d.x, dy = (float)d.y, dz = (float)d.z;
  return !(dx * dx + dy * dy + dz * dz > q.thr2);
}

// Everything after the first test, in exact FP64 as the reference. q comes
// from qa_cand_fetch_full. The residue data (types, sync budgets, n and h)
// comes through the geometry accessor: from HBM, or from LDS and positions.
template <typename R, typename G, typename A>
__device__ __forceinline__ double qa_cand_exact(const DevBatch &b, SysDescC &s, SysState &st,
                                                const G &g, A &acc, const QhFetch &q) {
#pragma clang fp contract(off)
  DevParamsC &P = *b.P;
  ParamsC &p = P.p;
  const int hk = q.hk, i1 = q.i1, i2 = q.i2;
  // independent of the geometry: issued first
  const uint32_t sy1 = g.sync4(b, i1), sy2 = g.sync4(b, i2);
  const int a1 = g.atype(b, i1), a2 = g.atype(b, i2);
  v3<double> d = xwrapd(xsub(g.r(i2), g.r(i1)), s);
  double rn = __ddiv_rn(1.0, xnorm3(d));
  v3<double> u{__dmul_rn(d.x, rn), __dmul_rn(d.y, rn), __dmul_rn(d.z, rn)};
  if (1.0 > __dmul_rn(rn, P.qa_max_req)) return 0.0;
  double V = 0.0;
  if (1.0 < __dmul_rn(rn, p.rep_cutoff)) {
    v3<R> r12 = cast3<R>(d);
    R rnR = rsq(nsq(r12)), Vr, dV;
    shifted_ljR<R>((R)p.pauli_depth, (R)p.rep_cutoff, rnR, Vr, dV);
    v3<R> f = (dV * rnR) * r12;
    acc.add(i1, f);
    acc.add(i2, -f);
    V = (double)Vr;
  }
  if (1.0 > __dmul_rn(rn, q.req)) return V;  // beyond every type's formation distance
  double nh1[6], nh2[6];
  g.nh(b, i1, nh1);
  g.nh(b, i2, nh2);
  double h1u = fabs(xdot(nh1[3], nh1[4], nh1[5], u));
  double h2u = fabs(xdot(nh2[3], nh2[4], nh2[5], u));
  double hh = fabs(__dadd_rn(__dadd_rn(__dmul_rn(nh1[3], nh2[3]), __dmul_rn(nh1[4], nh2[4])),
                             __dmul_rn(nh1[5], nh2[5])));
  double n1u = xdot(nh1[0], nh1[1], nh1[2], u);
  double n2u = xdot(nh2[0], nh2[1], nh2[2], u);
  bool back1 = h1u > p.qa_min_cos_hr, back2 = h2u > p.qa_min_cos_hr;
  bool bbx = hh > p.qa_min_cos_hh;
  bool side1 = n1u < p.qa_max_cos_nr, side2 = -n2u < p.qa_max_cos_nr;
  int type = 0;
  if (back1 && back2 && bbx) type = 1;
  else if (back1 && side2) type = 2;
  else if (side1 && back2) type = 3;
  else if (side1 && side2) type = 4 + a1 + a2 * 21;
  if (type == 0) return V;
  // r_high[type] * (1 + formation_tol)
  const double req = g.type_req(b, type);
  if (1.0 > __dmul_rn(rn, req)) return V;
  int pt1 = g.ptype(b, a1), pt2 = g.ptype(b, a2);
  int8_t s1[4] = {(int8_t)back1, (int8_t)side1, (int8_t)(side1 && pt2 == 1), (int8_t)(side1 && pt2 == 2)};
  int8_t s2[4] = {(int8_t)back2, (int8_t)side2, (int8_t)(side2 && pt1 == 1), (int8_t)(side2 && pt1 == 2)};
  for (int a = 0; a < 4; ++a)
    if ((int8_t)((int8_t)(sy1 >> (8 * a)) - s1[a]) < 0) return V;
  for (int a = 0; a < 4; ++a)
    if ((int8_t)((int8_t)(sy2 >> (8 * a)) - s2[a]) < 0) return V;
  int c = atomicAdd(&st.cand_n, 1);
  if (c >= s.nlcap) {
    st.overflow = 1;
    return V;
  }
  int cs = s.nl0 + c;
  int k = q.slot >= 0 ? q.slot : b.qh_slot[s.qf0 + hk];
  b.cand_i[0][cs] = i1;
  b.cand_i[1][cs] = i2;
  b.cand_od[cs] = b.qf_d[s.qf0 + k];
  b.cand_dist[cs] = __ddiv_rn(1.0, rn);
  b.cand_idx[cs] = k;
  b.cand_h[cs] = hk;
  b.cand_t[cs] = (int16_t)type;
  for (int a = 0; a < 4; ++a) {
    b.cand_sd[8 * cs + a] = s1[a];
    b.cand_sd[8 * cs + 4 + a] = s2[a];
  }
  return V;
}
template <typename R, typename G, typename A>
__device__ __forceinline__ double qa_cand_core(const DevBatch &b, SysDescC &s, SysState &st,
                                               const G &g, A &acc, const QhFetch &q) {
  if (!qa_cand_may_act(s, g, q)) return 0.0;
  return qa_cand_exact<R>(b, s, st, g, acc, q);
}
template <typename R, typename G, typename A>
__device__ __forceinline__ double qa_cand_item(const DevBatch &b, SysDescC &s, SysState &st,
                                               const G &g, A &acc, int hk) {
  return qa_cand_core<R>(b, s, st, g, acc, qa_cand_fetch_full(b, s, hk));
}

// qa::process_contacts::iter (cg/src/cg/qa/process_contacts.cpp:12-72)
template <typename R, typename G, typename A>
__device__ double qa_cont_item(const DevBatch &b, SysDescC &s, SysState &st, const G &g,
                               A &acc, int k, bool want_v = true) {
#pragma clang fp contract(off)
  DevParamsC &P = *b.P;
  ParamsC &p = P.p;
  int slot = s.qc0 + k;
  // every field loaded before the first branch
  int vac = b.qc_vac[slot], i1 = b.qc_i[0][slot], i2 = b.qc_i[1][slot], type = b.qc_t[slot];
  double sat = b.qc_sat[slot];
  keep(vac);
  keep(i1);
  keep(i2);
  keep(type);
  keep(sat);
  if (vac) return 0.0;
  const double sat_used = sat;
  v3<double> d = xwrapd(xsub(g.r(i2), g.r(i1)), s);
  double n2 = __dadd_rn(__dadd_rn(__dmul_rn(d.x, d.x), __dmul_rn(d.y, d.y)), __dmul_rn(d.z, d.z));
  double rn = __ddiv_rn(1.0, __dsqrt_rn(n2));
  if (__dmul_rn(rn, p.nl_cutoff) < 1.0) return 0.0;
  double r12n = __dmul_rn(rn, n2);
  v3<R> r12 = cast3<R>(d);
  R Vij, dV;
  sink_ljR<R>((R)p.qa_lj_depth[type], (R)p.qa_lj_r_low[type], (R)p.qa_lj_r_high[type], (R)r12n,
              (R)rn, Vij, dV);
  v3<R> f = ((R)sat * dV * (R)rn) * r12;
  acc.add(i1, f);
  acc.add(i2, -f);
  bool breaking = r12n > __dmul_rn(P.qa_factor, p.qa_lj_r_high[type]);
  double step = __dmul_rn(P.qa_sat_speed, p.dt);
  if (breaking) {
    sat = fmin(__dsub_rn(sat, step), 0.0);
    if (sat == 0.0) {
      int r = atomicAdd(&st.rem_n, 1);
      if (r < s.qccap) b.rem[s.qc0 + r] = k;
      else st.overflow = 1;
    }
  } else {
    sat = fmax(__dadd_rn(sat, step), 0.0);
  }
  b.qc_sat[slot] = sat;
  // the energy carries the saturation the force used (qa/process_contacts.cpp:40);
  // with want_v the FP32 form takes it in FP64: contacts in the repulsive
  // part of the well and in its tail have energies of both signs
  if constexpr (sizeof(R) == 4) {
    if (want_v) {
      double V64, dV64;
      sink_ljR<double>(p.qa_lj_depth[type], p.qa_lj_r_low[type], p.qa_lj_r_high[type], r12n, rn,
                       V64, dV64);
      return sat_used * V64;
    }
  }
  return (double)sat_used * (double)Vij;
}

} // namespace pcg
