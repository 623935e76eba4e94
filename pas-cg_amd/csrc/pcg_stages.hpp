// pcg_stages.hpp -- device stages shared by the grid-wide kernels and the
// persistent workgroup-per-system kernel: NL pair classification for the
// ordered list updates, the QA finish, and the Gear-5 update.
#pragma once
#include "pcg_kernels.hpp"

namespace pcg {

// partition classes of an NL pair; CL_NEAR marks a QA free pair within the
// QA maximum required distance (its near-list entry is written beside it)
enum { CL_NC = 1, CL_QA = 2, CL_PID = 4, CL_DH = 8, CL_PAULI = 16, CL_NEAR = 32 };

__device__ inline int find_contact(const DevBatch &b, SysDescC &s, int i1, int i2) {
  int lo = s.nc0, hi = s.nc0 + s.nnc;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    int a = b.nc_i[0][mid], c = b.nc_i[1][mid];
    if (a < i1 || (a == i1 && c < i2)) lo = mid + 1;
    else hi = mid;
  }
  if (lo < s.nc0 + s.nnc && b.nc_i[0][lo] == i1 && b.nc_i[1][lo] == i2) return lo;
  return -1;
}

// Slot q of the current native-contact list takes contact c of the list of
// all contacts, with a copy of its fields.
__device__ __forceinline__ void set_current_contact(const DevBatch &b, int q, int c, int res0) {
  b.ncur[q] = c;
  b.ncur_i[0][q] = b.nc_i[0][c];
  b.ncur_i[1][q] = b.nc_i[1][c];
  b.ncur_nat[q] = b.nc_nat[c];
  // the type, and the contact's state in bit 7 (kept equal to nc_active[c])
  b.ncur_t[q] = (uint8_t)(b.nc_type[c] | (b.nc_active[c] ? 0x80 : 0));
  double th = __dmul_rn(__dmul_rn(b.nc_nat[c], C216_INV), b.P->p.counting_factor);
  b.ncur_th2[q] = (float)(th * th);
  b.ncur_ijt[q] = (uint32_t)((b.nc_i[0][c] - res0) & 0xFFF) |
                  ((uint32_t)((b.nc_i[1][c] - res0) & 0xFFF) << 12) | ((uint32_t)b.ncur_t[q] << 24);
}

__device__ inline int classify(const DevBatch &b, SysDescC &s, const SysState &st,
                                        int slot) {
  DevParamsC &P = *b.P;
  ParamsC &p = P.p;
  int i1 = b.nl_i[0][slot], i2 = b.nl_i[1][slot];
  double d = b.nl_d[slot];
  double pad = st.orig_pad;
  bool four = b.chain[i1] == b.chain[i2] && abs(b.seq[i1] - b.seq[i2]) == 4;
  if (p.nat_cont_on && s.nnc > 0 && d < __dadd_rn(p.nl_cutoff, pad) && find_contact(b, s, i1, i2) >= 0)
    return CL_NC;
  if (p.qa_on && !(!p.qa_include4 && four) && d < __dadd_rn(p.nl_cutoff, pad))
    return CL_QA | (!(d > P.qa_max_req) ? CL_NEAR : 0);
  int cls = 0;
  if (p.pid_on && !(!p.pid_include4 && four) && d < __dadd_rn(p.nl_cutoff, pad)) cls |= CL_PID;
  if (p.dh_on) {
    double q = p.charge[b.atype[i1]] * p.charge[b.atype[i2]];
    if (q != 0.0 && __ddiv_rn(1.0, d) > __ddiv_rn(1.0, __dadd_rn(p.dh_cutoff, pad))) cls |= CL_DH;
  }
  if (P.pauli_on && d < __dadd_rn(p.rep_cutoff, pad)) cls |= CL_PAULI;
  return cls;
}

// qa::finish_processing (cg/src/cg/qa/finish_processing.cpp:5-74), one block
// per system: removals in contact order, candidates sorted by (distance,
// free-pair index), sequential acceptance against the sync budgets, and
// vect::set::compress, which moves items but not vacancy flags
// (cg/include/cg/vect/ind/set.h:62-71).
// IEEE-754 totalOrder as a signed integer key: equal to the numeric order
// on the non-negative distances ranked below, and still a strict total order
// when a diverged system produces NaN, so the ranks always form a permutation
__device__ __forceinline__ long long total_key(double d) {
  long long k = __double_as_longlong(d);
  return k ^ ((k >> 63) & 0x7fffffffffffffffLL);
}

// Filters of one near pair, from its residue types: req is the largest
// r_high * (1 + formation_tol) over the contact types the pair can form
// (bb, bs, sb and its side-side type), compared exactly as the per-type test
// of loop_over_candidates (multiplication by the inverse distance is
// monotonic, so rn * req < 1 implies rn * req_type < 1 for every type).
// Beyond min(max_req, max(rep_cutoff, req)) the pair has no repulsion and no
// candidate: thr2 is that radius squared in FP32, widened by 1e-5.
__device__ inline void qa_near_filters(const DevBatch &b, int i1, int i2, double &req, float &thr2) {
  ParamsC &p = b.P->p;
  double one_tol = __dadd_rn(1.0, p.qa_formation_tol);
  int ss = 4 + b.atype[i1] + 21 * b.atype[i2];
  req = __dmul_rn(p.qa_lj_r_high[1], one_tol);
  req = fmax(req, __dmul_rn(p.qa_lj_r_high[2], one_tol));
  req = fmax(req, __dmul_rn(p.qa_lj_r_high[3], one_tol));
  req = fmax(req, __dmul_rn(p.qa_lj_r_high[ss], one_tol));
  double thr = fmin(b.P->qa_max_req, fmax(p.rep_cutoff, req)) * (1.0 + 1e-5);
  thr2 = (float)(thr * thr);
}

// LDS staging of the candidates for a QA finish that runs alone in its
// workgroup: their sort keys, and in acceptance order their slots,
// residues and budget changes (8 bytes), for up to cap candidates
struct QfLds {
  long long *key;
  int *idx, *sc, *i1, *i2;
  uint64_t *sd;
  int cap;
};

// sy[4 * (i - sy_off) + c] is sync budget c of residue i (HBM or LDS).
__device__ inline void qa_finish_block(const DevBatch &b, SysDescC &sd, SysState &st,
                                       int8_t *sy, int sy_off, const QfLds *ql = nullptr) {
  // Nothing removed and no candidate: budgets, lists and counts stay as they
  // are (the compress condition changes only with removals or acceptances).
  // Every thread reads the same final counts, so the return is uniform.
  if (st.cand_n == 0 && st.rem_n == 0) return;
  int rem_n = st.rem_n < sd.qccap ? st.rem_n : sd.qccap;
  int cand_n = st.cand_n < sd.nlcap ? st.cand_n : sd.nlcap;
  // rank of every candidate by (dist, free-pair index) -> permutation
  const bool staged = ql != nullptr && cand_n <= ql->cap;
  if (staged) {
    for (int c = threadIdx.x; c < cand_n; c += blockDim.x) {
      ql->key[c] = total_key(b.cand_dist[sd.nl0 + c]);
      ql->idx[c] = b.cand_idx[sd.nl0 + c];
    }
    __syncthreads();
  }
  for (int c = threadIdx.x; c < cand_n; c += blockDim.x) {
    int sc = sd.nl0 + c;
    long long kc = staged ? ql->key[c] : total_key(b.cand_dist[sc]);
    int ic = staged ? ql->idx[c] : b.cand_idx[sc];
    int rank = 0;
    if (staged) {
      for (int o = 0; o < cand_n; ++o) {
        long long kd = ql->key[o];
        rank += (kd < kc || (kd == kc && ql->idx[o] < ic)) ? 1 : 0;
      }
      ql->sc[rank] = sc;
      ql->i1[rank] = b.cand_i[0][sc];
      ql->i2[rank] = b.cand_i[1][sc];
      ql->sd[rank] = *(const uint64_t *)(b.cand_sd + 8 * (size_t)sc);
    } else {
      for (int o = 0; o < cand_n; ++o) {
        int so = sd.nl0 + o;
        long long kd = total_key(b.cand_dist[so]);
        rank += (kd < kc || (kd == kc && b.cand_idx[so] < ic)) ? 1 : 0;
      }
      b.cand_perm[sd.nl0 + rank] = c;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    // removed contacts in increasing index order
    int *rem = b.rem + sd.qc0;
    for (int a = 1; a < rem_n; ++a) {
      int v = rem[a], j = a - 1;
      while (j >= 0 && rem[j] > v) {
        rem[j + 1] = rem[j];
        --j;
      }
      rem[j + 1] = v;
    }
    for (int a = 0; a < rem_n; ++a) {
      int q = sd.qc0 + rem[a];
      int i1 = b.qc_i[0][q], i2 = b.qc_i[1][q];
      for (int c = 0; c < 4; ++c) {
        sy[4 * (i1 - sy_off) + c] = (int8_t)(sy[4 * (i1 - sy_off) + c] + b.qc_sd[8 * q + c]);
        sy[4 * (i2 - sy_off) + c] = (int8_t)(sy[4 * (i2 - sy_off) + c] + b.qc_sd[8 * q + 4 + c]);
      }
      if (st.qf_store < sd.qfcap) {
        int f = sd.qf0 + st.qf_store;
        b.qf_i[0][f] = i1;
        b.qf_i[1][f] = i2;
        b.qf_d[f] = b.qc_d[q];
        b.qf_vac[f] = 0;
        // appended at the end, so the near list stays in slot order
        if (!(b.qf_d[f] > b.P->qa_max_req)) {
          int h = sd.qf0 + st.qh_n++;
          b.qh_slot[h] = st.qf_store;
          b.qh_i[0][h] = i1;
          b.qh_i[1][h] = i2;
          b.qh_vac[h] = 0;
          qa_near_filters(b, i1, i2, b.qh_req[h], b.qh_thr2[h]);
          b.qh_ij[h] = (uint32_t)(i1 - sd.res0) | ((uint32_t)(i2 - sd.res0) << 16);
        }
        st.qf_n += 1;
        st.qf_store += 1;
      } else {
        st.overflow = 1;
      }
      b.qc_vac[q] = 1;
      st.num_qa -= 1;
    }
    // sequential acceptance
    for (int a = 0; a < cand_n; ++a) {
      int sc, i1, i2;
      int8_t csd[8];
      if (staged) {
        sc = ql->sc[a];
        i1 = ql->i1[a];
        i2 = ql->i2[a];
        const uint64_t v = ql->sd[a];
        for (int c = 0; c < 8; ++c) csd[c] = (int8_t)(v >> (8 * c));
      } else {
        sc = sd.nl0 + b.cand_perm[sd.nl0 + a];
        i1 = b.cand_i[0][sc];
        i2 = b.cand_i[1][sc];
        for (int c = 0; c < 8; ++c) csd[c] = b.cand_sd[8 * sc + c];
      }
      bool ok = true;
      int8_t n1[4], n2[4];
      for (int c = 0; c < 4; ++c) {
        n1[c] = (int8_t)(sy[4 * (i1 - sy_off) + c] - csd[c]);
        n2[c] = (int8_t)(sy[4 * (i2 - sy_off) + c] - csd[4 + c]);
      }
      for (int c = 0; c < 4; ++c) ok = ok && n1[c] >= 0;
      if (!ok) continue;
      for (int c = 0; c < 4; ++c) ok = ok && n2[c] >= 0;
      if (!ok) continue;
      if (st.qc_n >= sd.qccap) {
        st.overflow = 1;
        continue;
      }
      int q = sd.qc0 + st.qc_n;
      b.qc_i[0][q] = i1;
      b.qc_i[1][q] = i2;
      b.qc_d[q] = b.cand_od[sc];
      b.qc_t[q] = b.cand_t[sc];
      b.qc_sat[q] = 0.0;
      for (int c = 0; c < 8; ++c) b.qc_sd[8 * q + c] = csd[c];
      b.qc_vac[q] = 0;
      st.qc_n += 1;
      st.num_qa += 1;
      for (int c = 0; c < 4; ++c) {
        sy[4 * (i1 - sy_off) + c] = n1[c];
        sy[4 * (i2 - sy_off) + c] = n2[c];
      }
      b.qf_vac[sd.qf0 + b.cand_idx[sc]] = 1;
      b.qh_vac[sd.qf0 + b.cand_h[sc]] = 1;
      b.qh_thr2[sd.qf0 + b.cand_h[sc]] = -1.0f;
    }
    // compress
    if (2 * st.num_qa < st.qc_n) {
      int slot = 0;
      for (int cur = 0; cur < st.qc_n; ++cur) {
        int qc = sd.qc0 + cur;
        if (b.qc_vac[qc]) continue;
        int qs = sd.qc0 + slot;
        if (qs != qc) {
          b.qc_i[0][qs] = b.qc_i[0][qc];
          b.qc_i[1][qs] = b.qc_i[1][qc];
          b.qc_d[qs] = b.qc_d[qc];
          b.qc_t[qs] = b.qc_t[qc];
          b.qc_sat[qs] = b.qc_sat[qc];
          for (int c = 0; c < 8; ++c) b.qc_sd[8 * qs + c] = b.qc_sd[8 * qc + c];
          // the vacancy flag of the destination slot stays as it was
        }
        ++slot;
      }
      st.qc_n = slot;
    }
    st.cand_n = 0;
    st.rem_n = 0;
  }
  __syncthreads();
}

// -------------------------------------------------------------- Langevin ----
// lang::step::omp_async (cg/src/cg/langevin/step.cpp:23-77): FP64 Gear-5
// predictor-corrector with the reference's float-computed coefficients
// (cg/include/cg/langevin/step.h:10-12). The displacement from the last NL
// rebuild is reduced here for the next step's verify.
// ya(k, c) is a reference to y_k, component c, of the residue (in HBM or LDS)
// minv = 1 / mass and sqminv = sqrt(minv) of the residue's type.
template <typename YA>
__device__ __forceinline__ void gear_core(const DevBatch &b, SysDescC &sd, const SysState &st,
                                          double minv, double sqminv, double Fx, double Fy,
                                          double Fz, const double nz[3], double out_r[3],
                                          double vel[3], YA ya) {
#pragma clang fp contract(off)
  DevParamsC &P = *b.P;
  const double f02 = (double)(3.0f / 16.0f), f12 = (double)(251.0f / 360.0f),
               f32 = (double)(11.0f / 18.0f), f42 = (double)(1.0f / 6.0f),
               f52 = (double)(1.0f / 60.0f);
  double gam2 = P.gamma2 * minv, con2 = sd.const2 * sqminv;
  double F[3] = {Fx, Fy, Fz};
  for (int c = 0; c < 3; ++c) {
    double y0 = ya(0, c), y1 = ya(1, c), y2 = ya(2, c), y3 = ya(3, c), y4 = ya(4, c),
           y5 = ya(5, c);
    if (st.t == 0) y2 = F[c] * P.deltsq;
    double f = F[c] - gam2 * y1;
    y1 = y1 + con2 * nz[c];
    double err = y2 - P.deltsq * f;
    y0 -= f02 * err;
    y1 -= f12 * err;
    y2 -= 1.0 * err;
    y3 -= f32 * err;
    y4 -= f42 * err;
    y5 -= f52 * err;
    y0 += y1 + y2 + y3 + y4 + y5;
    y1 += 2.0 * y2 + 3.0 * y3 + 4.0 * y4 + 5.0 * y5;
    y2 += 3.0 * y3 + 6.0 * y4 + 10.0 * y5;
    y3 += 4.0 * y4 + 10.0 * y5;
    y4 += 5.0 * y5;
    ya(0, c) = y0;
    ya(1, c) = y1;
    ya(2, c) = y2;
    ya(3, c) = y3;
    ya(4, c) = y4;
    ya(5, c) = y5;
    vel[c] = y1 * P.dt_inv;
    out_r[c] = y0;
  }
}

__device__ __forceinline__ void gear_update(const DevBatch &b, SysDescC &sd, const SysState &st,
                                            int i, double Fx, double Fy, double Fz,
                                            const double nz[3], double out_r[3]) {
  double vel[3];
  const double minv = 1.0 / b.P->p.mass[b.atype[i]];
  gear_core(b, sd, st, minv, sqrt(minv), Fx, Fy, Fz, nz, out_r, vel,
            [&](int k, int c) -> double & { return b.y[k][c][i]; });
  for (int c = 0; c < 3; ++c) b.vel[c][i] = vel[c];
}

// The square of disp_from, with the same operations before the root: since
// the correctly rounded square root is monotonic, the root of the largest
// square is exactly the largest disp_from.
__device__ __forceinline__ double disp2_from(SysDescC &sd, const double o[3], const double r[3]) {
#pragma clang fp contract(off)
  v3<double> d{__dsub_rn(o[0], r[0]), __dsub_rn(o[1], r[1]), __dsub_rn(o[2], r[2])};
  d = xwrapd(d, sd);
  return __dadd_rn(__dadd_rn(__dmul_rn(d.x, d.x), __dmul_rn(d.y, d.y)), __dmul_rn(d.z, d.z));
}
__device__ __forceinline__ double disp_from(SysDescC &sd, const double o[3], const double r[3]) {
#pragma clang fp contract(off)
  v3<double> d{__dsub_rn(o[0], r[0]), __dsub_rn(o[1], r[1]), __dsub_rn(o[2], r[2])};
  return xnorm3(xwrapd(d, sd));
}
__device__ __forceinline__ double disp_from_orig(const DevBatch &b, SysDescC &sd, int i,
                                                 const double r[3]) {
  const double o[3] = {b.orig_r[0][i], b.orig_r[1][i], b.orig_r[2][i]};
  return disp_from(sd, o, r);
}


} // namespace pcg
