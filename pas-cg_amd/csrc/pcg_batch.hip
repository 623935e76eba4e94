// pcg_batch.hip -- the persistent batched form: one workgroup owns one
// system and runs many whole steps of thread::advance_by_step
// (cg/src/cg/simul/thread_main.cpp:373-536) inside a single launch.
//
// Positions (FP64) and forces (R) of the system stay in LDS across steps;
// force terms accumulate with LDS atomics; the NL rebuild, the ordered list
// partition and the QA finish run inside the workgroup; the Gear-5 state is
// streamed from/to HBM (it is L2/MALL-resident for the batch sizes used).
// Systems are independent, so workgroups never synchronise with each other.
#include "pcg_launch.hpp"
#include "pcg_stages.hpp"

namespace pcg {

// Geometry and per-residue QA data of the workgroup's system, from LDS.
struct GeoL {
  const double *x, *y, *z;
  int res0;
  const uint8_t *at;       // residue type | has chain predecessor << 6 | successor << 7
  const uint32_t *sync;    // the four sync budgets of each residue, one byte each
  const int8_t *ptype_t;   // polarity class per residue type
  const double *req_t;     // r_high[type] * (1 + formation_tol) per QA type
  __device__ __forceinline__ v3<double> r(int i) const {
    int k = i - res0;
    return v3<double>{x[k], y[k], z[k]};
  }
  __device__ __forceinline__ int atype(const DevBatch &, int i) const { return at[i - res0] & 63; }
  __device__ __forceinline__ int ptype(const DevBatch &, int a) const { return ptype_t[a]; }
  __device__ __forceinline__ uint32_t sync4(const DevBatch &, int i) const { return sync[i - res0]; }
  __device__ __forceinline__ double type_req(const DevBatch &, int t) const { return req_t[t]; }
  // n and h of residue i as wg_prepare_nh left them for this step
  __device__ __forceinline__ void nh(const DevBatch &b, int i, double out[6]) const {
    for (int c = 0; c < 6; ++c) out[c] = b.nh[c][i];
  }
};

// qa::prepare_nh (cg/src/cg/qa/prepare_nh.cpp:10-19) over the workgroup's
// system, from the LDS positions, into the batch's n and h arrays (the chain
// neighbours of residue i are i - 1 and i + 1; zero at chain ends, where
// prepare_nh writes nothing). Run once per step after the positions move,
// so that a QA pair reads its residues' vectors instead of recomputing them.
__device__ __forceinline__ void wg_prepare_nh(const DevBatch &b, const GeoL &g, int n, int r0) {
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    const int i = r0 + k;
    double o[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if ((g.at[k] & 0xC0) == 0xC0) nh_from(g.r(i), g.r(i - 1), g.r(i + 1), o);
    for (int c = 0; c < 6; ++c) b.nh[c][i] = o[c];
  }
}

template <typename RF> struct AccL {
  RF *fx, *fy, *fz;
  int res0;
  template <typename T> __device__ __forceinline__ void add(int i, v3<T> f) {
    int k = i - res0;
    atomicAdd(fx + k, (RF)f.x);
    atomicAdd(fy + k, (RF)f.y);
    atomicAdd(fz + k, (RF)f.z);
  }
};

__device__ __forceinline__ int wave_incl_scan(int x) {
  int lane = threadIdx.x & 63;
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  return x;
}

// In-workgroup exclusive scan of the n row counts stored at cnt[0..n);
// returns the total (all threads).
__device__ int block_exscan(int *cnt, int n) {
  __shared__ int wsum[16];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  int nw = blockDim.x >> 6, w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int base = 0; base < n; base += blockDim.x) {
    int i = base + threadIdx.x;
    int v = i < n ? cnt[i] : 0;
    int x = wave_incl_scan(v);
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int off = carry;
    for (int k = 0; k < w; ++k) off += wsum[k];
    if (i < n) cnt[i] = off + x - v;
    __syncthreads();
    if (threadIdx.x == 0)
      for (int k = 0; k < nw; ++k) carry += wsum[k];
    __syncthreads();
  }
  int tot = carry;
  __syncthreads();
  return tot;
}

// Maximum over the workgroup, for thread 0 only: each wave's maximum goes to
// part[wave], the barrier, then thread 0 reads them. part is next written one
// step later, after the barrier that ends this step, so one barrier suffices.
__device__ __forceinline__ double block_max_to_first(double v, double *part) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) part[w] = v;
  __syncthreads();
  double m = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) m = fmax(m, part[k]);
  return m;
}

// ordered ranks for the partition, any block size up to 1024
__device__ __forceinline__ void block_ranks5(int cls, int valid, int counts[5], int ranks[5]) {
  __shared__ int wc[16][5];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  unsigned long long lt = (1ull << lane) - 1ull;
  int mine[5];
  for (int c = 0; c < 5; ++c) {
    unsigned long long m = __ballot(valid && (cls & (1 << c)));
    mine[c] = __popcll(m & lt);
    if (lane == 0) wc[w][c] = __popcll(m);
  }
  __syncthreads();
  for (int c = 0; c < 5; ++c) {
    int off = 0, tot = 0;
    for (int k = 0; k < nw; ++k) {
      if (k < w) off += wc[k][c];
      tot += wc[k][c];
    }
    ranks[c] = off + mine[c];
    counts[c] = tot;
  }
  __syncthreads();
}


// Tallies of the list rebuild: NL pairs, then the pairs of each list the
// ordered partition fills (native contacts, QA free pairs, the QA free pairs
// that can become candidates, PID, DH, Pauli).
enum { RB_NL = 0, RB_NC, RB_QA, RB_NEAR, RB_PID, RB_DH, RB_PAULI, RB_N };

// First row of wave w's block when nw waves split rows 0..n-1 of the
// upper-triangle pair rows (row k holds n - 1 - k pairs) into contiguous
// blocks of about equal pair counts: the smallest k with
// k (n - 1) - k (k - 1) / 2 >= w P / nw, P = n (n - 1) / 2. Monotone in w,
// 0 for w = 0 and n for w = nw.
__device__ __forceinline__ int rb_row_start(int w, int nw, int n) {
  if (w <= 0) return 0;
  if (w >= nw) return n;
  const double P = 0.5 * (double)n * (double)(n - 1);
  const double T = P * (double)w / (double)nw;
  const double h = (double)n - 0.5;
  const double disc = fmax(h * h - 2.0 * T, 0.0);
  int k = (int)ceil(h - sqrt(disc));
  k = max(0, min(n, k));
  // exact integer correction of the rounding
  auto C = [&](long long kk) { return kk * (long long)(n - 1) - kk * (kk - 1) / 2; };
  const long long Ti = (long long)ceil(T);
  while (k > 0 && C(k - 1) >= Ti) --k;
  while (k < n && C(k) < Ti) ++k;
  return k;
}

// nl::legacy_update (cg/src/cg/nl/legacy_update.cpp:11-55) fused with the
// ordered list updates of thread_main.cpp:409-423 and the QA near list,
// inside the workgroup and without writing the NL itself. The rows of the
// sorted NL (pairs i < j) are split into contiguous blocks of about equal
// pair counts, one per wave; a wave walks its rows in order, 64 columns j at
// a time, and classifies every pair on the spot as the partition would
// classify the NL pair: native contact (a taken pair of the row's native
// contacts), QA free pair (taken), PID bundle, DH pair, Pauli pair. A first
// pass counts each list's pairs per wave, a scan over the waves gives each
// wave its offsets, and a second pass writes every pair at its wave's
// offset plus its rank among the block's earlier pairs of that list (a
// running count and a ballot): the lists come out exactly in the order of
// partitioning the sorted NL. The counts are scalars; nothing per row goes
// to memory.
//
// Every decision is the FP64 one of the reference's expressions. The squared
// distance is formed in FP32 from the FP64 differences; when it is clear of
// every threshold by a relative 2e-5 (far beyond FP32's rounding of a sum of
// three squares) the FP32 comparison decides, else the exact FP64 distance
// does. The FP64 distance itself is formed for the pairs whose lists keep it
// (QA near pairs, PID bundles).
//
// QA free pairs beyond the QA maximum required distance are counted (qf_n,
// which sizes the contact loop as the reference's free-pair count does) but
// not stored: with total_disp reset after every verify
// (cg/src/cg/nl/verify.cpp:31-32) loop_over_candidates skips them until the
// next rebuild (qa/loop_over_candidates.cpp:20), and the stored pairs keep
// their relative order, so the (distance, free-pair index) order of the QA
// finish is unchanged.
//
// The chain and sequence indices of the residues, the row starts of the
// static native contacts and the contacts' second residues sit in the LDS
// region of the QA stacks (rb), which is free during a rebuild.
// (inlined: as a call it would make the kernel save every live register)
__device__ __forceinline__ void wg_fix_nl(const DevBatch &b, SysDescC &sd, SysState &st, const GeoL &g,
                                          void *rb) {
#pragma clang fp contract(off)
  DevParamsC &P = *b.P;
  ParamsC &p = P.p;
  const int n = sd.n, r0 = sd.res0;
  // the wave's index, known to the compiler as uniform: the row loops below
  // are then scalar loops and the running counts scalar registers
  const int nw = blockDim.x >> 6, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const unsigned long long below = (1ull << lane) - 1ull;
  const int nnc = p.nat_cont_on ? sd.nnc : 0;
  // each residue's sequence index, chain and type in one word (seq | chain
  // << 12 | type << 24: a persistent system has fewer than 4096 residues)
  uint32_t *csq = (uint32_t *)rb;
  int *rs = (int *)(csq + n);  // rs[0..n]
  uint16_t *pj = (uint16_t *)(rs + n + 1);               // [nnc]
  __shared__ int wcnt[16][RB_N];
  const bool cell = has_cell(sd);
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    csq[k] = (uint32_t)b.seq[r0 + k] | ((uint32_t)b.chain[r0 + k] << 12) |
             ((uint32_t)(g.at[k] & 63) << 24);
  }
  // contacts sd.nc0 + [rs[k], rs[k + 1]) of the list sorted by (i1, i2) are
  // those of residue r0 + k with later residues; pj holds their second
  // residues within the system
  if (nnc == 0) {
    for (int k = threadIdx.x; k <= n; k += blockDim.x) rs[k] = 0;
  } else {
    for (int c = threadIdx.x; c < nnc; c += blockDim.x) {
      const int a = b.nc_i[0][sd.nc0 + c] - r0;
      const int prev = c > 0 ? b.nc_i[0][sd.nc0 + c - 1] - r0 : -1;
      pj[c] = (uint16_t)(b.nc_i[1][sd.nc0 + c] - r0);
      for (int r = prev + 1; r <= a; ++r) rs[r] = c;
      if (c == nnc - 1)
        for (int r = a + 1; r <= n; ++r) rs[r] = nnc;
    }
  }
  const double pad = p.nl_pad;
  const double t_nl = __dadd_rn(P.nl_eff_cutoff, pad);
  const double t_nc = __dadd_rn(p.nl_cutoff, pad);
  const double t_dh = __dadd_rn(p.dh_cutoff, pad);
  const double t_rep = __dadd_rn(p.rep_cutoff, pad);
  const double t_near = P.qa_max_req;
  const double rt_dh = __ddiv_rn(1.0, t_dh);
  // squared thresholds in FP32, and their margins
  const float s_nl = (float)(t_nl * t_nl), s_nc = (float)(t_nc * t_nc), s_dh = (float)(t_dh * t_dh),
              s_rep = (float)(t_rep * t_rep), s_near = (float)(t_near * t_near);
  const float mlo = 1.0f - 2e-5f, mhi = 1.0f + 2e-5f;
  const float e_nl[2] = {s_nl * mlo, s_nl * mhi}, e_nc[2] = {s_nc * mlo, s_nc * mhi},
              e_dh[2] = {s_dh * mlo, s_dh * mhi}, e_rep[2] = {s_rep * mlo, s_rep * mhi},
              e_near[2] = {s_near * mlo, s_near * mhi};
  const float e_far = fmaxf(fmaxf(fmaxf(e_nl[1], e_nc[1]), fmaxf(e_dh[1], e_rep[1])), e_near[1]);
  const bool qa_on = p.qa_on, pid_on = p.pid_on, dh_on = p.dh_on, pauli_on = P.pauli_on;
  const bool qa4 = p.qa_include4, pid4 = p.pid_include4;
  // residue types with a charge, as bits (q1 q2 != 0 iff both are set)
  uint32_t chg = 0u;
  for (int t = 0; t < 21; ++t) chg |= (p.charge[t] != 0.0 ? 1u : 0u) << t;
  const int kbeg = rb_row_start(w, nw, n), kend = rb_row_start(w + 1, nw, n);
  // When every pair fits the NL capacity, the first pass also stores, in
  // order, the wave's pairs that the second pass writes somewhere (native
  // contacts, QA near pairs, PID, DH and Pauli pairs; QA free pairs beyond
  // the near distance are only counted): k | j << 16 in nl_i[0], the
  // classes in nl_cls, a native contact's index in nl_i[1], from the index
  // of its block's first pair, which no other wave's pairs reach. The
  // second pass walks those instead of every pair of the block.
  const bool compact = (long long)n * (n - 1) / 2 <= (long long)sd.nlcap;
  const int seg = sd.nl0 + (int)((long long)kbeg * (n - 1) - (long long)kbeg * (kbeg - 1) / 2);
  int my_nl = 0;
  // the second pass's writes of one pair: native contact c, QA free pair
  // and near-list entry, PID bundle, DH pair, Pauli pair, each at the wave's
  // running offset plus its rank among the chunk's earlier pairs of the list
  auto emit = [&](int k, int j, int c, bool is_nc, bool near, bool is_pid, bool is_dh,
                  bool is_pauli, unsigned long long m_nc, unsigned long long m_near,
                  unsigned long long m_pid, unsigned long long m_dh, unsigned long long m_pauli,
                  const int *off) {
    const int i = r0 + k, jg = r0 + j;
    const int ai = g.at[k] & 63, aj = g.at[j] & 63;
    if (is_nc) b.ncur[sd.nc0 + off[RB_NC] + __popcll(m_nc & below)] = c;
    double d64 = 0.0;
    if (near | is_pid) {
      v3<double> dd = xsub(v3<double>{g.x[j], g.y[j], g.z[j]}, v3<double>{g.x[k], g.y[k], g.z[k]});
      if (cell) dd = xwrapd(dd, sd);
      d64 = xnorm3(dd);
    }
    if (near) {
      const int f = off[RB_NEAR] + __popcll(m_near & below);
      if (f < sd.qfcap) {
        const int q = sd.qf0 + f;
        b.qf_i[0][q] = i;
        b.qf_i[1][q] = jg;
        b.qf_d[q] = d64;
        b.qf_vac[q] = 0;
        // near list entry f: the free pair in slot f
        const double req = fmax(fmax(fmax(g.req_t[1], g.req_t[2]), g.req_t[3]),
                                g.req_t[4 + ai + 21 * aj]);
        const double thr = fmin(P.qa_max_req, fmax(p.rep_cutoff, req)) * (1.0 + 1e-5);
        b.qh_slot[q] = f;
        b.qh_i[0][q] = i;
        b.qh_i[1][q] = jg;
        b.qh_vac[q] = 0;
        b.qh_req[q] = req;
        b.qh_thr2[q] = (float)(thr * thr);
        b.qh_ij[q] = (uint32_t)k | ((uint32_t)j << 16);
      }
    }
    if (is_pid) {
      const int pos = off[RB_PID] + __popcll(m_pid & below);
      if (pos < sd.nlcap) {
        const int q = sd.nl0 + pos;
        b.pid_i[0][q] = i;
        b.pid_i[1][q] = jg;
        b.pid_d[q] = d64;
        b.pid_t[q] = (int16_t)(ai * 21 + aj);
      }
    }
    if (is_dh) {
      const int pos = off[RB_DH] + __popcll(m_dh & below);
      if (pos < sd.nlcap) {
        const int q = sd.nl0 + pos;
        b.dh_i[0][q] = i;
        b.dh_i[1][q] = jg;
        b.dh_q[q] = p.charge[ai] * p.charge[aj];
      }
    }
    if (is_pauli) {
      const int pos = off[RB_PAULI] + __popcll(m_pauli & below);
      if (pos < sd.nlcap) {
        const int q = sd.nl0 + pos;
        b.pauli_i[0][q] = i;
        b.pauli_i[1][q] = jg;
      }
    }
  };
  __syncthreads();
  for (int pass = 0; pass < 2; ++pass) {
    int off[RB_N];
    for (int c = 0; c < RB_N; ++c) off[c] = pass ? wcnt[w][c] : 0;
    if (pass && compact) {
      // the wave's NL pairs, 64 at a time, in the first pass's order
      for (int t0 = 0; t0 < my_nl; t0 += 64) {
        const int t = t0 + lane;
        const bool v = t < my_nl;
        const uint32_t kw = v ? (uint32_t)b.nl_i[0][seg + t] : 0u;
        const int cls = v ? (int)b.nl_cls[seg + t] : 0;
        const int k = (int)(kw & 0xffffu), j = (int)(kw >> 16);
        const bool is_nc = cls & CL_NC, near = cls & CL_NEAR, is_pid = cls & CL_PID,
                   is_dh = cls & CL_DH, is_pauli = cls & CL_PAULI;
        const int c = is_nc ? b.nl_i[1][seg + t] : 0;
        const unsigned long long m_nc = __ballot(is_nc), m_near = __ballot(near),
                                 m_pid = __ballot(is_pid), m_dh = __ballot(is_dh),
                                 m_pauli = __ballot(is_pauli);
        if (v) emit(k, j, c, is_nc, near, is_pid, is_dh, is_pauli, m_nc, m_near, m_pid, m_dh,
                    m_pauli, off);
        off[RB_NC] += __popcll(m_nc);
        off[RB_NEAR] += __popcll(m_near);
        off[RB_PID] += __popcll(m_pid);
        off[RB_DH] += __popcll(m_dh);
        off[RB_PAULI] += __popcll(m_pauli);
      }
      continue;
    }
    for (int k = kbeg; k < kend; ++k) {
      const double xi = g.x[k], yi = g.y[k], zi = g.z[k];
      const uint32_t wk = csq[k];
      const int si = (int)(wk & 0xFFFu), ai = (int)(wk >> 24);
      const uint32_t ci = (wk >> 12) & 0xFFFu;
      const int c0 = rs[k], ce = rs[k + 1];
      int tp = c0;  // the row's next native partner
      for (int j0 = k + 1; j0 < n; j0 += 64) {
        const int j = j0 + lane;
        // native partners among columns j0 .. j0 + 63, as a lane mask
        unsigned long long natm = 0ull;
        const int tp0 = tp;
        while (tp < ce) {
          const int pt = pj[tp];
          if (pt > j0 + 63) break;
          natm |= 1ull << (pt - j0);
          ++tp;
        }
        // straight-line for every lane (column n - 1 stands in for those
        // past the row's end); the bitwise & and | keep it branch-free
        const bool valid = j < n;
        const int jj = valid ? j : n - 1;
        v3<double> dd = xsub(v3<double>{g.x[jj], g.y[jj], g.z[jj]}, v3<double>{xi, yi, zi});
        if (cell) dd = xwrapd(dd, sd);
        const float fx = (float)dd.x, fy = (float)dd.y, fz = (float)dd.z;
        const float d2 = fx * fx + fy * fy + fz * fz;
        // columns beyond every threshold and its margin join no list: a
        // chunk of only such columns adds nothing (its native partners are
        // already passed)
        if (!__ballot(valid & (d2 <= e_far))) continue;
        const uint32_t wj = csq[jj];
        const int dseq = abs((int)(wj & 0xFFFu) - si);
        const bool same = ((wj >> 12) & 0xFFFu) == ci;
        const bool four = same & (dseq == 4);
        const bool pair = valid & !(same & (dseq < 3));
        // below a threshold's lower margin the pair is in, above its upper
        // margin out; between the two (lo != hi) the exact distance decides
        bool b_nl = d2 < e_nl[0], b_nc = d2 < e_nc[0], b_dh = d2 < e_dh[0], b_rep = d2 < e_rep[0];
        bool b_near = d2 < e_near[0];
        const bool edge = pair & ((b_nl != (d2 <= e_nl[1])) | (b_nc != (d2 <= e_nc[1])) |
                                  (b_dh != (d2 <= e_dh[1])) | (b_rep != (d2 <= e_rep[1])) |
                                  (b_near != (d2 <= e_near[1])));
        if (__ballot(edge)) {  // rare: the exact FP64 decisions of the pairs in a band
          if (edge) {
            const double dist = xnorm3(dd);
            b_nl = dist < t_nl;
            b_nc = dist < t_nc;
            b_dh = __ddiv_rn(1.0, dist) > rt_dh;
            b_rep = dist < t_rep;
            b_near = !(dist > t_near);
          }
        }
        const bool in = pair & b_nl;
        const bool is_nc = in & b_nc & (bool)((natm >> lane) & 1ull);
        const bool is_qa = in & !is_nc & qa_on & (qa4 | !four) & b_nc;
        const bool near = is_qa & b_near;
        const bool rest = in & !is_nc & !is_qa;
        const bool is_pid = rest & pid_on & (pid4 | !four) & b_nc;
        const int aj = (int)(wj >> 24);
        const bool is_dh = rest & dh_on & (bool)((chg >> ai) & (chg >> aj) & 1u) & b_dh;
        const bool is_pauli = rest & pauli_on & b_rep;
        const unsigned long long m_nl = __ballot(in), m_nc = __ballot(is_nc),
                                 m_qa = __ballot(is_qa), m_near = __ballot(near),
                                 m_pid = __ballot(is_pid), m_dh = __ballot(is_dh),
                                 m_pauli = __ballot(is_pauli);
        if (pass) {
          const int c = is_nc ? sd.nc0 + tp0 + __popcll(natm & below) : 0;
          if (valid) emit(k, j, c, is_nc, near, is_pid, is_dh, is_pauli, m_nc, m_near, m_pid,
                          m_dh, m_pauli, off);
        } else if (compact) {
          const bool keep = is_nc | near | is_pid | is_dh | is_pauli;
          const unsigned long long m_keep = __ballot(keep);
          if (keep) {
            const int t = seg + my_nl + __popcll(m_keep & below);
            b.nl_i[0][t] = (int)((uint32_t)k | ((uint32_t)j << 16));
            b.nl_cls[t] = (uint8_t)((is_nc ? CL_NC : 0) | (near ? CL_NEAR : 0) |
                                    (is_pid ? CL_PID : 0) | (is_dh ? CL_DH : 0) |
                                    (is_pauli ? CL_PAULI : 0));
            if (is_nc) b.nl_i[1][t] = sd.nc0 + tp0 + __popcll(natm & below);
          }
          my_nl += __popcll(m_keep);
        }
        off[RB_NL] += __popcll(m_nl);
        off[RB_NC] += __popcll(m_nc);
        off[RB_QA] += __popcll(m_qa);
        off[RB_NEAR] += __popcll(m_near);
        off[RB_PID] += __popcll(m_pid);
        off[RB_DH] += __popcll(m_dh);
        off[RB_PAULI] += __popcll(m_pauli);
      }
    }
    if (!pass) {
      if (lane == 0)
        for (int c = 0; c < RB_N; ++c) wcnt[w][c] = off[c];
      __syncthreads();
      // the waves' offsets: an exclusive scan over the waves, in wave order
      if (threadIdx.x < RB_N) {
        const int c = threadIdx.x;
        int run = 0;
        for (int v = 0; v < nw; ++v) {
          const int x = wcnt[v][c];
          wcnt[v][c] = run;
          run += x;
        }
        switch (c) {
        case RB_NL: st.nl_n = run; break;
        case RB_NC: st.nc_n = run; break;
        case RB_QA: st.qf_n = run; break;
        case RB_NEAR:
          st.qf_store = run < sd.qfcap ? run : sd.qfcap;
          st.qh_n = st.qf_store;
          if (run > sd.qfcap) st.overflow = 1;
          break;
        case RB_PID:
          st.pid_n = run < sd.nlcap ? run : sd.nlcap;
          if (run > sd.nlcap) st.overflow = 1;
          break;
        case RB_DH:
          st.dh_n = run < sd.nlcap ? run : sd.nlcap;
          if (run > sd.nlcap) st.overflow = 1;
          break;
        case RB_PAULI:
          st.pauli_n = run < sd.nlcap ? run : sd.nlcap;
          if (run > sd.nlcap) st.overflow = 1;
          break;
        }
      }
      __syncthreads();
    }
  }
  __syncthreads();
  // the current native contacts' copied fields (set_current_contact), every
  // slot at once
  for (int q = threadIdx.x; q < st.nc_n; q += blockDim.x)
    set_current_contact(b, sd.nc0 + q, b.ncur[sd.nc0 + q], r0);
  // snapshot (legacy_update.cpp:80-89)
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    b.orig_r[0][r0 + k] = g.x[k];
    b.orig_r[1][r0 + k] = g.y[k];
    b.orig_r[2][r0 + k] = g.z[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    st.orig_pad = pad;
    for (int c = 0; c < 3; ++c) st.orig_cell[c] = sd.cell[c];
    st.nl_invalid = 0;
    st.rebuilds += 1;
  }
  __syncthreads();
}

// Shares of one chain-local item: slot s receives the force on residue
// base + s (s = 0..3), in registers.
template <typename RF> struct AccSlots {
  int base;
  RF c[4][3];
  template <typename T> __device__ __forceinline__ void add(int i, v3<T> f) {
    int s = i - base;
    // explicit selects keep c in registers
    RF fx = (RF)f.x, fy = (RF)f.y, fz = (RF)f.z;
    if (s == 0) { c[0][0] += fx; c[0][1] += fy; c[0][2] += fz; }
    else if (s == 1) { c[1][0] += fx; c[1][1] += fy; c[1][2] += fz; }
    else if (s == 2) { c[2][0] += fx; c[2][1] += fy; c[2][2] += fz; }
    else { c[3][0] += fx; c[3][1] += fy; c[3][2] += fz; }
  }
};

// Per-residue bonded parameters of a chain-local system in LDS: the item
// starting at residue k of each list exists if its flag bit is set.
enum { BF_TETHER = 1, BF_LREP = 2, BF_NAT_ANG = 4, BF_HEUR_ANG = 8, BF_NAT_DIH = 16,
       BF_HEUR_DIH = 32, BF_CHIR = 64 };
struct BondLds {
  // [6][n]: tether length, native angle, native dihedral (phi0 of the simple
  // variant, cos phi0 of the complex one), sin phi0, chiral f0, f1
  double *d;
  uint8_t *u;  // [3][n]: flags, heuristic angle type, heuristic dihedral type
  int n;
  __device__ double &teth(int k) const { return d[k]; }
  __device__ double &ang(int k) const { return d[n + k]; }
  __device__ double &dih(int k) const { return d[2 * n + k]; }
  __device__ double &dihs(int k) const { return d[3 * n + k]; }
  __device__ double &c0(int k) const { return d[4 * n + k]; }
  __device__ double &c1(int k) const { return d[5 * n + k]; }
  __device__ uint8_t &flag(int k) const { return u[k]; }
  __device__ uint8_t &at(int k) const { return u[n + k]; }
  __device__ uint8_t &dt(int k) const { return u[2 * n + k]; }
};

// Once per launch: the static bonded items of every base residue, from HBM
// into LDS (the per-step loop then reads no list of these terms from HBM).
__device__ __forceinline__ void stage_bonded(const DevBatch &b, SysDescC &sd, const BondLds &bl) {
  for (int k = threadIdx.x; k < sd.n; k += blockDim.x) {
    int r = sd.res0 + k, it;
    uint8_t f = 0;
    if ((it = b.bond_item[BOND_TETHER][r]) >= 0) f |= BF_TETHER, bl.teth(k) = b.teth_nat[it];
    if ((it = b.bond_item[BOND_LREP][r]) >= 0) f |= BF_LREP;
    if ((it = b.bond_item[BOND_NAT_ANG][r]) >= 0) f |= BF_NAT_ANG, bl.ang(k) = b.nang_th[it];
    if ((it = b.bond_item[BOND_HEUR_ANG][r]) >= 0) f |= BF_HEUR_ANG, bl.at(k) = b.hang_t[it];
    if ((it = b.bond_item[BOND_NAT_DIH][r]) >= 0) {
      f |= BF_NAT_DIH;
      const double phi0 = b.ndih_ph[it];
      if (b.P->p.nat_dih_variant == 0) {
        double s0, c0;
        sincos(phi0, &s0, &c0);
        bl.dih(k) = c0;
        bl.dihs(k) = s0;
      } else {
        bl.dih(k) = phi0;
      }
    }
    if ((it = b.bond_item[BOND_HEUR_DIH][r]) >= 0) f |= BF_HEUR_DIH, bl.dt(k) = b.hdih_t[it];
    if ((it = b.bond_item[BOND_CHIR][r]) >= 0) {
      f |= BF_CHIR;
      bl.c0(k) = b.chir_f0[it];
      bl.c1(k) = b.chir_f1[it];
    }
    bl.flag(k) = f;
  }
}

template <typename RF> __device__ __forceinline__ RF shfl_up_rf(RF v, int d) {
  return __shfl_up(v, (unsigned)d, 64);
}

// The static bonded terms of a system whose items are all chain-local. Each
// thread takes one base residue per pass (consecutive residues on consecutive
// lanes) and evaluates the items that start there; a residue's force is its
// own share plus the shares lanes 1..3 below it computed for it, gathered
// with cross-lane shifts, and one LDS add per component. Shares that would
// cross into the next wave go to LDS directly.
template <typename RF, bool STAGED>
__device__ __forceinline__ void bonded_by_base(const DevBatch &b, SysDescC &sd, SysState &st,
                                               const GeoL &g, const BondLds &bl, double *fx,
                                               double *fy, double *fz, double *vacc, bool want_v) {
  ParamsC &p = b.P->p;
  const int n = sd.n, r0 = sd.res0, lane = threadIdx.x & 63;
  // energies only on the step that reports them, added per item (no
  // accumulators live across the items: registers are the scarce resource)
  auto ev = [&](int term, double V) {
    if (want_v && V != 0.0) atomicAdd(&vacc[term], V);
  };
  const int dih_term = p.nat_dih_variant ? PCG_SND : PCG_CND;
  for (int base = 0; base < n; base += blockDim.x) {
    int k = base + (int)threadIdx.x;
    AccSlots<RF> a;
    a.base = r0 + k;
    for (int s = 0; s < 4; ++s) a.c[s][0] = a.c[s][1] = a.c[s][2] = (RF)0;
    if (STAGED && k < n) {
      int r = r0 + k;
      uint8_t f = bl.flag(k);
      if (p.tether_on && (f & BF_TETHER)) ev(PCG_TETHER, tether_core<double>(p, g, a, r, r + 1, bl.teth(k)));
      // local repulsion and chirality have no ill-conditioned step: in R
      if (p.lrep_on && (f & BF_LREP)) ev(PCG_LREP, lrep_core<RF>(p, g, a, r, r + 2));
      // the mixed form's FP32 angles on every step but the one that reports
      // energies (want_v, uniform), which takes FP64
      if (sizeof(RF) == 4 && !want_v) {
        if (p.nat_ang_on && (f & BF_NAT_ANG)) ev(PCG_NAT_ANG, nat_ang_mixed(p, g, a, r, r + 1, r + 2, bl.ang(k)));
        if (p.heur_ang_on && (f & BF_HEUR_ANG)) ev(PCG_HEUR_ANG, heur_ang_mixed(p, g, a, r, r + 1, r + 2, bl.at(k)));
      } else {
        if (p.nat_ang_on && (f & BF_NAT_ANG)) ev(PCG_NAT_ANG, nat_ang_core<double>(p, g, a, r, r + 1, r + 2, bl.ang(k)));
        if (p.heur_ang_on && (f & BF_HEUR_ANG)) ev(PCG_HEUR_ANG, heur_ang_core<double>(p, g, a, r, r + 1, r + 2, bl.at(k)));
      }
      if (p.nat_dih_on && (f & BF_NAT_DIH))
        ev(dih_term, nat_dih_core<double>(p, g, a, r, r + 1, r + 2, r + 3, bl.dih(k), bl.dih(k), bl.dihs(k),
                                    p.nat_dih_variant));
      if (p.heur_dih_on && (f & BF_HEUR_DIH))
        ev(PCG_HEUR_DIH, heur_dih_core<double>(p, g, a, r, r + 1, r + 2, r + 3, bl.dt(k)));
      if (p.chir_on && (f & BF_CHIR))
        ev(PCG_CHIR, chir_core<RF>(p, g, a, r, r + 1, r + 2, r + 3, bl.c0(k), bl.c1(k)));
    } else if (!STAGED && k < n) {
      int r = r0 + k, it;
      if (p.tether_on && (it = b.bond_item[BOND_TETHER][r]) >= 0) ev(PCG_TETHER, tether_item<double>(b, g, a, it));
      if (p.lrep_on && (it = b.bond_item[BOND_LREP][r]) >= 0) ev(PCG_LREP, lrep_item<double>(b, g, a, it));
      if (p.nat_ang_on && (it = b.bond_item[BOND_NAT_ANG][r]) >= 0) ev(PCG_NAT_ANG, nat_ang_item<double>(b, g, a, it));
      if (p.heur_ang_on && (it = b.bond_item[BOND_HEUR_ANG][r]) >= 0) ev(PCG_HEUR_ANG, heur_ang_item<double>(b, g, a, it));
      if (p.nat_dih_on && (it = b.bond_item[BOND_NAT_DIH][r]) >= 0)
        ev(dih_term, nat_dih_item<double>(b, g, a, it, p.nat_dih_variant));
      if (p.heur_dih_on && (it = b.bond_item[BOND_HEUR_DIH][r]) >= 0) ev(PCG_HEUR_DIH, heur_dih_item<double>(b, g, a, it));
      if (p.chir_on && (it = b.bond_item[BOND_CHIR][r]) >= 0) ev(PCG_CHIR, chir_item<double>(b, g, a, it));
    }
    RF t[3];
    for (int c = 0; c < 3; ++c) {
      RF s1 = shfl_up_rf(a.c[1][c], 1), s2 = shfl_up_rf(a.c[2][c], 2), s3 = shfl_up_rf(a.c[3][c], 3);
      t[c] = a.c[0][c];
      if (lane >= 1) t[c] += s1;
      if (lane >= 2) t[c] += s2;
      if (lane >= 3) t[c] += s3;
    }
    if (k < n) {
      atomicAdd(fx + k, (double)t[0]);
      atomicAdd(fy + k, (double)t[1]);
      atomicAdd(fz + k, (double)t[2]);
      // shares for residues handled by the next wave (they exist: an item
      // only reaches residues of its own system)
      for (int s = 1; s < 4; ++s)
        if (lane + s >= 64 && (a.c[s][0] != (RF)0 || a.c[s][1] != (RF)0 || a.c[s][2] != (RF)0)) {
          atomicAdd(fx + k + s, (double)a.c[s][0]);
          atomicAdd(fy + k + s, (double)a.c[s][1]);
          atomicAdd(fz + k + s, (double)a.c[s][2]);
        }
    }
  }
}

// Native contacts, NK items per thread per pass, with every item's list
// fields loaded before any is evaluated (their chains of list load -> LDS
// positions -> decisions overlap). The step takes NK = 1: on the ragged
// 256-protein workload NK = 1, 2, 3, 4 and 6 measured 34.9, 35.1, 35.6,
// 36.6 and 37.7 us a step (the larger NK keep more registers live across
// the loop, in a kernel already limited to 128 VGPRs).
template <int NK, typename R>
__device__ __forceinline__ void nat_cont_loop(const DevBatch &b, SysDescC &sd, SysState &st,
                                              const GeoL &g, AccL<double> &acc, const int *sizes,
                                              double *vacc, bool want_v) {
  const int lo = sizes[PCG_NAT_CONT], n = sizes[PCG_NAT_CONT + 1] - lo, B = blockDim.x;
  if (n <= 0) return;
  int k = ((int)threadIdx.x - lo % B + B) % B;
  double v = 0.0;
  ParamsC &p = b.P->p;
  if constexpr (sizeof(R) == 4) {
    if (p.nc_ss_kind == 0) {
      // the FP32 form of nat_cont_core's plain-LJ case, in stages over the NK
      // items so that the items' loads and arithmetic overlap: the geometry
      // and the FP32-first decisions of every item, straight-line; the exact
      // FP64 decisions of the few inside the margins; the forces of every
      // item, straight-line; then the adds and the state transitions
      const float cut2 = (float)(p.nl_cutoff * p.nl_cutoff), lo = 1.0f - 2e-5f, hi = 1.0f + 2e-5f;
      const float depth = (float)p.nat_cont_depth;
      const int r0 = sd.res0;
      for (; k < n; k += NK * B) {
        // three loads an item: the packed residues, type and state; the
        // native length; the FP32 squared transition distance
        uint32_t ijt[NK];
        double nat[NK];
        float th2[NK];
        for (int j = 0; j < NK; ++j) {
          const int kk = sd.nc0 + min(k + j * B, n - 1);
          ijt[j] = b.ncur_ijt[kk];
          nat[j] = b.ncur_nat[kk];
          th2[j] = b.ncur_th2[kk];
        }
        for (int j = 0; j < NK; ++j) {
          keep(ijt[j]);
          keep(nat[j]);
          keep(th2[j]);
        }
        v3<double> d[NK];
        v3<float> r12[NK];
        float rn[NK];
        int now[NK];
        bool act[NK], exact[NK];
        for (int j = 0; j < NK; ++j) {
          const int i1 = r0 + (int)(ijt[j] & 0xFFFu), i2 = r0 + (int)((ijt[j] >> 12) & 0xFFFu);
          d[j] = xwrapd(xsub(g.r(i2), g.r(i1)), sd);
          r12[j] = cast3<float>(d[j]);
          const float d2 = nsq(r12[j]);
          act[j] = k + j * B < n && !(d2 > cut2 * hi);
          exact[j] = !(d2 < cut2 * lo && (d2 < th2[j] * lo || d2 > th2[j] * hi));
          now[j] = d2 < th2[j] * lo ? 1 : 0;
          rn[j] = sq(d2);
        }
        for (int j = 0; j < NK; ++j)
          if (act[j] && exact[j]) {
            const double rnx = xnorm3(d[j]);
            act[j] = rnx < p.nl_cutoff;
            rn[j] = (float)rnx;
            now[j] = rnx <= __dmul_rn(__dmul_rn(nat[j], C216_INV), p.counting_factor) ? 1 : 0;
          }
        float Vj[NK];
        v3<float> f[NK];
        for (int j = 0; j < NK; ++j) {
          float dV;
          ljR<float>(depth, (float)nat[j], rcpR(rn[j]), Vj[j], dV);
          dV = clampR(dV, -1e3f, 1e3f);
          f[j] = (dV * rcpR(rn[j])) * r12[j];
        }
        for (int j = 0; j < NK; ++j)
          if (act[j]) {
            acc.add(r0 + (int)(ijt[j] & 0xFFFu), f[j]);
            acc.add(r0 + (int)((ijt[j] >> 12) & 0xFFFu), -f[j]);
            if (want_v) v += (double)Vj[j];
            if ((uint32_t)now[j] != (ijt[j] >> 31)) {
              // a transition (rare): the contact's index into the list of all
              const int kk = sd.nc0 + k + j * B, c = b.ncur[kk];
              b.nc_active[c] = (uint8_t)now[j];
              b.ncur_t[kk] = (uint8_t)(((ijt[j] >> 24) & 0x7Fu) | ((uint32_t)now[j] << 7));
              b.ncur_ijt[kk] = (ijt[j] & 0x7FFFFFFFu) | ((uint32_t)now[j] << 31);
              if (b.nc_change_t[c] < 0) {
                b.nc_change_t[c] = st.t;
                atomicAdd(&st.num_changed, 1);
              }
            }
          }
      }
      if (want_v && v != 0.0) atomicAdd(&vacc[PCG_NAT_CONT], v);
      return;
    }
  }
  for (; k < n; k += NK * B) {
    NcFetch q[NK];
    for (int j = 0; j < NK; ++j)
      if (k + j * B < n) q[j] = nat_cont_fetch(b, sd.nc0 + k + j * B);
    for (int j = 0; j < NK; ++j)
      if (k + j * B < n) keep_fetch(q[j]);
    for (int j = 0; j < NK; ++j)
      if (k + j * B < n) v += nat_cont_core<R>(b, sd, st, g, acc, q[j], st.t, true);
  }
  if (want_v && v != 0.0) atomicAdd(&vacc[PCG_NAT_CONT], v);
}

// QA near pairs without workgroup barriers. Thread t of the workgroup takes
// near pairs t, t + B, ...; every pair takes the cheap FP32 first test, and
// the ones that pass (about one near pair in seven) are stacked in the
// wave's own 128-entry LDS segment. Whenever 64 are stacked, the wave takes
// the newest 64 through the exact path, one a lane, so that path runs on
// full waves; the last few go after the loop. The next pair is fetched
// before the stack is drained, so its load overlaps the exact path. No wave
// waits for another: the force terms before and after this one flow on. The
// exact path reads nothing from HBM but n and h (a pair's formation
// distance bound comes from the LDS tables, its free-pair slot is loaded
// only if it becomes a candidate), and the order in which candidates are
// found does not matter: qa::finish ranks them by (distance, free-pair
// index).
template <typename R>
__device__ __forceinline__ void qa_cand_waves(const DevBatch &b, SysDescC &sd, SysState &st,
                                              const GeoL &g, AccL<double> &acc, const int *sizes,
                                              uint2 *queue, double *vacc, bool want_v) {
  constexpr int WCAP = 128;  // stacked pairs per wave: fewer than 64 + 64
  static_assert(16 * WCAP * sizeof(uint2) <= LdsLayout::QA_QUEUE * sizeof(int),
                "the QA queue holds every wave's stack");
  const int n = sizes[PCG_QA_CAND + 1] - sizes[PCG_QA_CAND], B = blockDim.x;
  if (n <= 0) return;
  const int tid = threadIdx.x, lane = tid & 63, r0 = sd.res0;
  const unsigned long long below = (1ull << lane) - 1ull;
  uint2 *wq = queue + (tid >> 6) * WCAP;
  auto fetch = [&](int hk) { return hk < n ? qa_cand_fetch_lite(b, sd, hk) : QhLite{0u, -1.0f}; };
  // the exact path's q: every field but the free-pair slot
  auto full = [&](int hk, int i1, int i2) {
    const int a1 = g.atype(b, i1), a2 = g.atype(b, i2);
    double req = fmax(fmax(fmax(g.req_t[1], g.req_t[2]), g.req_t[3]), g.req_t[4 + a1 + 21 * a2]);
    return QhFetch{hk, i1, i2, 0.f, 0, req, -1};
  };
  const int iters = (n + B - 1) / B;  // the same for every thread
  int pend = 0;                       // stacked pairs of this wave (uniform)
  double v = 0.0;
  QhLite cur = fetch(tid);
  for (int t = 0; t <= iters; ++t) {
    if (t < iters) {
      bool act = false;
      if (cur.thr2 >= 0.0f) {  // a vacant pair has thr2 < 0
        const v3<double> d = xwrapd(xsub(g.r(r0 + (int)(cur.ij >> 16)), g.r(r0 + (int)(cur.ij & 0xFFFFu))), sd);
        const float dx = (float)d.x, dy = (float)d.y, dz = (float)d.z;
        act = !(dx * dx + dy * dy + dz * dz > cur.thr2);
      }
      const unsigned long long m = __ballot(act);
      if (act) wq[pend + __popcll(m & below)] = make_uint2((unsigned)(tid + t * B), cur.ij);
      pend += __popcll(m);
      cur = fetch(tid + (t + 1) * B);
    }
    if (pend >= 64 || (t == iters && pend > 0)) {
      const int take = pend < 64 ? pend : 64;
      pend -= take;
      // the lanes' stack writes before the reads of other lanes' entries
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (lane < take) {
        const uint2 w = wq[pend + lane];
        v += qa_cand_exact<R>(b, sd, st, g, acc,
                              full((int)w.x, r0 + (int)(w.y & 0xFFFFu), r0 + (int)(w.y >> 16)));
      }
      // and those reads before the slots take new pairs
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }
  if (want_v && v != 0.0) atomicAdd(&vacc[PCG_QA_CAND], v);
}

// The items of one term: item k is flat index sizes[T] + k, taken by thread
// (sizes[T] + k) mod blockDim, as in one loop over all terms.
template <int T, typename R>
__device__ __forceinline__ void term_loop(const DevBatch &b, SysDescC &sd, SysState &st,
                                          const GeoL &g, AccL<double> &acc, const int *sizes,
                                          double *vacc, bool want_v) {
  const int lo = sizes[T], n = sizes[T + 1] - lo, B = blockDim.x;
  if (n <= 0) return;
  int k = ((int)threadIdx.x - lo % B + B) % B;
  double v = 0.0;
  for (; k < n; k += B) v += eval_item<R>(b, sd, st, g, acc, T, k, want_v);
  if (want_v && v != 0.0) atomicAdd(&vacc[T], v);
}

template <typename R, typename RF>
__device__ __forceinline__ void persistent_body(const DevBatch &b, int s, int nsteps,
                                                JobSlot *slot = nullptr) {
#pragma clang fp contract(off)
  SysDescC &sd = b.sys[s];
  if (!b.st[s].active) return;
  // the system's scalars live in LDS for the launch (only this workgroup
  // touches them) and go back to HBM at its end
  __shared__ SysState st;
  if (threadIdx.x == 0) st = b.st[s];
  __syncthreads();
  DevParamsC &P = *b.P;
  const int n = sd.n, r0 = sd.res0;
  const int64_t target = slot ? slot->target : sd.steps_target;
  extern __shared__ __align__(16) unsigned char lds[];
  const bool staged = sd.bond_regular && b.bond_lds;
  // forces accumulate in FP64 in LDS: on gfx950 an LDS FP64 atomic add
  // (ds_add_f64) runs at integer-atomic speed, an FP32 one far slower
  const bool gear_l = b.gear_lds != 0;
  const LdsLayout lay((size_t)n, sizeof(double), staged, gear_l);
  double *rx = (double *)lds, *ry = rx + n, *rz = ry + n;
  double *fx = (double *)(lds + lay.f), *fy = fx + n, *fz = fy + n;
  BondLds bl{(double *)(lds + lay.bd), (uint8_t *)(lds + lay.bu), n};
  uint8_t *atl = (uint8_t *)(lds + lay.at);      // residue types and chain flags
  uint32_t *syl = (uint32_t *)(lds + lay.sy);     // QA sync budgets
  // 1 / mass and its square root per residue type, for the Langevin step;
  // the QA polarity classes and per-type formation distances
  __shared__ double minv_t[21], sqminv_t[21], req_t[445];
  __shared__ int8_t ptype_t[21];
  if (threadIdx.x < 21) {
    minv_t[threadIdx.x] = 1.0 / P.p.mass[threadIdx.x];
    sqminv_t[threadIdx.x] = sqrt(minv_t[threadIdx.x]);
    ptype_t[threadIdx.x] = (int8_t)P.p.ptype[threadIdx.x];
  }
  for (int t = threadIdx.x; t < 445; t += blockDim.x)
    req_t[t] = __dmul_rn(P.p.qa_lj_r_high[t], __dadd_rn(1.0, P.p.qa_formation_tol));
  uint2 *qa_queue = (uint2 *)(lds + lay.q);
  double *ys = (double *)(lds + lay.g);  // [15][n]: y1..y5 x, y, z (gear_l)
  __shared__ double vacc[NT];
  __shared__ int sizes[NT + 1];
  __shared__ int flag_rebuild, stop;
  __shared__ double wave_max[16];

  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    rx[k] = b.y[0][0][r0 + k];
    ry[k] = b.y[0][1][r0 + k];
    rz[k] = b.y[0][2][r0 + k];
    // chain neighbours are always the adjacent residues (i - 1, i + 1)
    atl[k] = (uint8_t)(b.atype[r0 + k] | (b.prev[r0 + k] >= 0 ? 0x40 : 0) |
                       (b.next[r0 + k] >= 0 ? 0x80 : 0));
    syl[k] = *(const uint32_t *)(b.sync + 4 * (r0 + k));
  }
  GeoL g{rx, ry, rz, r0, atl, syl, ptype_t, req_t};
  AccL<double> acc{fx, fy, fz, r0};
  if (staged) stage_bonded(b, sd, bl);
  if (gear_l)
    for (int k = threadIdx.x; k < n; k += blockDim.x)
      for (int j = 0; j < 15; ++j) ys[j * n + k] = b.y[1 + j / 3][j % 3][r0 + k];
  __syncthreads();

  __shared__ int want_v, last;
  // term sizes of the flat index space
  auto set_sizes = [&]() {
    int acc_n = 0;
    for (int t = 0; t < NT; ++t) {
      sizes[t] = acc_n;
      acc_n += term_on(P, t) ? term_size(sd, st, t) : 0;
    }
    sizes[NT] = acc_n;
  };
  // The prologue of step it, by thread 0 at the end of the step before (one
  // barrier a step): the stop test, the energy flag, nl::verify and, without
  // a rebuild, the term sizes.
  auto prologue = [&](int it) {
    stop = st.step >= target;
    // the last step this launch makes for the system: energies (unless
    // off) and the write-back of its forces
    last = (it == nsteps - 1) || (st.step + 1 >= target);
    want_v = !b.energy_off && last;
    // nl::verify (cg/src/cg/nl/verify.cpp:3-34)
    if (!stop && P.nl_required) {
      if (st.verify_first) {
        st.verify_first = 0;
        st.nl_invalid = 1;
      } else {
        double maxd = __longlong_as_double((long long)st.max_disp_bits);
        double box = fabs(sd.cell[0] - st.orig_cell[0]) + fabs(sd.cell[1] - st.orig_cell[1]) +
                     fabs(sd.cell[2] - st.orig_cell[2]);
        st.nl_invalid = __dadd_rn(__dmul_rn(2.0, maxd), box) > st.orig_pad ? 1 : 0;
      }
    }
    flag_rebuild = st.nl_invalid;
    for (int k = 0; k < NT; ++k) vacc[k] = 0.0;
    if (!flag_rebuild) set_sizes();
  };
  // forces start at zero; afterwards each step's Langevin pass clears them
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    fx[k] = 0;
    fy[k] = 0;
    fz[k] = 0;
  }
  if (P.p.qa_on) wg_prepare_nh(b, g, n, r0);
  if (threadIdx.x == 0) prologue(0);
  __syncthreads();
  for (int it = 0; it < nsteps; ++it) {
    // a finished system keeps the forces of its last evaluation
    if (stop) break;
    // (n and h of qa::prepare_nh were computed after the positions last moved)
    if (flag_rebuild) {
      wg_fix_nl(b, sd, st, g, (void *)(lds + lay.q));
      if (threadIdx.x == 0) set_sizes();
      __syncthreads();
    }
    // every term's items over one flat index space (term order preserved,
    // a thread moves on to the next term without waiting); one loop per term
    // so each keeps only its own constants in registers
    const bool wv = want_v;
    if (staged) {
      bonded_by_base<RF, true>(b, sd, st, g, bl, fx, fy, fz, vacc, wv);
    } else if (sd.bond_regular) {
      bonded_by_base<RF, false>(b, sd, st, g, bl, fx, fy, fz, vacc, wv);
    } else {
      term_loop<PCG_CHIR, R>(b, sd, st, g, acc, sizes, vacc, wv);
      term_loop<PCG_TETHER, R>(b, sd, st, g, acc, sizes, vacc, wv);
      term_loop<PCG_LREP, R>(b, sd, st, g, acc, sizes, vacc, wv);
      term_loop<PCG_NAT_ANG, R>(b, sd, st, g, acc, sizes, vacc, wv);
      term_loop<PCG_HEUR_ANG, R>(b, sd, st, g, acc, sizes, vacc, wv);
      term_loop<PCG_HEUR_DIH, R>(b, sd, st, g, acc, sizes, vacc, wv);
      term_loop<PCG_CND, R>(b, sd, st, g, acc, sizes, vacc, wv);
      term_loop<PCG_SND, R>(b, sd, st, g, acc, sizes, vacc, wv);
    }
    // three slots a thread: the largest proteins (some 2600 contacts on 1024
    // threads) fill them in one pass
    nat_cont_loop<1, R>(b, sd, st, g, acc, sizes, vacc, wv);
    term_loop<PCG_PAULI, R>(b, sd, st, g, acc, sizes, vacc, wv);
    term_loop<PCG_CONST_DH, R>(b, sd, st, g, acc, sizes, vacc, wv);
    term_loop<PCG_REL_DH, R>(b, sd, st, g, acc, sizes, vacc, wv);
    if (P.p.qa_on) qa_cand_waves<R>(b, sd, st, g, acc, sizes, qa_queue, vacc, wv);
    term_loop<PCG_QA_CONT, R>(b, sd, st, g, acc, sizes, vacc, wv);
    term_loop<PCG_PID, R>(b, sd, st, g, acc, sizes, vacc, wv);
    term_loop<PCG_VEL_AFM, R>(b, sd, st, g, acc, sizes, vacc, wv);
    term_loop<PCG_FORCE_AFM, R>(b, sd, st, g, acc, sizes, vacc, wv);
    term_loop<PCG_SOLID_WALL, R>(b, sd, st, g, acc, sizes, vacc, wv);
    __syncthreads();
    if (P.p.qa_on) qa_finish_block(b, sd, st, (int8_t *)syl, r0);
    // lang::step (cg/src/cg/langevin/step.cpp:23-77); each residue's force,
    // once used, goes to HBM on the launch's last step and restarts at 0
    double maxd = 0.0;
    const bool wl = last;
    auto flush_force = [&](int k) {
      if (wl) {
        ((RF *)b.F[0])[r0 + k] = (RF)fx[k];
        ((RF *)b.F[1])[r0 + k] = (RF)fy[k];
        ((RF *)b.F[2])[r0 + k] = (RF)fz[k];
      }
      fx[k] = 0;
      fy[k] = 0;
      fz[k] = 0;
    };
    if (P.p.lang_on) {
      for (int k = threadIdx.x; k < n; k += blockDim.x) {
        int i = r0 + k;
        // the rebuild snapshot is loaded first: its latency hides behind
        // the noise and the corrector
        const double o[3] = {b.orig_r[0][i], b.orig_r[1][i], b.orig_r[2][i]};
        double nz[3];
        philox_normals(sd.philox_key, st.step, k, nz);
        double r[3], vel[3];
        const int aa = atl[k] & 63;  // the type, without the chain flags
        const double Fx = fx[k], Fy = fy[k], Fz = fz[k];
        flush_force(k);
        if (gear_l) {
          gear_core(b, sd, st, minv_t[aa], sqminv_t[aa], Fx, Fy, Fz, nz, r, vel,
                    [&](int kk, int c) -> double & {
                      return kk == 0 ? (c == 0 ? rx[k] : c == 1 ? ry[k] : rz[k])
                                     : ys[((kk - 1) * 3 + c) * n + k];
                    });
          // a force tip averages its residue's velocity in the next
          // evaluation, which reads it from HBM
          if (sd.nftip > 0)
            for (int c = 0; c < 3; ++c) b.vel[c][i] = vel[c];
        } else {
          gear_core(b, sd, st, minv_t[aa], sqminv_t[aa], Fx, Fy, Fz, nz, r, vel,
                    [&](int kk, int c) -> double & { return b.y[kk][c][i]; });
          for (int c = 0; c < 3; ++c) b.vel[c][i] = vel[c];
        }
        rx[k] = r[0];
        ry[k] = r[1];
        rz[k] = r[2];
        maxd = fmax(maxd, disp2_from(sd, o, r));  // squared; the root is taken once below
      }
    } else {
      for (int k = threadIdx.x; k < n; k += blockDim.x) flush_force(k);
    }
    maxd = block_max_to_first(maxd, wave_max);
    // every position has moved: the next step's n and h
    if (P.p.qa_on) wg_prepare_nh(b, g, n, r0);
    if (threadIdx.x == 0) {
      st.max_disp_bits = (unsigned long long)__double_as_longlong(__dsqrt_rn(maxd));
      if (want_v) {
        double v = 0.0;
        for (int k = 0; k < NT; ++k) {
          st.V[k] = vacc[k];
          v += vacc[k];
        }
        st.Vtot = v;
      }
      if (P.p.lang_on) {
        st.true_t += P.dt;
        st.t = st.true_t;
        st.step += 1;
      }
      prologue(it + 1);
    }
    __syncthreads();
  }
  // scalars and the forces of the last evaluation back to HBM
  __syncthreads();
  if (threadIdx.x == 0) {
    b.st[s] = st;
    if (slot) {
      slot->step = st.step;
      slot->Vtot = st.Vtot;
      slot->overflow = st.overflow;
    }
  }
  if (gear_l)
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
      b.y[0][0][r0 + k] = rx[k];
      b.y[0][1][r0 + k] = ry[k];
      b.y[0][2][r0 + k] = rz[k];
      for (int j = 0; j < 15; ++j) b.y[1 + j / 3][j % 3][r0 + k] = ys[j * n + k];
      for (int c = 0; c < 3; ++c) b.vel[c][r0 + k] = ys[c * n + k] * P.dt_inv;
    }
  if (P.p.qa_on)
    for (int k = threadIdx.x; k < n; k += blockDim.x) *(uint32_t *)(b.sync + 4 * (r0 + k)) = syl[k];
}

// The launch bound sets the register budget: 1024 threads leave 128 VGPRs
// per lane (the bonded items then spill), 768 leave 168, 512 leave 256.
template <typename R, typename RF, int MAXT>
__global__ __launch_bounds__(MAXT) void k_persistent(DevBatch b, int nsteps) {
  persistent_body<R, RF>(b, blockIdx.x, nsteps);
}

// served form: one resident job (a one-system batch) per workgroup
template <typename R, typename RF, int MAXT>
__global__ __launch_bounds__(MAXT) void k_persistent_jobs(const DevBatch *jobs, JobSlot *slots,
                                                          int nsteps) {
  persistent_body<R, RF>(jobs[blockIdx.x], 0, nsteps, slots + blockIdx.x);
}

template <int MAXT>
void launch_jobs_t(const DevBatch *jobs, JobSlot *slots, int njobs, int prec, int block, int nsteps,
                   size_t lds_bytes, hipStream_t s) {
  if (prec == PCG_FP64)
    hipLaunchKernelGGL((k_persistent_jobs<double, double, MAXT>), dim3(njobs), dim3(block),
                       lds_bytes, s, jobs, slots, nsteps);
  else
    hipLaunchKernelGGL((k_persistent_jobs<float, float, MAXT>), dim3(njobs), dim3(block),
                       lds_bytes, s, jobs, slots, nsteps);
}

template <int MAXT>
void launch_batch_t(const DevBatch &b, int prec, int block, int nsteps, size_t lds_bytes,
                    hipStream_t s) {
  if (prec == PCG_FP64)
    hipLaunchKernelGGL((k_persistent<double, double, MAXT>), dim3(b.nsys), dim3(block), lds_bytes,
                       s, b, nsteps);
  else
    hipLaunchKernelGGL((k_persistent<float, float, MAXT>), dim3(b.nsys), dim3(block), lds_bytes,
                       s, b, nsteps);
}

void launch_persistent_jobs(const DevBatch *jobs, JobSlot *slots, int njobs, int prec, int block,
                            int nsteps, size_t lds_bytes, hipStream_t s) {
  if (block > 768) launch_jobs_t<1024>(jobs, slots, njobs, prec, block, nsteps, lds_bytes, s);
  else if (block > 512) launch_jobs_t<768>(jobs, slots, njobs, prec, block, nsteps, lds_bytes, s);
  else launch_jobs_t<512>(jobs, slots, njobs, prec, block, nsteps, lds_bytes, s);
}

void launch_persistent(const DevBatch &b, int prec, int block, int nsteps, size_t lds_bytes,
                       hipStream_t s) {
  if (block > 768) launch_batch_t<1024>(b, prec, block, nsteps, lds_bytes, s);
  else if (block > 512) launch_batch_t<768>(b, prec, block, nsteps, lds_bytes, s);
  else launch_batch_t<512>(b, prec, block, nsteps, lds_bytes, s);
}

} // namespace pcg
