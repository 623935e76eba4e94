// pcg_kernels.hip -- grid-wide kernels of the per-step loop and their
// launchers. Each stage of thread::advance_by_step
// (cg/src/cg/simul/thread_main.cpp:373-536) is one launch over all systems of
// a batch; stages that the reference runs conditionally (NL rebuild and list
// updates) exit on the device when not needed, so that a step never waits on
// the host and can be captured in a hipGraph.
#include "pcg_launch.hpp"
#include "pcg_stages.hpp"

namespace pcg {

// ------------------------------------------------------------ begin step ----
// st.V zeroed; NL verify decision (nl::verify, cg/src/cg/nl/verify.cpp:3-34)
// from the displacement maximum the previous Langevin step left behind.
__global__ void k_begin(DevBatch b) {
#pragma clang fp contract(off)
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= b.nsys) return;
  SysState &st = b.st[s];
  SysDescC &d = b.sys[s];
  if (!st.active) return;
  for (int k = 0; k < NT; ++k) st.V[k] = 0.0;
  if (b.P->nl_required) {
    if (st.verify_first) {
      st.verify_first = 0;
      st.nl_invalid = 1;
    } else {
      double maxd = __longlong_as_double((long long)st.max_disp_bits);
      double box = fabs(d.cell[0] - st.orig_cell[0]) + fabs(d.cell[1] - st.orig_cell[1]) +
                   fabs(d.cell[2] - st.orig_cell[2]);
      double disp = __dadd_rn(__dmul_rn(2.0, maxd), box);
      st.nl_invalid = disp > st.orig_pad ? 1 : 0;
    }
  }
  st.max_disp_bits = 0ull;
}

template <typename RF> __global__ void k_zero_forces(DevBatch b) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b.ntot) return;
  if (!b.st[b.sysid[i]].active) return;
  ((RF *)b.F[0])[i] = 0;
  ((RF *)b.F[1])[i] = 0;
  ((RF *)b.F[2])[i] = 0;
}

// ------------------------------------------------------------ NL rebuild ----
// nl::legacy_update (cg/src/cg/nl/legacy_update.cpp:11-55): all pairs i < j
// of a system, same-chain |seq difference| < 3 skipped, kept if the wrapped
// distance is below max(nl.cutoff, dh.cutoff) + pad. One wavefront per row;
// rows come out in (i1, i2) order, which is the reference's sorted order.
__device__ __forceinline__ bool nl_pred(const DevBatch &b, SysDescC &s, const GeoG &g,
                                        int i, int j, double req, double &dist) {
  if (b.chain[i] == b.chain[j]) {
    int d = b.seq[i] - b.seq[j];
    if (d < 0) d = -d;
    if (d < 3) return false;
  }
  v3<double> dd = xwrapd(xsub(g.r(j), g.r(i)), s);
  dist = xnorm3(dd);
  return dist < req;
}

template <bool WRITE> __global__ void k_nl_rows(DevBatch b) {
  int row = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  int lane = threadIdx.x & 63;
  if (row >= b.ntot) return;
  int s = b.sysid[row];
  SysState &st = b.st[s];
  if (!st.active || !st.nl_invalid) return;
  SysDescC &sd = b.sys[s];
  GeoG g{b.y[0][0], b.y[0][1], b.y[0][2]};
  double req = b.P->nl_eff_cutoff + b.P->p.nl_pad;
  int end = sd.res0 + sd.n;
  int out = WRITE ? b.row_cnt[row] : 0;
  int cnt = 0;
  for (int j0 = row + 1; j0 < end; j0 += 64) {
    int j = j0 + lane;
    double dist = 0.0;
    bool keep = j < end && nl_pred(b, sd, g, row, j, req, dist);
    unsigned long long m = __ballot(keep);
    if (WRITE && keep) {
      int pos = out + cnt + __popcll(m & ((1ull << lane) - 1ull));
      if (pos < sd.nlcap) {
        int slot = sd.nl0 + pos;
        b.nl_i[0][slot] = row;
        b.nl_i[1][slot] = j;
        b.nl_d[slot] = dist;
      }
    }
    cnt += __popcll(m);
  }
  if (!WRITE && lane == 0) b.row_cnt[row] = cnt;
}

// Exclusive scan of the row counts of one system (one block per system).
__global__ void k_nl_scan(DevBatch b) {
  int s = blockIdx.x;
  SysState &st = b.st[s];
  if (!st.active || !st.nl_invalid) return;
  SysDescC &sd = b.sys[s];
  __shared__ int carry;
  __shared__ int wsum[BLK / 64];
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < sd.n; base += blockDim.x) {
    int i = base + threadIdx.x;
    int v = i < sd.n ? b.row_cnt[sd.res0 + i] : 0;
    // inclusive wave scan
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
      int y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int off = carry;
    for (int k = 0; k < w; ++k) off += wsum[k];
    if (i < sd.n) b.row_cnt[sd.res0 + i] = off + x - v;
    __syncthreads();
    if (threadIdx.x == 0) {
      int tot = 0;
      for (int k = 0; k < (int)(blockDim.x >> 6); ++k) tot += wsum[k];
      carry += tot;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    st.nl_n = carry;
    if (carry > sd.nlcap) {
      st.overflow = 1;
      st.nl_n = sd.nlcap;
    }
    // the list updates filter with the new pad (legacy_update.cpp:84)
    st.orig_pad = b.P->p.nl_pad;
    for (int c = 0; c < 3; ++c) st.orig_cell[c] = sd.cell[c];
  }
}

// Snapshot after a rebuild: orig_r, orig_pad, orig_pbc (legacy_update.cpp:80-89)
__global__ void k_nl_snapshot(DevBatch b) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b.ntot) return;
  int s = b.sysid[i];
  const SysState &st = b.st[s];
  if (!st.active || !st.nl_invalid) return;
  for (int c = 0; c < 3; ++c) b.orig_r[c][i] = b.y[0][c][i];
}

// ------------------------------------------------------------ list updates ----
// The ordered updates of thread::fix_nl_async (thread_main.cpp:409-423):
// native contacts take first (nat_cont/update_contacts.cpp:7-37), then QA
// free pairs (qa/update_free_pairs.cpp:4-25), PID bundles and DH pairs see the
// rest without taking (pid/update_bundles.cpp:4-34, dh/update_pairs.cpp:4-25),
// Pauli takes last (pauli/update_pairs.cpp:4-20). Each NL pair's destination
// depends only on the pair, so the partition is one classification pass plus
// an order-preserving compaction per list.
__device__ __forceinline__ void block_counts5(int cls, int valid, int counts[5], int ranks[5]) {
  // ordered rank of this thread within the block for each class, and totals
  __shared__ int wc[BLK / 64][5];
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

// blockIdx.x: chunk of BLK pairs, blockIdx.y: system
__global__ void k_part_count(DevBatch b) {
  int s = blockIdx.y;
  const SysState &st = b.st[s];
  if (!st.active || !st.nl_invalid) return;
  SysDescC &sd = b.sys[s];
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  int nchunks = (sd.nlcap + BLK - 1) / BLK;
  if ((int)blockIdx.x >= nchunks) return;
  int valid = k < st.nl_n;
  int cls = valid ? classify(b, sd, st, sd.nl0 + k) : 0;
  if (valid) b.nl_cls[sd.nl0 + k] = (uint8_t)cls;
  int counts[5], ranks[5];
  block_counts5(cls, valid, counts, ranks);
  if (threadIdx.x == 0) {
    int base = (sd.nl0 / BLK + blockIdx.x) * 5;
    for (int c = 0; c < 5; ++c) b.blk_cnt[base + c] = counts[c];
  }
}

__global__ void k_part_scan(DevBatch b) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= b.nsys) return;
  SysState &st = b.st[s];
  if (!st.active || !st.nl_invalid) return;
  SysDescC &sd = b.sys[s];
  int nchunks = (st.nl_n + BLK - 1) / BLK;
  int tot[5] = {0, 0, 0, 0, 0};
  for (int ch = 0; ch < nchunks; ++ch) {
    int base = (sd.nl0 / BLK + ch) * 5;
    for (int c = 0; c < 5; ++c) {
      int v = b.blk_cnt[base + c];
      b.blk_cnt[base + c] = tot[c];
      tot[c] += v;
    }
  }
  st.nc_n = tot[0];
  st.qf_n = tot[1] < sd.qfcap ? tot[1] : sd.qfcap;
  st.qf_store = st.qf_n;
  if (tot[1] > sd.qfcap) st.overflow = 1;
  st.pid_n = tot[2];
  st.dh_n = tot[3];
  st.pauli_n = tot[4];
}

__global__ void k_part_write(DevBatch b) {
  int s = blockIdx.y;
  SysState &st = b.st[s];
  if (!st.active || !st.nl_invalid) return;
  SysDescC &sd = b.sys[s];
  int nchunks = (sd.nlcap + BLK - 1) / BLK;
  if ((int)blockIdx.x >= nchunks) return;
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  int valid = k < st.nl_n;
  int slot = sd.nl0 + k;
  int cls = valid ? b.nl_cls[slot] : 0;
  int counts[5], ranks[5];
  block_counts5(cls, valid, counts, ranks);
  if (!valid || !cls) return;
  int base = (sd.nl0 / BLK + blockIdx.x) * 5;
  int i1 = b.nl_i[0][slot], i2 = b.nl_i[1][slot];
  double d = b.nl_d[slot];
  ParamsC &p = b.P->p;
  if (cls & CL_NC) {
    int pos = b.blk_cnt[base + 0] + ranks[0];
    set_current_contact(b, sd.nc0 + pos, find_contact(b, sd, i1, i2), sd.res0);
  }
  if (cls & CL_QA) {
    int pos = b.blk_cnt[base + 1] + ranks[1];
    if (pos < sd.qfcap) {
      int q = sd.qf0 + pos;
      b.qf_i[0][q] = i1;
      b.qf_i[1][q] = i2;
      b.qf_d[q] = d;
      b.qf_vac[q] = 0;
    }
  }
  if (cls & CL_PID) {
    int q = sd.nl0 + b.blk_cnt[base + 2] + ranks[2];
    b.pid_i[0][q] = i1;
    b.pid_i[1][q] = i2;
    b.pid_d[q] = d;
    b.pid_t[q] = (int16_t)(b.atype[i1] * 21 + b.atype[i2]);
  }
  if (cls & CL_DH) {
    int q = sd.nl0 + b.blk_cnt[base + 3] + ranks[3];
    b.dh_i[0][q] = i1;
    b.dh_i[1][q] = i2;
    b.dh_q[q] = p.charge[b.atype[i1]] * p.charge[b.atype[i2]];
  }
  if (cls & CL_PAULI) {
    int q = sd.nl0 + b.blk_cnt[base + 4] + ranks[4];
    b.pauli_i[0][q] = i1;
    b.pauli_i[1][q] = i2;
  }
}

// QA free pairs within the maximum required distance (one block per system).
__global__ void k_qa_near(DevBatch b) {
  int s = blockIdx.x;
  SysState &st = b.st[s];
  if (!st.active || !st.nl_invalid || !b.P->p.qa_on) return;
  build_qa_near(b, b.sys[s], st);
}

// The rebuild is complete: record it and clear nl_invalid.
__global__ void k_nl_done(DevBatch b) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= b.nsys) return;
  SysState &st = b.st[s];
  if (!st.active || !st.nl_invalid) return;
  st.nl_invalid = 0;
  st.rebuilds += 1;
}

// ------------------------------------------------------------------ terms ----
// Every enabled term in one launch: blockIdx.z is the term, blockIdx.y the
// system, blockIdx.x strides over the term's list.
template <typename R, typename RF> __global__ void k_terms(DevBatch b, int only_term, int from, int to) {
  int term = only_term >= 0 ? only_term : (int)blockIdx.z;
  int s = blockIdx.y;
  SysState &st = b.st[s];
  if (!st.active || !term_on(*b.P, term)) return;
  SysDescC &sd = b.sys[s];
  int n = term_size(sd, st, term);
  int lo = from > 0 ? from : 0, hi = (to >= 0 && to < n) ? to : n;
  GeoG g{b.y[0][0], b.y[0][1], b.y[0][2]};
  AccG<RF> acc{(RF *)b.F[0], (RF *)b.F[1], (RF *)b.F[2]};
  double v = 0.0;
  for (int k = lo + blockIdx.x * blockDim.x + threadIdx.x; k < hi; k += gridDim.x * blockDim.x)
    v += eval_item<R>(b, sd, st, g, acc, term, k);
  v = block_sum(v);
  if (threadIdx.x == 0 && v != 0.0) atomicAdd(&st.V[term], v);
}

// ------------------------------------------------------------------- QA ----
__global__ void k_prepare_nh(DevBatch b) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b.ntot) return;
  if (!b.st[b.sysid[i]].active) return;
  GeoG g{b.y[0][0], b.y[0][1], b.y[0][2]};
  prepare_nh_item(b, g, i);
}

__global__ void k_qa_finish(DevBatch b) {
  int s = blockIdx.x;
  SysState &st = b.st[s];
  if (!st.active || !b.P->p.qa_on) return;
  qa_finish_block(b, b.sys[s], st, b.sync, 0);
}

template <typename RF> __global__ void k_lang(DevBatch b, int noise_mode) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b.ntot) return;
  int s = b.sysid[i];
  SysState &st = b.st[s];
  if (!st.active || !b.P->p.lang_on) return;
  SysDescC &sd = b.sys[s];
  double nz[3];
  if (noise_mode == PCG_NOISE_RAN2) {
    nz[0] = b.noise[0][i];
    nz[1] = b.noise[1][i];
    nz[2] = b.noise[2][i];
  } else {
    philox_normals(sd.philox_key, st.step, b.local[i], nz);
  }
  double r[3];
  gear_update(b, sd, st, i, (double)((RF *)b.F[0])[i], (double)((RF *)b.F[1])[i],
              (double)((RF *)b.F[2])[i], nz, r);
  double disp = disp_from_orig(b, sd, i, r);
  atomicMax(&st.max_disp_bits, (unsigned long long)__double_as_longlong(disp));
}

// time and energy bookkeeping at the end of a step (step.cpp:157-162)
__global__ void k_tick(DevBatch b, int advance) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= b.nsys) return;
  SysState &st = b.st[s];
  if (!st.active) return;
  double v = 0.0;
  for (int k = 0; k < NT; ++k) v += st.V[k];
  st.Vtot = v;
  if (advance && b.P->p.lang_on) {
    st.true_t += b.P->dt;
    st.t = st.true_t;
    st.step += 1;
  }
}

// ------------------------------------------------------------- launchers ----
static inline dim3 grid1(int n, int blk = BLK) { return dim3((unsigned)((n + blk - 1) / blk)); }

void launch_begin(const DevBatch &b, hipStream_t s) {
  hipLaunchKernelGGL(k_begin, grid1(b.nsys, 64), dim3(64), 0, s, b);
}
void launch_zero_forces(const DevBatch &b, int prec, hipStream_t s) {
  if (prec == PCG_FP64)
    hipLaunchKernelGGL(k_zero_forces<double>, grid1(b.ntot), dim3(BLK), 0, s, b);
  else
    hipLaunchKernelGGL(k_zero_forces<float>, grid1(b.ntot), dim3(BLK), 0, s, b);
}
void launch_prepare_nh(const DevBatch &b, hipStream_t s) {
  hipLaunchKernelGGL(k_prepare_nh, grid1(b.ntot), dim3(BLK), 0, s, b);
}
void launch_fix_nl(const DevBatch &b, int max_nlcap, hipStream_t s) {
  int rows_per_block = BLK / 64;
  dim3 rg((unsigned)((b.ntot + rows_per_block - 1) / rows_per_block));
  hipLaunchKernelGGL(k_nl_rows<false>, rg, dim3(BLK), 0, s, b);
  hipLaunchKernelGGL(k_nl_scan, dim3(b.nsys), dim3(BLK), 0, s, b);
  hipLaunchKernelGGL(k_nl_rows<true>, rg, dim3(BLK), 0, s, b);
  hipLaunchKernelGGL(k_nl_snapshot, grid1(b.ntot), dim3(BLK), 0, s, b);
  dim3 pg((unsigned)((max_nlcap + BLK - 1) / BLK), (unsigned)b.nsys);
  hipLaunchKernelGGL(k_part_count, pg, dim3(BLK), 0, s, b);
  hipLaunchKernelGGL(k_part_scan, grid1(b.nsys, 64), dim3(64), 0, s, b);
  hipLaunchKernelGGL(k_part_write, pg, dim3(BLK), 0, s, b);
  hipLaunchKernelGGL(k_qa_near, dim3(b.nsys), dim3(BLK), 0, s, b);
  hipLaunchKernelGGL(k_nl_done, grid1(b.nsys, 64), dim3(64), 0, s, b);
}
void launch_terms(const DevBatch &b, int prec, int max_items, int only_term, int from, int to,
                  hipStream_t s) {
  int gx = (max_items + BLK - 1) / BLK;
  if (gx < 1) gx = 1;
  if (gx > 512) gx = 512;
  dim3 g((unsigned)gx, (unsigned)b.nsys, only_term >= 0 ? 1u : (unsigned)NT);
  if (prec == PCG_FP64)
    hipLaunchKernelGGL((k_terms<double, double>), g, dim3(BLK), 0, s, b, only_term, from, to);
  else
    hipLaunchKernelGGL((k_terms<float, float>), g, dim3(BLK), 0, s, b, only_term, from, to);
}
void launch_qa_finish(const DevBatch &b, hipStream_t s) {
  hipLaunchKernelGGL(k_qa_finish, dim3(b.nsys), dim3(BLK), 0, s, b);
}
void launch_lang(const DevBatch &b, int prec, int noise_mode, hipStream_t s) {
  if (prec == PCG_FP64)
    hipLaunchKernelGGL(k_lang<double>, grid1(b.ntot), dim3(BLK), 0, s, b, noise_mode);
  else
    hipLaunchKernelGGL(k_lang<float>, grid1(b.ntot), dim3(BLK), 0, s, b, noise_mode);
}
void launch_tick(const DevBatch &b, int advance, hipStream_t s) {
  hipLaunchKernelGGL(k_tick, grid1(b.nsys, 64), dim3(64), 0, s, b, advance);
}

} // namespace pcg
