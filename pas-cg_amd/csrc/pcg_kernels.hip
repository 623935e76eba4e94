// pcg_kernels.hip -- grid-wide kernels of the per-step loop and their
// launchers. Each stage of thread::advance_by_step
// (cg/src/cg/simul/thread_main.cpp:373-536) is one launch over all systems of
// a batch; stages that the reference runs conditionally (NL rebuild and list
// updates) exit on the device when not needed, so that a step never waits on
// the host and can be captured in a hipGraph.
#include <algorithm>

#include "pcg_launch.hpp"
#include "pcg_stages.hpp"

namespace pcg {

// ------------------------------------------------------------ begin step ----
// st.V zeroed; NL verify decision (nl::verify, cg/src/cg/nl/verify.cpp:3-34)
// from the displacement maximum the previous Langevin step left behind.
__device__ __forceinline__ void begin_system(const DevBatch &b, int s) {
#pragma clang fp contract(off)
  SysState &st = b.st[s];
  SysDescC &d = b.sys[s];
  if (!st.active) return;
  for (int k = 0; k < NT; ++k) st.V[k] = 0.0;
  if (b.wall)  // dynamics::reset (dynamics.cpp:11-19) of the wall forces
    for (int w = 0; w < 6; ++w)
      for (int c = 0; c < 3; ++c) b.wall->wall_F[w][c] = 0.0;
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

__global__ void k_begin(DevBatch b) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= b.nsys) return;
  begin_system(b, s);
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

// Completion count of a per-item stage kernel: every lane with `mine` (an
// item of system `sys`) counts itself toward the system's residue count in
// the SysState field `done`; a wave whose counted lanes share one system
// adds them with one atomic. Returns, in the lane whose add completed a
// system, that system (-1 elsewhere); what every lane wrote before the call
// is visible to it then.
__device__ __forceinline__ int count_done(const DevBatch &b, int SysState::*done, bool mine,
                                          int sys) {
  // a wave with nothing to count skips the fence: an agent-scope release
  // writes back the L2, and on a step without a rebuild every wave of the
  // cell count has nothing to count
  if (!__any(mine)) return -1;
  __threadfence();
  const int lane = threadIdx.x & 63;
  const unsigned long long m = __ballot(mine);
  const int sys0 = __shfl(sys, 0, 64);
  const bool uni = __all(!mine || sys == sys0) && ((m & 1ull) != 0ull);
  int who = -1, add = 0;
  if (uni) {
    if (lane == 0) who = sys0, add = __popcll(m);
  } else if (mine) {
    who = sys, add = 1;
  }
  if (who < 0) return -1;
  const int old = atomicAdd(&(b.st[who].*done), add);
  if (old + add != b.sys[who].n) return -1;
  __threadfence();
  return who;
}

// Inclusive scan of x over the wave's 64 lanes: four DPP row shifts give
// each 16-lane row its own prefix (a lane shifted in from outside the row
// reads 0), three readlanes carry the rows' totals across. Every lane of the
// wave must be active; `total` is the wave's sum.
__device__ __forceinline__ int wave_incl(int x, int &total) {
  x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);  // row_shr:1
  x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);  // row_shr:2
  x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);  // row_shr:4
  x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);  // row_shr:8
  const int r0 = __builtin_amdgcn_readlane(x, 15), r1 = __builtin_amdgcn_readlane(x, 31);
  const int r2 = __builtin_amdgcn_readlane(x, 47), r3 = __builtin_amdgcn_readlane(x, 63);
  const int row = (threadIdx.x & 63) >> 4;
  x += (row > 0 ? r0 : 0) + (row > 1 ? r1 : 0) + (row > 2 ? r2 : 0);
  total = r0 + r1 + r2 + r3;
  return x;
}

// One wave's exclusive scan of v[0..n) in place, also into also[0..n) when
// given; returns the total in every lane. The wave walks v in stripes of 64
// consecutive entries, loading 16 stripes at once (coalesced, all in flight
// together) and carrying the running sum from stripe to stripe. Every lane
// of the wave must be active.
__device__ __forceinline__ int wave_scan(int *v, int n, int *also) {
  constexpr int SB = 16;
  const int lane = threadIdx.x & 63;
  int carry = 0;
  for (int base = 0; base < n; base += 64 * SB) {
    int x[SB];
#pragma unroll
    for (int j = 0; j < SB; ++j) {
      const int k = base + j * 64 + lane;
      x[j] = k < n ? v[k] : 0;
    }
#pragma unroll
    for (int j = 0; j < SB; ++j) {
      int tot;
      const int inc = wave_incl(x[j], tot);
      const int k = base + j * 64 + lane;
      if (k < n) {
        v[k] = carry + inc - x[j];
        if (also) also[k] = carry + inc - x[j];
      }
      carry += tot;
    }
  }
  return carry;
}

// ------------------------------------------------------------- cell grid ----
// A large periodic system can take its NL rows from a cell grid instead of
// every later residue: the periodic cell cut into cg[a] >= 3 cells a side,
// each at least the NL radius wide (with a relative 1e-6 to spare), so every
// pair within the radius sits in neighbouring cells (the 27 of the stencil,
// periodic). A row's candidates from those cells go into a bitmask over the
// system's residues, which the row then walks in increasing column order:
// the same pairs, tested by the same predicate, in the same order as the
// scan over all columns. When the cell has shrunk below that (the box
// protocol), the rebuild scans every column.
__device__ __forceinline__ bool use_cells(const DevBatch &b, SysDescC &sd) {
  if (sd.cg[0] <= 0) return false;
  const double req = (b.P->nl_eff_cutoff + b.P->p.nl_pad) * (1.0 + 1e-6);
  for (int a = 0; a < 3; ++a)
    if (!(sd.cell[a] >= req * (double)sd.cg[a])) return false;
  return true;
}
__device__ __forceinline__ int cell_coord(SysDescC &sd, int a, double x) {
  // the position wrapped into [0, cell), then its cell along the axis
  const double u = x - sd.cell[a] * floor(x * sd.cell_inv[a]);
  const int k = (int)(u * sd.cell_inv[a] * (double)sd.cg[a]);
  return k < 0 ? 0 : (k >= sd.cg[a] ? sd.cg[a] - 1 : k);
}

// each residue's cell; cell_end counts the residues of each cell; the wave
// whose residue completes a system turns its counts into the cells' first
// slots (cell_start) and fill pointers (cell_end)
__global__ void k_cell_count(DevBatch b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool mine = false;
  int sys = -1;
  if (i < b.ntot) {
    sys = b.sysid[i];
    const SysState &st = b.st[sys];
    SysDescC &sd = b.sys[sys];
    if (st.active && st.nl_invalid && use_cells(b, sd)) {
      mine = true;
      const int c = (cell_coord(sd, 2, b.y[0][2][i]) * sd.cg[1] + cell_coord(sd, 1, b.y[0][1][i])) *
                        sd.cg[0] + cell_coord(sd, 0, b.y[0][0][i]);
      b.cell_of[i] = c;
      atomicAdd(&b.cell_end[sd.cell0 + c], 1);
    }
  }
  const int done = count_done(b, &SysState::done_cells, mine, sys);
  unsigned long long dm = __ballot(done >= 0);
  while (dm) {
    const int src = __ffsll((long long)dm) - 1;
    dm &= dm - 1ull;
    const int ds = __shfl(done, src, 64);
    SysDescC &sd = b.sys[ds];
    wave_scan(b.cell_end + sd.cell0, sd.cg[0] * sd.cg[1] * sd.cg[2], b.cell_start + sd.cell0);
    if ((threadIdx.x & 63) == 0) b.st[ds].done_cells = 0;
  }
}

// the residues into their cells' slots (any order within a cell: a row's
// candidates go through a bitmask); cell_end ends at each cell's end
__global__ void k_cell_fill(DevBatch b) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b.ntot) return;
  const int s = b.sysid[i];
  const SysState &st = b.st[s];
  if (!st.active || !st.nl_invalid) return;
  SysDescC &sd = b.sys[s];
  if (!use_cells(b, sd)) return;
  const int slot = atomicAdd(&b.cell_end[sd.cell0 + b.cell_of[i]], 1);
  b.cell_items[sd.res0 + slot] = i - sd.res0;
}

// cell_words: 64-bit mask words of the largest system on the grid (the
// dynamic LDS holds one row mask a wave), 0 without a grid
// The rows' counts of system s into their first slots, by one whole wave;
// then the list's length and the rebuild's snapshot of the pad and cell.
__device__ __forceinline__ void nl_scan_tail(const DevBatch &b, int s) {
  SysState &st = b.st[s];
  SysDescC &sd = b.sys[s];
  const int tot = wave_scan(b.row_cnt + sd.res0, sd.n, nullptr);
  if ((threadIdx.x & 63) == 0) {
    st.nl_n = tot > sd.nlcap ? sd.nlcap : tot;
    if (tot > sd.nlcap) st.overflow = 1;
    // the list updates filter with the new pad (legacy_update.cpp:84)
    st.orig_pad = b.P->p.nl_pad;
    for (int c = 0; c < 3; ++c) st.orig_cell[c] = sd.cell[c];
  }
}

template <bool WRITE> __global__ void k_nl_rows(DevBatch b, int cell_words, int fuse_scan) {
  extern __shared__ unsigned long long row_mask[];
  int row = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  int lane = threadIdx.x & 63;
  if (row >= b.ntot) return;
  int s = b.sysid[row];
  SysState &st = b.st[s];
  if (!st.active || !st.nl_invalid) return;
  SysDescC &sd = b.sys[s];
  GeoG g{b.y[0][0], b.y[0][1], b.y[0][2]};
  // the snapshot of the rebuild's positions (legacy_update.cpp:80-89)
  if (WRITE && lane == 0)
    for (int c = 0; c < 3; ++c) b.orig_r[c][row] = b.y[0][c][row];
  double req = b.P->nl_eff_cutoff + b.P->p.nl_pad;
  int end = sd.res0 + sd.n;
  int out = WRITE ? b.row_cnt[row] : 0;
  int cnt = 0;
  if (cell_words > 0 && use_cells(b, sd)) {
    unsigned long long *m = row_mask + (size_t)(threadIdx.x >> 6) * cell_words;
    const int il = row - sd.res0, W = (sd.n + 63) >> 6, w0 = (il + 1) >> 6;
    for (int w = w0 + lane; w < W; w += 64) m[w] = 0ull;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int c = b.cell_of[row], gx = sd.cg[0], gy = sd.cg[1], gz = sd.cg[2];
    const int cx = c % gx, cy = (c / gx) % gy, cz = c / (gx * gy);
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int nx = (cx + dx + gx) % gx, ny = (cy + dy + gy) % gy, nz = (cz + dz + gz) % gz;
          const int nc = sd.cell0 + (nz * gy + ny) * gx + nx;
          const int t1 = b.cell_end[nc];
          for (int t = b.cell_start[nc] + lane; t < t1; t += 64) {
            const int j = b.cell_items[sd.res0 + t];
            if (j > il) atomicOr(&m[j >> 6], 1ull << (j & 63));
          }
        }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int wb = w0; wb < W; wb += 64) {
      const int w = wb + lane;
      const unsigned long long bits = w < W ? m[w] : 0ull;
      unsigned long long nz = __ballot(bits != 0ull);
      while (nz) {
        const int k = __ffsll((long long)nz) - 1;
        nz &= nz - 1ull;
        const unsigned lo = __shfl((unsigned)bits, k, 64), hi = __shfl((unsigned)(bits >> 32), k, 64);
        const unsigned long long bw = (unsigned long long)lo | ((unsigned long long)hi << 32);
        const int j = sd.res0 + (wb + k) * 64 + lane;
        double dist = 0.0;
        const bool keep = ((bw >> lane) & 1ull) && nl_pred(b, sd, g, row, j, req, dist);
        const unsigned long long mm = __ballot(keep);
        if (WRITE && keep) {
          int pos = out + cnt + __popcll(mm & ((1ull << lane) - 1ull));
          if (pos < sd.nlcap) {
            int slot = sd.nl0 + pos;
            b.nl_i[0][slot] = row;
            b.nl_i[1][slot] = j;
            b.nl_d[slot] = dist;
          }
        }
        cnt += __popcll(mm);
      }
    }
  } else {
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
  }
  if (!WRITE) {
    // the row counts; with fuse_scan the wave of the system's last row to
    // finish scans them into the rows' first slots, else k_nl_scan does
    if (lane == 0) b.row_cnt[row] = cnt;
    if (!fuse_scan) return;
    int fin = 0;
    if (lane == 0) {
      __threadfence();
      fin = atomicAdd(&st.done_rows, 1) == sd.n - 1;
    }
    if (__shfl(fin, 0, 64)) {
      __threadfence();
      nl_scan_tail(b, s);
      if (lane == 0) st.done_rows = 0;
    }
  }
}

// The row scan as a launch of its own, one wave a system: for systems of
// ROWS_FUSE_MAX residues or more. Every row's count comes to its system's
// counter through an agent-scope fence, and on this device such a fence
// writes back and invalidates the L2: in the trace of a 6400-residue box a
// rebuild's count pass took 131 us with the fused scan and about 20 us
// without it.
__global__ void k_nl_scan(DevBatch b) {
  const int s = blockIdx.x;
  if (s >= b.nsys) return;
  const SysState &st = b.st[s];
  if (!st.active || !st.nl_invalid) return;
  nl_scan_tail(b, s);
}

// ------------------------------------------------------------ list updates ----
// The ordered updates of thread::fix_nl_async (thread_main.cpp:409-423):
// native contacts take first (nat_cont/update_contacts.cpp:7-37), then QA
// free pairs (qa/update_free_pairs.cpp:4-25), PID bundles and DH pairs see the
// rest without taking (pid/update_bundles.cpp:4-34, dh/update_pairs.cpp:4-25),
// Pauli takes last (pauli/update_pairs.cpp:4-20). Each NL pair's destination
// depends only on the pair, so the partition is one classification pass plus
// an order-preserving compaction per list.
__device__ __forceinline__ void block_counts6(int cls, int valid, int counts[NCL], int ranks[NCL]) {
  // ordered rank of this thread within the block for each class, and totals
  __shared__ int wc[BLK / 64][NCL];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  unsigned long long lt = (1ull << lane) - 1ull;
  int mine[NCL];
  for (int c = 0; c < NCL; ++c) {
    unsigned long long m = __ballot(valid && (cls & (1 << c)));
    mine[c] = __popcll(m & lt);
    if (lane == 0) wc[w][c] = __popcll(m);
  }
  __syncthreads();
  for (int c = 0; c < NCL; ++c) {
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

// One block of system s: every class's per-chunk counts become their exclusive
// prefix in chunk order, and the totals the lists' sizes. Each thread takes
// a contiguous run of chunks; a scan over the threads gives each run its
// offsets.
__device__ __forceinline__ void part_scan_block(const DevBatch &b, int s) {
  SysState &st = b.st[s];
  SysDescC &sd = b.sys[s];
  const int nchunks = (st.nl_n + BLK - 1) / BLK, T = blockDim.x;
  const int per = (nchunks + T - 1) / T;
  const int c0 = min((int)threadIdx.x * per, nchunks), c1 = min(c0 + per, nchunks);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = T >> 6;
  __shared__ int wsum[BLK / 64][NCL];
  int run[NCL], tot[NCL];
  for (int c = 0; c < NCL; ++c) run[c] = 0;
  for (int ch = c0; ch < c1; ++ch)
    for (int c = 0; c < NCL; ++c) run[c] += b.blk_cnt[(sd.nl0 / BLK + ch) * NCL + c];
  for (int c = 0; c < NCL; ++c) {
    int x = run[c];
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w][c] = x;
    run[c] = x - run[c];  // exclusive within the wave
  }
  __syncthreads();
  for (int c = 0; c < NCL; ++c) {
    int off = 0, t = 0;
    for (int k = 0; k < nw; ++k) {
      if (k < w) off += wsum[k][c];
      t += wsum[k][c];
    }
    run[c] += off;
    tot[c] = t;
  }
  for (int ch = c0; ch < c1; ++ch)
    for (int c = 0; c < NCL; ++c) {
      const int at = (sd.nl0 / BLK + ch) * NCL + c, v = b.blk_cnt[at];
      b.blk_cnt[at] = run[c];
      run[c] += v;
    }
  if (threadIdx.x == 0) {
    st.nc_n = tot[0];
    st.qf_n = tot[1] < sd.qfcap ? tot[1] : sd.qfcap;
    st.qf_store = st.qf_n;
    if (tot[1] > sd.qfcap) st.overflow = 1;
    st.pid_n = tot[2];
    st.dh_n = tot[3];
    st.pauli_n = tot[4];
    st.qh_n = tot[5] < sd.qfcap ? tot[5] : sd.qfcap;
  }
}

// One chunk of BLK pairs of system s: each pair's classes, and the chunk's
// count per class
__device__ __forceinline__ void part_count_chunk(DevBatch &b, SysDescC &sd, const SysState &st,
                                                 int ch) {
  int k = ch * BLK + (int)threadIdx.x;
  int valid = k < st.nl_n;
  int cls = valid ? classify(b, sd, st, sd.nl0 + k) : 0;
  if (valid) b.nl_cls[sd.nl0 + k] = (uint8_t)cls;
  int counts[NCL], ranks[NCL];
  block_counts6(cls, valid, counts, ranks);
  if (threadIdx.x == 0) {
    int base = (sd.nl0 / BLK + ch) * NCL;
    for (int c = 0; c < NCL; ++c) b.blk_cnt[base + c] = counts[c];
  }
}

// blockIdx.y: system; the blocks stride over the chunks the list fills, so
// a step without a rebuild costs a small grid's launch
__global__ void k_part_count(DevBatch b) {
  int s = blockIdx.y;
  SysState &st = b.st[s];
  if (!st.active || !st.nl_invalid) return;
  SysDescC &sd = b.sys[s];
  // only the blocks with chunks or cells to visit take part (and count
  // themselves below)
  const int nchunks = (st.nl_n + BLK - 1) / BLK;
  const int nc = sd.cg[0] > 0 ? sd.cg[0] * sd.cg[1] * sd.cg[2] : 0;
  const int nb = max(1, min((int)gridDim.x, max(nchunks, (nc + BLK - 1) / BLK)));
  if ((int)blockIdx.x >= nb) return;
  // the cell counts start the next rebuild at zero (the rows are written)
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += nb * blockDim.x)
    b.cell_end[sd.cell0 + c] = 0;
  for (int ch = blockIdx.x; ch < nchunks; ch += nb) part_count_chunk(b, sd, st, ch);
  // the system's last block to finish scans the chunk counts
  __shared__ int last;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    last = atomicAdd(&st.done_part, 1) == nb - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  part_scan_block(b, s);
  if (threadIdx.x == 0) st.done_part = 0;
}

// One chunk's pairs into their lists at the chunk's offsets (the scan of
// k_part_count)
__device__ __forceinline__ void part_write_chunk(DevBatch &b, SysDescC &sd, const SysState &st,
                                                 int ch) {
  int k = ch * BLK + (int)threadIdx.x;
  int valid = k < st.nl_n;
  int slot = sd.nl0 + k;
  int cls = valid ? b.nl_cls[slot] : 0;
  int counts[NCL], ranks[NCL];
  block_counts6(cls, valid, counts, ranks);
  if (!valid || !cls) return;
  int base = (sd.nl0 / BLK + ch) * NCL;
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
      if (cls & CL_NEAR) {
        // near list entry: the free pair in slot pos (its order is the
        // free pairs' order)
        const int h = sd.qf0 + b.blk_cnt[base + 5] + ranks[5];
        b.qh_slot[h] = pos;
        b.qh_i[0][h] = i1;
        b.qh_i[1][h] = i2;
        b.qh_vac[h] = 0;
        qa_near_filters(b, i1, i2, b.qh_req[h], b.qh_thr2[h]);
        b.qh_ij[h] = (uint32_t)(i1 - sd.res0) | ((uint32_t)(i2 - sd.res0) << 16);
      }
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


__global__ void k_part_write(DevBatch b) {
  int s = blockIdx.y;
  const SysState &st = b.st[s];
  if (!st.active || !st.nl_invalid) return;
  SysDescC &sd = b.sys[s];
  const int nchunks = (st.nl_n + BLK - 1) / BLK;
  for (int ch = blockIdx.x; ch < nchunks; ch += gridDim.x) part_write_chunk(b, sd, st, ch);
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

// The first kernel of a full step, over residues: each system's first
// residue runs the system's begin (energies and the walls' reaction forces
// cleared, nl::verify), every residue clears its force and, with QA on,
// computes its n and h (k_begin, k_zero_forces and k_prepare_nh in one
// launch; none reads what another writes)
template <typename RF> __global__ void k_step_begin(DevBatch b, int qa_on) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b.ntot) return;
  const int s = b.sysid[i];
  if (b.local[i] == 0) begin_system(b, s);
  if (!b.st[s].active) return;
  ((RF *)b.F[0])[i] = 0;
  ((RF *)b.F[1])[i] = 0;
  ((RF *)b.F[2])[i] = 0;
  if (qa_on) {
    GeoG g{b.y[0][0], b.y[0][1], b.y[0][2]};
    prepare_nh_item(b, g, i);
  }
}

// The serial part of the finish (removals, then acceptance in order) reads
// and writes the sync budgets of each candidate's residues: with
// sync_lds the system's budgets sit in LDS for the kernel (one word a
// residue, loaded and stored coalesced), and up to QF_CAND_LDS candidates
// are staged there in acceptance order, so the serial loop waits on LDS
// rather than on a chain of HBM loads.
__global__ void k_qa_finish(DevBatch b, int sync_lds) {
  extern __shared__ __align__(16) unsigned char qf_lds[];
  int s = blockIdx.x;
  SysState &st = b.st[s];
  if (!st.active || !b.P->p.qa_on) return;
  SysDescC &sd = b.sys[s];
  if (!sync_lds) {
    qa_finish_block(b, sd, st, b.sync, 0);
    return;
  }
  // nothing removed and no candidate: the block's own early return, before
  // any staging
  if (st.cand_n == 0 && st.rem_n == 0) return;
  const int n = sd.n, r0 = sd.res0;
  uint32_t *syl = (uint32_t *)qf_lds;
  QfLds ql;
  ql.key = (long long *)(qf_lds + 4 * (((size_t)sync_lds + 1) & ~(size_t)1));
  ql.sd = (uint64_t *)(ql.key + QF_CAND_LDS);
  ql.idx = (int *)(ql.sd + QF_CAND_LDS);
  ql.sc = ql.idx + QF_CAND_LDS;
  ql.i1 = ql.sc + QF_CAND_LDS;
  ql.i2 = ql.i1 + QF_CAND_LDS;
  ql.cap = QF_CAND_LDS;
  for (int k = threadIdx.x; k < n; k += blockDim.x) syl[k] = *(const uint32_t *)(b.sync + 4 * (r0 + k));
  __syncthreads();
  qa_finish_block(b, sd, st, (int8_t *)syl, r0, &ql);
  for (int k = threadIdx.x; k < n; k += blockDim.x) *(uint32_t *)(b.sync + 4 * (r0 + k)) = syl[k];
}

// time and energy bookkeeping at the end of a step (step.cpp:157-162)
__device__ __forceinline__ void tick_system(const DevBatch &b, int s, int advance, int nl_done) {
  SysState &st = b.st[s];
  // a full step's rebuild, if it made one, is complete (k_nl_done)
  if (nl_done && st.nl_invalid) {
    st.nl_invalid = 0;
    st.rebuilds += 1;
  }
  double v = 0.0;
  for (int k = 0; k < NT; ++k) v += st.V[k];
  st.Vtot = v;
  if (advance && b.P->p.lang_on) {
    st.true_t += b.P->dt;
    st.t = st.true_t;
    st.step += 1;
  }
}

__global__ void k_tick(DevBatch b, int advance, int nl_done) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= b.nsys) return;
  if (!b.st[s].active) return;
  tick_system(b, s, advance, nl_done);
}


// The displacement maximum goes to the system's state with one atomic per
// wave when the wave's residues belong to one system (a system of the plain
// context has thousands of residues: one atomic per residue on one address
// serialised the whole kernel), else one per residue.
template <typename RF> __global__ void k_lang(DevBatch b, int noise_mode, int tick) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int s = -1, sys = -1;
  bool mine = false;   // a residue of an active system
  double disp = -1.0;  // below every displacement: lanes without a residue
  if (i < b.ntot) {
    s = sys = b.sysid[i];
    SysState &st = b.st[s];
    mine = st.active != 0;
    if (st.active && b.P->p.lang_on) {
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
      disp = disp_from_orig(b, sd, i, r);
    } else {
      s = -1;
    }
  }
  const int s0 = __shfl(s, 0, 64);
  if (__all(s == s0)) {
    double m = disp;
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0 && s0 >= 0)
      atomicMax(&b.st[s0].max_disp_bits, (unsigned long long)__double_as_longlong(m));
  } else if (s >= 0) {
    atomicMax(&b.st[s].max_disp_bits, (unsigned long long)__double_as_longlong(disp));
  }
  // the step's tick (k_tick) by the residue that completes its system, once
  // every residue's Langevin work, and with it every read of the system's
  // time, is done
  if (tick) {
    const int done = count_done(b, &SysState::done_lang, mine, sys);
    if (done >= 0) {
      tick_system(b, done, 1, 1);
      b.st[done].done_lang = 0;
    }
  }
}


// ------------------------------------------------------------------ walls ----
// wall::lj::process_candidates (cg/src/cg/wall/lj/process_candidates.cpp:5-31)
// in one workgroup: the released slots leave (their residue unconnected, the
// plane's count down) and the live slots close up in order; then the
// candidates, sorted by distance (ties by residue, then plane: the order
// sift_free emits them in), connect one by one while their plane is below its
// limit and their residue is not connected, with the bead at the residue's
// projection on the plane. As in the reference a new connection does not
// mark its residue connected. The candidates are consumed.
constexpr int LJW_SORT_MAX = 4096;
__global__ void __launch_bounds__(1024) k_ljw_process(DevBatch b) {
#pragma clang fp contract(off)
  WallDev &W = *b.wall;
  SysState &st = b.st[0];
  ParamsC &p = b.P->p;
  const int tid = threadIdx.x, B = blockDim.x;
  __shared__ double key[LJW_SORT_MAX];
  __shared__ int val[LJW_SORT_MAX];
  __shared__ int scan[1024];
  __shared__ int base;
  const int n_lj = st.n_lj;
  // releases
  for (int k = tid; k < n_lj; k += B)
    if (W.lj_gone[k]) {
      W.is_conn[W.lj_res[k]] = 0;
      atomicSub(&W.num_conns[W.lj_wall[k]], 1);
    }
  if (tid == 0) base = 0;
  __syncthreads();
  // ordered compaction, one chunk of B slots at a time (reads before writes)
  for (int c0 = 0; c0 < n_lj; c0 += B) {
    const int k = c0 + tid;
    const bool keep = k < n_lj && !W.lj_gone[k];
    int res = 0, wl = 0;
    double o0 = 0, o1 = 0, o2 = 0, sat = 0;
    if (keep) {
      res = W.lj_res[k], wl = W.lj_wall[k], sat = W.lj_sat[k];
      o0 = W.lj_off[0][k], o1 = W.lj_off[1][k], o2 = W.lj_off[2][k];
    }
    scan[tid] = keep ? 1 : 0;
    __syncthreads();
    for (int off = 1; off < B; off <<= 1) {
      int v = tid >= off ? scan[tid - off] : 0;
      __syncthreads();
      scan[tid] += v;
      __syncthreads();
    }
    const int pos = base + scan[tid] - (keep ? 1 : 0);
    const int total = scan[B - 1];
    __syncthreads();
    if (keep) {
      W.lj_res[pos] = res, W.lj_wall[pos] = wl, W.lj_sat[pos] = sat;
      W.lj_off[0][pos] = o0, W.lj_off[1][pos] = o1, W.lj_off[2][pos] = o2;
      W.lj_gone[pos] = 0;
    }
    __syncthreads();
    if (tid == 0) base += total;
    __syncthreads();
  }
  int live = base;
  for (int k = live + tid; k < n_lj; k += B) W.lj_gone[k] = 0;
  // candidates by (distance, residue, plane)
  int nc = W.n_cand < W.cand_cap ? W.n_cand : W.cand_cap;
  if (nc > LJW_SORT_MAX) {
    if (tid == 0) W.overflow = 1;
    nc = LJW_SORT_MAX;
  }
  int m = 1;
  while (m < nc) m <<= 1;
  for (int k = tid; k < m; k += B) {
    key[k] = k < nc ? W.cand_d[k] : __builtin_inf();
    val[k] = k < nc ? k : -1;
  }
  __syncthreads();
  auto before = [&](int a, int c) {  // entry a sorts before entry c
    if (key[a] != key[c]) return key[a] < key[c];
    const int va = val[a], vc = val[c];
    if (va < 0 || vc < 0) return vc < 0 && va >= 0;
    const int ra = W.cand_res[va], rc = W.cand_res[vc];
    if (ra != rc) return ra < rc;
    return W.cand_wall[va] < W.cand_wall[vc];
  };
  for (int size = 2; size <= m; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = tid; i < m; i += B) {
        const int j = i ^ stride;
        if (j > i) {
          const bool up = (i & size) == 0;
          if (before(j, i) == up) {
            const double tk = key[i];
            key[i] = key[j], key[j] = tk;
            const int tv = val[i];
            val[i] = val[j], val[j] = tv;
          }
        }
      }
      __syncthreads();
    }
  if (tid == 0) {
    for (int a = 0; a < nc; ++a) {
      const int c = val[a], res = W.cand_res[c], w = W.cand_wall[c];
      if (W.is_conn[res] || W.num_conns[w] >= p.wall_limit[w]) continue;
      if (live >= W.lj_cap) {
        W.overflow = 1;
        break;
      }
      const double nx = p.wall_normal[w][0], ny = p.wall_normal[w][1], nz = p.wall_normal[w][2];
      const double ox = p.wall_origin[w][0], oy = p.wall_origin[w][1], oz = p.wall_origin[w][2];
      const double rx = b.y[0][0][res], ry = b.y[0][1][res], rz = b.y[0][2][res];
      const double sd = nx * (rx - ox) + ny * (ry - oy) + nz * (rz - oz);
      W.lj_res[live] = res, W.lj_wall[live] = w, W.lj_sat[live] = 0.0, W.lj_gone[live] = 0;
      W.lj_off[0][live] = (rx - sd * nx) - ox;
      W.lj_off[1][live] = (ry - sd * ny) - oy;
      W.lj_off[2][live] = (rz - sd * nz) - oz;
      ++live;
      ++W.num_conns[w];
    }
    W.n_cand = 0;
    W.n_gone = 0;
    st.n_lj = live;
  }
}

// ------------------------------------------------------------- box protocol ----
// state::reinit_wall_values (cg/src/cg/simul/state.cpp:708-729): each plane
// on a box side goes to the centre plus the projection on its normal of the
// vector to its face; shift = new origin - old; periodic axes take the box
// extent as their PBC cell
__device__ void box_reinit(WallDev &W, DevParams *P, SysDesc *sd) {
#pragma clang fp contract(off)
  double center[3], ext[3];
  for (int c = 0; c < 3; ++c) {
    center[c] = (W.box_max[c] + W.box_min[c]) / 2.0;
    ext[c] = W.box_max[c] - W.box_min[c];
  }
  for (int side = 0; side < 6; ++side) {
    const int w = W.side_wall[side];
    if (w < 0) continue;
    double d[3], *o = P->p.wall_origin[w];
    const double *n = P->p.wall_normal[w];
    for (int c = 0; c < 3; ++c) d[c] = (side % 2 == 1 ? W.box_max[c] : W.box_min[c]) - center[c];
    const double proj = d[0] * n[0] + d[1] * n[1] + d[2] * n[2];
    for (int c = 0; c < 3; ++c) {
      const double next = center[c] + n[c] * proj;
      W.shift[w][c] = next - o[c];
      o[c] = next;
    }
  }
  for (int c = 0; c < 3; ++c) {
    const double cell = W.bx.pbc[c] ? ext[c] : 0.0;
    sd->cell[c] = cell;
    sd->cell_inv[c] = cell != 0.0 ? 1.0 / cell : 0.0;
  }
}

// state::adjust_wall_pos (state.cpp:702-706) along z, or with the same
// change on every axis
__device__ __forceinline__ void box_adjust(WallDev &W, DevParams *P, SysDesc *sd, const double *size_change) {
#pragma clang fp contract(off)
  for (int c = 0; c < 3; ++c) {
    W.box_min[c] += -size_change[c] + 0.0;
    W.box_max[c] += size_change[c] + 0.0;
  }
  box_reinit(W, P, sd);
}

// state::setup_walls' box part: averages reset, planes to the faces
__global__ void k_box_setup(DevBatch b, DevParams *P, SysDesc *sd) {
  if (threadIdx.x != 0) return;
  WallDev &W = *b.wall;
  for (int w = 0; w < 6; ++w) {
    W.avg_F[w] = MAvg{};
    W.avg_F[w].window = W.bx.avg_forces_over;
    W.work[w] = 0.0;
  }
  W.avg_z = MAvg{};
  W.avg_z.window = W.bx.avg_forces_over;
  box_reinit(W, P, sd);
}

// The phase logic of thread_main.cpp:144-330 before a step: the phase
// changes of this time (which take no step in the reference) and the wall
// move of the phase that then advances.
__global__ void k_box_phase(DevBatch b, DevParams *P, SysDesc *sd) {
#pragma clang fp contract(off)
  if (threadIdx.x != 0) return;
  WallDev &W = *b.wall;
  const SysState &st = b.st[0];
  if (!st.active || W.phase == PCG_PHASE_NONE) return;
  const pcg_box &bx = W.bx;
  const double t = st.t, dt = P->p.dt, num_res = (double)sd->n;
  const double pi = 3.14159265358979323846, pi_2 = 1.57079632679489661923;
  for (int guard = 0; guard < 16; ++guard) {
    switch (W.phase) {
    case PCG_PHASE_SQUEEZING: {  // squeezing_step, thread_main.cpp:144-181
      if (!bx.squeezing) {
        W.since = t;
        W.phase = PCG_PHASE_FIND_FORCE_MIN;
        continue;
      }
      const double ex = W.box_max[0] - W.box_min[0], ey = W.box_max[1] - W.box_min[1],
                   ez = W.box_max[2] - W.box_min[2];
      const double cur_vol = ex * ey * ez, target_vol = num_res / bx.target_density;
      if (cur_vol <= target_vol) {
        W.until = fmin(t + bx.rest_period, bx.total_time);
        W.phase = PCG_PHASE_REST_AFTER_SQUEEZING;
        continue;
      }
      double vel;
      if (cur_vol > 2.0 * target_vol) {
        const double accel_time = bx.accel_dist / bx.vel_above_2V;
        vel = fmin((t - W.since) / accel_time, 1.0) * bx.vel_above_2V;
      } else {
        const double accel_time = bx.accel_dist / bx.target_vel;
        vel = fmin((t - W.since) / accel_time, 1.0) * bx.target_vel;
      }
      const double sh = 0.5 * vel * dt, ch[3] = {-sh, -sh, -sh};
      box_adjust(W, P, sd, ch);
      return;
    }
    case PCG_PHASE_REST_AFTER_SQUEEZING:  // :183-191
      if (t < W.until) return;
      W.phase = PCG_PHASE_FIND_FORCE_MIN;
      continue;
    case PCG_PHASE_FIND_FORCE_MIN: {  // :193-222
      if (!bx.force_min) {
        W.phase = PCG_PHASE_FREEFORM;
        continue;
      }
      if (W.avg_z.has && fabs(W.avg_z.avg[0]) < 5e-2 + 5e-5 * num_res) {
        W.until = t + bx.rest_period;
        W.phase = PCG_PHASE_REST_AFTER_FORCE_MIN;
        continue;
      }
      const double accel_time = bx.accel_dist / bx.target_vel;
      const double time_vel_frac = fmin((t - W.since) / accel_time, 1.0);
      double fvf = (W.avg_z.has ? W.avg_z.avg[0] : 0.0) / bx.force_for_max_vel;
      fvf = copysign(fmin(fabs(fvf), 1.0), fvf);
      const double vel = bx.target_vel * time_vel_frac * fvf;
      const double sh = 0.5 * vel * dt, ch[3] = {0.0, 0.0, sh};
      box_adjust(W, P, sd, ch);
      return;
    }
    case PCG_PHASE_REST_AFTER_FORCE_MIN:  // :224-243
      if (t < W.until) return;
      if (!bx.amp_relative) {
        W.amplitude = bx.amp_abs;
      } else {
        W.shear = bx.osc_shear != 0;
        const double base_val = W.shear ? W.box_max[0] - W.box_min[0] : W.box_max[2] - W.box_min[2];
        W.amplitude = bx.amp_rel * base_val;
      }
      W.since = t;
      W.phase = PCG_PHASE_MAX_AMPLITUDE;
      continue;
    case PCG_PHASE_MAX_AMPLITUDE: {  // :245-276
      if (!bx.oscillations) {
        W.phase = PCG_PHASE_FREEFORM;
        continue;
      }
      const double accel_time = bx.accel_dist / bx.target_vel;
      const double vel = fmin((t - W.since) / accel_time, 1.0) * bx.target_vel;
      const double sh = vel * dt;
      if (W.displacement > 0.5 * W.amplitude) {
        W.until = t + bx.rest_period;
        W.phase = PCG_PHASE_REST_AFTER_MAX_AMP;
        continue;
      }
      W.displacement += sh;
      if (W.shear) {
        for (int side = 4; side < 6; ++side)
          if (W.side_wall[side] >= 0) P->p.wall_normal[W.side_wall[side]][0] += sh;
      } else {
        const double ch[3] = {0.0, 0.0, -sh};
        box_adjust(W, P, sd, ch);
      }
      return;
    }
    case PCG_PHASE_REST_AFTER_MAX_AMP:  // :278-293
      if (t < W.until) return;
      W.since = t;
      W.omega = bx.angular_freq;
      W.until = t + (pi * 2 * bx.num_of_cycles + pi_2) / W.omega;
      W.phase = PCG_PHASE_OSCILLATIONS;
      continue;
    case PCG_PHASE_OSCILLATIONS: {  // :295-320
      if (t >= W.until) {
        W.until = t + bx.rest_period;
        W.phase = PCG_PHASE_REST_AFTER_OSCILLATIONS;
        continue;
      }
      const double vel = 0.5 * W.amplitude * W.omega * sin(W.omega * (t - W.since));
      const double sh = vel * dt;
      if (W.shear) {
        for (int side = 4; side < 6; ++side)
          if (W.side_wall[side] >= 0) P->p.wall_normal[W.side_wall[side]][0] += sh;
      } else {
        const double ch[3] = {0.0, 0.0, sh};
        box_adjust(W, P, sd, ch);
      }
      return;
    }
    case PCG_PHASE_REST_AFTER_OSCILLATIONS:  // :322-330
      if (t < W.until) return;
      W.phase = PCG_PHASE_FREEFORM;
      continue;
    default:
      return;
    }
  }
}

// wall::log_forces (cg/src/cg/wall/log_forces.cpp:4-17) after the forces of
// a step: every plane on a box side, then once more the -axis plane of each
// LJ axis (the reference's wall list holds it twice, state.cpp:680-681);
// work += F . shift; the z force averaged
__global__ void k_box_log(DevBatch b) {
#pragma clang fp contract(off)
  if (threadIdx.x != 0) return;
  WallDev &W = *b.wall;
  const SysState &st = b.st[0];
  if (!st.active || !W.log_on) return;
  ParamsC &p = b.P->p;
  const double t = st.t;
  auto one = [&](int w) {
    mavg_add(W.avg_F[w], t, W.wall_F[w], 3);
    W.work[w] += W.wall_F[w][0] * W.shift[w][0] + W.wall_F[w][1] * W.shift[w][1] +
                 W.wall_F[w][2] * W.shift[w][2];
  };
  for (int side = 0; side < 6; ++side)
    if (W.side_wall[side] >= 0) one(W.side_wall[side]);
  for (int side = 0; side < 6; side += 2) {
    const int w = W.side_wall[side];
    if (w >= 0 && p.wall_kind[w] == 2) one(w);
  }
  const int wn = W.side_wall[4], wp = W.side_wall[5];
  const double zf = (W.wall_F[wn][0] * p.wall_normal[wn][0] + W.wall_F[wn][1] * p.wall_normal[wn][1] +
                     W.wall_F[wn][2] * p.wall_normal[wn][2]) +
                    (W.wall_F[wp][0] * p.wall_normal[wp][0] + W.wall_F[wp][1] * p.wall_normal[wp][1] +
                     W.wall_F[wp][2] * p.wall_normal[wp][2]);
  mavg_add(W.avg_z, t, &zf, 1);
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
void launch_step_begin(const DevBatch &b, int prec, int qa_on, hipStream_t s) {
  if (prec == PCG_FP64)
    hipLaunchKernelGGL(k_step_begin<double>, grid1(b.ntot), dim3(BLK), 0, s, b, qa_on);
  else
    hipLaunchKernelGGL(k_step_begin<float>, grid1(b.ntot), dim3(BLK), 0, s, b, qa_on);
}
void launch_fix_nl(const DevBatch &b, int max_n, int max_nlcap, int cell_words, hipStream_t s,
                   bool done) {
  int rows_per_block = BLK / 64;
  dim3 rg((unsigned)((b.ntot + rows_per_block - 1) / rows_per_block));
  const size_t mask_lds = (size_t)rows_per_block * cell_words * sizeof(unsigned long long);
  // each count kernel ends with the scan of its counts, by the wave or
  // block that completes a system
  if (cell_words > 0) {
    hipLaunchKernelGGL(k_cell_count, grid1(b.ntot), dim3(BLK), 0, s, b);
    hipLaunchKernelGGL(k_cell_fill, grid1(b.ntot), dim3(BLK), 0, s, b);
  }
  const int fuse = max_n < ROWS_FUSE_MAX;
  hipLaunchKernelGGL(k_nl_rows<false>, rg, dim3(BLK), mask_lds, s, b, cell_words, fuse);
  if (!fuse) hipLaunchKernelGGL(k_nl_scan, dim3((unsigned)b.nsys), dim3(64), 0, s, b);
  hipLaunchKernelGGL(k_nl_rows<true>, rg, dim3(BLK), mask_lds, s, b, cell_words, fuse);
  // the partition strides over the filled chunks: at most PART_GRID blocks
  // a system (one per CU of a 256-CU device for one system)
  const int part_grid = std::max(1, std::min((max_nlcap + BLK - 1) / BLK, PART_GRID));
  dim3 pg((unsigned)part_grid, (unsigned)b.nsys);
  hipLaunchKernelGGL(k_part_count, pg, dim3(BLK), 0, s, b);
  hipLaunchKernelGGL(k_part_write, pg, dim3(BLK), 0, s, b);
  if (done) hipLaunchKernelGGL(k_nl_done, grid1(b.nsys, 64), dim3(64), 0, s, b);
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
size_t qa_finish_lds(int max_n) {
  return 4 * (((size_t)max_n + 1) & ~(size_t)1) + (size_t)QF_CAND_LDS * (8 + 8 + 4 * 4);
}
void launch_qa_finish(const DevBatch &b, int max_n, hipStream_t s) {
  // the budgets in LDS while they and the candidates fit 96 KB
  const size_t lds = qa_finish_lds(max_n);
  const int sync_lds = lds <= QF_LDS_MAX ? max_n : 0;
  hipLaunchKernelGGL(k_qa_finish, dim3(b.nsys), dim3(1024), sync_lds ? lds : 0, s, b, sync_lds);
}
void launch_lang(const DevBatch &b, int prec, int noise_mode, hipStream_t s, bool tick) {
  const int t = tick ? 1 : 0;
  if (prec == PCG_FP64)
    hipLaunchKernelGGL(k_lang<double>, grid1(b.ntot), dim3(BLK), 0, s, b, noise_mode, t);
  else
    hipLaunchKernelGGL(k_lang<float>, grid1(b.ntot), dim3(BLK), 0, s, b, noise_mode, t);
}
void launch_tick(const DevBatch &b, int advance, hipStream_t s, bool nl_done) {
  hipLaunchKernelGGL(k_tick, grid1(b.nsys, 64), dim3(64), 0, s, b, advance, nl_done ? 1 : 0);
}
void launch_ljw_process(const DevBatch &b, hipStream_t s) {
  hipLaunchKernelGGL(k_ljw_process, dim3(1), dim3(1024), 0, s, b);
}
void launch_box_setup(const DevBatch &b, DevParams *P, SysDesc *sd, hipStream_t s) {
  hipLaunchKernelGGL(k_box_setup, dim3(1), dim3(64), 0, s, b, P, sd);
}
void launch_box_phase(const DevBatch &b, DevParams *P, SysDesc *sd, hipStream_t s) {
  hipLaunchKernelGGL(k_box_phase, dim3(1), dim3(64), 0, s, b, P, sd);
}
void launch_box_log(const DevBatch &b, hipStream_t s) {
  hipLaunchKernelGGL(k_box_log, dim3(1), dim3(64), 0, s, b);
}

} // namespace pcg
