// pcg_report.hip -- the per-trajectory scalars of the reference's report
// (cg/src/cg/output/make_report.cpp:160-247): one workgroup per system sums
// what they need; the host turns the sums into the scalars (pcg_host.hip).
#include <mutex>

#include "pcg_launch.hpp"
#include "pcg_stages.hpp"

namespace pcg {

namespace {

constexpr int RB = 256;  // threads of a report workgroup
constexpr int NV = 26;   // values reduced per system (ReportSums after n)

// v[0..NV) summed over the block, in thread 0
__device__ void block_sums(double *v) {
  __shared__ double part[NV][RB / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int k = 0; k < NV; ++k) {
    double x = v[k];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
    if (lane == 0) part[k][w] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int j = 0; j < RB / 64; ++j) s += part[k][j];
      v[k] = s;
    }
  __syncthreads();
}

__global__ __launch_bounds__(RB) void k_report(DevBatch b, ReportSums *out) {
  const int s = blockIdx.x;
  SysDescC &sd = b.sys[s];
  const int n = sd.n, r0 = sd.res0;
  ParamsC &p = b.P->p;
  __shared__ double ctr[6];
  double v[NV];
  // centroids of the current and the original positions
  for (int k = 0; k < NV; ++k) v[k] = 0.0;
  for (int k = threadIdx.x; k < n; k += RB) {
    const int i = r0 + k;
    for (int c = 0; c < 3; ++c) {
      v[c] += b.y[0][c][i];
      v[3 + c] += b.nat_r[c][i];
    }
  }
  block_sums(v);
  if (threadIdx.x == 0)
    for (int c = 0; c < 6; ++c) ctr[c] = v[c] / (double)n;
  __syncthreads();
  // centred second moments, kinetic energy, contact counts
  for (int k = 0; k < NV; ++k) v[k] = 0.0;
  for (int k = threadIdx.x; k < n; k += RB) {
    const int i = r0 + k;
    double r[3], q[3], u2 = 0.0;
    for (int c = 0; c < 3; ++c) {
      r[c] = b.y[0][c][i] - ctr[c];
      q[c] = b.nat_r[c][i] - ctr[3 + c];
      u2 += b.vel[c][i] * b.vel[c][i];
    }
    v[0] += 0.5 * p.mass[b.atype[i]] * u2;  // make_report::kinetic_energy
    v[1] += r[0] * r[0];
    v[2] += r[1] * r[1];
    v[3] += r[2] * r[2];
    v[4] += r[0] * r[1];
    v[5] += r[0] * r[2];
    v[6] += r[1] * r[2];
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 3; ++c) v[7 + 3 * a + c] += q[a] * r[c];  // orig^T cur
    v[16] += r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    v[17] += q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
  }
  for (int c = threadIdx.x; c < sd.nnc; c += RB) {
    const int k = sd.nc0 + c;
    if (b.nc_active[k]) {
      v[18] += 1.0;
      if (b.nc_type[k] == 5) v[19] += 1.0;  // nat_cont::type::SSBOND
    }
  }
  const SysState &st = b.st[s];
  const int qn = st.qc_n < sd.qccap ? st.qc_n : sd.qccap;
  for (int c = threadIdx.x; c < qn; c += RB) {
    const int q = sd.qc0 + c;
    if (b.qc_vac[q]) continue;
    if (b.chain[b.qc_i[0][q]] == b.chain[b.qc_i[1][q]]) v[21] += 1.0;
    else v[20] += 1.0;
  }
  block_sums(v);
  if (threadIdx.x == 0) {
    ReportSums &o = out[s];
    o.n = n;
    o.K = v[0];
    for (int k = 0; k < 6; ++k) o.S[k] = v[1 + k];
    for (int k = 0; k < 9; ++k) o.H[k] = v[7 + k];
    o.pp = v[16];
    o.qq = v[17];
    o.icn = v[18];
    o.icnss = v[19];
    o.inter = v[20];
    o.intra = v[21];
    // end-to-end distance |r_last - r_first| (no wrap, as the reference)
    double L2 = 0.0;
    for (int c = 0; c < 3; ++c) {
      const double d = b.y[0][c][r0 + n - 1] - b.y[0][c][r0];
      L2 += d * d;
    }
    o.L = sqrt(L2);
    for (int k = 0; k < 6; ++k) o.pid[k] = 0;
  }
}

// ---------------------------------------------------------- PID counts ----
// pid::count_active (cg/src/cg/pid/count_active.cpp:4-164), in FP64 with the
// reference's operation order. The reference walks its PID bundle list; here
// every pair of the system is tested against the rules that put a pair on
// that list (nl::legacy_update's sequence gap, not a native contact, not a QA
// free pair, pid.include4; thread_main.cpp:403-427), so the counts do not
// depend on when the list was last rebuilt: a valid list holds every such
// pair within the cutoff, and only pairs within the cutoff are counted.

// improper dihedral (i, i+1, i-1, j) with raw i+-1 neighbours inside the
// system (count_active.cpp:36-77); skip when 1/|rm| or 1/|rn| exceeds 10
__device__ inline double count_psi(const DevBatch &b, int i, int j, int lo, int hi, bool &skip) {
  auto R = [&](int k) { return v3<double>{b.y[0][0][k], b.y[0][1][k], b.y[0][2][k]}; };
  const v3<double> z{0.0, 0.0, 0.0};
  const bool vp = i - 1 >= lo, vn = i + 1 < hi;
  const v3<double> r_in = vn ? R(i + 1) : z, r_ip = vp ? R(i - 1) : z;
  const v3<double> rij = xsub(R(i), r_in);
  const v3<double> rkj = vp ? xsub(r_ip, r_in) : z;
  const v3<double> rkl = xsub(r_ip, R(j));
  const v3<double> rm{rij.y * rkj.z - rij.z * rkj.y, rij.z * rkj.x - rij.x * rkj.z,
                      rij.x * rkj.y - rij.y * rkj.x};
  const v3<double> rn{rkj.y * rkl.z - rkj.z * rkl.y, rkj.z * rkl.x - rkj.x * rkl.z,
                      rkj.x * rkl.y - rkj.y * rkl.x};
  const double rm_ninv = 1.0 / __dsqrt_rn(rm.x * rm.x + rm.y * rm.y + rm.z * rm.z);
  const double rn_ninv = 1.0 / __dsqrt_rn(rn.x * rn.x + rn.y * rn.y + rn.z * rn.z);
  skip = rm_ninv > 10.0 || rn_ninv > 10.0;
  double psi = acos((rm.x * rn.x + rm.y * rn.y + rm.z * rn.z) * rm_ninv * rn_ninv);
  if (rij.x * rn.x + rij.y * rn.y + rij.z * rn.z < 0.0) psi = -psi;
  return psi;
}

// lambda value (cg/include/cg/pid/lambda.h:58-85); supp is checked before
__device__ inline double count_lam(int variant, double alpha, double psi0, double psi) {
  if (variant == 0) {
    const double s = alpha * (psi - psi0);
    return fabs(s) >= M_PI ? 0.0 : 0.5 * cos(s) + 0.5;
  }
  const double t = alpha * (psi - psi0) * M_1_PI, a = fabs(t);
  if (a >= M_PI) return 0.0;
  const double x = 2.0 * a * (a - 1.0);
  return 1.0 - (a * a) / (x + 1.0);
}
__device__ inline bool count_supp(double alpha, double psi0, double psi) {
  return fabs(alpha * (psi - psi0)) < M_PI;
}

// adds the pair's counts to c[6] (intra, intra bb, intra ss, inter, inter bb,
// inter ss)
__device__ inline void count_pair(const DevBatch &b, SysDescC &sd, int i1, int i2, int c[6]) {
  ParamsC &p = b.P->p;
  const bool same = b.chain[i1] == b.chain[i2];
  if (same) {
    const int d = b.seq[i1] > b.seq[i2] ? b.seq[i1] - b.seq[i2] : b.seq[i2] - b.seq[i1];
    if (d < 3) return;  // never on the NL
  }
  const v3<double> r1{b.y[0][0][i1], b.y[0][1][i1], b.y[0][2][i1]};
  const v3<double> r2{b.y[0][0][i2], b.y[0][1][i2], b.y[0][2][i2]};
  const double r12_n = xnorm3(xwrapd(xsub(r2, r1), sd));
  if (!(r12_n <= p.nl_cutoff)) return;  // count_active.cpp:17-18
  if (b.atype[i1] == 14 || b.atype[i2] == 14) return;  // PRO
  const bool four = same && (b.seq[i1] - b.seq[i2] == 4 || b.seq[i2] - b.seq[i1] == 4);
  if (p.nat_cont_on && sd.nnc > 0 && find_contact(b, sd, i1, i2) >= 0) return;
  if (p.qa_on && !(!p.qa_include4 && four)) return;  // a QA free pair
  if (!p.pid_include4 && four) return;
  const int lo = sd.res0, hi = sd.res0 + sd.n;
  bool skip;
  const double psi1 = count_psi(b, i1, i2, lo, hi, skip);
  if (skip) return;
  const double psi2 = count_psi(b, i2, i1, lo, hi, skip);
  if (skip) return;
  // lambda choice of count_active.cpp:67-75 and :110-118
  const bool sep3 = i2 - i1 == 3;
  const bool opt1 = count_supp(p.pid_bbp_alpha, p.pid_bbp_psi0, psi1) || sep3;
  const bool opt2 = !sep3 && count_supp(p.pid_bbp_alpha, p.pid_bbp_psi0, psi2);
  const double l1a = opt1 ? p.pid_bbp_alpha : p.pid_bbm_alpha, l1p = opt1 ? p.pid_bbp_psi0 : p.pid_bbm_psi0;
  const double l2a = opt2 ? p.pid_bbp_alpha : p.pid_bbm_alpha, l2p = opt2 ? p.pid_bbp_psi0 : p.pid_bbm_psi0;
  const double bb_r_high = opt2 ? p.pid_bbp_lj[2] : p.pid_bbm_lj[2];
  const double factor = p.counting_factor;
  const int var = p.pid_lambda_variant;
  if (count_supp(l1a, l1p, psi1) && count_supp(l2a, l2p, psi2)) {
    const double lam1 = count_lam(var, l1a, l1p, psi1), lam2 = count_lam(var, l2a, l2p, psi2);
    if (lam1 * lam2 > 5e-5 && r12_n < bb_r_high * C216_INV * factor) {
      ++c[same ? 0 : 3];
      ++c[same ? 1 : 4];
    }
  }
  const double sa = p.pid_ss_alpha, sp = p.pid_ss_psi0;
  if (count_supp(sa, sp, psi1) && count_supp(sa, sp, psi2)) {
    const int ty = b.atype[i1] * 21 + b.atype[i2];
    const double lam1 = count_lam(var, sa, sp, psi1), lam2 = count_lam(var, sa, sp, psi2);
    if (lam1 * lam2 > 5e-5 && p.pid_ss_depth[ty] > 5e-5 &&
        r12_n < p.pid_ss_r_high[ty] * C216_INV * factor) {
      ++c[same ? 0 : 3];
      ++c[same ? 2 : 5];
    }
  }
}

// one wavefront per residue i of the batch, its lanes over the residues
// j > i of the same system; the wave's counts go to its system's sums
__global__ __launch_bounds__(RB) void k_pid_count(DevBatch b, ReportSums *out) {
  const int row = blockIdx.x * (RB / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= b.ntot) return;  // uniform per wavefront
  const int s = b.sysid[row];
  SysDescC &sd = b.sys[s];
  const int end = sd.res0 + sd.n;
  int c[6] = {0, 0, 0, 0, 0, 0};
  for (int j = row + 1 + lane; j < end; j += 64) count_pair(b, sd, row, j, c);
  for (int k = 0; k < 6; ++k) {
    int x = c[k];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
    if (lane == 0 && x) atomicAdd(&out[s].pid[k], x);
  }
}

} // namespace

void launch_report(const DevBatch &b, int pid_on, ReportSums *out, hipStream_t s) {
  hipLaunchKernelGGL(k_report, dim3(b.nsys), dim3(RB), 0, s, b, out);
  if (pid_on && b.ntot > 0)
    hipLaunchKernelGGL(k_pid_count, dim3((unsigned)((b.ntot + RB / 64 - 1) / (RB / 64))), dim3(RB),
                       0, s, b, out);
}

// The report sums of the no-allocation report: a file-scope device array,
// one per GPU for the life of the process. k_report writes every field of a
// system's entry before k_pid_count adds into its counts, so nothing needs
// clearing between calls. Calls that overlap on one GPU share it and race.
namespace {
__device__ __align__(256) ReportSums scratch[REPORT_STATIC_MAX];
constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
ReportSums *scratch_ptr[MAX_DEVICES] = {};
}  // namespace

int report_static_sums(const DevBatch &b, int device, int pid_on, hipStream_t s,
                       ReportSums **sums) {
  if (b.nsys > REPORT_STATIC_MAX) return -1;
  if (device < 0 || device >= MAX_DEVICES) return -2;
  ReportSums *p;
  {
    std::lock_guard<std::mutex> lock(scratch_mu);
    p = scratch_ptr[device];
    if (!p) {
      // the array's address on this device (the host-side name is not one)
      int cur = 0;
      if (hipGetDevice(&cur) != hipSuccess || cur != device) return -3;
      void *ptr = nullptr;
      if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return -3;
      p = scratch_ptr[device] = (ReportSums *)ptr;
    }
  }
  launch_report(b, pid_on, p, s);
  if (hipGetLastError() != hipSuccess) return -3;
  *sums = p;
  return 0;
}

} // namespace pcg
