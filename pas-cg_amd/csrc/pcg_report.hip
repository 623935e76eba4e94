// pcg_report.hip -- the per-trajectory scalars of the reference's report
// (cg/src/cg/output/make_report.cpp:160-247): one workgroup per system sums
// what they need; the host turns the sums into the scalars (pcg_host.hip).
#include "pcg_launch.hpp"

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
  }
}

} // namespace

void launch_report(const DevBatch &b, ReportSums *out, hipStream_t s) {
  hipLaunchKernelGGL(k_report, dim3(b.nsys), dim3(RB), 0, s, b, out);
}

} // namespace pcg
