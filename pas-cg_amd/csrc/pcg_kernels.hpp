// pcg_kernels.hpp -- grid-wide kernels of the per-step loop over a batch:
// one launch per stage, items of every system in one grid (blockIdx.y is the
// system). Used by the plain drop-in context (a batch of one) and for systems
// too large for the persistent workgroup-per-system kernel.
#pragma once
#include "pcg_terms.hpp"

namespace pcg {

constexpr int BLK = 256;
// most blocks a system's partition kernels take (they stride over the chunks)
constexpr int PART_GRID = 256;
// systems of this many residues or more take the NL row scan as a launch of
// its own (k_nl_scan) instead of in the count pass's last row
constexpr int ROWS_FUSE_MAX = 2048;
// candidates the stage QA finish stages in LDS, and its LDS limit
constexpr int QF_CAND_LDS = 1024;
constexpr size_t QF_LDS_MAX = 96 * 1024;

// Geometry and per-residue QA data from HBM.
struct GeoG {
  const double *x, *y, *z;
  __device__ __forceinline__ v3<double> r(int i) const { return v3<double>{x[i], y[i], z[i]}; }
  __device__ __forceinline__ int atype(const DevBatch &b, int i) const { return b.atype[i]; }
  __device__ __forceinline__ int ptype(const DevBatch &b, int a) const { return b.P->p.ptype[a]; }
  // the four sync budgets of residue i, byte c = budget c
  __device__ __forceinline__ uint32_t sync4(const DevBatch &b, int i) const {
    return *(const uint32_t *)(b.sync + 4 * i);
  }
  __device__ __forceinline__ double type_req(const DevBatch &b, int t) const {
    return __dmul_rn(b.P->p.qa_lj_r_high[t], __dadd_rn(1.0, b.P->p.qa_formation_tol));
  }
  // n and h as prepare_nh left them
  __device__ __forceinline__ void nh(const DevBatch &b, int i, double out[6]) const {
    for (int c = 0; c < 6; ++c) out[c] = b.nh[c][i];
  }
};

template <typename RF> struct AccG {
  RF *fx, *fy, *fz;
  template <typename T> __device__ __forceinline__ void add(int i, v3<T> f) {
    atomicAdd(fx + i, (RF)f.x);
    atomicAdd(fy + i, (RF)f.y);
    atomicAdd(fz + i, (RF)f.z);
  }
};

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// Sum of v over the block, valid in thread 0.
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double part[BLK / 64];
  v = wave_sum(v);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) part[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += part[k];
  return s;
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), counter-based.
__device__ __forceinline__ void philox4x32(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0 = __umulhi(0xD2511F53u, c[0]), lo0 = 0xD2511F53u * c[0];
    uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]), lo1 = 0xCD9E8D57u * c[2];
    uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}
// Three standard normals for (system key, step, residue): Box-Muller in
// FP32 on the hardware's transcendental units. v_log_f32 is log2, and
// v_sin_f32 / v_cos_f32 take the angle in turns (sin(2 pi u)), so no range
// reduction is needed. u1, u3 lie in [2^-32, 1] (the radius is at most 6.66),
// u2, u4 in [0, 1]. A normal carries about 1e-7 relative error, far below
// what the noise's statistics can resolve.
__device__ __forceinline__ void philox_normals(uint64_t key, int64_t step, int res, double out[3]) {
  uint32_t c[4] = {(uint32_t)step, (uint32_t)((uint64_t)step >> 32), (uint32_t)res, 0x5043u};
  philox4x32(c, (uint32_t)key, (uint32_t)(key >> 32));
  const float inv = 0x1p-32f, m2ln2 = -1.3862943611198906f;  // -2 ln 2
  float u1 = fmaf((float)c[0], inv, inv), u2 = (float)c[1] * inv;
  float u3 = fmaf((float)c[2], inv, inv), u4 = (float)c[3] * inv;
  float r1 = __builtin_amdgcn_sqrtf(m2ln2 * __builtin_amdgcn_logf(u1));
  float r2 = __builtin_amdgcn_sqrtf(m2ln2 * __builtin_amdgcn_logf(u3));
  out[0] = (double)(r1 * __builtin_amdgcn_cosf(u2));
  out[1] = (double)(r1 * __builtin_amdgcn_sinf(u2));
  out[2] = (double)(r2 * __builtin_amdgcn_cosf(u4));
}

// ------------------------------------------------------------ per term ----
__device__ __forceinline__ int term_size(SysDescC &s, const SysState &st, int term) {
  switch (term) {
  case PCG_CHIR: return s.nchir;
  case PCG_TETHER: return s.nteth;
  case PCG_LREP: return s.nlrep;
  case PCG_NAT_CONT: return st.nc_n;
  case PCG_PAULI: return st.pauli_n;
  case PCG_CONST_DH: return st.dh_n;
  case PCG_REL_DH: return st.dh_n;
  case PCG_QA_CAND: return st.qh_n;
  case PCG_QA_CONT: return st.qc_n < st.qf_n ? st.qc_n : st.qf_n;
  case PCG_PID: return st.pid_n;
  case PCG_NAT_ANG: return s.nnang;
  case PCG_HEUR_ANG: return s.nhang;
  case PCG_HEUR_DIH: return s.nhdih;
  case PCG_CND: return s.nndih;
  case PCG_SND: return s.nndih;
  case PCG_VEL_AFM: return s.ntip;
  case PCG_FORCE_AFM: return s.nftip;
  case PCG_SOLID_WALL: return s.n;
  case PCG_HW_FREE: return s.n;
  case PCG_HW_CONN: return st.n_hw;
  case PCG_LJW_FREE: return s.n;
  case PCG_LJW_CONN: return st.n_lj;
  }
  return 0;
}

__device__ __forceinline__ bool term_on(DevParamsC &P, int term) {
  ParamsC &p = P.p;
  switch (term) {
  case PCG_CHIR: return p.chir_on;
  case PCG_TETHER: return p.tether_on;
  case PCG_LREP: return p.lrep_on;
  case PCG_NAT_CONT: return p.nat_cont_on;
  case PCG_PAULI: return P.pauli_on;
  case PCG_CONST_DH: return p.dh_on && p.dh_variant == 0;
  case PCG_REL_DH: return p.dh_on && p.dh_variant == 1;
  case PCG_QA_CAND: return p.qa_on;
  case PCG_QA_CONT: return p.qa_on;
  case PCG_PID: return p.pid_on;
  case PCG_NAT_ANG: return p.nat_ang_on;
  case PCG_HEUR_ANG: return p.heur_ang_on;
  case PCG_HEUR_DIH: return p.heur_dih_on;
  case PCG_CND: return p.nat_dih_on && p.nat_dih_variant == 0;
  case PCG_SND: return p.nat_dih_on && p.nat_dih_variant == 1;
  case PCG_VEL_AFM: return p.afm_on;
  case PCG_FORCE_AFM: return p.afm_on;
  case PCG_SOLID_WALL: return P.solid_on;
  case PCG_HW_FREE: return P.hw_terms;
  case PCG_HW_CONN: return P.hw_terms;
  case PCG_LJW_FREE: return P.ljw_terms;
  case PCG_LJW_CONN: return P.ljw_terms;
  }
  return false;
}

// One item k (local to the system's list) of one term. Pair terms run in R;
// the angular terms, tethers, local repulsion, chirality, PID and the AFM
// tips in FP64 (acos near 0/pi and the heuristic-angle polynomial lose too
// much in FP32).
template <typename R, typename G, typename A>
__device__ __forceinline__ double eval_item(const DevBatch &b, SysDescC &s, SysState &st,
                                            const G &g, A &acc, int term, int k,
                                            bool want_v = true) {
  switch (term) {
  case PCG_CHIR: return chir_item<double>(b, g, acc, s.chir0 + k);
  case PCG_TETHER: return tether_item<double>(b, g, acc, s.teth0 + k);
  case PCG_LREP: return lrep_item<double>(b, g, acc, s.lrep0 + k);
  case PCG_NAT_CONT: return nat_cont_item<R>(b, s, st, g, acc, s.nc0 + k, st.t);
  case PCG_PAULI: return pauli_item<R>(b, s, g, acc, s.nl0 + k);
  case PCG_CONST_DH: return dh_item<R>(b, s, g, acc, s.nl0 + k, 0, want_v);
  case PCG_REL_DH: return dh_item<R>(b, s, g, acc, s.nl0 + k, 1, want_v);
  case PCG_QA_CAND: return qa_cand_item<R>(b, s, st, g, acc, k);
  case PCG_QA_CONT: return qa_cont_item<R>(b, s, st, g, acc, k, want_v);
  case PCG_PID: return pid_item<double>(b, s, g, acc, s.nl0 + k);
  case PCG_NAT_ANG: return nat_ang_item<double>(b, g, acc, s.nang0 + k);
  case PCG_HEUR_ANG: return heur_ang_item<double>(b, g, acc, s.hang0 + k);
  case PCG_HEUR_DIH: return heur_dih_item<double>(b, g, acc, s.hdih0 + k);
  case PCG_CND: return nat_dih_item<double>(b, g, acc, s.ndih0 + k, 0);
  case PCG_SND: return nat_dih_item<double>(b, g, acc, s.ndih0 + k, 1);
  case PCG_VEL_AFM: return vel_afm_item<double>(b, g, acc, s.tip0 + k, st.t - s.afm_t0, st.t);
  case PCG_FORCE_AFM: return force_afm_item(b, acc, s.tip0 + s.ntip + k, st.t);
  case PCG_SOLID_WALL: return solid_wall_item(b, g, acc, s.res0 + k);
  case PCG_HW_FREE: return hw_free_item(b, g, acc, s.res0 + k);
  case PCG_HW_CONN: return hw_conn_item(b, g, acc, k);
  case PCG_LJW_FREE: return ljw_free_item(b, g, acc, s.res0 + k);
  case PCG_LJW_CONN: return ljw_conn_item(b, g, acc, k);
  }
  return 0.0;
}

} // namespace pcg
