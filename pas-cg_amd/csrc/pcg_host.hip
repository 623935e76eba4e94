// pcg_host.hip -- host side of the C ABI declared in include/pcg.h.
//
// Builds a batch (one or many systems) the way state::traj_init and
// thread::init_kernels build the reference's state (cg/src/cg/simul/state.cpp:
// 83-112, 144-568; thread_setup.cpp:51-720), lays it out in one device
// allocation, and drives the stage kernels (plain context), the persistent
// kernel (batch) or continuous rounds of it (server).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "pcg_launch.hpp"

using namespace pcg;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(PCG_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));       \
  } while (0)

// Numerical Recipes ran2 (L'Ecuyer with Bays-Durham shuffle) and its
// cos-branch Box-Muller, the generator of the reference's COMPAT_MODE build
// (cg/include/cg/utils/random/nr.h:24-91, cg/CMakeLists.txt:67-69).
struct Ran2 {
  int iy = 0, idum = -448, idum2 = 123456789;
  int iv[32] = {0};
  explicit Ran2(int seed = 0) {
    draw(seed);
    draw(seed);
  }
  double draw(int iseed = 0) {
    const int im1 = 2147483563, im2 = 2147483399, imm1 = im1 - 1, ia1 = 40014, ia2 = 40692,
              iq1 = 53668, iq2 = 52774, ir1 = 12211, ir2 = 3791, ntab = 32,
              ndiv = 1 + imm1 / ntab;
    const double eps = 1.2e-7f, rnmx = 1.f - eps, am = 1.f / (float)im1;
    int k, j;
    if (iseed != 0) idum = -iseed;
    if (idum <= 0) {
      idum2 = idum = std::max(-idum, 1);
      for (j = ntab + 7; j >= 0; --j) {
        k = idum / iq1;
        idum = ia1 * (idum - k * iq1) - k * ir1;
        if (idum < 0) idum += im1;
        if (j < ntab) iv[j] = idum;
      }
      iy = iv[0];
    }
    k = idum / iq1;
    idum = ia1 * (idum - k * iq1) - k * ir1;
    if (idum < 0) idum += im1;
    k = idum2 / iq2;
    idum2 = ia2 * (idum2 - k * iq2) - k * ir2;
    if (idum2 < 0) idum2 += im2;
    j = iy / ndiv;
    iy = iv[j] - idum2;
    iv[j] = idum;
    if (iy < 1) iy += imm1;
    return std::min(am * iy, rnmx);
  }
  double normal() {
    double r1 = draw(), r2 = draw();
    return std::sqrt(-2.0 * std::log(r1)) * std::cos((2.0 * M_PI) * r2);
  }
};

uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// ----------------------------------------------------------- host batch ----
struct HostSys {
  int n = 0;
  std::vector<double> pos, orig;
  std::vector<uint8_t> atype;
  std::vector<int> chain, seq;
  std::vector<int> teth[2];
  std::vector<double> teth_nat;
  std::vector<int> lrep[2];
  std::vector<int> chir[4];
  std::vector<double> chir_f0, chir_f1;
  std::vector<int> nang[3];
  std::vector<double> nang_th;
  std::vector<int> hang[3];
  std::vector<uint8_t> hang_t;
  std::vector<int> ndih[4];
  std::vector<double> ndih_ph;
  std::vector<int> hdih[4];
  std::vector<uint8_t> hdih_t;
  std::vector<int> nc[2];
  std::vector<double> nc_nat;
  std::vector<uint8_t> nc_type, nc_active;
  std::vector<int> tip_res;
  std::vector<double> tip_orig[3], tip_vel[3];  // velocity tips, then force tips
  int nftip = 0;                                // (tip_vel = force, tip_orig = 0)
  std::vector<double> y1;  // initial velocities * dt
  double temperature = 0.0;
  int nlcap = 0, qfcap = 0, qccap = 0;
};

double norm3(const double *a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

int heur_type(int a2, int a3) { // aa_heur_pair, cg/src/cg/angles/aa_heur_pair.cpp:11-27
  int o1 = a2 == 7 ? 0 : (a2 == 14 ? 1 : 2);
  int o2 = a3 == 7 ? 0 : (a3 == 14 ? 1 : 2);
  return 3 * o1 + o2;
}

int round_up(int x, int m) { return (x + m - 1) / m * m; }

// state::setup_* of cg/src/cg/simul/state.cpp for one system.
int build_system(const pcg_params &p, const pcg_model &m, double temperature, int seed,
                 HostSys &h, int64_t ran2_skip = 0) {
  int n = m.n;
  if (n <= 0 || !m.pos || !m.atype || !m.chain_idx || !m.seq_idx)
    return fail(PCG_E_ARG, "model needs n > 0, pos, atype, chain_idx, seq_idx");
  h.n = n;
  h.pos.assign(m.pos, m.pos + 3 * n);
  h.orig.assign(m.orig_pos ? m.orig_pos : m.pos, (m.orig_pos ? m.orig_pos : m.pos) + 3 * n);
  h.atype.assign(m.atype, m.atype + n);
  h.chain.assign(m.chain_idx, m.chain_idx + n);
  h.seq.assign(m.seq_idx, m.seq_idx + n);
  for (int i = 0; i < n; ++i)
    if (h.atype[i] > 20) return fail(PCG_E_ARG, "residue type out of range");
  h.temperature = temperature < 0 ? p.temperature : temperature;
  auto P = [&](int i) { return &h.pos[3 * i]; };
  auto chk = [&](int i) { return i >= 0 && i < n; };

  // initial velocities from ran2 (state::setup_langevin, state.cpp:270-335)
  h.y1.assign(3 * n, 0.0);
  if (p.lang_on) {
    Ran2 g(seed);
    for (int64_t k = 0; k < ran2_skip; ++k) g.draw();
    double sum[3] = {0, 0, 0};
    for (int i = 0; i < n; ++i) {
      double xx = 2.0 * (g.draw() - 0.5), yy = 2.0 * (g.draw() - 0.5), zz = 2.0 * (g.draw() - 0.5);
      double xyz = 1.0 / std::sqrt(xx * xx + yy * yy + zz * zz);
      h.y1[3 * i] = xx * xyz;
      h.y1[3 * i + 1] = yy * xyz;
      h.y1[3 * i + 2] = zz * xyz;
      for (int c = 0; c < 3; ++c) sum[c] += h.y1[3 * i + c];
    }
    double x = 0.0;
    for (int i = 0; i < n; ++i) {
      for (int c = 0; c < 3; ++c) h.y1[3 * i + c] -= sum[c] / (double)n;
      double *v = &h.y1[3 * i];
      x = x + v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    }
    double dt = p.dt, delsq = dt * dt;
    double heat = std::sqrt(delsq * 3.0 * n * h.temperature / x);
    for (double &v : h.y1) v *= heat;
  }
  // local repulsion: (i1, i3) of every angle (state.cpp:354-364)
  for (int a = 0; a < m.n_ang; ++a) {
    int i1 = m.ang[3 * a], i2 = m.ang[3 * a + 1], i3 = m.ang[3 * a + 2];
    if (!chk(i1) || !chk(i2) || !chk(i3)) return fail(PCG_E_ARG, "angle index out of range");
    if (p.lrep_on) {
      h.lrep[0].push_back(i1);
      h.lrep[1].push_back(i3);
    }
    if (m.ang_has && m.ang_has[a]) {
      if (p.nat_ang_on) {
        h.nang[0].push_back(i1);
        h.nang[1].push_back(i2);
        h.nang[2].push_back(i3);
        h.nang_th.push_back(m.ang_theta[a]);
      }
    } else if (p.heur_ang_on) {
      h.hang[0].push_back(i1);
      h.hang[1].push_back(i2);
      h.hang[2].push_back(i3);
      h.hang_t.push_back((uint8_t)heur_type(h.atype[i2], h.atype[i3]));
    }
  }
  for (int d = 0; d < m.n_dih; ++d) {
    int q[4] = {m.dih[4 * d], m.dih[4 * d + 1], m.dih[4 * d + 2], m.dih[4 * d + 3]};
    for (int k = 0; k < 4; ++k)
      if (!chk(q[k])) return fail(PCG_E_ARG, "dihedral index out of range");
    if (p.chir_on) {
      // state::setup_chir (state.cpp:366-387); the reference emplaces
      // (nat_factor, nat_chir) into fields declared (nat_chir, nat_factor)
      // (cg/include/cg/chir/chiral_quad.h:12-13), so the stored quad carries
      // them swapped: f0 (read as nat_factor) = nat_chir, f1 = nat_factor.
      double n12[3], n23[3], n34[3];
      for (int c = 0; c < 3; ++c) {
        n12[c] = P(q[1])[c] - P(q[0])[c];
        n23[c] = P(q[2])[c] - P(q[1])[c];
        n34[c] = P(q[3])[c] - P(q[2])[c];
      }
      double ni = 1.0 / std::sqrt(n23[0] * n23[0] + n23[1] * n23[1] + n23[2] * n23[2]);
      double nat_factor = ni * (ni * (ni * 1.0));
      double cx = n23[1] * n34[2] - n23[2] * n34[1], cy = n23[2] * n34[0] - n23[0] * n34[2],
             cz = n23[0] * n34[1] - n23[1] * n34[0];
      double nat_chir = (n12[0] * cx + n12[1] * cy + n12[2] * cz) * nat_factor;
      for (int k = 0; k < 4; ++k) h.chir[k].push_back(q[k]);
      h.chir_f0.push_back(nat_chir);
      h.chir_f1.push_back(nat_factor);
    }
    if (m.dih_has && m.dih_has[d]) {
      if (p.nat_dih_on) {
        for (int k = 0; k < 4; ++k) h.ndih[k].push_back(q[k]);
        h.ndih_ph.push_back(m.dih_phi[d]);
      }
    } else if (p.heur_dih_on) {
      for (int k = 0; k < 4; ++k) h.hdih[k].push_back(q[k]);
      h.hdih_t.push_back((uint8_t)heur_type(h.atype[q[1]], h.atype[q[2]]));
    }
  }
  // tethers: length from the orig positions (state.cpp:389-402)
  if (p.tether_on)
    for (int k = 0; k < m.n_teth; ++k) {
      int i1 = m.teth[2 * k], i2 = m.teth[2 * k + 1];
      if (!chk(i1) || !chk(i2)) return fail(PCG_E_ARG, "tether index out of range");
      double d[3];
      for (int c = 0; c < 3; ++c) d[c] = h.orig[3 * i2 + c] - h.orig[3 * i1 + c];
      h.teth[0].push_back(i1);
      h.teth[1].push_back(i2);
      h.teth_nat.push_back(norm3(d));
    }
  // native contacts (state.cpp:480-502), ordered by residue indices
  if (p.nat_cont_on && m.n_cont > 0) {
    std::vector<int> ord(m.n_cont);
    std::iota(ord.begin(), ord.end(), 0);
    auto key = [&](int k) {
      int a = m.cont[2 * k], b = m.cont[2 * k + 1];
      return a < b ? std::make_pair(a, b) : std::make_pair(b, a);
    };
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return key(a) < key(b); });
    for (int k : ord) {
      auto pr = key(k);
      if (!chk(pr.first) || !chk(pr.second)) return fail(PCG_E_ARG, "contact index out of range");
      double nat = m.cont_len[k];
      double d[3];
      for (int c = 0; c < 3; ++c) d[c] = P(pr.second)[c] - P(pr.first)[c];
      h.nc[0].push_back(pr.first);
      h.nc[1].push_back(pr.second);
      h.nc_nat.push_back(nat);
      h.nc_type.push_back((uint8_t)(m.cont_type ? m.cont_type[k] : 0));
      h.nc_active.push_back(norm3(d) < p.counting_factor * C216_INV * nat ? 1 : 0);
    }
  }
  for (int k = 0; k < m.n_tips; ++k) {
    if (!chk(m.tip_res[k])) return fail(PCG_E_ARG, "tip residue out of range");
    h.tip_res.push_back(m.tip_res[k]);
    for (int c = 0; c < 3; ++c) {
      h.tip_orig[c].push_back(m.tip_orig[3 * k + c]);
      h.tip_vel[c].push_back(m.tip_vel[3 * k + c]);
    }
  }
  if (m.n_ftips > 0 && (!m.ftip_res || !m.ftip_force))
    return fail(PCG_E_ARG, "force tips need ftip_res and ftip_force");
  for (int k = 0; k < m.n_ftips; ++k) {
    if (!chk(m.ftip_res[k])) return fail(PCG_E_ARG, "force tip residue out of range");
    h.tip_res.push_back(m.ftip_res[k]);
    for (int c = 0; c < 3; ++c) {
      h.tip_orig[c].push_back(0.0);
      h.tip_vel[c].push_back(m.ftip_force[3 * k + c]);
    }
    h.nftip += 1;
  }
  long long pairs = (long long)n * (n - 1) / 2;
  long long cap = std::min<long long>(pairs, 1200LL * n);
  h.nlcap = round_up((int)std::max<long long>(cap, 1), BLK);
  h.qccap = 12 * n + 64;
  h.qfcap = h.nlcap + h.qccap;
  return PCG_OK;
}

// an AFM tip's moving averages before its first sample
TipStat fresh_tip_stat(const pcg_params &p) {
  TipStat t;
  std::memset(&t, 0, sizeof(t));
  t.window = p.afm_stats_window;
  return t;
}

void fill_dev_params(const pcg_params &p, DevParams &d) {
  d.p = p;
  d.nl_eff_cutoff = std::max(p.nl_cutoff, p.dh_cutoff);
  d.V_factor = 1.0 / (4.0 * M_PI * p.dh_perm);
  d.screen_inv = 1.0 / p.dh_screen_dist;
  d.dt = p.dt;
  d.dt_inv = 1.0 / p.dt;
  d.gamma2 = p.gamma * d.dt_inv;
  d.deltsq = (p.dt * p.dt) / 2.0;
  double mx = 0.0;
  for (int k = 0; k < 445; ++k) mx = std::max(mx, p.qa_lj_r_high[k]);
  d.qa_max_req = mx;
  d.qa_factor = p.counting_factor * C216_INV;
  d.qa_sat_speed = 1.0 / p.qa_phase_dur;
  d.pauli_on = p.pauli_on && !p.qa_on;
  d.nl_required = d.pauli_on || p.dh_on || p.qa_on || p.pid_on;
  int kinds[3] = {0, 0, 0};
  for (int w = 0; w < p.n_walls && w < 6; ++w)
    if (p.wall_kind[w] >= 0 && p.wall_kind[w] < 3) kinds[p.wall_kind[w]] = 1;
  d.solid_on = kinds[0];
  d.hw_terms = p.hw_on && kinds[1];
  d.ljw_terms = p.ljw_on && kinds[2];
}

double const2_of(const pcg_params &p, double T) {
  double c = 2.0 * 1.0 * T * p.gamma * p.dt; // step.cpp:101-102
  return std::sqrt(c) * p.dt;
}

// One device allocation holding every array of a batch.
struct Arena {
  std::vector<std::pair<void **, size_t>> req;
  size_t total = 0;
  void *base = nullptr;
  template <typename T> void add(T **ptr, size_t count) {
    req.push_back({(void **)ptr, count * sizeof(T)});
    total += (count * sizeof(T) + 255) / 256 * 256;
  }
  bool borrowed = false;  // base came from the caller's pool, which keeps it
  int alloc(hipStream_t stream, void *pool_block = nullptr, size_t pool_bytes = 0) {
    if (pool_block && pool_bytes >= (total ? total : 256)) {
      base = pool_block;
      borrowed = true;
    } else if (hipMalloc(&base, total ? total : 256) != hipSuccess) {
      base = nullptr;
      return fail(PCG_E_NOMEM, "hipMalloc of " + std::to_string(total) + " bytes failed");
    }
    // every array starts at zero, as the reference's vectors do (e.g. the QA
    // n/h vectors of chain ends, which prepare_nh never writes)
    if (hipMemsetAsync(base, 0, total ? total : 256, stream) != hipSuccess)
      return fail(PCG_E_HIP, "hipMemsetAsync of the arena failed");
    size_t off = 0;
    for (auto &r : req) {
      *r.first = (char *)base + off;
      off += (r.second + 255) / 256 * 256;
    }
    return PCG_OK;
  }
};

} // namespace

// A device batch: the systems, their arena and the launch configuration.
struct DeviceBatch {
  int prec = PCG_MIXED;
  int noise = PCG_NOISE_PHILOX;
  int device = 0;
  hipStream_t stream = nullptr;
  DevBatch d{};
  DevParams hp{};
  DevParams *dP = nullptr;
  std::vector<SysDesc> sys;
  std::vector<HostSys> hs;
  Arena arena;
  int max_nlcap = 0, max_n = 0, max_items = 0;
  int max_nnc = 0;  // most native contacts of one system
  int cell_words = 0;  // row-mask words of the largest system on a cell grid
  size_t bytes = 0;
  std::vector<Ran2> ran2;  // per system, RAN2 noise mode
  std::vector<double> noise_h;
  double *noise_pinned = nullptr;
  bool own_stream = true;
  // plain context: walls and the box protocol (device copy dW; hW holds its
  // array pointers); box_set once pcg_box_setup ran (the step then runs the
  // box kernels)
  WallDev *dW = nullptr;
  WallDev hW{};
  bool box_set = false;

  ~DeviceBatch() {
    if (arena.base && !arena.borrowed) (void)hipFree(arena.base);
    if (noise_pinned) (void)hipHostFree(noise_pinned);
    if (stream && own_stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

int make_batch(const pcg_params &p, int nsys, const pcg_model *models, const double *temps,
               const int32_t *seeds, int prec, uint64_t philox_seed, int device,
               std::unique_ptr<DeviceBatch> &out, hipStream_t use_stream = nullptr,
               size_t *dry_run_bytes = nullptr, int64_t ran2_skip = 0,
               void *pool_block = nullptr, size_t pool_bytes = 0, bool plain = false) {
  if (nsys <= 0 || !models) return fail(PCG_E_ARG, "no systems");
  if (prec != PCG_MIXED && prec != PCG_FP64) return fail(PCG_E_ARG, "unknown precision");
  if (p.n_walls < 0 || p.n_walls > 6) return fail(PCG_E_ARG, "n_walls must be in [0, 6]");
  {
    bool side_used[6] = {false, false, false, false, false, false};
    for (int w = 0; w < p.n_walls; ++w) {
      if (p.wall_kind[w] < 0 || p.wall_kind[w] > 2)
        return fail(PCG_E_ARG, "wall_kind must be 0 (solid), 1 (harmonic) or 2 (LJ)");
      if (p.wall_side[w] < -1 || p.wall_side[w] > 5)
        return fail(PCG_E_ARG, "wall_side must be -1 or a box side 0..5");
      if (p.wall_side[w] >= 0) {
        if (side_used[p.wall_side[w]]) return fail(PCG_E_ARG, "two walls on one box side");
        side_used[p.wall_side[w]] = true;
      }
      if (p.wall_kind[w] != 0 && p.wall_limit[w] < 0)
        return fail(PCG_E_ARG, "wall_limit must not be negative");
      if (!plain && p.wall_kind[w] != 0)
        return fail(PCG_E_ARG, "harmonic and LJ walls run in the plain context (pcg_create) only");
    }
    if (plain && nsys != 1) return fail(PCG_E_ARG, "a plain context holds one system");
  }
  auto B = std::make_unique<DeviceBatch>();
  B->prec = prec;
  B->device = device;
  HIPCHK(hipSetDevice(device));
  if (use_stream) {
    B->stream = use_stream;
    B->own_stream = false;
  } else if (!dry_run_bytes) {
    HIPCHK(hipStreamCreate(&B->stream));
  }
  fill_dev_params(p, B->hp);
  B->hs.resize(nsys);
  for (int s = 0; s < nsys; ++s) {
    int rc = build_system(p, models[s], temps ? temps[s] : -1.0, seeds ? seeds[s] : 448, B->hs[s],
                          ran2_skip);
    if (rc) return rc;
  }
  // offsets
  long long ntot = 0, nteth = 0, nlrep = 0, nchir = 0, nnang = 0, nhang = 0, nndih = 0,
            nhdih = 0, nnc = 0, ntip = 0, nl = 0, qf = 0, qc = 0, ncells = 0;
  B->sys.resize(nsys);
  for (int s = 0; s < nsys; ++s) {
    HostSys &h = B->hs[s];
    SysDesc &d = B->sys[s];
    std::memset(&d, 0, sizeof(d));
    d.res0 = (int)ntot;
    d.n = h.n;
    d.teth0 = (int)nteth, d.nteth = (int)h.teth_nat.size();
    d.lrep0 = (int)nlrep, d.nlrep = (int)h.lrep[0].size();
    d.chir0 = (int)nchir, d.nchir = (int)h.chir_f0.size();
    d.nang0 = (int)nnang, d.nnang = (int)h.nang_th.size();
    d.hang0 = (int)nhang, d.nhang = (int)h.hang_t.size();
    d.ndih0 = (int)nndih, d.nndih = (int)h.ndih_ph.size();
    d.hdih0 = (int)nhdih, d.nhdih = (int)h.hdih_t.size();
    d.nc0 = (int)nnc, d.nnc = (int)h.nc_nat.size();
    d.tip0 = (int)ntip, d.ntip = (int)h.tip_res.size() - h.nftip, d.nftip = h.nftip;
    d.nl0 = (int)nl, d.nlcap = h.nlcap;
    d.qf0 = (int)qf, d.qfcap = h.qfcap;
    d.qc0 = (int)qc, d.qccap = h.qccap;
    const double *cell = models[s].cell;
    for (int c = 0; c < 3; ++c) {
      d.cell[c] = cell[c];
      d.cell_inv[c] = cell[c] != 0.0 ? 1.0 / cell[c] : 0.0;
    }
    // the stage rebuild's cell grid: a system of at least CELL_MIN_N
    // residues in a cell periodic along every axis with room for three
    // cells a side of at least the NL radius (with a relative 1e-6 to spare)
    d.cell0 = (int)ncells;
    d.cg[0] = d.cg[1] = d.cg[2] = 0;
    {
      const double req = (std::max(p.nl_cutoff, p.dh_cutoff) + p.nl_pad) * (1.0 + 1e-6);
      long long cells = 1;
      int gc[3];
      bool ok = h.n >= CELL_MIN_N;
      for (int c = 0; c < 3 && ok; ++c) {
        gc[c] = cell[c] > 0.0 ? (int)std::min(std::floor(cell[c] / req), 1024.0) : 0;
        ok = gc[c] >= 3;
        cells *= ok ? gc[c] : 1;
      }
      if (ok && cells <= CELL_MAX && (h.n + 63) / 64 <= CELL_WORDS_MAX) {
        for (int c = 0; c < 3; ++c) d.cg[c] = gc[c];
        ncells += cells;
        B->cell_words = std::max(B->cell_words, (h.n + 63) / 64);
      }
    }
    d.temperature = h.temperature;
    d.const2 = const2_of(p, h.temperature);
    d.philox_key = mix64(philox_seed ^ mix64((uint64_t)s + 1));
    d.steps_target = 0;
    ntot += h.n;
    nteth += d.nteth, nlrep += d.nlrep, nchir += d.nchir, nnang += d.nnang, nhang += d.nhang;
    nndih += d.nndih, nhdih += d.nhdih, nnc += d.nnc, ntip += d.ntip + d.nftip;
    nl += h.nlcap, qf += h.qfcap, qc += h.qccap;
    B->max_nlcap = std::max(B->max_nlcap, h.nlcap);
    B->max_n = std::max(B->max_n, h.n);
    B->max_nnc = std::max(B->max_nnc, d.nnc);
    int items = std::max({d.nteth, d.nlrep, d.nchir, d.nnang, d.nhang, d.nndih, d.nhdih, d.nnc,
                          h.nlcap, h.qfcap});
    B->max_items = std::max(B->max_items, items);
  }
  if (nl > INT32_MAX / 2 || qf > INT32_MAX / 2) return fail(PCG_E_NOMEM, "batch too large");
  DevBatch &D = B->d;
  D.nsys = nsys;
  D.ntot = (int)ntot;
  Arena &A = B->arena;
  size_t fsz = prec == PCG_FP64 ? sizeof(double) : sizeof(float);
  SysDesc *dsys;
  SysState *dst;
  A.add(&dsys, nsys);
  A.add(&dst, nsys);
  A.add(&B->dP, 1);
  for (int k = 0; k < 6; ++k)
    for (int c = 0; c < 3; ++c) A.add(&D.y[k][c], ntot);
  for (int c = 0; c < 3; ++c) {
    A.add(&D.vel[c], ntot);
    A.add((char **)&D.F[c], ntot * fsz);
    A.add(&D.orig_r[c], ntot);
    A.add(&D.nat_r[c], ntot);
    A.add(&D.noise[c], ntot);
  }
  for (int c = 0; c < 6; ++c) A.add(&D.nh[c], ntot);
  uint8_t *atype;
  int *chain, *seq, *prev, *next, *sysid, *local;
  A.add(&atype, ntot);
  A.add(&chain, ntot);
  A.add(&seq, ntot);
  A.add(&prev, ntot);
  A.add(&next, ntot);
  A.add(&sysid, ntot);
  A.add(&local, ntot);
  A.add(&D.sync, 4 * ntot);
  A.add(&D.row_cnt, ntot + 1);
  A.add(&D.cell_of, ntot);
  A.add(&D.cell_items, ntot);
  A.add(&D.cell_start, std::max(ncells, 1LL));
  A.add(&D.cell_end, std::max(ncells, 1LL));
  int *teth_i[2], *lrep_i[2], *chir_i[4], *nang_i[3], *hang_i[3], *ndih_i[4], *hdih_i[4], *nc_i[2];
  double *teth_nat, *chir_f0, *chir_f1, *nang_th, *ndih_ph, *nc_nat;
  uint8_t *hang_t, *hdih_t, *nc_type;
  int *tip_res;
  double *tip_orig[3], *tip_vel[3];
  for (int k = 0; k < 2; ++k) A.add(&teth_i[k], std::max(nteth, 1LL));
  A.add(&teth_nat, std::max(nteth, 1LL));
  for (int k = 0; k < 2; ++k) A.add(&lrep_i[k], std::max(nlrep, 1LL));
  for (int k = 0; k < 4; ++k) A.add(&chir_i[k], std::max(nchir, 1LL));
  A.add(&chir_f0, std::max(nchir, 1LL));
  A.add(&chir_f1, std::max(nchir, 1LL));
  for (int k = 0; k < 3; ++k) A.add(&nang_i[k], std::max(nnang, 1LL));
  A.add(&nang_th, std::max(nnang, 1LL));
  for (int k = 0; k < 3; ++k) A.add(&hang_i[k], std::max(nhang, 1LL));
  A.add(&hang_t, std::max(nhang, 1LL));
  for (int k = 0; k < 4; ++k) A.add(&ndih_i[k], std::max(nndih, 1LL));
  A.add(&ndih_ph, std::max(nndih, 1LL));
  for (int k = 0; k < 4; ++k) A.add(&hdih_i[k], std::max(nhdih, 1LL));
  A.add(&hdih_t, std::max(nhdih, 1LL));
  for (int k = 0; k < 2; ++k) A.add(&nc_i[k], std::max(nnc, 1LL));
  A.add(&nc_nat, std::max(nnc, 1LL));
  A.add(&nc_type, std::max(nnc, 1LL));
  A.add(&D.nc_active, std::max(nnc, 1LL));
  A.add(&D.nc_change_t, std::max(nnc, 1LL));
  A.add(&D.ncur, std::max(nnc, 1LL));
  A.add(&D.ncur_i[0], std::max(nnc, 1LL));
  A.add(&D.ncur_i[1], std::max(nnc, 1LL));
  A.add(&D.ncur_nat, std::max(nnc, 1LL));
  A.add(&D.ncur_t, std::max(nnc, 1LL));
  A.add(&D.ncur_th2, std::max(nnc, 1LL));
  A.add(&D.ncur_ijt, std::max(nnc, 1LL));
  // chain-local bonded items by base residue (see BOND_* in pcg_core.hpp)
  std::vector<int> bond_item((size_t)NBOND * std::max(ntot, 1LL), -1);
  for (int s = 0; s < nsys; ++s) {
    HostSys &h = B->hs[s];
    SysDesc &d = B->sys[s];
    bool ok = true;
    auto put = [&](int L, int base, const std::vector<int> *idx, const int *step, int slots) {
      size_t m = idx[0].size();
      for (size_t k = 0; k < m && ok; ++k) {
        int r = idx[0][k];
        for (int sl = 1; sl < slots; ++sl) ok = ok && idx[sl][k] == r + step[sl];
        size_t at = (size_t)L * ntot + d.res0 + r;
        ok = ok && r >= 0 && r + step[slots - 1] < h.n && bond_item[at] < 0;
        if (ok) bond_item[at] = base + (int)k;
      }
    };
    const int pair1[2] = {0, 1}, pair2[2] = {0, 2}, tri[3] = {0, 1, 2}, quad[4] = {0, 1, 2, 3};
    put(BOND_TETHER, d.teth0, h.teth, pair1, 2);
    put(BOND_LREP, d.lrep0, h.lrep, pair2, 2);
    put(BOND_NAT_ANG, d.nang0, h.nang, tri, 3);
    put(BOND_HEUR_ANG, d.hang0, h.hang, tri, 3);
    put(BOND_NAT_DIH, d.ndih0, h.ndih, quad, 4);
    put(BOND_HEUR_DIH, d.hdih0, h.hdih, quad, 4);
    put(BOND_CHIR, d.chir0, h.chir, quad, 4);
    d.bond_regular = ok ? 1 : 0;
  }
  int *d_bond_item;
  A.add(&d_bond_item, bond_item.size());
  A.add(&tip_res, std::max(ntip, 1LL));
  for (int c = 0; c < 3; ++c) {
    A.add(&tip_orig[c], std::max(ntip, 1LL));
    A.add(&tip_vel[c], std::max(ntip, 1LL));
  }
  A.add(&D.tip_stat, std::max(ntip, 1LL));
  for (int k = 0; k < 2; ++k) {
    A.add(&D.nl_i[k], nl);
    A.add(&D.pauli_i[k], nl);
    A.add(&D.dh_i[k], nl);
    A.add(&D.pid_i[k], nl);
    A.add(&D.cand_i[k], nl);
    A.add(&D.qf_i[k], qf);
    A.add(&D.qc_i[k], qc);
  }
  A.add(&D.nl_d, nl);
  A.add(&D.nl_cls, nl);
  A.add(&D.dh_q, nl);
  A.add(&D.pid_d, nl);
  A.add(&D.pid_t, nl);
  A.add(&D.cand_od, nl);
  A.add(&D.cand_dist, nl);
  A.add(&D.cand_idx, nl);
  A.add(&D.cand_t, nl);
  A.add(&D.cand_sd, 8 * nl);
  A.add(&D.cand_perm, nl);
  A.add(&D.qf_d, qf);
  A.add(&D.qf_vac, qf);
  A.add(&D.qh_slot, qf);
  A.add(&D.qh_i[0], qf);
  A.add(&D.qh_i[1], qf);
  A.add(&D.qh_vac, qf);
  A.add(&D.qh_req, qf);
  A.add(&D.qh_thr2, qf);
  A.add(&D.qh_ij, qf);
  A.add(&D.cand_h, nl);
  A.add(&D.qc_d, qc);
  A.add(&D.qc_t, qc);
  A.add(&D.qc_sat, qc);
  A.add(&D.qc_sd, 8 * qc);
  A.add(&D.qc_vac, qc);
  A.add(&D.rem, qc);
  A.add(&D.blk_cnt, (nl / BLK + 1) * NCL);
  // walls of a plain context: harmonic connections (state::setup_walls,
  // state.cpp:653-670, from the starting positions), LJ slots and candidates
  struct HwConn {
    int wall, res;
    double off[3];
  };
  std::vector<HwConn> hw;
  std::vector<uint8_t> conn;
  WallDev &W = B->hW;
  std::memset(&W, 0, sizeof(W));
  if (plain) {
    const HostSys &h = B->hs[0];
    const int n = h.n;
    conn.assign(n, 0);
    auto plane_dist = [&](int w, int i) {
      const double *o = p.wall_origin[w], *nv = p.wall_normal[w], *r = &h.pos[3 * i];
      return nv[0] * (r[0] - o[0]) + nv[1] * (r[1] - o[1]) + nv[2] * (r[2] - o[2]);
    };
    for (int w = 0; w < p.n_walls; ++w) {
      if (p.wall_kind[w] != 1) continue;
      if (p.wall_side[w] >= 0 && p.wall_side[w] % 2 == 0) continue;  // the axis' second plane ends it
      for (int hw_i = 0; hw_i <= w; ++hw_i) {
        if (p.wall_kind[hw_i] != 1) continue;
        std::vector<std::pair<double, int>> dist(n);
        for (int i = 0; i < n; ++i) dist[i] = std::make_pair(std::abs(plane_dist(hw_i, i)), i);
        std::sort(dist.begin(), dist.end());
        for (int idx = 0, made = 0; made < p.wall_limit[hw_i] && idx < n; ++idx) {
          const int res = dist[idx].second;
          if (conn[res]) continue;
          const double sd = plane_dist(hw_i, res), *o = p.wall_origin[hw_i], *nv = p.wall_normal[hw_i];
          HwConn c{hw_i, res, {0, 0, 0}};
          for (int k = 0; k < 3; ++k) c.off[k] = (h.pos[3 * res + k] - sd * nv[k]) - o[k];
          hw.push_back(c);
          conn[res] = 1;
          ++made;
        }
      }
    }
    long long lj_lim = 0;
    int lj_planes = 0;
    for (int w = 0; w < p.n_walls; ++w)
      if (p.wall_kind[w] == 2) lj_lim += p.wall_limit[w], ++lj_planes;
    W.lj_cap = (int)std::min<long long>(lj_lim, 16LL * n) + 1;
    W.cand_cap = std::max(1, lj_planes * n);
    A.add(&B->dW, 1);
    A.add((int **)&W.hw_res, std::max<size_t>(hw.size(), 1));
    A.add((int **)&W.hw_wall, std::max<size_t>(hw.size(), 1));
    for (int k = 0; k < 3; ++k) A.add((double **)&W.hw_off[k], std::max<size_t>(hw.size(), 1));
    A.add((double **)&W.hw_nat, std::max<size_t>(hw.size(), 1));
    A.add(&W.lj_res, W.lj_cap);
    A.add(&W.lj_wall, W.lj_cap);
    for (int k = 0; k < 3; ++k) A.add(&W.lj_off[k], W.lj_cap);
    A.add(&W.lj_sat, W.lj_cap);
    A.add(&W.lj_gone, W.lj_cap);
    A.add(&W.cand_res, W.cand_cap);
    A.add(&W.cand_wall, W.cand_cap);
    A.add(&W.cand_d, W.cand_cap);
    A.add(&W.is_conn, std::max(n, 1));
  }
  if (dry_run_bytes) {
    *dry_run_bytes = A.total;
    return PCG_OK;
  }
  int rc = A.alloc(B->stream, pool_block, pool_bytes);
  if (rc) return rc;
  B->bytes = A.total;
  D.sys = (SysDescC *)dsys;
  D.st = dst;
  D.P = (DevParamsC *)B->dP;
  D.wall = B->dW;
  D.atype = atype, D.chain = chain, D.seq = seq, D.prev = prev, D.next = next;
  D.sysid = sysid, D.local = local;
  for (int k = 0; k < 2; ++k) D.teth_i[k] = teth_i[k], D.lrep_i[k] = lrep_i[k], D.nc_i[k] = nc_i[k];
  for (int k = 0; k < 4; ++k) D.chir_i[k] = chir_i[k], D.ndih_i[k] = ndih_i[k], D.hdih_i[k] = hdih_i[k];
  for (int k = 0; k < 3; ++k) D.nang_i[k] = nang_i[k], D.hang_i[k] = hang_i[k];
  D.teth_nat = teth_nat, D.chir_f0 = chir_f0, D.chir_f1 = chir_f1, D.nang_th = nang_th;
  D.ndih_ph = ndih_ph, D.nc_nat = nc_nat, D.hang_t = hang_t, D.hdih_t = hdih_t, D.nc_type = nc_type;
  D.tip_res = tip_res;
  for (int c = 0; c < 3; ++c) D.tip_orig[c] = tip_orig[c], D.tip_vel[c] = tip_vel[c];
  for (int L = 0; L < NBOND; ++L) D.bond_item[L] = d_bond_item + (size_t)L * std::max(ntot, 1LL);

  // host staging, then one copy per array
  std::vector<SysState> st(nsys);
  for (int s = 0; s < nsys; ++s) {
    std::memset(&st[s], 0, sizeof(SysState));
    st[s].nl_invalid = 1;
    st[s].verify_first = 1;
    st[s].active = 1;
  }
  if (plain) st[0].n_hw = (int)hw.size();
  // Every array set from the host but the walls' sits before the first
  // NL-sized array: those go into one host image of that prefix of the
  // arena, zeros elsewhere as the arena starts, which goes up in one copy
  // (a job's admission is then one copy and one synchronisation); writes
  // past the prefix are copied on their own.
  const size_t prefix = (size_t)((char *)D.nl_i[0] - (char *)A.base);
  std::vector<char> image(prefix, 0);
  auto up = [&](void *dst_, const void *src, size_t bytes) -> int {
    if (!bytes) return PCG_OK;
    const size_t off = (size_t)((char *)dst_ - (char *)A.base);
    if ((char *)dst_ >= (char *)A.base && off + bytes <= prefix) {
      std::memcpy(image.data() + off, src, bytes);
      return PCG_OK;
    }
    HIPCHK(hipMemcpyAsync(dst_, src, bytes, hipMemcpyHostToDevice, B->stream));
    return PCG_OK;
  };
  std::vector<double> col(ntot);
  std::vector<int> ci(ntot);
  std::vector<uint8_t> cu(ntot);
  std::vector<int8_t> sync(4 * ntot);
  std::vector<double> ybuf[6][3], vel[3], natb[3];
  for (int c = 0; c < 3; ++c) natb[c].assign(ntot, 0.0);
  for (int k = 0; k < 6; ++k)
    for (int c = 0; c < 3; ++c) ybuf[k][c].assign(ntot, 0.0);
  std::vector<int> vchain(ntot), vseq(ntot), vprev(ntot), vnext(ntot), vsys(ntot), vloc(ntot);
  std::vector<uint8_t> vat(ntot);
  for (int s = 0; s < nsys; ++s) {
    HostSys &h = B->hs[s];
    int r0 = B->sys[s].res0;
    for (int i = 0; i < h.n; ++i) {
      int g = r0 + i;
      for (int c = 0; c < 3; ++c) {
        ybuf[0][c][g] = h.pos[3 * i + c];
        ybuf[1][c][g] = h.y1[3 * i + c];
        natb[c][g] = h.orig[3 * i + c];
      }
      vat[g] = h.atype[i];
      vchain[g] = h.chain[i];
      vseq[g] = h.seq[i];
      vprev[g] = (i > 0 && h.chain[i - 1] == h.chain[i]) ? g - 1 : -1;
      vnext[g] = (i + 1 < h.n && h.chain[i + 1] == h.chain[i]) ? g + 1 : -1;
      vsys[g] = s;
      vloc[g] = i;
      for (int k = 0; k < 4; ++k) sync[4 * g + k] = (int8_t)p.base_sync[h.atype[i]][k];
    }
  }
  for (int k = 0; k < 6; ++k)
    for (int c = 0; c < 3; ++c)
      if ((rc = up(D.y[k][c], ybuf[k][c].data(), ntot * sizeof(double)))) return rc;
  for (int c = 0; c < 3; ++c)
    if ((rc = up(D.nat_r[c], natb[c].data(), ntot * sizeof(double)))) return rc;
  if ((rc = up(atype, vat.data(), ntot))) return rc;
  if ((rc = up(chain, vchain.data(), ntot * 4))) return rc;
  if ((rc = up(seq, vseq.data(), ntot * 4))) return rc;
  if ((rc = up(prev, vprev.data(), ntot * 4))) return rc;
  if ((rc = up(next, vnext.data(), ntot * 4))) return rc;
  if ((rc = up(sysid, vsys.data(), ntot * 4))) return rc;
  if ((rc = up(local, vloc.data(), ntot * 4))) return rc;
  if ((rc = up(D.sync, sync.data(), 4 * ntot))) return rc;
  if ((rc = up(d_bond_item, bond_item.data(), bond_item.size() * 4))) return rc;
  // static lists, concatenated with residue indices shifted to the batch
  auto cat_i = [&](int *dst_, auto getv, long long total) -> int {
    std::vector<int> v;
    v.reserve(total);
    for (int s = 0; s < nsys; ++s) {
      const std::vector<int> &x = getv(B->hs[s]);
      for (int a : x) v.push_back(a + B->sys[s].res0);
    }
    return up(dst_, v.data(), v.size() * 4);
  };
  auto cat_d = [&](double *dst_, auto getv) -> int {
    std::vector<double> v;
    for (int s = 0; s < nsys; ++s) {
      const std::vector<double> &x = getv(B->hs[s]);
      v.insert(v.end(), x.begin(), x.end());
    }
    return up(dst_, v.data(), v.size() * 8);
  };
  auto cat_u = [&](uint8_t *dst_, auto getv) -> int {
    std::vector<uint8_t> v;
    for (int s = 0; s < nsys; ++s) {
      const std::vector<uint8_t> &x = getv(B->hs[s]);
      v.insert(v.end(), x.begin(), x.end());
    }
    return up(dst_, v.data(), v.size());
  };
#define CATI(dst_, field) \
  if ((rc = cat_i(dst_, [&](HostSys &h) -> const std::vector<int> & { return h.field; }, 0))) return rc
#define CATD(dst_, field) \
  if ((rc = cat_d(dst_, [&](HostSys &h) -> const std::vector<double> & { return h.field; }))) return rc
#define CATU(dst_, field) \
  if ((rc = cat_u(dst_, [&](HostSys &h) -> const std::vector<uint8_t> & { return h.field; }))) return rc
  for (int k = 0; k < 2; ++k) {
    CATI(teth_i[k], teth[k]);
    CATI(lrep_i[k], lrep[k]);
    CATI(nc_i[k], nc[k]);
  }
  for (int k = 0; k < 4; ++k) {
    CATI(chir_i[k], chir[k]);
    CATI(ndih_i[k], ndih[k]);
    CATI(hdih_i[k], hdih[k]);
  }
  for (int k = 0; k < 3; ++k) {
    CATI(nang_i[k], nang[k]);
    CATI(hang_i[k], hang[k]);
  }
  CATI(tip_res, tip_res);
  CATD(teth_nat, teth_nat);
  CATD(chir_f0, chir_f0);
  CATD(chir_f1, chir_f1);
  CATD(nang_th, nang_th);
  CATD(ndih_ph, ndih_ph);
  CATD(nc_nat, nc_nat);
  for (int c = 0; c < 3; ++c) {
    CATD(tip_orig[c], tip_orig[c]);
    CATD(tip_vel[c], tip_vel[c]);
  }
  CATU(hang_t, hang_t);
  CATU(hdih_t, hdih_t);
  CATU(nc_type, nc_type);
  CATU(D.nc_active, nc_active);
  {
    std::vector<double> ct(std::max(nnc, 1LL), -1.0);
    if ((rc = up(D.nc_change_t, ct.data(), ct.size() * 8))) return rc;
  }
  {
    // every tip's moving averages start empty, with the parameters' window
    std::vector<TipStat> ts(std::max(ntip, 1LL), fresh_tip_stat(p));
    if ((rc = up(D.tip_stat, ts.data(), ts.size() * sizeof(TipStat)))) return rc;
  }
#undef CATI
#undef CATD
#undef CATU
  if (plain) {
    const size_t nh = hw.size();
    std::vector<int> hr(nh), hwl(nh);
    std::vector<double> ho[3], hn(nh, 0.0);  // the natural distance: 0 (state.cpp:664-665)
    for (int k = 0; k < 3; ++k) ho[k].resize(nh);
    for (size_t a = 0; a < nh; ++a) {
      hr[a] = hw[a].res, hwl[a] = hw[a].wall;
      for (int k = 0; k < 3; ++k) ho[k][a] = hw[a].off[k];
    }
    if ((rc = up((void *)W.hw_res, hr.data(), nh * 4))) return rc;
    if ((rc = up((void *)W.hw_wall, hwl.data(), nh * 4))) return rc;
    for (int k = 0; k < 3; ++k)
      if ((rc = up((void *)W.hw_off[k], ho[k].data(), nh * 8))) return rc;
    if ((rc = up((void *)W.hw_nat, hn.data(), nh * 8))) return rc;
    if ((rc = up(W.is_conn, conn.data(), conn.size()))) return rc;
    std::vector<uint8_t> zero8(W.lj_cap, 0);
    if ((rc = up(W.lj_gone, zero8.data(), zero8.size()))) return rc;
    for (int side = 0; side < 6; ++side) W.side_wall[side] = -1;
    for (int w = 0; w < p.n_walls; ++w)
      if (p.wall_side[w] >= 0) W.side_wall[p.wall_side[w]] = w;
    if ((rc = up(B->dW, &W, sizeof(WallDev)))) return rc;
  }
  if ((rc = up(dsys, B->sys.data(), nsys * sizeof(SysDesc)))) return rc;
  if ((rc = up(dst, st.data(), nsys * sizeof(SysState)))) return rc;
  if ((rc = up(B->dP, &B->hp, sizeof(DevParams)))) return rc;
  // forces start at zero so that the first Langevin step (t == 0) reads them
  std::vector<char> zero(ntot * fsz, 0);
  for (int c = 0; c < 3; ++c)
    if ((rc = up(D.F[c], zero.data(), zero.size()))) return rc;
  if (prefix) HIPCHK(hipMemcpyAsync(A.base, image.data(), prefix, hipMemcpyHostToDevice, B->stream));
  HIPCHK(hipStreamSynchronize(B->stream));
  out = std::move(B);
  return PCG_OK;
}

int read_state(DeviceBatch &B, int s, SysState &st) {
  HIPCHK(hipMemcpyAsync(&st, B.d.st + s, sizeof(SysState), hipMemcpyDeviceToHost, B.stream));
  HIPCHK(hipStreamSynchronize(B.stream));
  return PCG_OK;
}
int write_state(DeviceBatch &B, int s, const SysState &st) {
  HIPCHK(hipMemcpyAsync(B.d.st + s, &st, sizeof(SysState), hipMemcpyHostToDevice, B.stream));
  HIPCHK(hipStreamSynchronize(B.stream));
  return PCG_OK;
}

// copies 3 device columns of system s into xyz-interleaved host memory
int get_cols(DeviceBatch &B, int s, double *const cols[3], double *out) {
  const SysDesc &d = B.sys[s];
  std::vector<double> tmp(3 * (size_t)d.n);
  for (int c = 0; c < 3; ++c)
    HIPCHK(hipMemcpyAsync(tmp.data() + c * d.n, cols[c] + d.res0, d.n * 8, hipMemcpyDeviceToHost,
                          B.stream));
  HIPCHK(hipStreamSynchronize(B.stream));
  for (int i = 0; i < d.n; ++i)
    for (int c = 0; c < 3; ++c) out[3 * i + c] = tmp[c * d.n + i];
  return PCG_OK;
}
int get_forces(DeviceBatch &B, int s, double *out) {
  const SysDesc &d = B.sys[s];
  size_t fsz = B.prec == PCG_FP64 ? 8 : 4;
  std::vector<char> tmp(3 * fsz * d.n);
  for (int c = 0; c < 3; ++c)
    HIPCHK(hipMemcpyAsync(tmp.data() + c * fsz * d.n, (char *)B.d.F[c] + fsz * d.res0, fsz * d.n,
                          hipMemcpyDeviceToHost, B.stream));
  HIPCHK(hipStreamSynchronize(B.stream));
  for (int i = 0; i < d.n; ++i)
    for (int c = 0; c < 3; ++c)
      out[3 * i + c] = B.prec == PCG_FP64 ? ((double *)tmp.data())[c * d.n + i]
                                          : (double)((float *)tmp.data())[c * d.n + i];
  return PCG_OK;
}

// The same three columns to or from caller arrays of one component each,
// element i at comp[c][i * stride]; unit stride copies straight between the
// device columns and the caller's arrays.
int get_cols3(DeviceBatch &B, int s, const void *const cols[3], bool f32, double *const out[3],
              int64_t stride) {
  const SysDesc &d = B.sys[s];
  const size_t esz = f32 ? 4 : 8;
  if (stride == 1 && !f32) {
    for (int c = 0; c < 3; ++c)
      HIPCHK(hipMemcpyAsync(out[c], (const char *)cols[c] + esz * d.res0, d.n * esz,
                            hipMemcpyDeviceToHost, B.stream));
    HIPCHK(hipStreamSynchronize(B.stream));
    return PCG_OK;
  }
  std::vector<char> tmp(3 * esz * d.n);
  for (int c = 0; c < 3; ++c)
    HIPCHK(hipMemcpyAsync(tmp.data() + c * esz * d.n, (const char *)cols[c] + esz * d.res0,
                          esz * d.n, hipMemcpyDeviceToHost, B.stream));
  HIPCHK(hipStreamSynchronize(B.stream));
  for (int c = 0; c < 3; ++c)
    for (int i = 0; i < d.n; ++i)
      out[c][(size_t)i * stride] = f32 ? (double)((const float *)tmp.data())[c * d.n + i]
                                       : ((const double *)tmp.data())[c * d.n + i];
  return PCG_OK;
}
int set_cols3(DeviceBatch &B, int s, double *const cols[3], const double *const in[3],
              int64_t stride) {
  const SysDesc &d = B.sys[s];
  if (stride == 1) {
    for (int c = 0; c < 3; ++c)
      HIPCHK(hipMemcpyAsync(cols[c] + d.res0, in[c], d.n * 8, hipMemcpyHostToDevice, B.stream));
  } else {
    std::vector<double> tmp(3 * (size_t)d.n);
    for (int c = 0; c < 3; ++c)
      for (int i = 0; i < d.n; ++i) tmp[c * d.n + i] = in[c][(size_t)i * stride];
    for (int c = 0; c < 3; ++c)
      HIPCHK(hipMemcpyAsync(cols[c] + d.res0, tmp.data() + c * d.n, d.n * 8,
                            hipMemcpyHostToDevice, B.stream));
    HIPCHK(hipStreamSynchronize(B.stream));
  }
  HIPCHK(hipStreamSynchronize(B.stream));
  return PCG_OK;
}

// Reference-order noise for RAN2 mode: all x, then all y, then all z per
// system (cg/src/cg/langevin/step.cpp:119-127).
int upload_ran2_noise(DeviceBatch &B) {
  int ntot = B.d.ntot;
  if (!B.noise_pinned) HIPCHK(hipHostMalloc((void **)&B.noise_pinned, 3 * ntot * sizeof(double)));
  for (size_t s = 0; s < B.sys.size(); ++s) {
    const SysDesc &d = B.sys[s];
    for (int c = 0; c < 3; ++c)
      for (int i = 0; i < d.n; ++i) B.noise_pinned[c * ntot + d.res0 + i] = B.ran2[s].normal();
  }
  for (int c = 0; c < 3; ++c)
    HIPCHK(hipMemcpyAsync(B.d.noise[c], B.noise_pinned + c * ntot, ntot * 8, hipMemcpyHostToDevice,
                          B.stream));
  // the pinned buffer is rewritten next step: wait for the copy
  HIPCHK(hipStreamSynchronize(B.stream));
  return PCG_OK;
}

// Eigenvalues of a symmetric 3x3 matrix, in descending order (cyclic Jacobi
// rotations until the off-diagonal part vanishes).
void eig_sym3(double a[3][3], double e[3]) {
  static const int P[3][2] = {{0, 1}, {0, 2}, {1, 2}};
  for (int sweep = 0; sweep < 64; ++sweep) {
    const double off = std::fabs(a[0][1]) + std::fabs(a[0][2]) + std::fabs(a[1][2]);
    const double dia = std::fabs(a[0][0]) + std::fabs(a[1][1]) + std::fabs(a[2][2]);
    if (off == 0.0 || off <= 1e-300 + 1e-18 * dia) break;
    for (auto &pq : P) {
      const int p = pq[0], q = pq[1];
      if (a[p][q] == 0.0) continue;
      const double th = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
      const double t = std::fabs(th) > 1e150 ? 0.5 / th
                                             : (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
      const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
      for (int k = 0; k < 3; ++k) {
        const double kp = a[k][p], kq = a[k][q];
        a[k][p] = c * kp - s * kq;
        a[k][q] = s * kp + c * kq;
      }
      for (int k = 0; k < 3; ++k) {
        const double pk = a[p][k], qk = a[q][k];
        a[p][k] = c * pk - s * qk;
        a[q][k] = s * pk + c * qk;
      }
    }
  }
  for (int k = 0; k < 3; ++k) e[k] = a[k][k];
  std::sort(e, e + 3, [](double x, double y) { return x > y; });
}

// make_report::add_cur_scalars (cg/src/cg/output/make_report.cpp:196-247)
// from the device sums of one system
void report_scalars(const ReportSums &o, const SysState &st, double *out) {
  const double N = o.n;
  out[PCG_REP_TIME] = st.t;
  out[PCG_REP_EPOT] = st.Vtot;
  out[PCG_REP_EKIN] = o.K;
  out[PCG_REP_ETOT] = st.Vtot + o.K;
  // gyration_radius: sqrt(sum |r - mean|^2 / (float)N) (make_report.cpp:168-179)
  out[PCG_REP_RG] = std::sqrt(o.pp / (double)(float)N);
  out[PCG_REP_L] = o.L;
  // rmsd (make_report.cpp:132-158): the optimal superposition's residual from
  // the singular values s1 >= s2 >= s3 of H = orig^T cur and d = sign det H:
  // sum |cur R - orig|^2 = |cur|^2 + |orig|^2 - 2 (s1 + s2 + d s3)
  double HtH[3][3], e[3];
  const double *Hm = o.H;
  for (int a = 0; a < 3; ++a)
    for (int c = 0; c < 3; ++c) {
      double x = 0.0;
      for (int k = 0; k < 3; ++k) x += Hm[3 * k + a] * Hm[3 * k + c];
      HtH[a][c] = x;
    }
  eig_sym3(HtH, e);
  const double detH = Hm[0] * (Hm[4] * Hm[8] - Hm[5] * Hm[7]) - Hm[1] * (Hm[3] * Hm[8] - Hm[5] * Hm[6]) +
                      Hm[2] * (Hm[3] * Hm[7] - Hm[4] * Hm[6]);
  const double s1 = std::sqrt(std::max(e[0], 0.0)), s2 = std::sqrt(std::max(e[1], 0.0)),
               s3 = std::sqrt(std::max(e[2], 0.0));
  const double r2 = (o.pp + o.qq - 2.0 * (s1 + s2 + (detH < 0 ? -s3 : s3))) / N;
  out[PCG_REP_RMSD] = std::sqrt(std::max(r2, 0.0));
  // asphericity (make_report.cpp:181-194): the singular values of the centred
  // positions are the square roots of the eigenvalues of sum (r - c)(r - c)^T;
  // W = 1.5 s1 - 0.5 (s1 + s2 + s3)
  double S[3][3] = {{o.S[0], o.S[3], o.S[4]}, {o.S[3], o.S[1], o.S[5]}, {o.S[4], o.S[5], o.S[2]}};
  eig_sym3(S, e);
  const double g1 = std::sqrt(std::max(e[0], 0.0)), g2 = std::sqrt(std::max(e[1], 0.0)),
               g3 = std::sqrt(std::max(e[2], 0.0));
  out[PCG_REP_W] = 1.5 * g1 - 0.5 * (g1 + g2 + g3);
  out[PCG_REP_ICN] = o.icn;
  out[PCG_REP_ICNSS] = o.icnss;
  out[PCG_REP_INTEHC] = o.inter;
  out[PCG_REP_INTRHC] = o.intra;
  for (int k = 0; k < 6; ++k) out[PCG_REP_PID_INTRHC + k] = o.pid[k];
}

// Replaces the AFM tips of system s (state::setup_afm, state.cpp:548-568):
// velocity tips, then force tips, in the slots the system's model reserved;
// their reference time becomes the system's current time. Zero tips turn
// the system's AFM off.
int set_tips(DeviceBatch &B, int s, int32_t nv, const int32_t *vres, const double *vorig,
             const double *vvel, int32_t nf, const int32_t *fres, const double *fforce) {
  if (s < 0 || s >= (int)B.sys.size()) return fail(PCG_E_ARG, "system index out of range");
  if (nv < 0 || nf < 0 || (nv > 0 && (!vres || !vorig || !vvel)) || (nf > 0 && (!fres || !fforce)))
    return fail(PCG_E_ARG, "bad tip arrays");
  SysDesc &d = B.sys[s];
  const int cap = (int)B.hs[s].tip_res.size();
  if (nv + nf > cap)
    return fail(PCG_E_ARG, "more tips than the model reserved (" + std::to_string(cap) + ")");
  std::vector<int32_t> res(nv + nf);
  std::vector<double> orig[3], vel[3];
  for (int c = 0; c < 3; ++c) orig[c].assign(nv + nf, 0.0), vel[c].assign(nv + nf, 0.0);
  for (int k = 0; k < nv + nf; ++k) {
    const bool v = k < nv;
    const int j = v ? k : k - nv;
    res[k] = v ? vres[j] : fres[j];
    if (res[k] < 0 || res[k] >= d.n) return fail(PCG_E_ARG, "tip residue out of range");
    res[k] += d.res0;
    for (int c = 0; c < 3; ++c) {
      orig[c][k] = v ? vorig[3 * j + c] : 0.0;
      vel[c][k] = v ? vvel[3 * j + c] : fforce[3 * j + c];
    }
  }
  SysState st;
  int rc = read_state(B, s, st);
  if (rc) return rc;
  if (nv + nf > 0) {
    const std::vector<TipStat> ts(nv + nf, fresh_tip_stat(B.hp.p));
    HIPCHK(hipMemcpyAsync(B.d.tip_stat + d.tip0, ts.data(), ts.size() * sizeof(TipStat),
                          hipMemcpyHostToDevice, B.stream));
    // the device views are const for the kernels; the arena itself is writable
    HIPCHK(hipMemcpyAsync((void *)(B.d.tip_res + d.tip0), res.data(), res.size() * sizeof(int32_t),
                          hipMemcpyHostToDevice, B.stream));
    for (int c = 0; c < 3; ++c) {
      HIPCHK(hipMemcpyAsync((void *)(B.d.tip_orig[c] + d.tip0), orig[c].data(), res.size() * 8,
                            hipMemcpyHostToDevice, B.stream));
      HIPCHK(hipMemcpyAsync((void *)(B.d.tip_vel[c] + d.tip0), vel[c].data(), res.size() * 8,
                            hipMemcpyHostToDevice, B.stream));
    }
  }
  d.ntip = nv;
  d.nftip = nf;
  d.afm_t0 = st.t;
  HIPCHK(hipMemcpyAsync((void *)(B.d.sys + s), &d, sizeof(SysDesc), hipMemcpyHostToDevice, B.stream));
  HIPCHK(hipStreamSynchronize(B.stream));
  return PCG_OK;
}

int report_into(DeviceBatch &B, double *out) {
  const size_t ns = B.sys.size();
  ReportSums *d = nullptr;
  if (hipMalloc(&d, ns * sizeof(ReportSums)) != hipSuccess)
    return fail(PCG_E_NOMEM, "no device memory for the report sums");
  launch_report(B.d, B.hp.p.pid_on, d, B.stream);
  std::vector<ReportSums> h(ns);
  std::vector<SysState> st(ns);
  hipError_t e1 = hipMemcpyAsync(h.data(), d, ns * sizeof(ReportSums), hipMemcpyDeviceToHost, B.stream);
  hipError_t e2 = hipMemcpyAsync(st.data(), B.d.st, ns * sizeof(SysState), hipMemcpyDeviceToHost, B.stream);
  hipError_t e3 = hipStreamSynchronize(B.stream);
  (void)hipFree(d);
  if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) return fail(PCG_E_HIP, "report copy failed");
  for (size_t s = 0; s < ns; ++s) report_scalars(h[s], st[s], out + s * PCG_REPORT_NUM);
  return PCG_OK;
}

// pcg_batch_report without allocation: the sums go to the library's device
// array, and the host reads them back in chunks through buffers on its stack
int report_into_static(DeviceBatch &B, double *out) {
  const int ns = (int)B.sys.size();
  if (ns > REPORT_STATIC_MAX)
    return fail(PCG_E_ARG, "pcg_batch_report_static: " + std::to_string(ns) +
                               " systems, more than its maximum of " +
                               std::to_string(REPORT_STATIC_MAX));
  ReportSums *d = nullptr;
  int rc = report_static_sums(B.d, B.device, B.hp.p.pid_on, B.stream, &d);
  if (rc == -2) return fail(PCG_E_ARG, "pcg_batch_report_static: device index beyond 63");
  if (rc) return fail(PCG_E_HIP, "pcg_batch_report_static: report kernels or symbol lookup failed");
  constexpr int CH = 16;
  ReportSums h[CH];
  SysState st[CH];
  for (int s0 = 0; s0 < ns; s0 += CH) {
    const int k = std::min(CH, ns - s0);
    HIPCHK(hipMemcpyAsync(h, d + s0, k * sizeof(ReportSums), hipMemcpyDeviceToHost, B.stream));
    HIPCHK(hipMemcpyAsync(st, B.d.st + s0, k * sizeof(SysState), hipMemcpyDeviceToHost, B.stream));
    HIPCHK(hipStreamSynchronize(B.stream));
    for (int j = 0; j < k; ++j) report_scalars(h[j], st[j], out + (size_t)(s0 + j) * PCG_REPORT_NUM);
  }
  return PCG_OK;
}

int check_overflow(DeviceBatch &B) {
  // every system's scalars in one copy
  std::vector<SysState> st(B.sys.size());
  HIPCHK(hipMemcpyAsync(st.data(), B.d.st, st.size() * sizeof(SysState), hipMemcpyDeviceToHost,
                        B.stream));
  HIPCHK(hipStreamSynchronize(B.stream));
  for (size_t s = 0; s < st.size(); ++s)
    if (st[s].overflow)
      return fail(PCG_E_OVERFLOW, "a neighbour or QA list of system " + std::to_string(s) +
                                      " outgrew its capacity");
  if (B.dW && B.hp.ljw_terms) {
    int ovf = 0;
    HIPCHK(hipMemcpyAsync(&ovf, (char *)B.dW + offsetof(WallDev, overflow), sizeof(int),
                          hipMemcpyDeviceToHost, B.stream));
    HIPCHK(hipStreamSynchronize(B.stream));
    if (ovf)
      return fail(PCG_E_OVERFLOW, "the LJ wall candidates (at most 4096 a step) or connections "
                                  "outgrew their capacity");
  }
  return PCG_OK;
}

size_t lds_for(const DeviceBatch &B, bool staged) {
  return LdsLayout((size_t)B.max_n, sizeof(double), staged, false, (size_t)B.max_nnc).total;  // FP64 LDS forces
}

// Launch configuration of the persistent kernel for a batch: per-residue
// bonded data joins positions and forces in LDS when it fits, the Gear
// derivatives too when they also fit; systems whose positions and forces do
// not fit take the stage kernels (ok false). The default workgroup is the
// smallest power of two of at least twice the largest system, in [128, 512].
struct PersistentCfg {
  bool ok;
  int block;
  size_t lds;
};
PersistentCfg persistent_cfg(DeviceBatch &B) {
  bool staged = lds_for(B, true) <= LDS_DYN_MAX;
  bool gear = staged && LdsLayout((size_t)B.max_n, sizeof(double), true, true,
                                  (size_t)B.max_nnc).total <= LDS_DYN_MAX;
  PersistentCfg c;
  c.lds = LdsLayout((size_t)B.max_n, sizeof(double), staged, gear, (size_t)B.max_nnc).total;
  B.d.bond_lds = staged ? 1 : 0;
  B.d.gear_lds = gear ? 1 : 0;
  // (the rebuild packs a residue's sequence index and chain in 12 bits each)
  c.ok = c.lds <= LDS_DYN_MAX && B.max_n < 4096;
  int want = 64;
  while (want < 1024 && want < B.max_n * 2) want *= 2;
  c.block = std::min(std::max(want, 128), 512);
  return c;
}

} // namespace

// ------------------------------------------------------------------ ctx ----
// A plain context steps in one of two forms: the stage kernels (one launch
// per stage of the reference's step, optionally replayed as a hipGraph), or
// the persistent kernel of the batched form for this one system (one
// workgroup, many steps a launch, everything in LDS). The persistent form is
// taken when the system fits its LDS, the noise is Philox (the reference's
// ran2 noise is drawn on the host every step) and there are no walls or box
// protocol (which the stage path carries); a switch between the forms makes
// the next step rebuild the neighbour list, so that each form works on lists
// it built itself.
enum { FORM_NONE = 0, FORM_STAGES = 1, FORM_PERSISTENT = 2 };
struct pcg_ctx {
  std::unique_ptr<DeviceBatch> B;
  bool graph_on = false;
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  int64_t steps_since_check = 0;
  int form = PCG_FORM_AUTO;  // requested
  int last_form = FORM_NONE;
  PersistentCfg pcfg{false, 0, 0};
  int spl = 500;             // persistent form: steps per launch
  ~pcg_ctx() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

namespace {

SysDesc *sys_rw(DeviceBatch &B) { return const_cast<SysDesc *>((const SysDesc *)B.d.sys); }

void enqueue_step(DeviceBatch &B) {
  const DevBatch &d = B.d;
  const pcg_params &p = B.hp.p;
  if (B.box_set) launch_box_phase(d, B.dP, sys_rw(B), B.stream);
  launch_step_begin(d, B.prec, p.qa_on, B.stream);
  launch_fix_nl(d, B.max_n, B.max_nlcap, B.cell_words, B.stream, false);
  launch_terms(d, B.prec, B.max_items, -1, 0, -1, B.stream);
  if (p.qa_on) launch_qa_finish(d, B.max_n, B.stream);
  if (B.hp.ljw_terms) launch_ljw_process(d, B.stream);
  if (B.box_set) launch_box_log(d, B.stream);
  launch_lang(d, B.prec, B.noise, B.stream, true);
}

// the reaction forces of the walls start a term evaluation at zero
int zero_wall_F(DeviceBatch &B) {
  if (B.dW)
    HIPCHK(hipMemsetAsync((char *)B.dW + offsetof(WallDev, wall_F), 0, sizeof(B.hW.wall_F),
                          B.stream));
  return PCG_OK;
}

// the host copy of the system descriptor after the box protocol moved the
// PBC cell on the device
int refresh_sys(DeviceBatch &B) {
  if (!B.box_set) return PCG_OK;
  HIPCHK(hipMemcpyAsync(B.sys.data(), (const void *)B.d.sys, B.sys.size() * sizeof(SysDesc),
                        hipMemcpyDeviceToHost, B.stream));
  HIPCHK(hipStreamSynchronize(B.stream));
  return PCG_OK;
}

bool persistent_eligible(const pcg_ctx &c) {
  const DeviceBatch &B = *c.B;
  return c.pcfg.ok && B.noise == PCG_NOISE_PHILOX && B.hp.p.n_walls == 0 && !B.box_set;
}
bool use_persistent(const pcg_ctx &c) {
  return c.form == PCG_FORM_PERSISTENT || (c.form == PCG_FORM_AUTO && persistent_eligible(c));
}
// the first step in a form after steps in the other rebuilds the lists (the
// device state records the form, so a restored checkpoint carries it)
int switch_form(pcg_ctx &c, int to) {
  DeviceBatch &B = *c.B;
  if (c.last_form == to) return PCG_OK;
  SysState st;
  int rc = read_state(B, 0, st);
  if (rc) return rc;
  if (st.form != FORM_NONE && st.form != to) {
    st.verify_first = 1;
    st.nl_invalid = 1;
  }
  st.form = to;
  if ((rc = write_state(B, 0, st))) return rc;
  c.last_form = to;
  return PCG_OK;
}

int read_wall(DeviceBatch &B, WallDev &W) {
  HIPCHK(hipMemcpyAsync(&W, B.dW, sizeof(WallDev), hipMemcpyDeviceToHost, B.stream));
  HIPCHK(hipStreamSynchronize(B.stream));
  return PCG_OK;
}
// writes the state part of W, with this context's array pointers
int write_wall(DeviceBatch &B, WallDev W) {
  const WallDev &H = B.hW;
  W.hw_res = H.hw_res, W.hw_wall = H.hw_wall, W.hw_nat = H.hw_nat;
  W.lj_res = H.lj_res, W.lj_wall = H.lj_wall, W.lj_sat = H.lj_sat, W.lj_gone = H.lj_gone;
  for (int k = 0; k < 3; ++k) W.hw_off[k] = H.hw_off[k], W.lj_off[k] = H.lj_off[k];
  W.cand_res = H.cand_res, W.cand_wall = H.cand_wall, W.cand_d = H.cand_d;
  W.is_conn = H.is_conn;
  W.lj_cap = H.lj_cap, W.cand_cap = H.cand_cap;
  HIPCHK(hipMemcpyAsync(B.dW, &W, sizeof(WallDev), hipMemcpyHostToDevice, B.stream));
  HIPCHK(hipStreamSynchronize(B.stream));
  return PCG_OK;
}

} // namespace

extern "C" {

const char *pcg_last_error(void) { return g_err.c_str(); }

int pcg_device_count(int *count) {
  if (!count) return fail(PCG_E_ARG, "null count");
  HIPCHK(hipGetDeviceCount(count));
  return PCG_OK;
}

int pcg_create(const pcg_params *p, const pcg_model *m, const pcg_options *o, pcg_ctx **out) {
  if (!p || !m || !out) return fail(PCG_E_ARG, "null argument");
  pcg_options opt{};
  if (o) opt = *o;
  double T = opt.temperature > 0 ? opt.temperature : -1.0;
  int32_t seed = opt.seed;
  auto c = std::make_unique<pcg_ctx>();
  int rc = make_batch(*p, 1, m, &T, &seed, opt.precision, opt.philox_seed, opt.device, c->B,
                      nullptr, nullptr, opt.ran2_skip, nullptr, 0, true);
  if (rc) return rc;
  c->B->noise = opt.noise;
  {
    DeviceBatch &B = *c->B;
    c->pcfg = persistent_cfg(B);
    // the persistent kernel stops a system at its step target: never here
    B.sys[0].steps_target = INT64_MAX;
    HIPCHK(hipMemcpyAsync((void *)B.d.sys, B.sys.data(), sizeof(SysDesc), hipMemcpyHostToDevice,
                          B.stream));
    HIPCHK(hipStreamSynchronize(B.stream));
  }
  if (opt.noise == PCG_NOISE_RAN2) {
    // the noise stream continues the one the initial velocities came from
    Ran2 g(seed);
    for (int64_t k = 0; k < opt.ran2_skip; ++k) g.draw();
    if (p->lang_on)
      for (int k = 0; k < 3 * m->n; ++k) g.draw();
    c->B->ran2.push_back(g);
  }
  *out = c.release();
  return PCG_OK;
}

int pcg_destroy(pcg_ctx *c) {
  delete c;
  return PCG_OK;
}

int pcg_step(pcg_ctx *c, int32_t nsteps) {
  if (!c || nsteps < 0) return fail(PCG_E_ARG, "bad arguments");
  DeviceBatch &B = *c->B;
  HIPCHK(hipSetDevice(B.device));
  int rc;
  if (use_persistent(*c)) {
    if ((rc = switch_form(*c, FORM_PERSISTENT))) return rc;
    for (int done = 0; done < nsteps; done += c->spl)
      launch_persistent(B.d, B.prec, c->pcfg.block, std::min(c->spl, nsteps - done),
                        c->pcfg.lds, B.stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(B.stream));
    return check_overflow(B);
  }
  if ((rc = switch_form(*c, FORM_STAGES))) return rc;
  for (int k = 0; k < nsteps; ++k) {
    if (B.noise == PCG_NOISE_RAN2 && (rc = upload_ran2_noise(B))) return rc;
    if (c->graph_on) {
      if (!c->gexec) {
        HIPCHK(hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal));
        enqueue_step(B);
        HIPCHK(hipStreamEndCapture(B.stream, &c->graph));
        HIPCHK(hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
      }
      HIPCHK(hipGraphLaunch(c->gexec, B.stream));
    } else {
      enqueue_step(B);
    }
  }
  HIPCHK(hipStreamSynchronize(B.stream));
  return check_overflow(B);
}

int pcg_use_graph(pcg_ctx *c, int32_t on) {
  if (!c) return fail(PCG_E_ARG, "null ctx");
  c->graph_on = on != 0;
  return PCG_OK;
}

int pcg_step_form(pcg_ctx *c, int32_t form, int32_t *active) {
  if (!c || form < PCG_FORM_AUTO || form > PCG_FORM_PERSISTENT)
    return fail(PCG_E_ARG, "bad arguments");
  if (form == PCG_FORM_PERSISTENT && !persistent_eligible(*c))
    return fail(PCG_E_STATE, "pcg_step_form: the persistent form needs a system that fits its "
                             "LDS, Philox noise and no walls or box");
  c->form = form;
  if (active) *active = use_persistent(*c) ? PCG_FORM_PERSISTENT : PCG_FORM_STAGES;
  return PCG_OK;
}

int pcg_nl_verify(pcg_ctx *c, int32_t *invalid) {
  if (!c || !invalid) return fail(PCG_E_ARG, "null argument");
  DeviceBatch &B = *c->B;
  int rc = switch_form(*c, FORM_STAGES);
  if (rc) return rc;
  launch_begin(B.d, B.stream);
  SysState st;
  rc = read_state(B, 0, st);
  if (rc) return rc;
  *invalid = st.nl_invalid;
  return PCG_OK;
}

int pcg_fix_nl(pcg_ctx *c) {
  if (!c) return fail(PCG_E_ARG, "null ctx");
  DeviceBatch &B = *c->B;
  SysState st;
  int rc = switch_form(*c, FORM_STAGES);
  if (rc || (rc = read_state(B, 0, st))) return rc;
  if (rc) return rc;
  st.nl_invalid = 1;
  st.verify_first = 0;
  if ((rc = write_state(B, 0, st))) return rc;
  launch_fix_nl(B.d, B.max_n, B.max_nlcap, B.cell_words, B.stream);
  HIPCHK(hipStreamSynchronize(B.stream));
  return check_overflow(B);
}

int pcg_prepare_nh(pcg_ctx *c) {
  if (!c) return fail(PCG_E_ARG, "null ctx");
  int rc = switch_form(*c, FORM_STAGES);
  if (rc) return rc;
  launch_prepare_nh(c->B->d, c->B->stream);
  HIPCHK(hipStreamSynchronize(c->B->stream));
  return PCG_OK;
}

int pcg_eval_forces(pcg_ctx *c) {
  if (!c) return fail(PCG_E_ARG, "null ctx");
  DeviceBatch &B = *c->B;
  SysState st;
  int rc = switch_form(*c, FORM_STAGES);
  if (rc || (rc = read_state(B, 0, st))) return rc;
  if (rc) return rc;
  for (int k = 0; k < NT; ++k) st.V[k] = 0.0;
  if ((rc = write_state(B, 0, st))) return rc;
  if ((rc = zero_wall_F(B))) return rc;
  launch_zero_forces(B.d, B.prec, B.stream);
  launch_terms(B.d, B.prec, B.max_items, -1, 0, -1, B.stream);
  launch_tick(B.d, 0, B.stream);
  HIPCHK(hipStreamSynchronize(B.stream));
  return check_overflow(B);
}

static int host_term_size(const SysDesc &s, const SysState &st, int term) {
  int v = 0;
  switch (term) {
  case PCG_CHIR: v = s.nchir; break;
  case PCG_TETHER: v = s.nteth; break;
  case PCG_LREP: v = s.nlrep; break;
  case PCG_NAT_CONT: v = st.nc_n; break;
  case PCG_PAULI: v = st.pauli_n; break;
  case PCG_CONST_DH: case PCG_REL_DH: v = st.dh_n; break;
  case PCG_QA_CAND: v = st.qh_n; break;
  case PCG_QA_CONT: v = std::min(st.qc_n, st.qf_n); break;
  case PCG_PID: v = st.pid_n; break;
  case PCG_NAT_ANG: v = s.nnang; break;
  case PCG_HEUR_ANG: v = s.nhang; break;
  case PCG_HEUR_DIH: v = s.nhdih; break;
  case PCG_CND: case PCG_SND: v = s.nndih; break;
  case PCG_VEL_AFM: v = s.ntip; break;
  case PCG_FORCE_AFM: v = s.nftip; break;
  case PCG_SOLID_WALL: v = s.n; break;
  case PCG_HW_FREE: case PCG_LJW_FREE: v = s.n; break;
  case PCG_HW_CONN: v = st.n_hw; break;
  case PCG_LJW_CONN: v = st.n_lj; break;
  }
  return v;
}

int pcg_term_size(pcg_ctx *c, int32_t term, int32_t *size) {
  if (!c || !size || term < 0 || term >= NT) return fail(PCG_E_ARG, "bad arguments");
  DeviceBatch &B = *c->B;
  SysState st;
  int rc = read_state(B, 0, st);
  if (rc) return rc;
  *size = host_term_size(B.sys[0], st, term);
  return PCG_OK;
}

static int eval_term_only(pcg_ctx *c, int32_t term, int32_t from, int32_t to, double *V) {
  if (!c || term < 0 || term >= NT) return fail(PCG_E_ARG, "bad arguments");
  DeviceBatch &B = *c->B;
  SysState st;
  int rc = switch_form(*c, FORM_STAGES);
  if (rc || (rc = read_state(B, 0, st))) return rc;
  if (rc) return rc;
  for (int k = 0; k < NT; ++k) st.V[k] = 0.0;
  if ((rc = write_state(B, 0, st))) return rc;
  if ((rc = zero_wall_F(B))) return rc;
  launch_zero_forces(B.d, B.prec, B.stream);
  launch_terms(B.d, B.prec, B.max_items, term, from, to, B.stream);
  HIPCHK(hipStreamSynchronize(B.stream));
  if ((rc = read_state(B, 0, st))) return rc;
  if (V) *V = st.V[term];
  return PCG_OK;
}

int pcg_eval_term(pcg_ctx *c, int32_t term, int32_t from, int32_t to, double *F, double *V) {
  int rc = eval_term_only(c, term, from, to, V);
  if (rc) return rc;
  if (F && (rc = get_forces(*c->B, 0, F))) return rc;
  return check_overflow(*c->B);
}

int pcg_eval_term3(pcg_ctx *c, int32_t term, int32_t from, int32_t to, double *Fx, double *Fy,
                   double *Fz, int64_t stride, double *V) {
  if ((Fx || Fy || Fz) && (!Fx || !Fy || !Fz || stride < 1))
    return fail(PCG_E_ARG, "pcg_eval_term3: all three components and a positive stride");
  int rc = eval_term_only(c, term, from, to, V);
  if (rc) return rc;
  DeviceBatch &B = *c->B;
  if (Fx) {
    const void *cols[3] = {B.d.F[0], B.d.F[1], B.d.F[2]};
    double *const out[3] = {Fx, Fy, Fz};
    if ((rc = get_cols3(B, 0, cols, B.prec != PCG_FP64, out, stride))) return rc;
  }
  return check_overflow(B);
}

int pcg_qa_finish(pcg_ctx *c) {
  if (!c) return fail(PCG_E_ARG, "null ctx");
  int rc = switch_form(*c, FORM_STAGES);
  if (rc) return rc;
  launch_qa_finish(c->B->d, c->B->max_n, c->B->stream);
  HIPCHK(hipStreamSynchronize(c->B->stream));
  return check_overflow(*c->B);
}

int pcg_lang_step(pcg_ctx *c) {
  if (!c) return fail(PCG_E_ARG, "null ctx");
  DeviceBatch &B = *c->B;
  int rc = switch_form(*c, FORM_STAGES);
  if (rc) return rc;
  if (B.noise == PCG_NOISE_RAN2 && (rc = upload_ran2_noise(B))) return rc;
  launch_lang(B.d, B.prec, B.noise, B.stream);
  launch_tick(B.d, 1, B.stream);
  HIPCHK(hipStreamSynchronize(B.stream));
  return PCG_OK;
}

int pcg_get(pcg_ctx *c, int32_t which, double *out) {
  if (!c || !out || which < 0 || which > 7) return fail(PCG_E_ARG, "bad arguments");
  DeviceBatch &B = *c->B;
  if (which <= 5) return get_cols(B, 0, B.d.y[which], out);
  if (which == 6) return get_forces(B, 0, out);
  return get_cols(B, 0, B.d.vel, out);
}

int pcg_get3(pcg_ctx *c, int32_t which, double *x, double *y, double *z, int64_t stride) {
  if (!c || !x || !y || !z || which < 0 || which > 7 || stride < 1)
    return fail(PCG_E_ARG, "bad arguments");
  DeviceBatch &B = *c->B;
  double *const out[3] = {x, y, z};
  if (which == 6) {
    const void *cols[3] = {B.d.F[0], B.d.F[1], B.d.F[2]};
    return get_cols3(B, 0, cols, B.prec != PCG_FP64, out, stride);
  }
  double *const *src = which <= 5 ? B.d.y[which] : B.d.vel;
  const void *cols[3] = {src[0], src[1], src[2]};
  return get_cols3(B, 0, cols, false, out, stride);
}

int pcg_set3(pcg_ctx *c, int32_t which, const double *x, const double *y, const double *z,
             int64_t stride) {
  if (!c || !x || !y || !z || which < 0 || which > 5 || stride < 1)
    return fail(PCG_E_ARG, "bad arguments");
  const double *const in[3] = {x, y, z};
  return set_cols3(*c->B, 0, c->B->d.y[which], in, stride);
}

int pcg_set(pcg_ctx *c, int32_t which, const double *in) {
  if (!c || !in || which < 0 || which > 5) return fail(PCG_E_ARG, "bad arguments");
  DeviceBatch &B = *c->B;
  const SysDesc &d = B.sys[0];
  std::vector<double> tmp(3 * (size_t)d.n);
  for (int i = 0; i < d.n; ++i)
    for (int c2 = 0; c2 < 3; ++c2) tmp[c2 * d.n + i] = in[3 * i + c2];
  for (int c2 = 0; c2 < 3; ++c2)
    HIPCHK(hipMemcpyAsync(B.d.y[which][c2] + d.res0, tmp.data() + c2 * d.n, d.n * 8,
                          hipMemcpyHostToDevice, B.stream));
  HIPCHK(hipStreamSynchronize(B.stream));
  return PCG_OK;
}

int pcg_report(pcg_ctx *c, double *out) {
  if (!c || !out) return fail(PCG_E_ARG, "null argument");
  HIPCHK(hipSetDevice(c->B->device));
  return report_into(*c->B, out);
}

int pcg_get_energy(pcg_ctx *c, double *V, double *term_V) {
  if (!c) return fail(PCG_E_ARG, "null ctx");
  SysState st;
  int rc = read_state(*c->B, 0, st);
  if (rc) return rc;
  if (V) *V = st.Vtot;
  if (term_V)
    for (int k = 0; k < NT; ++k) term_V[k] = st.V[k];
  return PCG_OK;
}

int pcg_get_time(pcg_ctx *c, double *t, int64_t *step) {
  if (!c) return fail(PCG_E_ARG, "null ctx");
  SysState st;
  int rc = read_state(*c->B, 0, st);
  if (rc) return rc;
  if (t) *t = st.t;
  if (step) *step = st.step;
  return PCG_OK;
}

int pcg_set_time(pcg_ctx *c, double t) {
  if (!c) return fail(PCG_E_ARG, "null ctx");
  SysState st;
  int rc = read_state(*c->B, 0, st);
  if (rc) return rc;
  st.t = st.true_t = t;
  return write_state(*c->B, 0, st);
}

int pcg_set_tips(pcg_ctx *c, int32_t n_tips, const int32_t *tip_res, const double *tip_orig,
                 const double *tip_vel, int32_t n_ftips, const int32_t *ftip_res,
                 const double *ftip_force) {
  if (!c) return fail(PCG_E_ARG, "null ctx");
  int rc = refresh_sys(*c->B);
  if (rc) return rc;
  return set_tips(*c->B, 0, n_tips, tip_res, tip_orig, tip_vel, n_ftips, ftip_res, ftip_force);
}

int pcg_afm_stats(pcg_ctx *c, int32_t cap, int32_t *size, double *avg, double *perp,
                  uint8_t *has) {
  if (!c || !size || cap < 0) return fail(PCG_E_ARG, "bad arguments");
  DeviceBatch &B = *c->B;
  const SysDesc &d = B.sys[0];
  const int nt = d.ntip + d.nftip, m = std::min(nt, (int)cap);
  *size = nt;
  if (m <= 0) return PCG_OK;
  HIPCHK(hipSetDevice(B.device));
  std::vector<TipStat> ts(m);
  HIPCHK(hipStreamSynchronize(B.stream));
  HIPCHK(hipMemcpy(ts.data(), B.d.tip_stat + d.tip0, m * sizeof(TipStat), hipMemcpyDeviceToHost));
  for (int k = 0; k < m; ++k) {
    for (int a = 0; a < 3; ++a)
      if (avg) avg[3 * k + a] = ts[k].avg[a];
    if (perp) perp[k] = k < d.ntip ? ts[k].avg[3] : 0.0;
    if (has) has[k] = ts[k].has ? 1 : 0;
  }
  return PCG_OK;
}

int pcg_set_noise_scale(pcg_ctx *c, double scale) {
  if (!c || scale < 0) return fail(PCG_E_ARG, "bad arguments");
  DeviceBatch &B = *c->B;
  int rc = refresh_sys(B);
  if (rc) return rc;
  B.sys[0].const2 = const2_of(B.hp.p, B.sys[0].temperature) * scale;
  HIPCHK(hipMemcpyAsync((void *)B.d.sys, B.sys.data(), sizeof(SysDesc), hipMemcpyHostToDevice,
                        B.stream));
  HIPCHK(hipStreamSynchronize(B.stream));
  return PCG_OK;
}

int pcg_get_list(pcg_ctx *c, int32_t kind, int32_t cap, int32_t *size, int32_t *i1, int32_t *i2,
                 double *val, uint8_t *flag) {
  if (!c || !size || kind < 0 || kind > 7) return fail(PCG_E_ARG, "bad arguments");
  // the persistent form keeps the per-term lists but not the NL itself
  if (kind == 0 && c->last_form == FORM_PERSISTENT)
    return fail(PCG_E_STATE, "the persistent step form keeps no neighbour list; "
                             "pcg_fix_nl rebuilds it in the stage form");
  DeviceBatch &B = *c->B;
  SysState st;
  int rc = read_state(B, 0, st);
  if (rc) return rc;
  const SysDesc &d = B.sys[0];
  const DevBatch &D = B.d;
  int n = 0, off = 0;
  const int *a = nullptr, *b2 = nullptr;
  const double *v = nullptr;
  const uint8_t *f = nullptr;
  switch (kind) {
  case 0: n = st.nl_n, off = d.nl0, a = D.nl_i[0], b2 = D.nl_i[1], v = D.nl_d; break;
  case 1: n = st.nc_n, off = d.nc0; break;
  case 2: n = st.pauli_n, off = d.nl0, a = D.pauli_i[0], b2 = D.pauli_i[1]; break;
  case 3: n = st.dh_n, off = d.nl0, a = D.dh_i[0], b2 = D.dh_i[1], v = D.dh_q; break;
  case 4: n = st.pid_n, off = d.nl0, a = D.pid_i[0], b2 = D.pid_i[1], v = D.pid_d; break;
  case 5: n = st.qf_n, off = d.qf0, a = D.qf_i[0], b2 = D.qf_i[1], v = D.qf_d, f = D.qf_vac; break;
  case 6: n = st.qc_n, off = d.qc0, a = D.qc_i[0], b2 = D.qc_i[1], v = D.qc_sat, f = D.qc_vac; break;
  case 7: n = st.qc_n, off = d.qc0, a = D.qc_i[0], b2 = D.qc_i[1], f = D.qc_vac; break;
  }
  *size = n;
  int m = std::min(n, cap);
  if (m <= 0) return PCG_OK;
  std::vector<int> ta(m), tb(m);
  std::vector<double> tv(m, 0.0);
  std::vector<uint8_t> tf(m, 0);
  if (kind == 1) {
    std::vector<int> idx(m);
    HIPCHK(hipMemcpy(idx.data(), D.ncur + off, m * 4, hipMemcpyDeviceToHost));
    std::vector<int> A(d.nnc), Bv(d.nnc);
    std::vector<double> nat(d.nnc);
    std::vector<uint8_t> act(d.nnc);
    HIPCHK(hipMemcpy(A.data(), D.nc_i[0] + d.nc0, d.nnc * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(Bv.data(), D.nc_i[1] + d.nc0, d.nnc * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(nat.data(), D.nc_nat + d.nc0, d.nnc * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(act.data(), D.nc_active + d.nc0, d.nnc, hipMemcpyDeviceToHost));
    for (int k = 0; k < m; ++k) {
      int q = idx[k] - d.nc0;
      ta[k] = A[q] - d.res0;
      tb[k] = Bv[q] - d.res0;
      tv[k] = nat[q];
      tf[k] = act[q];
    }
  } else {
    HIPCHK(hipMemcpy(ta.data(), a + off, m * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(tb.data(), b2 + off, m * 4, hipMemcpyDeviceToHost));
    if (v) HIPCHK(hipMemcpy(tv.data(), v + off, m * 8, hipMemcpyDeviceToHost));
    if (f) HIPCHK(hipMemcpy(tf.data(), f + off, m, hipMemcpyDeviceToHost));
    if (kind == 7) {
      std::vector<int16_t> ty(m);
      HIPCHK(hipMemcpy(ty.data(), D.qc_t + off, m * sizeof(int16_t), hipMemcpyDeviceToHost));
      for (int k = 0; k < m; ++k) tv[k] = (double)ty[k];
    }
    for (int k = 0; k < m; ++k) {
      ta[k] -= d.res0;
      tb[k] -= d.res0;
    }
  }
  if (i1) std::memcpy(i1, ta.data(), m * 4);
  if (i2) std::memcpy(i2, tb.data(), m * 4);
  if (val) std::memcpy(val, tv.data(), m * 8);
  if (flag) std::memcpy(flag, tf.data(), m);
  return PCG_OK;
}

int pcg_nat_cont_state(pcg_ctx *c, uint8_t *active, double *change_t, int32_t *num_changed) {
  if (!c) return fail(PCG_E_ARG, "null ctx");
  DeviceBatch &B = *c->B;
  const SysDesc &d = B.sys[0];
  if (active && d.nnc)
    HIPCHK(hipMemcpy(active, B.d.nc_active + d.nc0, d.nnc, hipMemcpyDeviceToHost));
  if (change_t && d.nnc)
    HIPCHK(hipMemcpy(change_t, B.d.nc_change_t + d.nc0, d.nnc * 8, hipMemcpyDeviceToHost));
  if (num_changed) {
    SysState st;
    int rc = read_state(B, 0, st);
    if (rc) return rc;
    *num_changed = st.num_changed;
  }
  return PCG_OK;
}

int pcg_qa_state(pcg_ctx *c, int8_t *sync, int32_t *num_contacts) {
  if (!c) return fail(PCG_E_ARG, "null ctx");
  DeviceBatch &B = *c->B;
  const SysDesc &d = B.sys[0];
  if (sync) HIPCHK(hipMemcpy(sync, B.d.sync + 4 * d.res0, 4 * d.n, hipMemcpyDeviceToHost));
  if (num_contacts) {
    SysState st;
    int rc = read_state(B, 0, st);
    if (rc) return rc;
    *num_contacts = st.num_qa;
  }
  return PCG_OK;
}

int pcg_nl_rebuilds(pcg_ctx *c, int32_t *count) {
  if (!c || !count) return fail(PCG_E_ARG, "null argument");
  SysState st;
  int rc = read_state(*c->B, 0, st);
  if (rc) return rc;
  *count = st.rebuilds;
  return PCG_OK;
}

int pcg_stream(pcg_ctx *c, void **stream) {
  if (!c || !stream) return fail(PCG_E_ARG, "null argument");
  *stream = (void *)c->B->stream;
  return PCG_OK;
}

} // extern "C"

// ---------------------------------------------------------------- batch ----
struct pcg_batch {
  std::unique_ptr<DeviceBatch> B;
  int block = 0, spl = 0;  // threads per workgroup, steps per launch
  bool persistent = true;
  bool use_graph = false;
  size_t lds = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  int last_launches = 0;
  bool tuned = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int g_block = 0, g_spl = 0;
  ~pcg_batch() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
  }
};

namespace {

// steps owed by every system = current + nsteps
// The batch's systems never stop on their own (steps_target is INT64_MAX), so
// a run is a fixed sequence of launches: captured once per (nsteps, block,
// steps per launch) as a hipGraph when the batch asks for it.
int run_persistent(pcg_batch *b, int nsteps) {
  DeviceBatch &B = *b->B;
  int launches = (nsteps + b->spl - 1) / b->spl;
  const int full = nsteps / b->spl, rem = nsteps % b->spl;
  if (b->use_graph && full > 0) {
    // the graph holds one launch of steps_per_launch steps, replayed; a
    // shorter last launch goes directly
    if (!b->gexec || b->g_block != b->block || b->g_spl != b->spl) {
      if (b->gexec) HIPCHK(hipGraphExecDestroy(b->gexec));
      if (b->graph) HIPCHK(hipGraphDestroy(b->graph));
      b->gexec = nullptr;
      b->graph = nullptr;
      HIPCHK(hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal));
      launch_persistent(B.d, B.prec, b->block, b->spl, b->lds, B.stream);
      HIPCHK(hipStreamEndCapture(B.stream, &b->graph));
      HIPCHK(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
      b->g_block = b->block;
      b->g_spl = b->spl;
    }
    HIPCHK(hipEventRecord(b->ev0, B.stream));
    for (int k = 0; k < full; ++k) HIPCHK(hipGraphLaunch(b->gexec, B.stream));
    if (rem) launch_persistent(B.d, B.prec, b->block, rem, b->lds, B.stream);
  } else {
    HIPCHK(hipEventRecord(b->ev0, B.stream));
    for (int done = 0; done < nsteps; done += b->spl)
      launch_persistent(B.d, B.prec, b->block, std::min(b->spl, nsteps - done), b->lds, B.stream);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(b->ev1, B.stream));
  HIPCHK(hipEventSynchronize(b->ev1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->last_ms = ms;
  b->last_launches = launches;
  return PCG_OK;
}

int run_stages(pcg_batch *b, int nsteps) {
  DeviceBatch &B = *b->B;
  HIPCHK(hipEventRecord(b->ev0, B.stream));
  for (int k = 0; k < nsteps; ++k) enqueue_step(B);
  HIPCHK(hipEventRecord(b->ev1, B.stream));
  HIPCHK(hipEventSynchronize(b->ev1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->last_ms = ms;
  b->last_launches = nsteps * 8;
  return PCG_OK;
}

} // namespace

extern "C" {

int pcg_batch_create(const pcg_params *p, int32_t nsys, const pcg_model *models,
                     const double *temps, const int32_t *seeds, const pcg_batch_options *o,
                     pcg_batch **out) {
  if (!p || !models || !out || nsys <= 0) return fail(PCG_E_ARG, "bad arguments");
  pcg_batch_options opt{};
  if (o) opt = *o;
  auto b = std::make_unique<pcg_batch>();
  int rc = make_batch(*p, nsys, models, temps, seeds, opt.precision, opt.philox_seed, opt.device,
                      b->B);
  if (rc) return rc;
  DeviceBatch &B = *b->B;
  // per-residue bonded data joins positions and forces in LDS when it fits;
  // systems whose positions and forces do not fit take the stage kernels
  const PersistentCfg pc = persistent_cfg(B);
  b->lds = pc.lds;
  b->persistent = pc.ok;
  if (opt.block != 0 && (opt.block < 64 || opt.block > 1024 || opt.block % 64 != 0))
    return fail(PCG_E_ARG, "block must be a multiple of 64 in [64, 1024]");
  if (opt.steps_per_launch < 0) return fail(PCG_E_ARG, "steps_per_launch must be >= 0");
  b->block = opt.block > 0 ? opt.block : pc.block;
  b->spl = opt.steps_per_launch > 0 ? opt.steps_per_launch : 100;
  b->use_graph = opt.use_graph != 0;
  // batch systems run until told otherwise: each launch makes its steps
  for (auto &d : B.sys) d.steps_target = INT64_MAX;
  HIPCHK(hipMemcpyAsync((void *)B.d.sys, B.sys.data(), B.sys.size() * sizeof(SysDesc),
                        hipMemcpyHostToDevice, B.stream));
  HIPCHK(hipStreamSynchronize(B.stream));
  HIPCHK(hipEventCreate(&b->ev0));
  HIPCHK(hipEventCreate(&b->ev1));
  *out = b.release();
  return PCG_OK;
}

int pcg_batch_destroy(pcg_batch *b) {
  delete b;
  return PCG_OK;
}

int pcg_batch_run(pcg_batch *b, int32_t nsteps) {
  if (!b || nsteps < 0) return fail(PCG_E_ARG, "bad arguments");
  HIPCHK(hipSetDevice(b->B->device));
  int rc = b->persistent ? run_persistent(b, nsteps) : run_stages(b, nsteps);
  if (rc) return rc;
  return check_overflow(*b->B);
}

int pcg_batch_sync(pcg_batch *b) {
  if (!b) return fail(PCG_E_ARG, "null batch");
  HIPCHK(hipStreamSynchronize(b->B->stream));
  return PCG_OK;
}

// Checkpoint of a batch: a header, then the device arena byte for byte. The
// arena holds every piece of device state (Gear y0..y5, velocities, time and
// step counters, which are also the Philox position, lists, native-contact
// and QA state, sync budgets, system descriptors, parameters) and no absolute
// addresses, so it restores into any batch built from the same models,
// parameters and precision.
namespace {
struct CkptHeader {
  char magic[8];
  uint32_t version;
  int32_t nsys;
  int64_t ntot;
  int32_t prec;
  int32_t nran2;  // host ran2 streams after the arena
  uint64_t arena_bytes;
  uint64_t fingerprint;
};
extern "C++" {
// Digest of byte ranges and of the parameter fields one by one (the struct
// has padding, whose bytes a caller need not have cleared).
struct Digest {
  uint64_t h = 0x5043474350434743ull;
  void bytes(const void *p, size_t n) {
    const unsigned char *c = (const unsigned char *)p;
    for (size_t k = 0; k < n; ++k) h = (h ^ c[k]) * 0x100000001B3ull;
    h = mix64(h ^ n);
  }
  template <typename T> void vec(const std::vector<T> &v) { bytes(v.data(), v.size() * sizeof(T)); }
  template <typename T> void val(const T &x) { bytes(&x, sizeof x); }
};
void digest_params(Digest &d, const pcg_params &p) {
#define PCG_D(f) d.val(p.f)
  PCG_D(rep_cutoff); PCG_D(counting_factor); PCG_D(nl_cutoff); PCG_D(nl_pad);
  PCG_D(tether_on); PCG_D(tether_H1); PCG_D(tether_H2); PCG_D(lrep_on); PCG_D(lrep_depth);
  PCG_D(chir_on); PCG_D(e_chi); PCG_D(nat_ang_on); PCG_D(CBA); PCG_D(heur_ang_on);
  PCG_D(heur_ang_poly); PCG_D(nat_dih_on); PCG_D(nat_dih_variant); PCG_D(CDA); PCG_D(CDB);
  PCG_D(CDH); PCG_D(heur_dih_on); PCG_D(heur_dih_coeffs); PCG_D(pauli_on); PCG_D(pauli_depth);
  PCG_D(nat_cont_on); PCG_D(nat_cont_depth); PCG_D(nc_ss_kind); PCG_D(nc_ss_H1); PCG_D(nc_ss_H2);
  PCG_D(nc_ss_nat_r); PCG_D(nc_ss_depth); PCG_D(nc_ss_r_min); PCG_D(dh_on); PCG_D(dh_variant);
  PCG_D(dh_screen_dist); PCG_D(dh_cutoff); PCG_D(dh_perm); PCG_D(qa_on); PCG_D(qa_include4);
  PCG_D(qa_phase_dur); PCG_D(qa_formation_tol); PCG_D(qa_min_cos_hr); PCG_D(qa_min_cos_hh);
  PCG_D(qa_max_cos_nr); PCG_D(qa_lj_depth); PCG_D(qa_lj_r_low); PCG_D(qa_lj_r_high);
  PCG_D(pid_on); PCG_D(pid_include4); PCG_D(pid_lambda_variant); PCG_D(pid_bbp_alpha);
  PCG_D(pid_bbp_psi0); PCG_D(pid_bbm_alpha); PCG_D(pid_bbm_psi0); PCG_D(pid_ss_alpha);
  PCG_D(pid_ss_psi0); PCG_D(pid_bbp_lj); PCG_D(pid_bbm_lj); PCG_D(pid_ss_depth);
  PCG_D(pid_ss_r_low); PCG_D(pid_ss_r_high); PCG_D(afm_on); PCG_D(afm_H1); PCG_D(afm_H2);
  PCG_D(lang_on); PCG_D(gamma); PCG_D(temperature); PCG_D(dt); PCG_D(mass); PCG_D(charge);
  PCG_D(ptype); PCG_D(base_sync); PCG_D(n_walls); PCG_D(wall_origin); PCG_D(wall_normal);
  PCG_D(wall_depth); PCG_D(wall_min_dist); PCG_D(wall_kind); PCG_D(wall_side); PCG_D(wall_limit);
  PCG_D(hw_on); PCG_D(ljw_on); PCG_D(wall_hh1); PCG_D(wall_lj_depth); PCG_D(ljw_breaking_dist);
  PCG_D(ljw_saturation_diff); PCG_D(afm_stats_window);
#undef PCG_D
}
}  // extern "C++"
// What a checkpoint must be restored into: the same system layout, the same
// parameters, the same models (types, chains, every static list and its
// native values, tips) and the same precision.
uint64_t ckpt_fingerprint(const DeviceBatch &B) {
  Digest d;
  d.val(B.sys.size());
  d.val(B.arena.total);
  d.val(B.prec);
  digest_params(d, B.hp.p);
  for (size_t s = 0; s < B.sys.size(); ++s) {
    const HostSys &h = B.hs[s];
    d.val(B.sys[s].n);
    d.val(B.sys[s].res0);
    d.vec(h.atype); d.vec(h.chain); d.vec(h.seq);
    for (auto &v : h.teth) d.vec(v);
    d.vec(h.teth_nat);
    for (auto &v : h.lrep) d.vec(v);
    for (auto &v : h.chir) d.vec(v);
    d.vec(h.chir_f0); d.vec(h.chir_f1);
    for (auto &v : h.nang) d.vec(v);
    d.vec(h.nang_th);
    for (auto &v : h.hang) d.vec(v);
    d.vec(h.hang_t);
    for (auto &v : h.ndih) d.vec(v);
    d.vec(h.ndih_ph);
    for (auto &v : h.hdih) d.vec(v);
    d.vec(h.hdih_t);
    for (auto &v : h.nc) d.vec(v);
    d.vec(h.nc_nat); d.vec(h.nc_type);
    d.vec(h.tip_res);
    d.val(h.nftip);
    d.val(h.temperature);
  }
  return d.h;
}

// A checkpoint is the header, the arena, then the host ran2 streams of the
// RAN2 noise mode (one per system; none in the Philox mode, whose position is
// the step counter in the arena).
size_t ckpt_bytes(const DeviceBatch &B) {
  return sizeof(CkptHeader) + B.arena.total + B.ran2.size() * sizeof(Ran2);
}

int ckpt_write(DeviceBatch &B, const char magic[8], void *buf, int64_t bytes) {
  if (!buf) return fail(PCG_E_ARG, "null argument");
  if (bytes < (int64_t)ckpt_bytes(B)) return fail(PCG_E_ARG, "checkpoint buffer too small");
  HIPCHK(hipSetDevice(B.device));
  CkptHeader h{};
  std::memcpy(h.magic, magic, 8);
  h.version = 6;
  h.nsys = (int32_t)B.sys.size();
  h.ntot = B.d.ntot;
  h.prec = B.prec;
  h.nran2 = (int32_t)B.ran2.size();
  h.arena_bytes = B.arena.total;
  h.fingerprint = ckpt_fingerprint(B);
  std::memcpy(buf, &h, sizeof h);
  HIPCHK(hipMemcpyAsync((char *)buf + sizeof h, B.arena.base, B.arena.total, hipMemcpyDeviceToHost,
                        B.stream));
  HIPCHK(hipStreamSynchronize(B.stream));
  if (!B.ran2.empty())
    std::memcpy((char *)buf + sizeof h + B.arena.total, B.ran2.data(), B.ran2.size() * sizeof(Ran2));
  return PCG_OK;
}

int ckpt_read(DeviceBatch &B, const char magic[8], const void *buf, int64_t bytes) {
  if (!buf) return fail(PCG_E_ARG, "null argument");
  CkptHeader h;
  if (bytes < (int64_t)sizeof h) return fail(PCG_E_ARG, "not a checkpoint");
  std::memcpy(&h, buf, sizeof h);
  if (std::memcmp(h.magic, magic, 8) != 0 || h.version != 6) return fail(PCG_E_ARG, "not a checkpoint");
  if (h.nsys != (int32_t)B.sys.size() || h.ntot != B.d.ntot || h.prec != B.prec ||
      h.arena_bytes != B.arena.total || h.fingerprint != ckpt_fingerprint(B))
    return fail(PCG_E_ARG, "checkpoint of a different batch (models, parameters or precision)");
  if (h.nran2 != (int32_t)B.ran2.size())
    return fail(PCG_E_ARG, "checkpoint of another noise source (ran2 and Philox differ)");
  if (bytes < (int64_t)ckpt_bytes(B)) return fail(PCG_E_ARG, "truncated checkpoint");
  HIPCHK(hipSetDevice(B.device));
  HIPCHK(hipMemcpyAsync(B.arena.base, (const char *)buf + sizeof h, B.arena.total,
                        hipMemcpyHostToDevice, B.stream));
  // the host copy of the system descriptors follows the restored ones
  HIPCHK(hipMemcpyAsync(B.sys.data(), (const void *)B.d.sys, B.sys.size() * sizeof(SysDesc),
                        hipMemcpyDeviceToHost, B.stream));
  HIPCHK(hipStreamSynchronize(B.stream));
  if (!B.ran2.empty())
    std::memcpy(B.ran2.data(), (const char *)buf + sizeof h + B.arena.total,
                B.ran2.size() * sizeof(Ran2));
  return PCG_OK;
}

}  // namespace

int pcg_batch_checkpoint_size(pcg_batch *b, int64_t *bytes) {
  if (!b || !bytes) return fail(PCG_E_ARG, "null argument");
  *bytes = (int64_t)ckpt_bytes(*b->B);
  return PCG_OK;
}

int pcg_batch_checkpoint(pcg_batch *b, void *buf, int64_t bytes) {
  if (!b) return fail(PCG_E_ARG, "null argument");
  return ckpt_write(*b->B, "PCGCKPTB", buf, bytes);
}

int pcg_batch_restore(pcg_batch *b, const void *buf, int64_t bytes) {
  if (!b) return fail(PCG_E_ARG, "null argument");
  return ckpt_read(*b->B, "PCGCKPTB", buf, bytes);
}

// The plain context: the same layout for its one system, with the ran2
// stream of its RAN2 noise (state_io.cpp:117-156 saves the state but not
// the generator; here a resumed run continues the same noise).
int pcg_checkpoint_size(pcg_ctx *c, int64_t *bytes) {
  if (!c || !bytes) return fail(PCG_E_ARG, "null argument");
  *bytes = (int64_t)ckpt_bytes(*c->B);
  return PCG_OK;
}

int pcg_checkpoint(pcg_ctx *c, void *buf, int64_t bytes) {
  if (!c) return fail(PCG_E_ARG, "null argument");
  return ckpt_write(*c->B, "PCGCKPTC", buf, bytes);
}

int pcg_restore(pcg_ctx *c, const void *buf, int64_t bytes) {
  if (!c) return fail(PCG_E_ARG, "null argument");
  DeviceBatch &B = *c->B;
  int rc = ckpt_read(B, "PCGCKPTC", buf, bytes);
  c->last_form = FORM_NONE;  // the restored state says which form built its lists
  if (rc || !B.dW) return rc;
  // the wall state came with the arena; its array pointers are this context's
  WallDev W;
  if ((rc = read_wall(B, W)) || (rc = write_wall(B, W))) return rc;
  B.box_set = W.box_on != 0;
  if (c->gexec) {  // the captured step may lack (or hold) the box kernels
    HIPCHK(hipGraphExecDestroy(c->gexec));
    HIPCHK(hipGraphDestroy(c->graph));
    c->gexec = nullptr;
    c->graph = nullptr;
  }
  return PCG_OK;
}

int pcg_box_setup(pcg_ctx *c, const pcg_box *box) {
  if (!c || !box) return fail(PCG_E_ARG, "null argument");
  DeviceBatch &B = *c->B;
  for (int k = 0; k < 3; ++k)
    if (!std::isfinite(box->box_min[k]) || !std::isfinite(box->box_max[k]) ||
        box->box_max[k] < box->box_min[k])
      return fail(PCG_E_ARG, "pcg_box_setup: box_max must not be below box_min");
  if (!(box->avg_forces_over >= 0.0)) return fail(PCG_E_ARG, "pcg_box_setup: negative averaging time");
  HIPCHK(hipSetDevice(B.device));
  WallDev W;
  int rc = read_wall(B, W);
  if (rc) return rc;
  W.bx = *box;
  for (int k = 0; k < 3; ++k) W.box_min[k] = box->box_min[k], W.box_max[k] = box->box_max[k];
  W.box_on = 1;
  W.log_on = W.side_wall[4] >= 0 && W.side_wall[5] >= 0;
  W.phase = PCG_PHASE_NONE;
  W.shear = 0;
  W.since = W.until = W.amplitude = W.omega = W.displacement = 0.0;
  if ((rc = write_wall(B, W))) return rc;
  launch_box_setup(B.d, B.dP, sys_rw(B), B.stream);
  HIPCHK(hipStreamSynchronize(B.stream));
  B.box_set = true;
  if ((rc = refresh_sys(B))) return rc;
  if (c->gexec) {  // the next step captures the box kernels
    HIPCHK(hipGraphExecDestroy(c->gexec));
    HIPCHK(hipGraphDestroy(c->graph));
    c->gexec = nullptr;
    c->graph = nullptr;
  }
  return PCG_OK;
}

int pcg_box_start(pcg_ctx *c) {
  if (!c) return fail(PCG_E_ARG, "null ctx");
  DeviceBatch &B = *c->B;
  if (!B.box_set) return fail(PCG_E_STATE, "pcg_box_start: no box (pcg_box_setup)");
  HIPCHK(hipSetDevice(B.device));
  WallDev W;
  SysState st;
  int rc = read_wall(B, W);
  if (rc || (rc = read_state(B, 0, st))) return rc;
  W.until = W.bx.total_time;
  W.since = st.t;
  W.phase = PCG_PHASE_SQUEEZING;
  return write_wall(B, W);
}

int pcg_box_state(pcg_ctx *c, double *out) {
  if (!c || !out) return fail(PCG_E_ARG, "null argument");
  DeviceBatch &B = *c->B;
  HIPCHK(hipSetDevice(B.device));
  WallDev W;
  DevParams P;
  SysDesc sd;
  int rc = read_wall(B, W);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(&P, B.dP, sizeof(DevParams), hipMemcpyDeviceToHost, B.stream));
  HIPCHK(hipMemcpyAsync(&sd, (const void *)B.d.sys, sizeof(SysDesc), hipMemcpyDeviceToHost,
                        B.stream));
  HIPCHK(hipStreamSynchronize(B.stream));
  for (int k = 0; k < PCG_BOX_STATE_NUM; ++k) out[k] = 0.0;
  out[PCG_BOX_PHASE] = W.phase;
  out[PCG_BOX_SINCE] = W.since;
  out[PCG_BOX_UNTIL] = W.until;
  for (int k = 0; k < 3; ++k) {
    out[PCG_BOX_MIN + k] = W.box_min[k];
    out[PCG_BOX_MAX + k] = W.box_max[k];
    out[PCG_BOX_CELL + k] = sd.cell[k];
  }
  out[PCG_BOX_AMPLITUDE] = W.amplitude;
  out[PCG_BOX_OMEGA] = W.omega;
  out[PCG_BOX_DISPLACEMENT] = W.displacement;
  out[PCG_BOX_SHEAR] = W.shear;
  out[PCG_BOX_AVG_Z_HAS] = W.avg_z.has;
  out[PCG_BOX_AVG_Z] = W.avg_z.has ? W.avg_z.avg[0] : 0.0;
  for (int w = 0; w < 6; ++w) {
    for (int k = 0; k < 3; ++k) {
      out[PCG_BOX_ORIGIN + 3 * w + k] = P.p.wall_origin[w][k];
      out[PCG_BOX_NORMAL + 3 * w + k] = P.p.wall_normal[w][k];
      out[PCG_BOX_WALL_F + 3 * w + k] = W.wall_F[w][k];
      out[PCG_BOX_SHIFT + 3 * w + k] = W.shift[w][k];
      out[PCG_BOX_AVG_F + 3 * w + k] = W.avg_F[w].has ? W.avg_F[w].avg[k] : 0.0;
    }
    out[PCG_BOX_WORK + w] = W.work[w];
    out[PCG_BOX_AVG_F_HAS + w] = W.avg_F[w].has;
  }
  return PCG_OK;
}

int pcg_ljw_process(pcg_ctx *c) {
  if (!c) return fail(PCG_E_ARG, "null ctx");
  DeviceBatch &B = *c->B;
  if (!B.hp.ljw_terms) return fail(PCG_E_STATE, "pcg_ljw_process: no LJ walls (ljw_on and a plane of kind 2)");
  HIPCHK(hipSetDevice(B.device));
  launch_ljw_process(B.d, B.stream);
  HIPCHK(hipStreamSynchronize(B.stream));
  return check_overflow(B);
}

int pcg_wall_conns(pcg_ctx *c, int32_t kind, int32_t cap, int32_t *size, int32_t *res,
                   int32_t *wall, double *offset, double *sat) {
  if (!c || !size || (kind != 0 && kind != 1) || cap < 0) return fail(PCG_E_ARG, "bad arguments");
  DeviceBatch &B = *c->B;
  HIPCHK(hipSetDevice(B.device));
  SysState st;
  int rc = read_state(B, 0, st);
  if (rc) return rc;
  const int k = kind == 0 ? st.n_hw : st.n_lj;
  *size = k;
  const int m = std::min(k, cap);
  if (m == 0) return PCG_OK;
  const WallDev &H = B.hW;
  std::vector<double> o[3];
  for (int a = 0; a < 3; ++a) o[a].resize(m);
  std::vector<int> r(m), w(m);
  std::vector<double> sv(m);
  auto get = [&](void *dst, const void *src, size_t bytes) -> int {
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, B.stream));
    return PCG_OK;
  };
  if ((rc = get(r.data(), kind == 0 ? (const void *)H.hw_res : H.lj_res, m * 4)) ||
      (rc = get(w.data(), kind == 0 ? (const void *)H.hw_wall : H.lj_wall, m * 4)) ||
      (rc = get(sv.data(), kind == 0 ? (const void *)H.hw_nat : H.lj_sat, m * 8)))
    return rc;
  for (int a = 0; a < 3; ++a)
    if ((rc = get(o[a].data(), kind == 0 ? (const void *)H.hw_off[a] : H.lj_off[a], m * 8)))
      return rc;
  HIPCHK(hipStreamSynchronize(B.stream));
  for (int a = 0; a < m; ++a) {
    if (res) res[a] = r[a];
    if (wall) wall[a] = w[a];
    if (sat) sat[a] = sv[a];
    if (offset)
      for (int q = 0; q < 3; ++q) offset[3 * a + q] = o[q][a];
  }
  return PCG_OK;
}

int pcg_batch_report(pcg_batch *b, double *out) {
  if (!b || !out) return fail(PCG_E_ARG, "null argument");
  HIPCHK(hipSetDevice(b->B->device));
  return report_into(*b->B, out);
}

int pcg_batch_report_static(pcg_batch *b, double *out) {
  if (!b || !out) return fail(PCG_E_ARG, "null argument");
  HIPCHK(hipSetDevice(b->B->device));
  return report_into_static(*b->B, out);
}

int pcg_batch_energies(pcg_batch *b, int32_t on) {
  if (!b) return fail(PCG_E_ARG, "null batch");
  b->B->d.energy_off = on ? 0 : 1;
  if (b->gexec) {  // the captured launch holds the old flag
    HIPCHK(hipGraphExecDestroy(b->gexec));
    HIPCHK(hipGraphDestroy(b->graph));
    b->gexec = nullptr;
    b->graph = nullptr;
  }
  return PCG_OK;
}

int pcg_batch_set_tips(pcg_batch *b, int32_t sys, int32_t n_tips, const int32_t *tip_res,
                       const double *tip_orig, const double *tip_vel, int32_t n_ftips,
                       const int32_t *ftip_res, const double *ftip_force) {
  if (!b) return fail(PCG_E_ARG, "null batch");
  return set_tips(*b->B, sys, n_tips, tip_res, tip_orig, tip_vel, n_ftips, ftip_res, ftip_force);
}

int pcg_batch_set_noise_scale(pcg_batch *b, double scale) {
  if (!b || scale < 0) return fail(PCG_E_ARG, "bad arguments");
  DeviceBatch &B = *b->B;
  for (size_t s = 0; s < B.sys.size(); ++s)
    B.sys[s].const2 = const2_of(B.hp.p, B.sys[s].temperature) * scale;
  HIPCHK(hipMemcpyAsync((void *)B.d.sys, B.sys.data(), B.sys.size() * sizeof(SysDesc),
                        hipMemcpyHostToDevice, B.stream));
  HIPCHK(hipStreamSynchronize(B.stream));
  return PCG_OK;
}

int pcg_batch_get(pcg_batch *b, int32_t sys, int32_t which, double *out) {
  if (!b || !out || sys < 0 || sys >= (int)b->B->sys.size() || which < 0 || which > 7)
    return fail(PCG_E_ARG, "bad arguments");
  DeviceBatch &B = *b->B;
  if (which <= 5) return get_cols(B, sys, B.d.y[which], out);
  if (which == 6) return get_forces(B, sys, out);
  return get_cols(B, sys, B.d.vel, out);
}

int pcg_batch_set(pcg_batch *b, int32_t sys, int32_t which, const double *in) {
  if (!b || !in || sys < 0 || sys >= (int)b->B->sys.size() || which < 0 || which > 5)
    return fail(PCG_E_ARG, "bad arguments");
  DeviceBatch &B = *b->B;
  const SysDesc &d = B.sys[sys];
  std::vector<double> tmp(3 * (size_t)d.n);
  for (int i = 0; i < d.n; ++i)
    for (int c2 = 0; c2 < 3; ++c2) tmp[c2 * d.n + i] = in[3 * i + c2];
  for (int c2 = 0; c2 < 3; ++c2)
    HIPCHK(hipMemcpyAsync(B.d.y[which][c2] + d.res0, tmp.data() + c2 * d.n, d.n * 8,
                          hipMemcpyHostToDevice, B.stream));
  if (which == 0) {  // moved residues: the next step rebuilds the neighbour list
    SysState st;
    int rc = read_state(B, sys, st);
    if (rc) return rc;
    st.verify_first = 1;
    if ((rc = write_state(B, sys, st))) return rc;
  }
  HIPCHK(hipStreamSynchronize(B.stream));
  return PCG_OK;
}

int pcg_batch_energy(pcg_batch *b, int32_t sys, double *V, double *term_V) {
  if (!b || sys < 0 || sys >= (int)b->B->sys.size()) return fail(PCG_E_ARG, "bad arguments");
  SysState st;
  int rc = read_state(*b->B, sys, st);
  if (rc) return rc;
  if (V) *V = st.Vtot;
  if (term_V)
    for (int k = 0; k < NT; ++k) term_V[k] = st.V[k];
  return PCG_OK;
}

int pcg_batch_info(pcg_batch *b, int32_t *block, int32_t *spl, int64_t *bytes) {
  if (!b) return fail(PCG_E_ARG, "null batch");
  if (block) *block = b->persistent ? b->block : 0;
  if (spl) *spl = b->spl;
  if (bytes) *bytes = (int64_t)b->B->bytes;
  return PCG_OK;
}

int pcg_batch_autotune(pcg_batch *b, int32_t trial_steps, int32_t *block, double *ms_per_step) {
  if (!b || trial_steps <= 0) return fail(PCG_E_ARG, "bad arguments");
  HIPCHK(hipSetDevice(b->B->device));
  DeviceBatch &B = *b->B;
  if (!b->persistent) {  // stage kernels: fixed launch shapes, nothing to tune
    if (block) *block = 0;
    if (ms_per_step) *ms_per_step = 0.0;
    return PCG_OK;
  }
  // trials run on the live state, which is restored after each one, so the
  // trajectory does not depend on whether (or how) the batch was tuned
  size_t total = B.arena.total ? B.arena.total : 256;
  void *snap = nullptr;
  if (hipMalloc(&snap, total) != hipSuccess)
    return fail(PCG_E_NOMEM, "no device memory for the autotuner's state snapshot");
  int rc = PCG_OK;
  int spl0 = b->spl, block0 = b->block;
  b->spl = std::min(b->spl, (int)trial_steps);
  const int cands[] = {128, 256, 512, 768, 1024};
  int best = block0;
  double best_ms = 1e300;
  if (hipMemcpyAsync(snap, B.arena.base, total, hipMemcpyDeviceToDevice, B.stream) != hipSuccess)
    rc = fail(PCG_E_HIP, "snapshot copy failed");
  for (int pass = 0; pass < 2 && rc == PCG_OK; ++pass) {
    for (int c : cands) {
      if (pass == 0 && c != cands[0]) continue;  // pass 0 warms up the code objects
      b->block = c;
      if ((rc = run_persistent(b, trial_steps)) != PCG_OK) break;
      double ms = b->last_ms;
      if (hipMemcpyAsync(B.arena.base, snap, total, hipMemcpyDeviceToDevice, B.stream) != hipSuccess ||
          hipStreamSynchronize(B.stream) != hipSuccess) {
        rc = fail(PCG_E_HIP, "snapshot restore failed");
        break;
      }
      if (pass == 1 && ms < best_ms) {
        best_ms = ms;
        best = c;
      }
    }
  }
  (void)hipFree(snap);
  b->spl = spl0;
  b->block = rc == PCG_OK ? best : block0;
  if (rc) return rc;
  b->tuned = true;
  if (block) *block = best;
  if (ms_per_step) *ms_per_step = best_ms / trial_steps;
  return PCG_OK;
}

int pcg_batch_sizes(pcg_batch *b, int32_t sys, int32_t *sizes) {
  if (!b || !sizes || sys < 0 || sys >= (int)b->B->sys.size()) return fail(PCG_E_ARG, "bad arguments");
  SysState st;
  int rc = read_state(*b->B, sys, st);
  if (rc) return rc;
  const SysDesc &d = b->B->sys[sys];
  for (int t = 0; t < NT; ++t) sizes[t] = host_term_size(d, st, t);
  sizes[NT] = st.nl_n;
  sizes[NT + 1] = d.n;
  sizes[NT + 2] = st.rebuilds;
  return PCG_OK;
}

int pcg_batch_last_ms(pcg_batch *b, double *ms, int32_t *launches) {
  if (!b) return fail(PCG_E_ARG, "null batch");
  if (ms) *ms = b->last_ms;
  if (launches) *launches = b->last_launches;
  return PCG_OK;
}

} // extern "C"

// --------------------------------------------------------------- server ----
// Continuous batching: every job owns a one-system arena sized to it, which
// is what admission charges against the HBM budget. A round launches one
// persistent kernel over the descriptors of all resident jobs; jobs finish,
// fail or are cancelled independently and their slots refill from the queue
// in the next round.
namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct OwnedModel {
  int n = 0;
  std::vector<double> pos, orig, ang_theta, dih_phi, cont_len, tip_orig, tip_vel, ftip_force;
  std::vector<uint8_t> atype, ang_has, dih_has;
  std::vector<int32_t> chain, seq, teth, ang, dih, cont, cont_type, tip_res, ftip_res;
  double cell[3] = {0, 0, 0};
  pcg_model view() const {
    pcg_model m{};
    m.n = n;
    m.pos = pos.data();
    m.orig_pos = orig.data();
    m.atype = atype.data();
    m.chain_idx = chain.data();
    m.seq_idx = seq.data();
    m.n_teth = (int)teth.size() / 2;
    m.teth = teth.data();
    m.n_ang = (int)ang.size() / 3;
    m.ang = ang.data();
    m.ang_has = ang_has.data();
    m.ang_theta = ang_theta.data();
    m.n_dih = (int)dih.size() / 4;
    m.dih = dih.data();
    m.dih_has = dih_has.data();
    m.dih_phi = dih_phi.data();
    m.n_cont = (int)cont.size() / 2;
    m.cont = cont.data();
    m.cont_len = cont_len.data();
    m.cont_type = cont_type.data();
    for (int c = 0; c < 3; ++c) m.cell[c] = cell[c];
    m.n_tips = (int)tip_res.size();
    m.tip_res = tip_res.data();
    m.tip_orig = tip_orig.data();
    m.tip_vel = tip_vel.data();
    m.n_ftips = (int)ftip_res.size();
    m.ftip_res = ftip_res.data();
    m.ftip_force = ftip_force.data();
    return m;
  }
};

template <typename T> std::vector<T> cp(const T *p, size_t n) {
  return p ? std::vector<T>(p, p + n) : std::vector<T>(n, T{});
}

OwnedModel own(const pcg_model &m) {
  OwnedModel o;
  o.n = m.n;
  o.pos = cp(m.pos, 3 * (size_t)m.n);
  o.orig = m.orig_pos ? cp(m.orig_pos, 3 * (size_t)m.n) : o.pos;
  o.atype = cp(m.atype, m.n);
  o.chain = cp(m.chain_idx, m.n);
  o.seq = cp(m.seq_idx, m.n);
  o.teth = cp(m.teth, 2 * (size_t)m.n_teth);
  o.ang = cp(m.ang, 3 * (size_t)m.n_ang);
  o.ang_has = cp(m.ang_has, m.n_ang);
  o.ang_theta = cp(m.ang_theta, m.n_ang);
  o.dih = cp(m.dih, 4 * (size_t)m.n_dih);
  o.dih_has = cp(m.dih_has, m.n_dih);
  o.dih_phi = cp(m.dih_phi, m.n_dih);
  o.cont = cp(m.cont, 2 * (size_t)m.n_cont);
  o.cont_len = cp(m.cont_len, m.n_cont);
  o.cont_type = cp(m.cont_type, m.n_cont);
  o.tip_res = cp(m.tip_res, m.n_tips);
  o.tip_orig = cp(m.tip_orig, 3 * (size_t)m.n_tips);
  o.tip_vel = cp(m.tip_vel, 3 * (size_t)m.n_tips);
  o.ftip_res = cp(m.ftip_res, m.n_ftips);
  o.ftip_force = cp(m.ftip_force, 3 * (size_t)m.n_ftips);
  for (int c = 0; c < 3; ++c) o.cell[c] = m.cell[c];
  return o;
}

struct Job {
  int64_t id = 0;
  OwnedModel model;
  double T = -1;
  int32_t seed = 448;
  int64_t nsteps = 0, done = 0;
  int status = PCG_JOB_QUEUED;
  size_t bytes = 0;        // arena bytes (dry run)
  size_t block_bytes = 0;  // the pooled device block it runs in (>= bytes)
  std::unique_ptr<DeviceBatch> B;
  std::vector<double> result, result_F, result_vel;  // [3n] each, at retirement
  double V = 0;
  double t_submit = 0, t_start = 0, t_end = 0;
};

} // namespace

// Size classes of the served form: each runs its own launches on its own
// stream, with a workgroup size and an LDS layout of its own, so that small
// jobs never wait for the largest one (at most four streams: the hardware
// queues a process gets).
constexpr int SERVE_BUCKETS = 4;
constexpr int SERVE_BUCKET_MAX_N[SERVE_BUCKETS] = {128, 320, 640, 1 << 30};
constexpr int SERVE_BUCKET_BLOCK[SERVE_BUCKETS] = {256, 512, 1024, 1024};

struct ServeBucket {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  std::vector<int64_t> resident;  // jobs resident in this class
  std::vector<int64_t> inflight;  // the jobs of the launch in flight, in slot order
  bool busy = false;
  DevBatch *d_jobs = nullptr;
  JobSlot *d_slots = nullptr;
  DevBatch *h_jobs = nullptr;  // pinned: the copies run asynchronously
  JobSlot *h_slots = nullptr;
  int cap = 0;
};

struct pcg_server {
  pcg_params p{};
  pcg_server_options o{};
  hipStream_t stream = nullptr;  // size checks at submission
  std::map<int64_t, std::unique_ptr<Job>> jobs;
  std::deque<int64_t> queue;
  ServeBucket bucket[SERVE_BUCKETS];
  int64_t next_id = 1;
  size_t resident_bytes = 0;
  // device blocks of retired jobs, by size, reused for the next admissions
  // (no hipMalloc/hipFree per job); sizes are multiples of 256 KiB
  std::multimap<size_t, void *> pool;
  size_t pooled_bytes = 0;
  double noise_scale = 1.0;  // scales the Langevin noise of every job
  ~pcg_server() {
    for (auto &bk : bucket)
      if (bk.stream) (void)hipStreamSynchronize(bk.stream);
    for (auto &kv : jobs)
      if (kv.second->B && kv.second->B->arena.base)
        pool.emplace(kv.second->block_bytes, kv.second->B->arena.base);
    jobs.clear();
    for (auto &kv : pool) (void)hipFree(kv.second);
    for (auto &bk : bucket) {
      if (bk.d_jobs) (void)hipFree(bk.d_jobs);
      if (bk.d_slots) (void)hipFree(bk.d_slots);
      if (bk.h_jobs) (void)hipHostFree(bk.h_jobs);
      if (bk.h_slots) (void)hipHostFree(bk.h_slots);
      if (bk.done) (void)hipEventDestroy(bk.done);
      if (bk.stream) (void)hipStreamDestroy(bk.stream);
    }
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

size_t block_size(size_t arena_bytes) {
  const size_t g = (size_t)256 << 10;
  return (arena_bytes + g - 1) / g * g;
}

int retire(pcg_server *s, Job &j, int status) {
  if (j.B) {
    SysState st;
    if (status == PCG_JOB_DONE) {
      const size_t n3 = 3 * (size_t)j.model.n;
      j.result.assign(n3, 0.0);
      j.result_F.assign(n3, 0.0);
      j.result_vel.assign(n3, 0.0);
      int rc = get_cols(*j.B, 0, j.B->d.y[0], j.result.data());
      if (rc) return rc;
      if ((rc = get_forces(*j.B, 0, j.result_F.data()))) return rc;
      if ((rc = get_cols(*j.B, 0, j.B->d.vel, j.result_vel.data()))) return rc;
      if ((rc = read_state(*j.B, 0, st))) return rc;
      j.V = st.Vtot;
      j.done = st.step;
    }
    s->resident_bytes -= j.block_bytes;
    // the block goes back to the pool (the batch only borrowed it)
    s->pool.emplace(j.block_bytes, j.B->arena.base);
    s->pooled_bytes += j.block_bytes;
    j.B.reset();
  }
  j.status = status;
  j.t_end = now_s();
  return PCG_OK;
}

} // namespace

extern "C" {

int pcg_server_create(const pcg_params *p, const pcg_server_options *o, pcg_server **out) {
  if (!p || !out) return fail(PCG_E_ARG, "null argument");
  auto s = std::make_unique<pcg_server>();
  s->p = *p;
  if (o) s->o = *o;
  if (s->o.max_resident <= 0) s->o.max_resident = 1024;
  if (s->o.chunk_steps <= 0) s->o.chunk_steps = 1000;
  if (s->o.hbm_budget_bytes <= 0) s->o.hbm_budget_bytes = (int64_t)64 << 30;
  HIPCHK(hipSetDevice(s->o.device));
  HIPCHK(hipStreamCreate(&s->stream));
  for (auto &bk : s->bucket) {
    HIPCHK(hipStreamCreate(&bk.stream));
    HIPCHK(hipEventCreateWithFlags(&bk.done, hipEventDisableTiming));
  }
  *out = s.release();
  return PCG_OK;
}

int pcg_server_destroy(pcg_server *s) {
  delete s;
  return PCG_OK;
}

int pcg_server_submit(pcg_server *s, const pcg_model *m, double temperature, int32_t seed,
                      int64_t nsteps, int64_t *job_id) {
  if (!s || !m || nsteps < 0 || m->n <= 0) return fail(PCG_E_ARG, "bad arguments");
  auto j = std::make_unique<Job>();
  j->id = s->next_id++;
  j->model = own(*m);
  j->T = temperature;
  j->seed = seed;
  j->nsteps = nsteps;
  j->t_submit = now_s();
  // size check up front: a job that can never fit is refused
  std::unique_ptr<DeviceBatch> probe;
  pcg_model v = j->model.view();
  size_t bytes = 0;
  int rc = make_batch(s->p, 1, &v, &j->T, &j->seed, s->o.precision, 0, s->o.device, probe,
                      s->stream, &bytes);
  if (rc) return rc;
  if ((int64_t)block_size(bytes) > s->o.hbm_budget_bytes)
    return fail(PCG_E_NOMEM, "job needs " + std::to_string(block_size(bytes)) +
                                 " bytes, over the budget");
  const size_t nnc_bound = s->p.nat_cont_on ? (size_t)std::max(m->n_cont, 0) : 0;
  if (LdsLayout((size_t)m->n, sizeof(double), false, false, nnc_bound).total > LDS_DYN_MAX ||
      m->n >= 4096)
    return fail(PCG_E_ARG, "job of " + std::to_string(m->n) +
                               " residues does not fit the persistent kernel's LDS");
  j->bytes = bytes;
  if (job_id) *job_id = j->id;
  s->queue.push_back(j->id);
  s->jobs[j->id] = std::move(j);
  return PCG_OK;
}

namespace {

int bucket_of(int n) {
  int k = 0;
  while (k + 1 < SERVE_BUCKETS && n > SERVE_BUCKET_MAX_N[k]) ++k;
  return k;
}

// A finished launch of bucket k: every job's progress from its slot; jobs
// that are done, failed (non-finite energy or a list overflow) or cancelled
// leave the bucket.
int serve_collect(pcg_server *s, int k) {
  ServeBucket &bk = s->bucket[k];
  bk.busy = false;
  std::vector<int64_t> keep;
  for (size_t q = 0; q < bk.inflight.size(); ++q) {
    Job &j = *s->jobs[bk.inflight[q]];
    const JobSlot &sl = bk.h_slots[q];
    j.done = sl.step;
    int rc = PCG_OK;
    if (j.status == PCG_JOB_CANCELLED) {
      rc = retire(s, j, PCG_JOB_CANCELLED);
    } else if (sl.overflow || !std::isfinite(sl.Vtot)) {
      rc = retire(s, j, PCG_JOB_FAILED);
    } else if (j.done >= j.nsteps) {
      rc = retire(s, j, PCG_JOB_DONE);
    } else {
      keep.push_back(j.id);
    }
    if (rc) return rc;
  }
  // jobs admitted while the launch ran stay resident too
  for (int64_t id : bk.resident)
    if (std::find(bk.inflight.begin(), bk.inflight.end(), id) == bk.inflight.end())
      keep.push_back(id);
  bk.resident.swap(keep);
  bk.inflight.clear();
  return PCG_OK;
}

// The next launch of bucket k (idle, with resident jobs): chunk_steps steps of
// every resident job, the job descriptors and step targets copied in and the
// slots copied out on the bucket's stream, its event recorded after them.
int serve_launch(pcg_server *s, int k) {
  ServeBucket &bk = s->bucket[k];
  const int nres = (int)bk.resident.size();
  if (nres == 0 || bk.busy) return PCG_OK;
  if (nres > bk.cap) {
    if (bk.d_jobs) HIPCHK(hipFree(bk.d_jobs));
    if (bk.d_slots) HIPCHK(hipFree(bk.d_slots));
    if (bk.h_jobs) HIPCHK(hipHostFree(bk.h_jobs));
    if (bk.h_slots) HIPCHK(hipHostFree(bk.h_slots));
    bk.d_jobs = nullptr;
    bk.d_slots = nullptr;
    bk.h_jobs = nullptr;
    bk.h_slots = nullptr;
    bk.cap = std::max(nres, 2 * bk.cap);
    HIPCHK(hipMalloc(&bk.d_jobs, bk.cap * sizeof(DevBatch)));
    HIPCHK(hipMalloc(&bk.d_slots, bk.cap * sizeof(JobSlot)));
    HIPCHK(hipHostMalloc(&bk.h_jobs, bk.cap * sizeof(DevBatch)));
    HIPCHK(hipHostMalloc(&bk.h_slots, bk.cap * sizeof(JobSlot)));
  }
  const size_t fsz = sizeof(double);  // the persistent kernel's LDS forces are FP64
  // The launch's length: chunk_steps; while no job waits for admission,
  // up to four chunks, but no longer than the class's nearest job end (so
  // no workgroup idles to the launch's end and no finished job waits)
  int64_t round = s->o.chunk_steps;
  if (s->queue.empty()) {
    int64_t rem = INT64_MAX;
    for (int q = 0; q < nres; ++q) {
      const Job &j = *s->jobs[bk.resident[q]];
      rem = std::min<int64_t>(rem, j.nsteps - j.done);
    }
    round = std::max<int64_t>(round, std::min<int64_t>(rem, 4 * (int64_t)s->o.chunk_steps));
  }
  size_t lds = 0;
  for (int q = 0; q < nres; ++q) {
    Job &j = *s->jobs[bk.resident[q]];
    DeviceBatch &B = *j.B;
    bk.h_slots[q] = JobSlot{std::min<int64_t>(j.done + round, j.nsteps), 0, 0.0, 0, 0};
    // each job stages its bonded data and its Gear state in LDS when that
    // fits on its own
    const size_t nj = (size_t)B.max_n, cj = (size_t)B.max_nnc;
    B.d.bond_lds = LdsLayout(nj, fsz, true, false, cj).total <= LDS_DYN_MAX ? 1 : 0;
    B.d.gear_lds = B.d.bond_lds && LdsLayout(nj, fsz, true, true, cj).total <= LDS_DYN_MAX ? 1 : 0;
    lds = std::max(lds, LdsLayout(nj, fsz, B.d.bond_lds != 0, B.d.gear_lds != 0, cj).total);
    bk.h_jobs[q] = B.d;
  }
  if (lds > LDS_DYN_MAX) return fail(PCG_E_ARG, "job too large for the persistent kernel");
  HIPCHK(hipMemcpyAsync(bk.d_jobs, bk.h_jobs, nres * sizeof(DevBatch), hipMemcpyHostToDevice,
                        bk.stream));
  HIPCHK(hipMemcpyAsync(bk.d_slots, bk.h_slots, nres * sizeof(JobSlot), hipMemcpyHostToDevice,
                        bk.stream));
  launch_persistent_jobs(bk.d_jobs, bk.d_slots, nres, s->o.precision, SERVE_BUCKET_BLOCK[k],
                         (int)round, lds, bk.stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(bk.h_slots, bk.d_slots, nres * sizeof(JobSlot), hipMemcpyDeviceToHost,
                        bk.stream));
  HIPCHK(hipEventRecord(bk.done, bk.stream));
  bk.inflight = bk.resident;
  bk.busy = true;
  return PCG_OK;
}

// Admission, first fit in arrival order, within the HBM budget and the slot
// count; each job goes to the size class of its length, its state set up on
// that class's stream.
int serve_admit(pcg_server *s) {
  int nres = 0;
  for (auto &bk : s->bucket) nres += (int)bk.resident.size();
  for (auto it = s->queue.begin(); it != s->queue.end() && nres < s->o.max_resident;) {
    Job &j = *s->jobs[*it];
    if (j.status == PCG_JOB_CANCELLED) {
      it = s->queue.erase(it);
      continue;
    }
    const size_t need = block_size(j.bytes);
    if ((int64_t)(s->resident_bytes + need) > s->o.hbm_budget_bytes) {
      ++it;
      continue;
    }
    // a pooled block that fits, else a new one; idle pooled blocks are given
    // back when resident plus pooled memory would pass the budget
    void *blk = nullptr;
    size_t blk_bytes = need;
    auto pit = s->pool.lower_bound(need);
    if (pit != s->pool.end() && pit->first <= need + need / 4) {
      blk = pit->second;
      blk_bytes = pit->first;
      s->pooled_bytes -= pit->first;
      s->pool.erase(pit);
    } else {
      while (!s->pool.empty() &&
             (int64_t)(s->resident_bytes + s->pooled_bytes + need) > s->o.hbm_budget_bytes) {
        auto last = std::prev(s->pool.end());
        (void)hipFree(last->second);
        s->pooled_bytes -= last->first;
        s->pool.erase(last);
      }
      if (hipMalloc(&blk, need) != hipSuccess) {
        blk = nullptr;
        ++it;
        continue;
      }
    }
    const int k = bucket_of(j.model.n);
    ServeBucket &bk = s->bucket[k];
    pcg_model v = j.model.view();
    uint64_t key = mix64(s->o.philox_seed ^ mix64((uint64_t)(uint32_t)j.seed));
    // set up on the server's own stream: a class's stream may be busy with
    // its launch, which the set-up's synchronisation would wait for
    int rc = make_batch(s->p, 1, &v, &j.T, &j.seed, s->o.precision, key, s->o.device, j.B,
                        s->stream, nullptr, 0, blk, blk_bytes);
    if (rc) {
      s->pool.emplace(blk_bytes, blk);
      s->pooled_bytes += blk_bytes;
      if (rc == PCG_E_NOMEM) {
        ++it;
        continue;
      }
      return rc;
    }
    if (s->noise_scale != 1.0) {
      SysDesc &d = j.B->sys[0];
      d.const2 = const2_of(j.B->hp.p, d.temperature) * s->noise_scale;
      HIPCHK(hipMemcpyAsync((void *)j.B->d.sys, &d, sizeof(SysDesc), hipMemcpyHostToDevice,
                            bk.stream));
    }
    j.status = PCG_JOB_RUNNING;
    j.t_start = now_s();
    j.block_bytes = blk_bytes;
    s->resident_bytes += blk_bytes;
    bk.resident.push_back(j.id);
    ++nres;
    it = s->queue.erase(it);
  }
  return PCG_OK;
}

} // namespace

int pcg_server_round(pcg_server *s, int32_t *active) {
  if (!s) return fail(PCG_E_ARG, "null server");
  HIPCHK(hipSetDevice(s->o.device));
  int rc;
  // launches that have finished
  bool progressed = false;
  for (int k = 0; k < SERVE_BUCKETS; ++k) {
    ServeBucket &bk = s->bucket[k];
    if (!bk.busy) continue;
    hipError_t e = hipEventQuery(bk.done);
    if (e == hipErrorNotReady) continue;
    if (e != hipSuccess) return fail(PCG_E_HIP, std::string("HIP: ") + hipGetErrorString(e));
    if ((rc = serve_collect(s, k))) return rc;
    progressed = true;
  }
  if ((rc = serve_admit(s))) return rc;
  for (int k = 0; k < SERVE_BUCKETS; ++k)
    if (!s->bucket[k].busy && !s->bucket[k].resident.empty()) {
      if ((rc = serve_launch(s, k))) return rc;
      progressed = true;
    }
  // nothing finished and nothing to start: wait until some launch finishes
  // (whichever does first), so that a caller's loop of rounds does not spin
  while (!progressed) {
    bool any_busy = false;
    for (int k = 0; k < SERVE_BUCKETS; ++k) {
      ServeBucket &bk = s->bucket[k];
      if (!bk.busy) continue;
      any_busy = true;
      hipError_t e = hipEventQuery(bk.done);
      if (e == hipErrorNotReady) continue;
      if (e != hipSuccess) return fail(PCG_E_HIP, std::string("HIP: ") + hipGetErrorString(e));
      if ((rc = serve_collect(s, k))) return rc;
      progressed = true;
    }
    if (!any_busy) break;
    if (!progressed) {
      std::this_thread::sleep_for(std::chrono::microseconds(20));
      continue;
    }
    if ((rc = serve_admit(s))) return rc;
    for (int k = 0; k < SERVE_BUCKETS; ++k)
      if (!s->bucket[k].busy && !s->bucket[k].resident.empty() && (rc = serve_launch(s, k)))
        return rc;
  }
  int n = (int)s->queue.size();
  for (auto &bk : s->bucket) n += (int)bk.resident.size();
  if (active) *active = n;
  return PCG_OK;
}

int pcg_server_status(pcg_server *s, int64_t job, int32_t *status, int64_t *steps_done) {
  if (!s) return fail(PCG_E_ARG, "null server");
  auto it = s->jobs.find(job);
  if (it == s->jobs.end()) return fail(PCG_E_NOTFOUND, "no such job");
  if (status) *status = it->second->status;
  if (steps_done) *steps_done = it->second->done;
  return PCG_OK;
}

int pcg_server_result(pcg_server *s, int64_t job, double *pos, double *V) {
  if (!s) return fail(PCG_E_ARG, "null server");
  auto it = s->jobs.find(job);
  if (it == s->jobs.end()) return fail(PCG_E_NOTFOUND, "no such job");
  Job &j = *it->second;
  if (j.status != PCG_JOB_DONE) return fail(PCG_E_STATE, "job not finished");
  if (pos) std::memcpy(pos, j.result.data(), j.result.size() * sizeof(double));
  if (V) *V = j.V;
  return PCG_OK;
}

int pcg_server_get(pcg_server *s, int64_t job, int32_t which, double *out) {
  if (!s || !out || (which != 0 && which != 6 && which != 7)) return fail(PCG_E_ARG, "bad arguments");
  auto it = s->jobs.find(job);
  if (it == s->jobs.end()) return fail(PCG_E_NOTFOUND, "no such job");
  Job &j = *it->second;
  if (j.status != PCG_JOB_DONE) return fail(PCG_E_STATE, "job not finished");
  const std::vector<double> &v = which == 0 ? j.result : which == 6 ? j.result_F : j.result_vel;
  std::memcpy(out, v.data(), v.size() * sizeof(double));
  return PCG_OK;
}

int pcg_server_set_noise_scale(pcg_server *s, double scale) {
  if (!s || !(scale >= 0)) return fail(PCG_E_ARG, "bad arguments");
  HIPCHK(hipSetDevice(s->o.device));
  s->noise_scale = scale;
  // resident jobs change from their next launch on (queued on their class's
  // stream, after any launch in flight)
  for (auto &bk : s->bucket) {
    for (int64_t id : bk.resident) {
      Job &j = *s->jobs[id];
      SysDesc &d = j.B->sys[0];
      d.const2 = const2_of(j.B->hp.p, d.temperature) * scale;
      HIPCHK(hipMemcpyAsync((void *)j.B->d.sys, &d, sizeof(SysDesc), hipMemcpyHostToDevice,
                            bk.stream));
    }
    HIPCHK(hipStreamSynchronize(bk.stream));
  }
  return PCG_OK;
}

int pcg_server_cancel(pcg_server *s, int64_t job) {
  if (!s) return fail(PCG_E_ARG, "null server");
  auto it = s->jobs.find(job);
  if (it == s->jobs.end()) return fail(PCG_E_NOTFOUND, "no such job");
  Job &j = *it->second;
  if (j.status == PCG_JOB_QUEUED) {
    j.status = PCG_JOB_CANCELLED;
    j.t_end = now_s();
  } else if (j.status == PCG_JOB_RUNNING) {
    // its memory may be in use by a launch in flight: the job leaves its
    // class once the class's launch has finished
    ServeBucket &bk = s->bucket[bucket_of(j.model.n)];
    if (bk.busy) HIPCHK(hipEventSynchronize(bk.done));
    if (bk.busy) {
      int rc = serve_collect(s, bucket_of(j.model.n));
      if (rc) return rc;
    }
    if (j.status == PCG_JOB_RUNNING) {
      int rc = retire(s, j, PCG_JOB_CANCELLED);
      if (rc) return rc;
      bk.resident.erase(std::remove(bk.resident.begin(), bk.resident.end(), job), bk.resident.end());
    }
  }
  return PCG_OK;
}

int pcg_server_stats(pcg_server *s, int64_t *resident_bytes, int32_t *resident, int32_t *queued) {
  if (!s) return fail(PCG_E_ARG, "null server");
  if (resident_bytes) *resident_bytes = (int64_t)s->resident_bytes;
  if (resident) {
    int n = 0;
    for (auto &bk : s->bucket) n += (int)bk.resident.size();
    *resident = n;
  }
  if (queued) *queued = (int)s->queue.size();
  return PCG_OK;
}

int pcg_server_job_times(pcg_server *s, int64_t job, double *submit, double *start, double *end) {
  if (!s) return fail(PCG_E_ARG, "null server");
  auto it = s->jobs.find(job);
  if (it == s->jobs.end()) return fail(PCG_E_NOTFOUND, "no such job");
  if (submit) *submit = it->second->t_submit;
  if (start) *start = it->second->t_start;
  if (end) *end = it->second->t_end;
  return PCG_OK;
}


// ------------------------------------------------------ input preparation ----
// model::morph_into_saw (cg/src/cg/input/model.cpp:108-234). Each attempt
// draws the chain's bend and torsion angles, chains the local frames
// T_0 .. T_{n-2} (rotation i applied to the bond of direction (theta0, phi0)
// after rotations i-1 .. 0), places the chain at a point drawn in the start
// box and accepts it when no pair i, j >= i + 3 is closer than the
// intersection distance.
int pcg_morph_saw(const pcg_saw_options *o, int32_t n, const int32_t *chain_idx, int32_t seed,
                  double *pos, int64_t *draws) {
  if (!o || (n > 0 && (!chain_idx || !pos))) return fail(PCG_E_ARG, "null argument");
  if (n < 0) return fail(PCG_E_ARG, "negative residue count");
  if (!(o->bond > 0.0)) return fail(PCG_E_ARG, "SAW bond length must be positive");
  for (int i = 1; i < n; ++i)
    if (chain_idx[i] < chain_idx[i - 1])
      return fail(PCG_E_ARG, "chains must be contiguous and in chain-index order");
  double side = 0.0;
  if (o->box_size > 0.0) side = o->box_size;
  else if (o->density > 0.0) side = std::cbrt((double)n / o->density);
  if (o->pbc && !(side > 0.0))
    return fail(PCG_E_ARG, "SAW with periodic boundaries needs a non-zero start box");
  const int tries = o->max_tries > 0 ? o->max_tries : 9000;
  const double lo = -side / 2.0, hi = side / 2.0, span = hi - lo;
  const double cell = o->pbc ? span : 0.0, cell_inv = o->pbc ? 1.0 / span : 0.0;
  const double min_d2 = std::pow(o->intersection, 2.0);
  const double pi = std::acos(-1.0);
  Ran2 g(seed);
  int64_t used = 0;
  auto u = [&]() {
    ++used;
    return g.draw();
  };
  auto wrap = [&](double d) { return d - cell * std::round(d * cell_inv); };
  std::vector<double> phi, theta, T, R;
  for (int a = 0; a < n;) {
    int b = a;
    while (b < n && chain_idx[b] == chain_idx[a]) ++b;
    const int m = b - a;
    phi.assign(m + 1, 0.0);
    theta.assign(m + 1, 0.0);
    T.assign(9 * (size_t)m, 0.0);
    R.assign(3 * (size_t)m, 0.0);
    for (int attempt = 0; attempt < tries; ++attempt) {
      phi[0] = pi / 2.0;
      theta[0] = 0.0;
      phi[1] = 0.0;
      theta[1] = u() * pi / 3.0;  // drawn for every chain, one residue long too
      for (int i = 2; i < m - 1; ++i) {
        phi[i] = (2.0 * u() - 1.0) * pi;
        theta[i] = u() * pi / 3.0;
      }
      for (int i = 0; i < m - 1; ++i) {
        const double ct = std::cos(theta[i]), st = std::sin(theta[i]);
        const double cp = std::cos(phi[i]), sp = std::sin(phi[i]);
        double *M = &T[9 * (size_t)i];
        M[0] = ct, M[1] = st, M[2] = 0.0;
        M[3] = st * cp, M[4] = -ct * cp, M[5] = sp;
        M[6] = st * sp, M[7] = -ct * sp, M[8] = -cp;
      }
      const double theta0 = std::acos(1.0 - 2.0 * u());
      const double phi0 = 2.0 * pi * u();
      for (int i = 0; i < m - 1; ++i) {
        double v[3] = {o->bond * std::sin(theta0) * std::cos(phi0),
                       o->bond * std::sin(theta0) * std::sin(phi0), o->bond * std::cos(theta0)};
        for (int j = i; j >= 0; --j) {
          const double *M = &T[9 * (size_t)j];
          const double w0 = M[0] * v[0] + M[1] * v[1] + M[2] * v[2];
          const double w1 = M[3] * v[0] + M[4] * v[1] + M[5] * v[2];
          const double w2 = M[6] * v[0] + M[7] * v[1] + M[8] * v[2];
          v[0] = w0, v[1] = w1, v[2] = w2;
        }
        R[3 * (size_t)i] = v[0], R[3 * (size_t)i + 1] = v[1], R[3 * (size_t)i + 2] = v[2];
      }
      double start[3];
      for (int c = 0; c < 3; ++c) start[c] = span * u() + lo;
      for (int i = 0; i < m; ++i)
        for (int c = 0; c < 3; ++c) {
          double x = start[c];
          for (int j = 0; j < i; ++j) x += R[3 * (size_t)j + c];
          pos[3 * (size_t)(a + i) + c] = x;
        }
      bool ok = true;
      for (int i = 0; i + 3 < m && ok; ++i)
        for (int j = i + 3; j < m && ok; ++j) {
          const double *p1 = &pos[3 * (size_t)(a + i)], *p2 = &pos[3 * (size_t)(a + j)];
          const double dx = wrap(p2[0] - p1[0]), dy = wrap(p2[1] - p1[1]), dz = wrap(p2[2] - p1[2]);
          if (dx * dx + dy * dy + dz * dz < min_d2) ok = false;
        }
      if (ok) break;
    }
    a = b;
  }
  if (draws) *draws = used;
  return PCG_OK;
}

} // extern "C"
