/*
 * oracle.cpp -- FP64 CPU restatement of pas-cg's per-step simulation loop.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, to check results against. The product never uses it.
 *
 * Build with -O2 -ffp-contract=off and without fast-math so that the
 * arithmetic follows the expression order written here, which in turn follows
 * the reference expressions cited next to each function. Every expression
 * that the reference writes as a vector expression is evaluated per component
 * in the same operand order.
 *
 * Pinning (how each function is checked; see tests/):
 *   - every force term, NL rebuild, list updates, QA stages, PID, AFM, the
 *     Langevin step: per-call comparison against the reference's own kernel
 *     sources compiled into oracle/_ref (tests/test_oracle_vs_ref.py), plus
 *     finite-difference F = -dV/dr checks where the reference force is a
 *     gradient (tests/test_oracle_math.py);
 *   - ran2, initial velocities, Gear-5 + noise, tether/lrep/angles/Pauli/PID/
 *     DH/heuristic dihedrals/SAW: the golden V traces of the reference's
 *     parity fixtures (tests/test_golden.py);
 *   - walls and the box protocol: the reference's wall kernels through
 *     oracle/_ref and closed forms of the phases (tests/test_walls.py);
 *   - the AFM tips' moving averages: the reference's afm::vel and
 *     afm::force eval_forces fed the same samples, and a zero window's last
 *     sample (tests/test_afm_stats.py).
 */
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

constexpr double C216_INV = 0.89089871814033927; /* cg/include/cg/utils/math.h:6 */

struct v3 {
  double x, y, z;
};
inline v3 mk(double x, double y, double z) { return v3{x, y, z}; }
inline v3 operator+(v3 a, v3 b) { return v3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline v3 operator-(v3 a, v3 b) { return v3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline v3 operator-(v3 a) { return v3{-a.x, -a.y, -a.z}; }
inline v3 operator*(v3 a, double s) { return v3{a.x * s, a.y * s, a.z * s}; }
inline v3 operator*(double s, v3 a) { return v3{s * a.x, s * a.y, s * a.z}; }
inline v3 operator/(v3 a, double s) { return v3{a.x / s, a.y / s, a.z / s}; }
inline v3 &operator+=(v3 &a, v3 b) {
  a.x += b.x;
  a.y += b.y;
  a.z += b.z;
  return a;
}
inline v3 &operator-=(v3 &a, v3 b) {
  a.x -= b.x;
  a.y -= b.y;
  a.z -= b.z;
  return a;
}
/* cg/include/cg/types/vec3_ops_expr.h:6-33,197-226 */
inline double norm_sq(v3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline double norm(v3 a) { return std::sqrt(norm_sq(a)); }
inline double norm_inv(v3 a) { return 1.0 / std::sqrt(norm_sq(a)); }
inline double dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double l1(v3 a) { return std::abs(a.x) + std::abs(a.y) + std::abs(a.z); }
inline v3 cross(v3 a, v3 b) {
  return v3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z,
            a.x * b.y - a.y * b.x};
}
inline v3 unit(v3 a) { return norm_inv(a) * a; }

/* cg/include/cg/utils/math.h:307-311,366-369 */
inline double clampd(double v, double lo, double hi) {
  return v < lo ? lo : (v < hi ? v : hi);
}
inline double acos_c(double x) { return std::acos(clampd(x, -1.0, 1.0)); }
inline double ipow6(double s) { return s * (s * (s * (s * (s * (s * 1.0))))); }

/* Minimum-image wrap, cg/include/cg/sbox/pbc.h:37-49 */
struct pbc_t {
  v3 cell{0, 0, 0}, cell_inv{0, 0, 0};
  void set(v3 c) {
    cell = c;
    cell_inv = v3{0, 0, 0};
    if (c.x != 0.0) cell_inv.x = 1.0 / c.x;
    if (c.y != 0.0) cell_inv.y = 1.0 / c.y;
    if (c.z != 0.0) cell_inv.z = 1.0 / c.z;
  }
  v3 wrap(v3 v) const {
    v.x -= cell.x * std::round(v.x * cell_inv.x);
    v.y -= cell.y * std::round(v.y * cell_inv.y);
    v.z -= cell.z * std::round(v.z * cell_inv.z);
    return v;
  }
  v3 wrap(v3 e, v3 f) const { return wrap(f - e); }
};

/* Base potentials: cg/include/cg/base_forces/{harmonic,lj,shifted_lj,sink_lj}.h */
inline void harmonic(double H1, double H2, double nat_r, double r, double &V,
                     double &dV) {
  double dr = r - nat_r, dr2 = dr * dr;
  V = 0.5 * dr2 * (H1 + H2 * dr2);
  dV = dr * (H1 + 2.0 * H2 * dr2);
  dV = clampd(dV, -1.0e3, 1.0e3);
}
inline void lj(double depth, double r_min, double r_inv, double &V, double &dV) {
  double s = r_inv * r_min, s6 = ipow6(s), s12 = s6 * s6;
  V = depth * (s12 - 2.0 * s6);
  dV = 12.0 * depth * r_inv * (s6 - s12);
}
inline void shifted_lj(double depth, double r_min, double r_inv, double &V,
                       double &dV) {
  double s = r_inv * r_min, s6 = ipow6(s), s12 = s6 * s6;
  V = depth * (s12 - 2.0 * s6 + 1.0);
  dV = 12.0 * depth * r_inv * (s6 - s12);
  dV = clampd(dV, -1.0e3, 1.0e3);
}
struct sink_lj_t {
  double depth = 0, r_low = 0, r_high = 0;
  void operator()(double r, double r_inv, double &V, double &dV) const {
    double r_eff = r < r_low ? r_low : (r < r_high ? r : r_high);
    double s = r_inv * r_eff, s6 = ipow6(s), s12 = s6 * s6;
    V = depth * (s12 - 2.0 * s6);
    dV = 12.0 * depth * r_inv * (s6 - s12);
    /* sink_lj.h:20 zeroes dV when (r_low < r) && (r < r_low): never true. */
  }
};

/* Numerical Recipes ran2 (L'Ecuyer + Bays-Durham), nr.h:10-91 (COMPAT_MODE). */
struct ran2_t {
  static constexpr int im1 = 2147483563, im2 = 2147483399, imm1 = im1 - 1,
                       ia1 = 40014, ia2 = 40692, iq1 = 53668, iq2 = 52774,
                       ir1 = 12211, ir2 = 3791, ntab = 32,
                       ndiv = 1 + imm1 / ntab;
  static constexpr double eps = 1.2e-7f, rnmx = 1.f - eps,
                          am = 1.f / (float)im1;
  int iy = 0, idum = -448, idum2 = 123456789;
  int iv[ntab] = {0};
  long long draws = 0; /* draws after seeding */
  explicit ran2_t(int seed = 0) {
    next(seed);
    next(seed);
    draws = 0;
  }
  double next(int iseed = 0) {
    ++draws;
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
  double uniform() { return next(); }
  double normal() {
    double r1 = uniform(), r2 = uniform();
    double r = std::sqrt(-2.0 * std::log(r1));
    double t = (2.0 * M_PI) * r2;
    return r * std::cos(t);
  }
};

/* sync_data, cg/include/cg/amino/sync_data.h:98-104 */
struct sync_t {
  int8_t v[4] = {0, 0, 0, 0};
  bool valid() const { return v[0] >= 0 && v[1] >= 0 && v[2] >= 0 && v[3] >= 0; }
};
inline sync_t sub(sync_t a, sync_t b) {
  sync_t r;
  for (int k = 0; k < 4; ++k) r.v[k] = (int8_t)(a.v[k] - b.v[k]);
  return r;
}
inline sync_t add(sync_t a, sync_t b) {
  sync_t r;
  for (int k = 0; k < 4; ++k) r.v[k] = (int8_t)(a.v[k] + b.v[k]);
  return r;
}

struct nl_pair {
  int i1, i2;
  double orig_dist;
  bool taken;
};
struct nat_cont_t {
  int i1, i2;
  double nat_dist;
  int type;
  bool active;
  double change_t;
  int all_idx;
};
struct qa_free {
  int i1, i2;
  double orig_dist;
  bool vacant;
};
struct qa_cont {
  int i1, i2;
  double orig_dist;
  int type;
  double sat;
  sync_t d1, d2;
  bool vacant;
};
struct qa_cand {
  int i1, i2;
  double orig_dist, dist;
  int free_idx, type;
  sync_t d1, d2;
};
struct tip_t {
  int res;
  v3 orig;
  double ref_t;
  v3 vel;
};

} // namespace

struct ora_sys {
  ora_params p;
  int n = 0;
  std::vector<v3> r, F, v, orig_r_nl, n_vec, h_vec, noise;
  std::vector<v3> y[6];
  std::vector<int> atype, chain, seq, prev, next;
  pbc_t pbc, nl_orig_pbc;
  ran2_t gen;
  double noise_scale = 1.0;
  long long saw_draws = 0;

  /* time */
  double t = 0.0, true_t = 0.0;
  int step_idx = 0;

  /* energy */
  double V = 0.0, term_V[ORA_NUM_TERMS] = {0};

  /* flags (state.cpp setup_*) */
  bool nl_required = false, nl_invalid = true, verify_first = true;
  bool pauli_on = false, nat_cont_on = false;
  double total_disp = 0.0, nl_orig_pad = 0.0, nl_eff_cutoff = 0.0;
  int nl_rebuilds = 0;

  /* static lists */
  struct teth_t {
    int i1, i2;
    double nat;
  };
  std::vector<teth_t> tethers;
  std::vector<std::pair<int, int>> lrep;
  struct chir_t {
    int i1, i2, i3, i4;
    double nat_factor, nat_chir;
  };
  std::vector<chir_t> chir;
  struct ang_t {
    int i1, i2, i3;
    double theta;
    int type;
  };
  std::vector<ang_t> nat_ang, heur_ang;
  struct dih_t {
    int i1, i2, i3, i4;
    double phi;
    int type;
  };
  std::vector<dih_t> nat_dih, heur_dih;
  std::vector<tip_t> tips;
  std::vector<std::pair<int, v3>> ftips; /* residue, pulling force */

  /* dynamic lists */
  std::vector<nl_pair> nl;
  std::vector<nat_cont_t> all_nc, cur_nc;
  int num_changed = 0;
  std::vector<std::pair<int, int>> pauli_pairs;
  struct dh_pair {
    int i1, i2;
    double q1q2;
  };
  std::vector<dh_pair> dh_pairs;
  struct pid_bundle {
    int i1, i2;
    double orig_dist;
    int type;
  };
  std::vector<pid_bundle> pid_bundles;
  std::vector<qa_free> qa_free_pairs;
  std::vector<qa_cont> qa_contacts;
  std::vector<qa_cand> qa_cands;
  std::vector<int> qa_removed;
  std::vector<sync_t> sync;
  int num_qa_contacts = 0;

  /* walls (state::setup_walls, state.cpp:570-700): reaction force of each
   * plane in the last evaluation, residues connected to a wall, harmonic
   * connections (fixed at set-up), LJ connections in slot order (vect::set:
   * removal leaves a vacant slot, emplace appends), the LJ candidates and
   * releases since the last processing */
  v3 wall_F[6]{};
  std::vector<uint8_t> is_conn;
  struct hw_conn_t {
    int wall, res;
    v3 off;
    double nat;
  };
  std::vector<hw_conn_t> hw_conns;
  struct ljw_conn_t {
    int res, wall;
    v3 off;
    double sat;
    bool vacant;
  };
  std::vector<ljw_conn_t> ljw_conns;
  struct ljw_cand_t {
    int res, wall;
    double dist;
  };
  std::vector<ljw_cand_t> ljw_cands;
  std::vector<int> ljw_removed;
  int ljw_num_conns[6] = {0, 0, 0, 0, 0, 0};

  /* simulation box and its protocol (state.cpp:702-730, thread_main.cpp:144-330) */
  template <typename T> struct mavg_t { /* moving_avg, cg/include/cg/types/avg.h:5-38 */
    bool has = false;
    T avg{}, sum{};
    double init_time = 0.0, window = 0.0;
    int num = 0;
    void add(double time, T sample) {
      if (num == 0) {
        init_time = time;
        sum = sample;
        num = 1;
      } else {
        sum += sample;
        ++num;
      }
      if (init_time + window <= time) {
        avg = sum / (double)num;
        has = true;
        num = 0;
      }
    }
  };
  /* the tips' moving averages (afm::vel::tip avg_force and avg_perp_force,
   * afm::force::tip avg_vel), window afm.average stats over, fresh with every
   * set of tips (state::setup_afm, state.cpp:552-565) */
  std::vector<mavg_t<v3>> tip_avg_F, ftip_avg_v;
  std::vector<mavg_t<double>> tip_avg_perp;
  void reset_tip_stats() {
    mavg_t<v3> a;
    a.window = p.afm_stats_window;
    mavg_t<double> b;
    b.window = p.afm_stats_window;
    tip_avg_F.assign(tips.size(), a);
    tip_avg_perp.assign(tips.size(), b);
    ftip_avg_v.assign(ftips.size(), a);
  }
  bool box_set = false, log_on = false, shear = false;
  ora_box bx{};
  v3 box_min{0, 0, 0}, box_max{0, 0, 0};
  int side_wall[6] = {-1, -1, -1, -1, -1, -1};
  int phase = 0;
  double since = 0.0, until = 0.0, amplitude = 0.0, omega = 0.0, displacement = 0.0;
  v3 shift[6]{};
  double work[6] = {0, 0, 0, 0, 0, 0};
  mavg_t<v3> avg_F[6];
  mavg_t<double> avg_z;

  /* derived parameters */
  double V_factor = 0.0, screen_inv = 0.0;
  double gamma2 = 0, const2 = 0, dt = 0, dt_inv = 0, deltsq = 0;
  sink_lj_t qa_ljs[445];
  double qa_req_min_dist[445];
  double qa_max_req = 0.0, qa_factor = 0.0, qa_sat_speed = 0.0;
  sink_lj_t pid_bbp, pid_bbm, pid_ss[441];

  /* ---- force terms ---------------------------------------------------- */
  void add_F(double *Fo, int i, v3 f) const {
    Fo[3 * i] += f.x;
    Fo[3 * i + 1] += f.y;
    Fo[3 * i + 2] += f.z;
  }
  void sub_F(double *Fo, int i, v3 f) const {
    Fo[3 * i] -= f.x;
    Fo[3 * i + 1] -= f.y;
    Fo[3 * i + 2] -= f.z;
  }

  /* chir::eval_forces::iter, cg/src/cg/chir/eval_forces.cpp:10-31 */
  void eval_chir(double *Fo, double &Vo) {
    for (auto const &q : chir) {
      v3 r1 = r[q.i1], r2 = r[q.i2], r3 = r[q.i3], r4 = r[q.i4];
      v3 r12 = r2 - r1, r23 = r3 - r2, r34 = r4 - r3;
      v3 x12_23 = cross(r12, r23), x12_34 = cross(r12, r34),
         x23_34 = cross(r23, r34);
      double c = dot(r12, x23_34) * q.nat_factor;
      double cd = c - q.nat_chir;
      Vo += 0.5 * p.e_chi * cd * cd;
      double f = p.e_chi * cd * q.nat_factor;
      add_F(Fo, q.i1, f * x12_23);
      sub_F(Fo, q.i2, f * (x12_34 + x23_34));
      add_F(Fo, q.i3, f * (x12_23 + x12_34));
      sub_F(Fo, q.i4, f * x12_23);
    }
  }

  /* tether::eval_forces::iter, cg/src/cg/tether/eval_forces.cpp:18-33 */
  void eval_tether(double *Fo, double &Vo) {
    for (auto const &t_ : tethers) {
      v3 r12 = r[t_.i2] - r[t_.i1];
      double r12_n = norm(r12), V_, dV;
      harmonic(p.tether_H1, p.tether_H2, t_.nat, r12_n, V_, dV);
      dV = clampd(dV, -1e3, 1e3);
      v3 u = r12 / r12_n;
      Vo += V_;
      add_F(Fo, t_.i1, dV * u);
      sub_F(Fo, t_.i2, dV * u);
    }
  }

  /* local_rep::eval_forces::iter, cg/src/cg/local_rep/eval_forces.cpp:11-28 */
  void eval_lrep(double *Fo, double &Vo) {
    double cutoff = p.rep_cutoff;
    for (auto const &pr : lrep) {
      v3 r12 = r[pr.second] - r[pr.first];
      double rn = norm_inv(r12);
      if (1.0 > rn * cutoff) continue;
      v3 u = r12 * rn;
      double V_, dV;
      shifted_lj(p.lrep_depth, cutoff, rn, V_, dV);
      dV = clampd(dV, -1e3, 1e3);
      Vo += V_;
      add_F(Fo, pr.first, u * dV);
      sub_F(Fo, pr.second, u * dV);
    }
  }

  /* disulfide_force::operator() is called with a distance where it expects
   * an inverse distance (nat_cont/eval_forces.cpp:27, disulfide.h:12-20). */
  void nc_disulfide(double norm_inv_arg, double &V_, double &dV) const {
    if (p.nc_ss_kind == 1) {
      harmonic(p.nc_ss_H1, p.nc_ss_H2, p.nc_ss_nat_r, 1.0 / norm_inv_arg, V_,
               dV);
    } else if (p.nc_ss_kind == 2) {
      lj(p.nc_ss_depth, p.nc_ss_r_min, norm_inv_arg, V_, dV);
    } else {
      V_ = 0.0;
      dV = 0.0;
    }
  }

  /* nat_cont::eval_forces::iter, cg/src/cg/nat_cont/eval_forces.cpp:11-50 */
  void eval_nat_cont(double *Fo, double &Vo) {
    double cutoff = p.nl_cutoff;
    for (auto &c : cur_nc) {
      v3 r12 = pbc.wrap(r[c.i1], r[c.i2]);
      double r12_n = norm(r12);
      if (r12_n >= cutoff) continue;
      v3 u = r12 / r12_n;
      double V_, dV;
      if (p.nc_ss_kind != 0 && c.type == 5) {
        nc_disulfide(r12_n, V_, dV);
      } else {
        double rn = 1.0 / r12_n;
        lj(p.nat_cont_depth, c.nat_dist, rn, V_, dV);
      }
      dV = clampd(dV, -1e3, 1e3);
      Vo += V_;
      add_F(Fo, c.i1, u * dV);
      sub_F(Fo, c.i2, u * dV);

      bool active_now = r12_n <= c.nat_dist * C216_INV * p.counting_factor;
      if (active_now != c.active) {
        nat_cont_t &ref = all_nc[c.all_idx];
        c.active = ref.active = active_now;
        if (c.change_t < 0) {
          c.change_t = ref.change_t = t;
          ++num_changed;
        }
      }
    }
  }

  /* pauli::eval_forces::iter, cg/src/cg/pauli/eval_forces.cpp:12-30 */
  void eval_pauli(double *Fo, double &Vo) {
    double r_excl = p.rep_cutoff;
    for (auto const &pr : pauli_pairs) {
      v3 r12 = pbc.wrap(r[pr.first], r[pr.second]);
      double rn = norm_inv(r12);
      if (1.0 > rn * r_excl) continue;
      v3 u = r12 * rn;
      double V_, dV;
      shifted_lj(p.pauli_depth, r_excl, rn, V_, dV);
      dV = clampd(dV, -1e3, 1e3);
      Vo += V_;
      add_F(Fo, pr.first, u * dV);
      sub_F(Fo, pr.second, u * dV);
    }
  }

  /* const_dh / rel_dh ::iter, cg/src/cg/dh/const.cpp:18-40, rel.cpp:17-40 */
  void eval_dh(double *Fo, double &Vo, int variant) {
    if (!p.dh_on || p.dh_variant != variant) return;
    for (auto const &e : dh_pairs) {
      v3 r12 = pbc.wrap(r[e.i1], r[e.i2]);
      double r12_n = norm(r12);
      if (r12_n > p.dh_cutoff) continue;
      double rn = 1.0 / r12_n;
      v3 u = r12 * rn;
      double Vij, dV;
      if (variant == 0) {
        Vij = V_factor * e.q1q2 * std::exp(-r12_n * screen_inv) * rn;
        dV = -Vij * (screen_inv + rn);
      } else {
        Vij = V_factor * e.q1q2 * std::exp(-r12_n * screen_inv) * rn * rn;
        dV = -Vij * (screen_inv + 2.0 * rn);
      }
      Vo += Vij;
      v3 f = u * dV;
      add_F(Fo, e.i1, f);
      sub_F(Fo, e.i2, f);
    }
  }

  /* nat_ang::eval_forces::iter (angles/nat_ang/eval_forces.cpp:10-38) and
   * heur_ang::eval_forces::iter (angles/heur_ang/eval_forces.cpp:10-51). */
  void angle_geom(int i1, int i2, int i3, double &theta, v3 &d1, v3 &d2,
                  v3 &d3) const {
    v3 r12 = r[i2] - r[i1], r23 = r[i3] - r[i2];
    v3 x12_23 = cross(r12, r23);
    double r12_rn = norm_inv(r12), r23_rn = norm_inv(r23);
    d1 = unit(cross(r12, x12_23)) * r12_rn;
    d3 = unit(cross(r23, x12_23)) * r23_rn;
    d2 = -d1 - d3;
    double cos_theta = -dot(r12, r23) * r12_rn * r23_rn;
    cos_theta = clampd(cos_theta, -1.0, 1.0);
    if (cos_theta == 0.0) d1 = d2 = d3 = v3{0, 0, 0};
    theta = acos_c(cos_theta);
  }
  void eval_nat_ang(double *Fo, double &Vo) {
    for (auto const &a : nat_ang) {
      double theta;
      v3 d1, d2, d3;
      angle_geom(a.i1, a.i2, a.i3, theta, d1, d2, d3);
      double dth = theta - a.theta;
      Vo += p.CBA * dth * dth;
      double dV = 2.0 * p.CBA * dth;
      sub_F(Fo, a.i1, dV * d1);
      sub_F(Fo, a.i2, dV * d2);
      sub_F(Fo, a.i3, dV * d3);
    }
  }
  void eval_heur_ang(double *Fo, double &Vo) {
    for (auto const &a : heur_ang) {
      double theta;
      v3 d1, d2, d3;
      angle_geom(a.i1, a.i2, a.i3, theta, d1, d2, d3);
      int ty = a.type;
      const double(*c)[9] = p.heur_ang_poly;
      double t2 = theta * theta, t3 = t2 * theta, t4 = t3 * theta,
             t5 = t4 * theta, t6 = t3 * t3;
      double V_ = c[0][ty] + c[1][ty] * theta + c[2][ty] * t2 + c[3][ty] * t3 +
                  c[4][ty] * t4 + c[5][ty] * t5 + c[6][ty] * t6;
      double dV = c[1][ty] + (2.0 * c[2][ty]) * theta + (3.0 * c[3][ty]) * t2 +
                  (4.0 * c[4][ty]) * t3 + (5.0 * c[5][ty]) * t4 +
                  (6.0 * c[6][ty]) * t5;
      Vo += V_;
      sub_F(Fo, a.i1, dV * d1);
      sub_F(Fo, a.i2, dV * d2);
      sub_F(Fo, a.i3, dV * d3);
    }
  }

  /* cnd / snd ::iter, cg/src/cg/angles/nat_dih/complex.cpp:10-47,
   * simple.cpp:10-42 */
  void eval_nat_dih(double *Fo, double &Vo, int variant) {
    if (p.nat_dih_variant != variant) return;
    for (auto const &d : nat_dih) {
      v3 r1 = r[d.i1], r2 = r[d.i2], r3 = r[d.i3], r4 = r[d.i4];
      v3 r12 = r2 - r1, r23 = r3 - r2, r34 = r4 - r3;
      v3 x12_23 = cross(r12, r23), x23_34 = cross(r23, r34);
      double x12_23_rn = norm_inv(x12_23), x23_34_rn = norm_inv(x23_34);
      v3 x12_23_u = x12_23 * x12_23_rn, x23_34_u = x23_34 * x23_34_rn;
      double cos_phi = dot(x12_23_u, x23_34_u);
      double phi = acos_c(cos_phi);
      if (dot(x12_23, r34) < 0.0) phi = -phi;
      double diff = phi - d.phi, dV_dphi;
      if (variant == 0) {
        double sd = std::sin(diff), cd = std::cos(diff);
        double s3 = std::sin(3.0 * diff), c3 = std::cos(3.0 * diff);
        Vo += p.CDA * (1.0 - cd) + p.CDB * (1.0 - c3);
        dV_dphi = p.CDA * sd + p.CDB * 3.0 * s3;
      } else {
        Vo += 0.5 * p.CDH * diff * diff;
        dV_dphi = p.CDH * diff;
      }
      double r23_n = norm(r23);
      v3 dr1 = -x12_23_u * r23_n * x12_23_rn;
      v3 dr4 = x23_34_u * r23_n * x23_34_rn;
      v3 df = (-dr1 * dot(r12, r23) + dr4 * dot(r23, r34)) / (r23_n * r23_n);
      v3 dr2 = -dr1 + df;
      v3 dr3 = -dr4 - df;
      sub_F(Fo, d.i1, dV_dphi * dr1);
      sub_F(Fo, d.i2, dV_dphi * dr2);
      sub_F(Fo, d.i3, dV_dphi * dr3);
      sub_F(Fo, d.i4, dV_dphi * dr4);
    }
  }

  /* heur_dih::eval_forces::iter, cg/src/cg/angles/heur_dih/eval_forces.cpp:10-55 */
  void eval_heur_dih(double *Fo, double &Vo) {
    for (auto const &d : heur_dih) {
      v3 r1 = r[d.i1], r2 = r[d.i2], r3 = r[d.i3], r4 = r[d.i4];
      v3 r12 = r2 - r1, r23 = r3 - r2, r34 = r4 - r3;
      v3 x23_12 = cross(r23, r12), x34_23 = cross(r34, r23);
      double x23_12_rn = norm_inv(x23_12), x34_23_rn = norm_inv(x34_23);
      v3 x23_12_u = x23_12 * x23_12_rn, x34_23_u = x34_23 * x34_23_rn;
      double cos_phi = dot(x23_12_u, x34_23_u);
      double phi = acos_c(cos_phi);
      if (-dot(r12, x34_23_u) < 0.0) phi = -phi;
      double sin_phi = std::sin(phi);
      double s2 = sin_phi * sin_phi, c2 = cos_phi * cos_phi,
             sc = sin_phi * cos_phi;
      int ty = d.type;
      const double(*c)[9] = p.heur_dih_coeffs;
      Vo += c[0][ty] + c[1][ty] * sin_phi + c[2][ty] * cos_phi + c[3][ty] * s2 +
            c[4][ty] * c2 + c[5][ty] * sc;
      double dV = c[1][ty] * cos_phi - c[2][ty] * sin_phi +
                  2.0 * (c[3][ty] - c[4][ty]) * sc + c[5][ty] * (c2 - s2);
      double r23_n = norm(r23);
      v3 dr1 = x23_12_u * r23_n * x23_12_rn;
      v3 dr4 = -x34_23_u * r23_n * x34_23_rn;
      v3 df = (dr1 * (-dot(r12, r23)) - dr4 * (-dot(r34, r23))) /
              (r23_n * r23_n);
      v3 dr2 = -dr1 + df;
      v3 dr3 = -dr4 - df;
      sub_F(Fo, d.i1, dV * dr1);
      sub_F(Fo, d.i2, dV * dr2);
      sub_F(Fo, d.i3, dV * dr3);
      sub_F(Fo, d.i4, dV * dr4);
    }
  }

  /* pid::eval_forces, vectorised semantics (pid/eval_forces.cpp:337-538):
   * no orig_dist pruning, r12_n = sqrt, cos(psi) clamped, bb_lj chosen by the
   * second improper dihedral's block. Lambda: pid/lambda.h:58-85. */
  static bool lam_supp(double alpha, double psi0, double psi) {
    return std::abs(alpha * (psi - psi0)) < M_PI;
  }
  void lam_eval(double alpha, double psi0, double psi, double &val,
                double &deriv) const {
    if (p.pid_lambda_variant == 0) {
      double s = alpha * (psi - psi0);
      if (std::abs(s) >= M_PI) {
        val = deriv = 0.0;
      } else {
        val = 0.5 * std::cos(s) + 0.5;
        deriv = -0.5 * alpha * std::sin(s);
      }
    } else {
      double tt = alpha * (psi - psi0) * M_1_PI, a = std::abs(tt);
      if (a >= M_PI) {
        val = deriv = 0.0;
      } else {
        double x = 2.0 * a * (a - 1.0);
        val = 1.0 - (a * a) / (x + 1.0);
        double dval_da = x / ((x + 1.0) * (x + 1.0));
        deriv = alpha * M_1_PI * dval_da;
        if (tt < 0.0) deriv = -deriv;
      }
    }
  }
  void pid_psi(int i1, int i2, double &psi, v3 &d_self, v3 &d_next,
               v3 &d_prev, v3 &d_other, bool &ok) const {
    /* improper dihedral (i, i+1, i-1, j) with raw i+-1 neighbours */
    int ip = i1 - 1, in = i1 + 1;
    v3 zero{0, 0, 0};
    v3 r_in = in < n ? r[in] : zero, r_ip = ip >= 0 ? r[ip] : zero;
    v3 rij = r[i1] - r_in;
    v3 rkj = ip >= 0 ? (r_ip - r_in) : zero;
    v3 rkl = r_ip - r[i2];
    v3 rm = cross(rij, rkj), rn = cross(rkj, rkl);
    double rm_n = std::sqrt(norm_sq(rm)), rn_n = std::sqrt(norm_sq(rn));
    ok = !(rm_n < 0.1 || rn_n < 0.1);
    double rm_ninv = 1.0 / rm_n, rn_ninv = 1.0 / rn_n,
           rkj_ninv = norm_inv(rkj), rkj_n = 1.0 / rkj_ninv;
    v3 fi = rm * rkj_n * rm_ninv * rm_ninv;
    v3 fl = -rn * rkj_n * rn_ninv * rn_ninv;
    v3 df = (fi * dot(rij, rkj) - fl * dot(rkl, rkj)) * rkj_ninv * rkj_ninv;
    d_self = fi;
    d_next = -fi + df;
    d_prev = -fl - df;
    d_other = fl;
    double cos_psi = dot(rm, rn) * rm_ninv * rn_ninv;
    cos_psi = std::max(std::min(cos_psi, 1.0), -1.0);
    psi = std::acos(cos_psi);
    if (dot(rij, rn) < 0.0) psi = -psi;
  }
  void eval_pid(double *Fo, double &Vo) {
    double cutoff = p.nl_cutoff;
    for (auto const &b : pid_bundles) {
      int i1 = b.i1, i2 = b.i2;
      if (atype[i1] == 14 || atype[i2] == 14) continue; /* PRO */
      v3 r12 = pbc.wrap(r[i1], r[i2]);
      double r12_n = std::sqrt(norm_sq(r12));
      double r12_rn = 1.0 / r12_n;
      if (!(r12_n < cutoff)) continue;
      double psi1, psi2;
      v3 d1_1, d1_1n, d1_1p, d1_2, d2_2, d2_2n, d2_2p, d2_1;
      bool ok1, ok2;
      pid_psi(i1, i2, psi1, d1_1, d1_1n, d1_1p, d1_2, ok1);
      pid_psi(i2, i1, psi2, d2_2, d2_2n, d2_2p, d2_1, ok2);
      if (!ok1 || !ok2) continue;

      bool sep3 = (i2 - i1 == 3);
      /* Lambda choice of pid/eval_forces.cpp:86-97 (m = 1) and :135-146
       * (m = 2), "opt" meaning bb+. Block 1 starts at bb+ and drops to bb-
       * when psi1 is outside bb+ support and the pair is not (i, i+3).
       * Block 2 starts at bb+ unless the pair is (i, i+3), then drops to bb-
       * when psi2 is outside the support; it also fixes the bb LJ. */
      bool opt1 = lam_supp(p.pid_bbp_alpha, p.pid_bbp_psi0, psi1) || sep3;
      bool opt2 = !sep3 && lam_supp(p.pid_bbp_alpha, p.pid_bbp_psi0, psi2);
      double l1a = opt1 ? p.pid_bbp_alpha : p.pid_bbm_alpha;
      double l1p = opt1 ? p.pid_bbp_psi0 : p.pid_bbm_psi0;
      double l2a = opt2 ? p.pid_bbp_alpha : p.pid_bbm_alpha;
      double l2p = opt2 ? p.pid_bbp_psi0 : p.pid_bbm_psi0;
      const sink_lj_t &bb_lj = opt2 ? pid_bbp : pid_bbm;

      double dV1 = 0, dV2 = 0, dVr = 0, V_ = 0;
      if (lam_supp(l1a, l1p, psi1) && lam_supp(l2a, l2p, psi2)) {
        double lam1, der1, lam2, der2;
        lam_eval(l1a, l1p, psi1, lam1, der1);
        lam_eval(l2a, l2p, psi2, lam2, der2);
        if (lam1 * lam2 > 5e-5) {
          double ljV, ljdV;
          bb_lj(r12_n, r12_rn, ljV, ljdV);
          V_ += lam1 * lam2 * ljV;
          dV1 += der1 * lam2 * ljV;
          dV2 += der2 * lam1 * ljV;
          dVr += lam1 * lam2 * ljdV;
        }
      }
      const sink_lj_t &ss = pid_ss[b.type];
      double sa = p.pid_ss_alpha, sp = p.pid_ss_psi0;
      if (lam_supp(sa, sp, psi1) && lam_supp(sa, sp, psi2)) {
        double lam1, der1, lam2, der2;
        lam_eval(sa, sp, psi1, lam1, der1);
        lam_eval(sa, sp, psi2, lam2, der2);
        if (lam1 * lam2 > 5e-5 && ss.depth > 5e-5) {
          double ljV, ljdV;
          ss(r12_n, r12_rn, ljV, ljdV);
          V_ += lam1 * lam2 * ljV;
          dV1 += der1 * lam2 * ljV;
          dV2 += der2 * lam1 * ljV;
          dVr += lam1 * lam2 * ljdV;
        }
      }
      Vo += V_;
      v3 u = r12 * r12_rn;
      int i1p = i1 - 1, i1n = i1 + 1, i2p = i2 - 1, i2n = i2 + 1;
      if (i1p >= 0) sub_F(Fo, i1p, dV1 * d1_1p);
      sub_F(Fo, i1, dV1 * d1_1 + dV2 * d2_1 - dVr * u);
      if (i1n < n) sub_F(Fo, i1n, dV1 * d1_1n);
      if (i2p >= 0) sub_F(Fo, i2p, dV2 * d2_2p);
      sub_F(Fo, i2, dV1 * d1_2 + dV2 * d2_2 + dVr * u);
      if (i2n < n) sub_F(Fo, i2n, dV2 * d2_2n);
    }
  }

  /* pid::count_active (cg/src/cg/pid/count_active.cpp:4-164), the report's
   * PID counts over the current bundles: out = {intra, intra_bb, intra_ss,
   * inter, inter_bb, inter_ss}. Unlike the force evaluation it skips a
   * bundle when 1/|rm| or 1/|rn| exceeds 10, keeps pairs with
   * r12 <= cutoff, takes acos of the unclamped cosine, and counts a lambda
   * pair whose product exceeds 5e-5 when r12 < r_high * 2^(-1/6) * factor. */
  void improper_psi(int i, int j, double &psi, bool &skip) const {
    int ip = i - 1, in = i + 1;
    v3 zero{0, 0, 0};
    v3 r_in = in < n ? r[in] : zero, r_ip = ip >= 0 ? r[ip] : zero;
    v3 rij = r[i] - r_in;
    v3 rkj = ip >= 0 ? (r_ip - r_in) : zero;
    v3 rkl = r_ip - r[j];
    v3 rm = cross(rij, rkj), rn = cross(rkj, rkl);
    double rm_ninv = norm_inv(rm), rn_ninv = norm_inv(rn);
    skip = rm_ninv > 10.0 || rn_ninv > 10.0;
    psi = std::acos(dot(rm, rn) * rm_ninv * rn_ninv);
    if (dot(rij, rn) < 0.0) psi = -psi;
  }
  void pid_counts(int32_t out[6]) const {
    for (int k = 0; k < 6; ++k) out[k] = 0;
    const double factor = p.counting_factor, cutoff = p.nl_cutoff;
    for (auto const &b : pid_bundles) {
      int i1 = b.i1, i2 = b.i2;
      if (atype[i1] == 14 || atype[i2] == 14) continue; /* PRO */
      double r12_n = std::sqrt(norm_sq(pbc.wrap(r[i1], r[i2])));
      if (r12_n > cutoff) continue;
      double psi1, psi2;
      bool skip1, skip2;
      improper_psi(i1, i2, psi1, skip1);
      if (skip1) continue;
      improper_psi(i2, i1, psi2, skip2);
      if (skip2) continue;
      bool sep3 = (i2 - i1 == 3);
      bool opt1 = lam_supp(p.pid_bbp_alpha, p.pid_bbp_psi0, psi1) || sep3;
      bool opt2 = !sep3 && lam_supp(p.pid_bbp_alpha, p.pid_bbp_psi0, psi2);
      double l1a = opt1 ? p.pid_bbp_alpha : p.pid_bbm_alpha;
      double l1p = opt1 ? p.pid_bbp_psi0 : p.pid_bbm_psi0;
      double l2a = opt2 ? p.pid_bbp_alpha : p.pid_bbm_alpha;
      double l2p = opt2 ? p.pid_bbp_psi0 : p.pid_bbm_psi0;
      double bb_r_high = opt2 ? p.pid_bbp_lj[2] : p.pid_bbm_lj[2];
      bool same = chain[i1] == chain[i2];
      if (lam_supp(l1a, l1p, psi1) && lam_supp(l2a, l2p, psi2)) {
        double lam1, der1, lam2, der2;
        lam_eval(l1a, l1p, psi1, lam1, der1);
        lam_eval(l2a, l2p, psi2, lam2, der2);
        if (lam1 * lam2 > 5e-5 && r12_n < bb_r_high * C216_INV * factor) {
          ++out[same ? 0 : 3];
          ++out[same ? 1 : 4];
        }
      }
      double sa = p.pid_ss_alpha, sp = p.pid_ss_psi0;
      if (lam_supp(sa, sp, psi1) && lam_supp(sa, sp, psi2)) {
        double lam1, der1, lam2, der2;
        lam_eval(sa, sp, psi1, lam1, der1);
        lam_eval(sa, sp, psi2, lam2, der2);
        if (lam1 * lam2 > 5e-5 && p.pid_ss_depth[b.type] > 5e-5 &&
            r12_n < p.pid_ss_r_high[b.type] * C216_INV * factor) {
          ++out[same ? 0 : 3];
          ++out[same ? 2 : 5];
        }
      }
    }
  }

  /* qa::prepare_nh::iter, cg/src/cg/qa/prepare_nh.cpp:10-19 */
  void prepare_nh() {
    for (int i = 0; i < n; ++i) {
      int ip = prev[i], inx = next[i];
      if (ip < 0 || inx < 0) continue;
      v3 v1 = r[i] - r[ip], v2 = r[inx] - r[i];
      n_vec[i] = unit(v2 - v1);
      h_vec[i] = unit(cross(v2, v1));
    }
  }

  /* qa::loop_over_candidates::iter, cg/src/cg/qa/loop_over_candidates.cpp:12-99 */
  void eval_qa_cand(double *Fo, double &Vo) {
    for (int idx = 0; idx < (int)qa_free_pairs.size(); ++idx) {
      qa_free const &fp = qa_free_pairs[idx];
      if (fp.vacant) continue;
      if (fp.orig_dist - total_disp > qa_max_req) continue;
      int i1 = fp.i1, i2 = fp.i2;
      v3 r12 = pbc.wrap(r[i1], r[i2]);
      double rn = norm_inv(r12);
      v3 u = r12 * rn;
      if (1.0 > rn * qa_max_req) continue;
      if (1.0 < rn * p.rep_cutoff) {
        double V_, dV;
        shifted_lj(p.pauli_depth, p.rep_cutoff, rn, V_, dV);
        Vo += V_;
        add_F(Fo, i1, u * dV);
        sub_F(Fo, i2, u * dV);
      }
      v3 n1 = n_vec[i1], n2 = n_vec[i2], h1 = h_vec[i1], h2 = h_vec[i2];
      bool back1 = std::abs(dot(h1, u)) > p.qa_min_cos_hr;
      bool back2 = std::abs(dot(h2, u)) > p.qa_min_cos_hr;
      bool bb_extra = std::abs(dot(h1, h2)) > p.qa_min_cos_hh;
      bool side1 = dot(n1, u) < p.qa_max_cos_nr;
      bool side2 = -dot(n2, u) < p.qa_max_cos_nr;
      int a1 = atype[i1], a2 = atype[i2];
      int type = 0;
      if (back1 && back2 && bb_extra)
        type = 1;
      else if (back1 && side2)
        type = 2;
      else if (side1 && back2)
        type = 3;
      else if (side1 && side2)
        type = 4 + a1 + a2 * 21;
      if (type == 0) continue;
      double req = qa_req_min_dist[type] * (1.0 + p.qa_formation_tol);
      if (1.0 > rn * req) continue;
      int pt2 = p.ptype[a2];
      sync_t sd1;
      sd1.v[0] += back1 ? 1 : 0;
      sd1.v[1] += side1 ? 1 : 0;
      sd1.v[2] += (side1 && pt2 == 1) ? 1 : 0;
      sd1.v[3] += (side1 && pt2 == 2) ? 1 : 0;
      if (!sub(sync[i1], sd1).valid()) continue;
      int pt1 = p.ptype[a1];
      sync_t sd2;
      sd2.v[0] += back2 ? 1 : 0;
      sd2.v[1] += side2 ? 1 : 0;
      sd2.v[2] += (side2 && pt1 == 1) ? 1 : 0;
      sd2.v[3] += (side2 && pt1 == 2) ? 1 : 0;
      if (!sub(sync[i2], sd2).valid()) continue;
      qa_cands.push_back(qa_cand{i1, i2, fp.orig_dist, 1.0 / rn, idx, type,
                                 sd1, sd2});
    }
  }

  /* qa::process_contacts::iter, cg/src/cg/qa/process_contacts.cpp:12-72.
   * The reference sizes this loop by free_pairs->size() (line 90-92); the
   * defined part of that is min(free pairs, contacts). */
  void eval_qa_cont(double *Fo, double &Vo) {
    int cnt = (int)std::min(qa_free_pairs.size(), qa_contacts.size());
    for (int idx = 0; idx < cnt; ++idx) {
      qa_cont &c = qa_contacts[idx];
      if (c.vacant) continue;
      int i1 = c.i1, i2 = c.i2;
      double sat = c.sat;
      v3 r12 = pbc.wrap(r[i1], r[i2]);
      double rn = norm_inv(r12);
      if (rn * p.nl_cutoff < 1.0) continue;
      v3 u = r12 * rn;
      const sink_lj_t &ljf = qa_ljs[c.type];
      double r12_n = rn * norm_sq(r12);
      double Vij, dVij;
      ljf(r12_n, rn, Vij, dVij);
      Vo += sat * Vij;
      v3 f = sat * dVij * u;
      add_F(Fo, i1, f);
      sub_F(Fo, i2, f);
      bool breaking = r12_n > qa_factor * ljf.r_high;
      if (breaking) {
        sat = std::min(sat - qa_sat_speed * p.dt, 0.0);
        if (sat == 0.0) qa_removed.push_back(idx);
      } else {
        sat = std::max(sat + qa_sat_speed * p.dt, 0.0);
      }
      c.sat = sat;
    }
  }

  /* vel AFM tips, cg/src/cg/afm/vel/eval_forces.cpp:10-31 */
  void eval_vel_afm(double *Fo, double &Vo) {
    for (size_t k = 0; k < tips.size(); ++k) {
      tip_t const &tp = tips[k];
      v3 cur = tp.orig + (t - tp.ref_t) * tp.vel;
      v3 ra = cur - r[tp.res];
      double ra_n = norm(ra);
      if (ra_n > 1e-10) {
        v3 u = ra / ra_n;
        double V_, dV;
        harmonic(p.afm_H1, p.afm_H2, 0.0, ra_n, V_, dV);
        Vo += V_;
        v3 force = dV * u;
        add_F(Fo, tp.res, force);
        /* the force on the tip, and its part perpendicular to the pulling
         * direction in the x-z plane (eval_forces.cpp:23-29) */
        tip_avg_F[k].add(t, -force);
        v3 w = unit(tp.vel);
        double perp = force.x * w.z / (w.x + w.z * w.z / w.x) -
                      force.z / (1.0 + (w.z / w.x) * (w.z / w.x));
        tip_avg_perp[k].add(t, perp);
      }
    }
  }

  /* force AFM tips, cg/src/cg/afm/force/eval_forces.cpp:10-13 (no energy).
   * The average velocity reads the state's velocities (v = y1 / dt as
   * lang::step leaves them, step.cpp:67, zero before the first step) at
   * the current time: thread_setup.cpp:711-713 binds neither v nor t of
   * this task, so the reference reads unset views there. */
  void eval_force_afm(double *Fo) {
    for (size_t k = 0; k < ftips.size(); ++k) {
      add_F(Fo, ftips[k].first, ftips[k].second);
      ftip_avg_v[k].add(t, v[ftips[k].first]);
    }
  }

  v3 plane_n(int w) const { return v3{p.wall_normal[w][0], p.wall_normal[w][1], p.wall_normal[w][2]}; }
  v3 plane_o(int w) const { return v3{p.wall_origin[w][0], p.wall_origin[w][1], p.wall_origin[w][2]}; }
  static double sgn(double x) { return (double)((0.0 < x) - (x < 0.0)); } /* math.h:119-121 */

  /* wall::solid::eval_forces::iter, cg/src/cg/wall/solid/eval_forces.cpp:4-17 */
  void eval_solid_walls(double *Fo) {
    for (int i = 0; i < n; ++i)
      for (int w = 0; w < p.n_walls; ++w) {
        if (p.wall_kind[w] != 0) continue;
        v3 nrm = plane_n(w), org = plane_o(w);
        double sd = dot(nrm, r[i] - org), dist = std::abs(sd);
        if (dist >= p.wall_min_dist) continue;
        double dV = p.wall_depth * std::pow(dist, -10.0) * sgn(sd);
        add_F(Fo, i, dV * nrm);
        wall_F[w] -= dV * nrm;
      }
  }

  /* wall::harmonic::eval_free::iter, cg/src/cg/wall/harmonic/eval_free.cpp:15-31:
   * the sign is the reference's (the opposite of the solid wall's) */
  void eval_hw_free(double *Fo) {
    for (int i = 0; i < n; ++i) {
      if (is_conn[i]) continue;
      for (int w = 0; w < p.n_walls; ++w) {
        if (p.wall_kind[w] != 1) continue;
        v3 nrm = plane_n(w), org = plane_o(w);
        double sd = dot(nrm, r[i] - org), dist = std::abs(sd);
        if (dist >= p.wall_min_dist) continue;
        double dV = p.wall_depth * std::pow(dist, -10.0) * sgn(sd);
        v3 force = dV * nrm;
        sub_F(Fo, i, force);
        wall_F[w] += force;
      }
    }
  }

  /* wall::harmonic::eval_connected::iter,
   * cg/src/cg/wall/harmonic/eval_connected.cpp:17-29 (the force sign is the
   * reference's: F[res] -= dV_dr * u with u pointing from the residue to the bead) */
  void eval_hw_conn(double *Fo, double &Vo) {
    for (auto const &c : hw_conns) {
      v3 bead = c.off + plane_o(c.wall);
      v3 r12 = bead - r[c.res];
      double r12_n = norm(r12), r12_rn = 1.0 / r12_n;
      v3 u = r12 * r12_rn;
      double V_, dV;
      harmonic(p.wall_hh1, 0.0, c.nat, r12_n, V_, dV);
      Vo += V_;
      sub_F(Fo, c.res, dV * u);
      wall_F[c.wall] += dV * u;
    }
  }

  /* wall::lj::sift_free::iter, cg/src/cg/wall/lj/sift_free.cpp:15-37 */
  void eval_ljw_free(double *Fo, double &Vo) {
    for (int i = 0; i < n; ++i) {
      if (is_conn[i]) continue;
      for (int w = 0; w < p.n_walls; ++w) {
        if (p.wall_kind[w] != 2) continue;
        v3 nrm = plane_n(w), org = plane_o(w);
        double sd = dot(nrm, r[i] - org), dist = std::abs(sd);
        if (dist > p.wall_min_dist) continue;
        double V_, dV;
        shifted_lj(p.wall_lj_depth, p.wall_min_dist, 1.0 / dist, V_, dV);
        Vo += V_;
        v3 F_ = (dV * nrm) * sgn(sd);
        sub_F(Fo, i, F_);
        wall_F[w] += F_;
        if (ljw_num_conns[w] < p.wall_limit[w]) ljw_cands.push_back(ljw_cand_t{i, w, dist});
      }
    }
  }

  /* wall::lj::eval_connected::iter, cg/src/cg/wall/lj/eval_connected.cpp:15-41
   * (force sign as the reference's) */
  void eval_ljw_conn(double *Fo, double &Vo) {
    for (int k = 0; k < (int)ljw_conns.size(); ++k) {
      auto &c = ljw_conns[k];
      if (c.vacant) continue;
      v3 bead = plane_o(c.wall) + c.off;
      v3 r12 = bead - r[c.res];
      double r12_n = norm(r12), r12_rn = 1.0 / r12_n;
      v3 u = r12 * r12_rn;
      double V_, dV;
      lj(p.wall_lj_depth, p.wall_min_dist, r12_rn, V_, dV);
      Vo += c.sat * V_;
      v3 F_ = (c.sat * dV) * u;
      sub_F(Fo, c.res, F_);
      wall_F[c.wall] += F_;
      if (r12_n < p.ljw_breaking_dist) {
        c.sat = std::min(c.sat + p.ljw_saturation_diff, 1.0);
      } else {
        c.sat = std::max(c.sat - p.ljw_saturation_diff, 0.0);
        if (c.sat == 0.0) ljw_removed.push_back(k);
      }
    }
  }

  /* wall::lj::process_candidates, cg/src/cg/wall/lj/process_candidates.cpp:5-31:
   * releases first, then the candidates nearest first; a new connection does
   * not mark its residue connected (as in the reference). Candidates with
   * equal distances keep their (residue, wall) order. The reference keeps its
   * candidates for ever; here each call consumes its own. */
  void ljw_process() {
    for (int k : ljw_removed) {
      auto &c = ljw_conns[k];
      is_conn[c.res] = 0;
      --ljw_num_conns[c.wall];
      c.vacant = true;
    }
    ljw_removed.clear();
    std::stable_sort(ljw_cands.begin(), ljw_cands.end(),
                     [](ljw_cand_t const &a, ljw_cand_t const &b) { return a.dist < b.dist; });
    for (auto const &cd : ljw_cands) {
      int w = cd.wall;
      if (is_conn[cd.res] || ljw_num_conns[w] >= p.wall_limit[w]) continue;
      v3 nrm = plane_n(w), org = plane_o(w), rr = r[cd.res];
      v3 proj = rr - dot(nrm, rr - org) * nrm;
      ljw_conns.push_back(ljw_conn_t{cd.res, w, proj - org, 0.0, false});
      ++ljw_num_conns[w];
    }
    ljw_cands.clear();
  }

  /* state::setup_walls' harmonic connections (state.cpp:653-670): after each
   * axis with harmonic walls, every harmonic plane so far takes up to its
   * limit of the unconnected residues nearest to it */
  void setup_harmonic() {
    for (int w = 0; w < p.n_walls; ++w) {
      if (p.wall_kind[w] != 1) continue;
      if (p.wall_side[w] >= 0 && p.wall_side[w] % 2 == 0) continue; /* the axis' second plane ends it */
      for (int hw = 0; hw <= w; ++hw) {
        if (p.wall_kind[hw] != 1) continue;
        v3 nrm = plane_n(hw), org = plane_o(hw);
        std::vector<std::pair<double, int>> dist(n);
        for (int i = 0; i < n; ++i) dist[i] = std::make_pair(std::abs(dot(nrm, r[i] - org)), i);
        std::sort(dist.begin(), dist.end());
        for (int idx = 0, conn = 0; conn < p.wall_limit[hw] && idx < n; ++idx) {
          int res = dist[idx].second;
          if (is_conn[res]) continue;
          v3 proj = r[res] - dot(nrm, r[res] - org) * nrm;
          hw_conns.push_back(hw_conn_t{hw, res, proj - org, 0.0});
          is_conn[res] = 1;
          ++conn;
        }
      }
    }
  }

  /* state::reinit_wall_values, state.cpp:708-729 */
  void reinit_walls() {
    v3 center = (box_max + box_min) / 2.0, ext = box_max - box_min;
    for (int side = 0; side < 6; ++side) {
      int w = side_wall[side];
      if (w < 0) continue;
      v3 prev = plane_o(w), nrm = plane_n(w);
      v3 d = side % 2 == 1 ? box_max - center : box_min - center;
      v3 next = center + nrm * dot(d, nrm);
      p.wall_origin[w][0] = next.x, p.wall_origin[w][1] = next.y, p.wall_origin[w][2] = next.z;
      shift[w] = next - prev;
    }
    pbc.set(v3{bx.pbc[0] ? ext.x : 0.0, bx.pbc[1] ? ext.y : 0.0, bx.pbc[2] ? ext.z : 0.0});
  }
  /* state::adjust_wall_pos, state.cpp:702-706 */
  void adjust_walls(v3 size_change, v3 translation) {
    box_min += -size_change + translation;
    box_max += size_change + translation;
    reinit_walls();
  }
  void shear_walls(double sh) {
    for (int side : {4, 5})
      if (side_wall[side] >= 0) p.wall_normal[side_wall[side]][0] += sh;
  }

  /* The protocol phases of thread_main.cpp:144-330 up to the next step that
   * advances (a step() that only changes the phase takes no time). */
  void box_phase() {
    const double dt_ = p.dt;
    const double num_res = (double)n;
    for (;;) {
      switch (phase) {
      case 1: { /* squeezing_step, :144-181 */
        if (!bx.squeezing) {
          since = t;
          phase = 3;
          continue;
        }
        v3 ext = box_max - box_min;
        double cur_vol = ext.x * ext.y * ext.z;
        double target_vol = num_res / bx.target_density;
        if (cur_vol <= target_vol) {
          until = std::min(t + bx.rest_period, bx.total_time);
          phase = 2;
          continue;
        }
        double vel;
        if (cur_vol > 2.0 * target_vol) {
          double accel_time = bx.accel_dist / bx.vel_above_2V;
          double time_frac = std::min((t - since) / accel_time, 1.0);
          vel = time_frac * bx.vel_above_2V;
        } else {
          double accel_time = bx.accel_dist / bx.target_vel;
          double time_frac = std::min((t - since) / accel_time, 1.0);
          vel = time_frac * bx.target_vel;
        }
        double sh = 0.5 * vel * dt_;
        adjust_walls(v3{-sh, -sh, -sh}, v3{0, 0, 0});
        return;
      }
      case 2: /* rest_after_squeezing_step, :183-191 */
        if (t < until) return;
        phase = 3;
        continue;
      case 3: { /* find_force_min_step, :193-222 */
        if (!bx.force_min) {
          phase = 9;
          continue;
        }
        if (avg_z.has && std::abs(avg_z.avg) < 5e-2 + 5e-5 * num_res) {
          until = t + bx.rest_period;
          phase = 4;
          continue;
        }
        double accel_time = bx.accel_dist / bx.target_vel;
        double time_vel_frac = std::min((t - since) / accel_time, 1.0);
        double fvf = (avg_z.has ? avg_z.avg : 0.0) / bx.force_for_max_vel;
        fvf = std::copysign(std::min(std::abs(fvf), 1.0), fvf);
        double vel = bx.target_vel * time_vel_frac * fvf;
        double sh = 0.5 * vel * dt_;
        adjust_walls(v3{0, 0, sh}, v3{0, 0, 0});
        return;
      }
      case 4: /* rest_after_force_min_step, :224-243 */
        if (t < until) return;
        if (!bx.amp_relative) {
          amplitude = bx.amp_abs;
        } else {
          shear = bx.osc_shear != 0;
          v3 ext = box_max - box_min;
          double base_val = shear ? ext.x : ext.z;
          amplitude = bx.amp_rel * base_val;
        }
        since = t;
        phase = 5;
        continue;
      case 5: { /* max_amplitude_step, :245-276 */
        if (!bx.oscillations) {
          phase = 9;
          continue;
        }
        double accel_time = bx.accel_dist / bx.target_vel;
        double time_frac = std::min((t - since) / accel_time, 1.0);
        double vel = time_frac * bx.target_vel;
        double sh = vel * dt_;
        if (displacement > 0.5 * amplitude) {
          until = t + bx.rest_period;
          phase = 6;
          continue;
        }
        displacement += sh;
        if (shear)
          shear_walls(sh);
        else
          adjust_walls(v3{0, 0, -sh}, v3{0, 0, 0});
        return;
      }
      case 6: /* rest_after_max_amp_step, :278-293 */
        if (t < until) return;
        since = t;
        omega = bx.angular_freq;
        until = t + (M_PI * 2 * bx.num_of_cycles + M_PI_2) / omega;
        phase = 7;
        continue;
      case 7: { /* oscillations_step, :295-320 */
        if (t >= until) {
          until = t + bx.rest_period;
          phase = 8;
          continue;
        }
        double vel = 0.5 * amplitude * omega * std::sin(omega * (t - since));
        double sh = vel * dt_;
        if (shear)
          shear_walls(sh);
        else
          adjust_walls(v3{0, 0, sh}, v3{0, 0, 0});
        return;
      }
      case 8: /* rest_after_oscillations_step, :322-330 */
        if (t < until) return;
        phase = 9;
        continue;
      default: /* NONE, FREEFORM */
        return;
      }
    }
  }

  /* wall::log_forces, cg/src/cg/wall/log_forces.cpp:4-17, over the
   * reference's wall pointer list: the six sides, then once more the -axis
   * plane of every LJ axis (setup_walls pushes it twice, state.cpp:680-681) */
  void log_forces() {
    auto one = [&](int w) {
      avg_F[w].add(t, wall_F[w]);
      work[w] += dot(wall_F[w], shift[w]);
    };
    for (int side = 0; side < 6; ++side)
      if (side_wall[side] >= 0) one(side_wall[side]);
    for (int side = 0; side < 6; side += 2) {
      int w = side_wall[side];
      if (w >= 0 && p.wall_kind[w] == 2) one(w);
    }
    int wn = side_wall[4], wp = side_wall[5];
    double z_force = dot(wall_F[wn], plane_n(wn)) + dot(wall_F[wp], plane_n(wp));
    avg_z.add(t, z_force);
  }

  bool has_kind(int kind) const {
    for (int w = 0; w < p.n_walls; ++w)
      if (p.wall_kind[w] == kind) return true;
    return false;
  }

  bool term_enabled(int term) const {
    switch (term) {
    case ORA_CHIR: return p.chir_on;
    case ORA_TETHER: return p.tether_on;
    case ORA_LREP: return p.lrep_on;
    case ORA_NAT_CONT: return nat_cont_on;
    case ORA_PAULI: return pauli_on;
    case ORA_CONST_DH: return p.dh_on && p.dh_variant == 0;
    case ORA_REL_DH: return p.dh_on && p.dh_variant == 1;
    case ORA_QA_CAND: return p.qa_on;
    case ORA_QA_CONT: return p.qa_on;
    case ORA_PID: return p.pid_on;
    case ORA_NAT_ANG: return p.nat_ang_on;
    case ORA_HEUR_ANG: return p.heur_ang_on;
    case ORA_HEUR_DIH: return p.heur_dih_on;
    case ORA_CND: return p.nat_dih_on && p.nat_dih_variant == 0;
    case ORA_SND: return p.nat_dih_on && p.nat_dih_variant == 1;
    case ORA_VEL_AFM: return p.afm_on;
    case ORA_FORCE_AFM: return p.afm_on;
    case ORA_SOLID_WALL: return has_kind(0);
    case ORA_HW_FREE: return p.hw_on && has_kind(1);
    case ORA_HW_CONN: return p.hw_on && has_kind(1);
    case ORA_LJW_FREE: return p.ljw_on && has_kind(2);
    case ORA_LJW_CONN: return p.ljw_on && has_kind(2);
    }
    return false;
  }

  void run_term(int term, double *Fo, double &Vo) {
    switch (term) {
    case ORA_CHIR: eval_chir(Fo, Vo); break;
    case ORA_TETHER: eval_tether(Fo, Vo); break;
    case ORA_LREP: eval_lrep(Fo, Vo); break;
    case ORA_NAT_CONT: eval_nat_cont(Fo, Vo); break;
    case ORA_PAULI: eval_pauli(Fo, Vo); break;
    case ORA_CONST_DH: eval_dh(Fo, Vo, 0); break;
    case ORA_REL_DH: eval_dh(Fo, Vo, 1); break;
    case ORA_QA_CAND: eval_qa_cand(Fo, Vo); break;
    case ORA_QA_CONT: eval_qa_cont(Fo, Vo); break;
    case ORA_PID: eval_pid(Fo, Vo); break;
    case ORA_NAT_ANG: eval_nat_ang(Fo, Vo); break;
    case ORA_HEUR_ANG: eval_heur_ang(Fo, Vo); break;
    case ORA_HEUR_DIH: eval_heur_dih(Fo, Vo); break;
    case ORA_CND: eval_nat_dih(Fo, Vo, 0); break;
    case ORA_SND: eval_nat_dih(Fo, Vo, 1); break;
    case ORA_VEL_AFM: eval_vel_afm(Fo, Vo); break;
    case ORA_FORCE_AFM: eval_force_afm(Fo); break;
    case ORA_SOLID_WALL: eval_solid_walls(Fo); break;
    case ORA_HW_FREE: eval_hw_free(Fo); break;
    case ORA_HW_CONN: eval_hw_conn(Fo, Vo); break;
    case ORA_LJW_FREE: eval_ljw_free(Fo, Vo); break;
    case ORA_LJW_CONN: eval_ljw_conn(Fo, Vo); break;
    }
  }

  /* thread::eval_forces (thread_main.cpp:435-502) with one thread: V and F
   * accumulate in the term order of thread_setup.cpp:27-48. */
  void eval_forces() {
    std::vector<double> Fb(3 * (size_t)n, 0.0);
    V = 0.0;
    for (auto &f : wall_F) f = v3{0, 0, 0}; /* dynamics::reset, dynamics.cpp:11-19 */
    for (int k = 0; k < ORA_NUM_TERMS; ++k) {
      term_V[k] = 0.0;
      if (!term_enabled(k)) continue;
      double before = V;
      run_term(k, Fb.data(), V);
      term_V[k] = V - before;
    }
    for (int i = 0; i < n; ++i) F[i] = v3{Fb[3 * i], Fb[3 * i + 1], Fb[3 * i + 2]};
  }

  /* nl::verify::omp_async, cg/src/cg/nl/verify.cpp:3-34 (one thread). */
  void nl_verify() {
    if (verify_first) {
      verify_first = false;
      nl_invalid = true;
      return;
    }
    double max_r_disp = 0.0;
    for (int i = 0; i < n; ++i) {
      v3 dr = pbc.wrap(r[i], orig_r_nl[i]);
      max_r_disp = std::max(max_r_disp, norm(dr));
    }
    double box_disp = l1(pbc.cell - nl_orig_pbc.cell);
    double thread_disp = 2.0 * max_r_disp + box_disp;
    total_disp = std::max(total_disp, thread_disp);
    nl_invalid = total_disp > nl_orig_pad;
    total_disp = 0.0;
  }

  /* nl::legacy_update::omp_async, cg/src/cg/nl/legacy_update.cpp:11-55 */
  void nl_rebuild() {
    double req_r = nl_eff_cutoff + p.nl_pad;
    nl.clear();
    for (int i1 = 0; i1 < n; ++i1) {
      for (int i2 = i1 + 1; i2 < n; ++i2) {
        int diff = seq[i1] > seq[i2] ? seq[i1] - seq[i2] : seq[i2] - seq[i1];
        if (chain[i1] == chain[i2] && diff < 3) continue;
        double d = norm(pbc.wrap(r[i1], r[i2]));
        if (d < req_r) nl.push_back(nl_pair{i1, i2, d, false});
      }
    }
    /* generated in (i1, i2) lexicographic order already == std::sort result */
    nl_orig_pad = p.nl_pad;
    orig_r_nl = r;
    nl_orig_pbc = pbc;
    nl_invalid = false;
    ++nl_rebuilds;
  }

  /* nat_cont::update_contacts, cg/src/cg/nat_cont/update_contacts.cpp:7-37 */
  void update_nat_contacts() {
    cur_nc.clear();
    size_t it = 0;
    for (auto &pr : nl) {
      if (pr.taken) continue;
      while (it < all_nc.size()) {
        auto const &c = all_nc[it];
        if (c.i1 < pr.i1 || (c.i1 == pr.i1 && c.i2 < pr.i2)) {
          ++it;
        } else {
          if (c.i1 == pr.i1 && c.i2 == pr.i2) {
            double d = norm(pbc.wrap(r[pr.i1], r[pr.i2]));
            if (d < p.nl_cutoff + nl_orig_pad) {
              cur_nc.push_back(c);
              pr.taken = true;
            }
          }
          break;
        }
      }
    }
  }
  /* qa::update_free_pairs, cg/src/cg/qa/update_free_pairs.cpp:4-25 */
  void update_qa_pairs() {
    qa_free_pairs.clear();
    for (auto &pr : nl) {
      if (pr.taken) continue;
      if (!p.qa_include4 && chain[pr.i1] == chain[pr.i2] &&
          std::abs(seq[pr.i1] - seq[pr.i2]) == 4)
        continue;
      double d = norm(pbc.wrap(r[pr.i1], r[pr.i2]));
      if (d < p.nl_cutoff + nl_orig_pad) {
        qa_free_pairs.push_back(qa_free{pr.i1, pr.i2, pr.orig_dist, false});
        pr.taken = true;
      }
    }
  }
  /* pid::update_bundles, cg/src/cg/pid/update_bundles.cpp:4-34 */
  void update_pid_bundles() {
    pid_bundles.clear();
    for (auto &pr : nl) {
      if (pr.taken) continue;
      if (!p.pid_include4 && chain[pr.i1] == chain[pr.i2] &&
          std::abs(seq[pr.i1] - seq[pr.i2]) == 4)
        continue;
      double d = norm(pbc.wrap(r[pr.i1], r[pr.i2]));
      if (d < p.nl_cutoff + nl_orig_pad) {
        int ty = atype[pr.i1] * 21 + atype[pr.i2];
        pid_bundles.push_back(pid_bundle{pr.i1, pr.i2, pr.orig_dist, ty});
      }
    }
  }
  /* dh::update_pairs, cg/src/cg/dh/update_pairs.cpp:4-25 */
  void update_dh_pairs() {
    dh_pairs.clear();
    double min_norm_inv = 1.0 / (p.dh_cutoff + nl_orig_pad);
    for (auto &pr : nl) {
      if (pr.taken) continue;
      double q = p.charge[atype[pr.i1]] * p.charge[atype[pr.i2]];
      if (q == 0) continue;
      if (norm_inv(pbc.wrap(r[pr.i1], r[pr.i2])) > min_norm_inv)
        dh_pairs.push_back(dh_pair{pr.i1, pr.i2, q});
    }
  }
  /* pauli::update_pairs, cg/src/cg/pauli/update_pairs.cpp:4-20 */
  void update_pauli_pairs() {
    pauli_pairs.clear();
    for (auto &pr : nl) {
      if (pr.taken) continue;
      double d = norm(pbc.wrap(r[pr.i1], r[pr.i2]));
      if (d < p.rep_cutoff + nl_orig_pad) {
        pauli_pairs.emplace_back(pr.i1, pr.i2);
        pr.taken = true;
      }
    }
  }

  /* thread::fix_nl_async, cg/src/cg/simul/thread_main.cpp:403-427 */
  void fix_nl() {
    nl_rebuild();
    if (nat_cont_on) update_nat_contacts();
    if (p.qa_on) update_qa_pairs();
    if (p.pid_on) update_pid_bundles();
    if (p.dh_on) update_dh_pairs();
    if (pauli_on) update_pauli_pairs();
  }

  /* qa::finish_processing, cg/src/cg/qa/finish_processing.cpp:5-74; ties of
   * the distance sort are broken by candidate order (free-pair index). */
  void qa_finish() {
    for (int ridx : qa_removed) {
      qa_cont &c = qa_contacts[ridx];
      sync[c.i1] = add(sync[c.i1], c.d1);
      sync[c.i2] = add(sync[c.i2], c.d2);
      qa_free_pairs.push_back(qa_free{c.i1, c.i2, c.orig_dist, false});
      c.vacant = true;
      --num_qa_contacts;
    }
    qa_removed.clear();
    std::stable_sort(qa_cands.begin(), qa_cands.end(),
                     [](qa_cand const &a, qa_cand const &b) {
                       return a.dist < b.dist;
                     });
    for (auto const &cd : qa_cands) {
      sync_t ns1 = sub(sync[cd.i1], cd.d1);
      if (!ns1.valid()) continue;
      sync_t ns2 = sub(sync[cd.i2], cd.d2);
      if (!ns2.valid()) continue;
      qa_contacts.push_back(
          qa_cont{cd.i1, cd.i2, cd.orig_dist, cd.type, 0.0, cd.d1, cd.d2, false});
      ++num_qa_contacts;
      sync[cd.i1] = ns1;
      sync[cd.i2] = ns2;
      qa_free_pairs[cd.free_idx].vacant = true;
    }
    qa_cands.clear();
    /* vect::set::compress (vect/ind/set.h:62-71) moves items without their
     * vacancy flags. */
    if (2 * num_qa_contacts < (int)qa_contacts.size()) {
      int slot = 0, items = 0;
      for (size_t cur = 0; cur < qa_contacts.size(); ++cur) {
        if (!qa_contacts[cur].vacant) {
          bool keep_vac = qa_contacts[slot].vacant;
          qa_contacts[slot] = qa_contacts[cur];
          qa_contacts[slot].vacant = keep_vac;
          ++slot;
          ++items;
        }
      }
      qa_contacts.resize(items);
    }
  }

  /* lang::step::omp_async, cg/src/cg/langevin/step.cpp:23-77 */
  void lang_step() {
    static const double f02 = 3.0f / 16.0f, f12 = 251.0f / 360.0f,
                        f32 = 11.0f / 18.0f, f42 = 1.0f / 6.0f,
                        f52 = 1.0f / 60.0f;
    if (t == 0) {
      for (int i = 0; i < n; ++i) y[2][i] = F[i] * deltsq;
    }
    for (int i = 0; i < n; ++i) noise[i].x = gen.normal();
    for (int i = 0; i < n; ++i) noise[i].y = gen.normal();
    for (int i = 0; i < n; ++i) noise[i].z = gen.normal();
    for (int i = 0; i < n; ++i) {
      int aa = atype[i];
      double gam2 = gamma2 * (1.0 / p.mass[aa]);
      double con2 = const2 * std::sqrt(1.0 / p.mass[aa]) * noise_scale;
      v3 f = F[i] - gam2 * y[1][i];
      y[1][i] = y[1][i] + con2 * noise[i];
      v3 err = y[2][i] - deltsq * f;
      y[0][i] -= f02 * err;
      y[1][i] -= f12 * err;
      y[2][i] -= 1.0 * err;
      y[3][i] -= f32 * err;
      y[4][i] -= f42 * err;
      y[5][i] -= f52 * err;
      y[0][i] += y[1][i] + y[2][i] + y[3][i] + y[4][i] + y[5][i];
      y[1][i] += 2.0 * y[2][i] + 3.0 * y[3][i] + 4.0 * y[4][i] + 5.0 * y[5][i];
      y[2][i] += 3.0 * y[3][i] + 6.0 * y[4][i] + 10.0 * y[5][i];
      y[3][i] += 4.0 * y[4][i] + 10.0 * y[5][i];
      y[4][i] += 5.0 * y[5][i];
      r[i] = y[0][i];
      v[i] = y[1][i] * dt_inv;
    }
    true_t += dt;
    t = true_t;
    ++step_idx;
  }

  /* thread::advance_by_step (thread_main.cpp:373-536) */
  void step() {
    if (p.qa_on) prepare_nh();
    if (nl_required) nl_verify();
    if (nl_invalid) fix_nl();
    eval_forces();
    if (p.qa_on) qa_finish();
    if (p.ljw_on && has_kind(2)) ljw_process();
    if (log_on) log_forces();
    if (p.lang_on) lang_step();
  }
};

namespace {
/* model::morph_into_saw, cg/src/cg/input/model.cpp:108-234 */
void morph_saw(ran2_t &gen, const ora_saw &sp, int n, const int *chain,
               std::vector<v3> &pos) {
  double min_dist_sq = std::pow(sp.intersection, 2.0);
  double bond = sp.bond;
  pbc_t saw_pbc;
  v3 bmin{0, 0, 0}, bmax{0, 0, 0};
  if (sp.with_box) {
    double a = sp.box_size;
    bmin = v3{-a / 2.0, -a / 2.0, -a / 2.0};
    bmax = -bmin;
    if (sp.pbc) saw_pbc.set(bmax - bmin);
  }
  /* chains in order of chain index; residues of a chain are contiguous */
  int start = 0;
  while (start < n) {
    int end = start;
    while (end < n && chain[end] == chain[start]) ++end;
    int cn = end - start;
    for (int kter = 0; kter < 9000; ++kter) {
      std::vector<double> phi(cn), theta(cn);
      double pi = std::acos(-1.0);
      phi[0] = pi / 2.0;
      theta[0] = 0.0;
      double theta1 = gen.uniform() * pi / 3.0; /* drawn even for cn == 1 */
      if (cn > 1) {
        phi[1] = 0.0;
        theta[1] = theta1;
      }
      for (int i = 2; i < cn - 1; ++i) {
        phi[i] = (2.0 * gen.uniform() - 1.0) * pi;
        theta[i] = gen.uniform() * pi / 3.0;
      }
      std::vector<double> T(9 * (size_t)cn);
      for (int i = 0; i < cn - 1; ++i) {
        double ct = std::cos(theta[i]), st = std::sin(theta[i]),
               cp = std::cos(phi[i]), spp = std::sin(phi[i]);
        double *M = &T[9 * i];
        M[0] = ct;
        M[1] = st;
        M[2] = 0.0;
        M[3] = st * cp;
        M[4] = -ct * cp;
        M[5] = spp;
        M[6] = st * spp;
        M[7] = -ct * spp;
        M[8] = -cp;
      }
      double theta0 = std::acos(1.0 - 2.0 * gen.uniform());
      double phi0 = 2.0 * pi * gen.uniform();
      std::vector<v3> R(cn);
      for (int i = 0; i < cn - 1; ++i) {
        v3 rr{bond * std::sin(theta0) * std::cos(phi0),
              bond * std::sin(theta0) * std::sin(phi0), bond * std::cos(theta0)};
        for (int j = i; j >= 0; --j) {
          const double *M = &T[9 * j];
          /* Eigen 3x3 * vector: row . vector, summed left to right */
          R[i] = v3{M[0] * rr.x + M[1] * rr.y + M[2] * rr.z,
                    M[3] * rr.x + M[4] * rr.y + M[5] * rr.z,
                    M[6] * rr.x + M[7] * rr.y + M[8] * rr.z};
          rr = R[i];
        }
      }
      v3 ran;
      ran.x = (bmax.x - bmin.x) * gen.uniform() + bmin.x;
      ran.y = (bmax.y - bmin.y) * gen.uniform() + bmin.y;
      ran.z = (bmax.z - bmin.z) * gen.uniform() + bmin.z;
      for (int i = 0; i < cn; ++i) {
        v3 rr = ran;
        for (int j = 0; j < i; ++j) rr += R[j];
        pos[start + i] = rr;
      }
      bool ok = true;
      for (int i = 0; i < cn - 3 && ok; ++i)
        for (int j = i + 3; j < cn && ok; ++j)
          if (norm_sq(saw_pbc.wrap(pos[start + i], pos[start + j])) < min_dist_sq)
            ok = false;
      if (ok) break;
    }
    start = end;
  }
}
} // namespace

extern "C" {

ora_sys *ora_create(const ora_params *p, const ora_model *m, int32_t seed,
                    const ora_saw *saw, double *pos_out) {
  ora_sys *s = new ora_sys();
  s->p = *p;
  int n = s->n = m->n;
  s->gen = ran2_t(seed);

  std::vector<v3> pos(n), orig(n);
  for (int i = 0; i < n; ++i) {
    pos[i] = mk(m->pos[3 * i], m->pos[3 * i + 1], m->pos[3 * i + 2]);
    orig[i] = mk(m->orig_pos[3 * i], m->orig_pos[3 * i + 1], m->orig_pos[3 * i + 2]);
  }
  if (saw && saw->perform) {
    morph_saw(s->gen, *saw, n, m->chain_idx, pos);
    s->saw_draws = s->gen.draws;
    /* for sequence-file inputs orig_model follows the SAW (state.cpp:134-135) */
    if (saw->update_orig) orig = pos;
    if (pos_out)
      for (int i = 0; i < n; ++i) {
        pos_out[3 * i] = pos[i].x;
        pos_out[3 * i + 1] = pos[i].y;
        pos_out[3 * i + 2] = pos[i].z;
      }
  }
  s->r = pos;
  s->atype.assign(m->atype, m->atype + n);
  s->chain.assign(m->chain_idx, m->chain_idx + n);
  s->seq.assign(m->seq_idx, m->seq_idx + n);
  s->prev.assign(n, -1);
  s->next.assign(n, -1);
  for (int i = 0; i < n; ++i) {
    if (i > 0 && s->chain[i - 1] == s->chain[i]) s->prev[i] = i - 1;
    if (i + 1 < n && s->chain[i + 1] == s->chain[i]) s->next[i] = i + 1;
  }
  s->F.assign(n, v3{0, 0, 0});
  s->v.assign(n, v3{0, 0, 0});
  s->noise.assign(n, v3{0, 0, 0});
  s->n_vec.assign(n, v3{0, 0, 0});
  s->h_vec.assign(n, v3{0, 0, 0});
  s->orig_r_nl.assign(n, v3{0, 0, 0});
  s->pbc.set(mk(m->cell[0], m->cell[1], m->cell[2]));

  /* state::setup_langevin, state.cpp:270-335 */
  for (int k = 0; k < 6; ++k) s->y[k].assign(n, v3{0, 0, 0});
  s->y[0] = s->r;
  s->t = s->true_t = 0.0;
  if (p->lang_on) {
    v3 sum_vel{0, 0, 0};
    for (int i = 0; i < n; ++i) {
      double xx = 2.0 * (s->gen.uniform() - 0.5);
      double yy = 2.0 * (s->gen.uniform() - 0.5);
      double zz = 2.0 * (s->gen.uniform() - 0.5);
      double xyz = 1.0 / std::sqrt(xx * xx + yy * yy + zz * zz);
      s->y[1][i] = mk(xx, yy, zz) * xyz;
      sum_vel += s->y[1][i];
    }
    double x = 0.0;
    for (int i = 0; i < n; ++i) {
      s->y[1][i] -= sum_vel / (double)n;
      v3 vv = s->y[1][i];
      x = x + vv.x * vv.x + vv.y * vv.y + vv.z * vv.z;
    }
    double dt = p->dt, delsq = dt * dt;
    double aheat = delsq * 3.0 * n * p->temperature;
    double heat = std::sqrt(aheat / x);
    for (int i = 0; i < n; ++i) s->y[1][i] = s->y[1][i] * heat;

    /* lang::step::set_params, langevin/step.cpp:7-17 */
    s->dt = p->dt;
    s->dt_inv = 1.0 / s->dt;
    s->gamma2 = p->gamma * s->dt_inv;
    s->const2 = 2.0 * 1.0 * p->temperature * p->gamma * s->dt;
    s->const2 = std::sqrt(s->const2) * s->dt;
    s->deltsq = (s->dt * s->dt) / 2.0;
  }

  /* state::setup_nl, state.cpp:345-352 and thread::setup_dh cutoff max,
   * thread_setup.cpp:368 */
  s->nl_invalid = true;
  s->verify_first = true;
  s->nl_eff_cutoff = std::max(p->nl_cutoff, p->dh_cutoff);

  /* state::setup_local_rep (state.cpp:354-364): (i1, i3) of every angle */
  if (p->lrep_on)
    for (int a = 0; a < m->n_ang; ++a)
      s->lrep.emplace_back(m->ang[3 * a], m->ang[3 * a + 2]);

  /* state::setup_chir (state.cpp:366-387): natives from model positions */
  if (p->chir_on)
    for (int d = 0; d < m->n_dih; ++d) {
      int i1 = m->dih[4 * d], i2 = m->dih[4 * d + 1], i3 = m->dih[4 * d + 2],
          i4 = m->dih[4 * d + 3];
      v3 n12 = pos[i2] - pos[i1], n23 = pos[i3] - pos[i2], n34 = pos[i4] - pos[i3];
      double ni = norm_inv(n23);
      double nat_factor = ni * (ni * (ni * 1.0));
      double nat_chir = dot(n12, cross(n23, n34)) * nat_factor;
      /* The reference emplaces (.., nat_factor, nat_chir) into a struct whose
       * fields are declared (.., nat_chir, nat_factor)
       * (state.cpp:384, cg/include/cg/chir/chiral_quad.h:12-13), so the
       * evaluated quad carries the two values swapped. */
      s->chir.push_back(ora_sys::chir_t{i1, i2, i3, i4, nat_chir, nat_factor});
    }

  /* state::setup_tether (state.cpp:389-402): length from orig positions;
   * the PBC of the first trajectory is not set yet at this point. */
  if (p->tether_on)
    for (int k = 0; k < m->n_teth; ++k) {
      int i1 = m->teth[2 * k], i2 = m->teth[2 * k + 1];
      double nat = norm(orig[i2] - orig[i1]);
      s->tethers.push_back(ora_sys::teth_t{i1, i2, nat});
    }

  /* aa_heur_pair, cg/src/cg/angles/aa_heur_pair.cpp:11-27 */
  auto heur_type = [&](int a2, int a3) {
    int o1 = a2 == 7 ? 0 : (a2 == 14 ? 1 : 2);
    int o2 = a3 == 7 ? 0 : (a3 == 14 ? 1 : 2);
    return 3 * o1 + o2;
  };
  for (int a = 0; a < m->n_ang; ++a) {
    int i1 = m->ang[3 * a], i2 = m->ang[3 * a + 1], i3 = m->ang[3 * a + 2];
    if (m->ang_has[a]) {
      if (p->nat_ang_on)
        s->nat_ang.push_back(ora_sys::ang_t{i1, i2, i3, m->ang_theta[a], 0});
    } else if (p->heur_ang_on) {
      s->heur_ang.push_back(
          ora_sys::ang_t{i1, i2, i3, 0.0, heur_type(s->atype[i2], s->atype[i3])});
    }
  }
  for (int d = 0; d < m->n_dih; ++d) {
    int i1 = m->dih[4 * d], i2 = m->dih[4 * d + 1], i3 = m->dih[4 * d + 2],
        i4 = m->dih[4 * d + 3];
    if (m->dih_has[d]) {
      if (p->nat_dih_on)
        s->nat_dih.push_back(ora_sys::dih_t{i1, i2, i3, i4, m->dih_phi[d], 0});
    } else if (p->heur_dih_on) {
      s->heur_dih.push_back(ora_sys::dih_t{
          i1, i2, i3, i4, 0.0, heur_type(s->atype[i2], s->atype[i3])});
    }
  }

  /* state::setup_pauli (state.cpp:472-478) */
  s->pauli_on = p->pauli_on && !p->qa_on;
  if (s->pauli_on) s->nl_required = true;

  /* state::setup_nat_cont (state.cpp:480-502); contacts sorted as in
   * state::morph_model (state.cpp:116-128) */
  s->nat_cont_on = m->n_cont > 0 && p->nat_cont_on;
  if (s->nat_cont_on) {
    /* The reference sorts by (min, max) of seq_idx, which equals the
     * residue index for one chain; with several chains its ties have no
     * specified order, so contacts are ordered by residue indices here
     * (stable), which makes the merge-join of update_contacts a membership
     * test in every case. */
    std::vector<int> order(m->n_cont);
    for (int k = 0; k < m->n_cont; ++k) order[k] = k;
    auto key = [&](int k) {
      int a = m->cont[2 * k], b = m->cont[2 * k + 1];
      return a < b ? std::make_pair(a, b) : std::make_pair(b, a);
    };
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return key(a) < key(b); });
    for (int idx = 0; idx < m->n_cont; ++idx) {
      int k = order[idx];
      int i1 = m->cont[2 * k], i2 = m->cont[2 * k + 1];
      if (i1 > i2) std::swap(i1, i2);
      double nat = m->cont_len[k];
      bool active = norm(pos[i2] - pos[i1]) < p->counting_factor * C216_INV * nat;
      s->all_nc.push_back(
          nat_cont_t{i1, i2, nat, m->cont_type[k], active, -1.0, idx});
    }
  }

  /* state::setup_dh (state.cpp:504-510), V factor thread_setup.cpp:370-391 */
  if (p->dh_on) s->nl_required = true;
  s->V_factor = 1.0 / (4.0 * M_PI * p->dh_perm);
  s->screen_inv = 1.0 / p->dh_screen_dist;

  /* state::setup_qa (state.cpp:512-538), thread::setup_qa
   * (thread_setup.cpp:394-587) */
  if (p->qa_on) {
    s->sync.resize(n);
    for (int i = 0; i < n; ++i)
      for (int k = 0; k < 4; ++k)
        s->sync[i].v[k] = (int8_t)p->base_sync[s->atype[i]][k];
    s->nl_required = true;
    for (int k = 0; k < 445; ++k) {
      s->qa_ljs[k].depth = p->qa_lj_depth[k];
      s->qa_ljs[k].r_low = p->qa_lj_r_low[k];
      s->qa_ljs[k].r_high = p->qa_lj_r_high[k];
      s->qa_req_min_dist[k] = p->qa_lj_r_high[k];
    }
    s->qa_max_req = 0.0;
    for (int k = 0; k < 445; ++k)
      s->qa_max_req = std::max(s->qa_max_req, s->qa_req_min_dist[k]);
    s->qa_factor = p->counting_factor * C216_INV;
    s->qa_sat_speed = 1.0 / p->qa_phase_dur;
  }

  /* state::setup_pid (state.cpp:540-546), thread::setup_pid */
  if (p->pid_on) {
    s->nl_required = true;
    s->pid_bbp = sink_lj_t{p->pid_bbp_lj[0], p->pid_bbp_lj[1], p->pid_bbp_lj[2]};
    s->pid_bbm = sink_lj_t{p->pid_bbm_lj[0], p->pid_bbm_lj[1], p->pid_bbm_lj[2]};
    for (int k = 0; k < 441; ++k)
      s->pid_ss[k] = sink_lj_t{p->pid_ss_depth[k], p->pid_ss_r_low[k],
                               p->pid_ss_r_high[k]};
  }

  /* velocity AFM tips (state::setup_afm, state.cpp:548-568) */
  for (int k = 0; k < m->n_tips; ++k) {
    tip_t tp;
    tp.res = m->tip_res[k];
    tp.orig = mk(m->tip_orig[3 * k], m->tip_orig[3 * k + 1], m->tip_orig[3 * k + 2]);
    tp.ref_t = 0.0;
    tp.vel = mk(m->tip_vel[3 * k], m->tip_vel[3 * k + 1], m->tip_vel[3 * k + 2]);
    s->tips.push_back(tp);
  }
  for (int k = 0; k < m->n_ftips; ++k)
    s->ftips.emplace_back(m->ftip_res[k], mk(m->ftip_force[3 * k], m->ftip_force[3 * k + 1],
                                              m->ftip_force[3 * k + 2]));
  s->reset_tip_stats();
  /* walls: harmonic connections from the starting positions */
  s->is_conn.assign(n, 0);
  s->setup_harmonic();
  return s;
}

void ora_destroy(ora_sys *s) { delete s; }

void ora_step(ora_sys *s, int32_t nsteps, double *V_trace) {
  for (int k = 0; k < nsteps; ++k) {
    if (s->phase) s->box_phase();
    s->step();
    if (V_trace) V_trace[k] = s->V;
  }
}
double ora_last_V(const ora_sys *s) { return s->V; }
void ora_last_term_V(const ora_sys *s, double *out) {
  for (int k = 0; k < ORA_NUM_TERMS; ++k) out[k] = s->term_V[k];
}
int32_t ora_nl_verify(ora_sys *s) {
  s->nl_verify();
  return s->nl_invalid ? 1 : 0;
}
void ora_fix_nl(ora_sys *s) { s->fix_nl(); }
void ora_eval_forces(ora_sys *s) { s->eval_forces(); }
void ora_eval_term(ora_sys *s, int32_t term, double *F, double *V) {
  for (int i = 0; i < 3 * s->n; ++i) F[i] = 0.0;
  *V = 0.0;
  for (auto &f : s->wall_F) f = v3{0, 0, 0};
  if (term < 0 || term >= ORA_NUM_TERMS) return;
  s->run_term(term, F, *V);
}
void ora_qa_finish(ora_sys *s) { s->qa_finish(); }
void ora_lang_step(ora_sys *s) { s->lang_step(); }
void ora_prepare_nh(ora_sys *s) { s->prepare_nh(); }

void ora_get_r(const ora_sys *s, double *r) {
  for (int i = 0; i < s->n; ++i) {
    r[3 * i] = s->r[i].x;
    r[3 * i + 1] = s->r[i].y;
    r[3 * i + 2] = s->r[i].z;
  }
}
void ora_set_r(ora_sys *s, const double *r) {
  for (int i = 0; i < s->n; ++i) {
    s->r[i] = mk(r[3 * i], r[3 * i + 1], r[3 * i + 2]);
    s->y[0][i] = s->r[i];
  }
}
void ora_get_F(const ora_sys *s, double *F) {
  for (int i = 0; i < s->n; ++i) {
    F[3 * i] = s->F[i].x;
    F[3 * i + 1] = s->F[i].y;
    F[3 * i + 2] = s->F[i].z;
  }
}
void ora_get_y(const ora_sys *s, int32_t k, double *y) {
  for (int i = 0; i < s->n; ++i) {
    y[3 * i] = s->y[k][i].x;
    y[3 * i + 1] = s->y[k][i].y;
    y[3 * i + 2] = s->y[k][i].z;
  }
}
void ora_set_y(ora_sys *s, int32_t k, const double *y) {
  for (int i = 0; i < s->n; ++i) s->y[k][i] = mk(y[3 * i], y[3 * i + 1], y[3 * i + 2]);
  if (k == 0) s->r = s->y[0];
}
double ora_get_t(const ora_sys *s) { return s->t; }
void ora_set_t(ora_sys *s, double t) { s->t = s->true_t = t; }
void ora_set_noise_scale(ora_sys *s, double scale) { s->noise_scale = scale; }

int32_t ora_list(const ora_sys *s, int32_t kind, int32_t cap, int32_t *i1,
                 int32_t *i2, double *val, uint8_t *flag) {
  int32_t size = 0;
  auto put = [&](int k, int a, int b, double v, int f) {
    if (k < cap) {
      if (i1) i1[k] = a;
      if (i2) i2[k] = b;
      if (val) val[k] = v;
      if (flag) flag[k] = (uint8_t)f;
    }
  };
  switch (kind) {
  case 0:
    size = (int)s->nl.size();
    for (int k = 0; k < size; ++k)
      put(k, s->nl[k].i1, s->nl[k].i2, s->nl[k].orig_dist, s->nl[k].taken);
    break;
  case 1:
    size = (int)s->cur_nc.size();
    for (int k = 0; k < size; ++k)
      put(k, s->cur_nc[k].i1, s->cur_nc[k].i2, s->cur_nc[k].nat_dist,
          s->cur_nc[k].active);
    break;
  case 2:
    size = (int)s->pauli_pairs.size();
    for (int k = 0; k < size; ++k)
      put(k, s->pauli_pairs[k].first, s->pauli_pairs[k].second, 0.0, 0);
    break;
  case 3:
    size = (int)s->dh_pairs.size();
    for (int k = 0; k < size; ++k)
      put(k, s->dh_pairs[k].i1, s->dh_pairs[k].i2, s->dh_pairs[k].q1q2, 0);
    break;
  case 4:
    size = (int)s->pid_bundles.size();
    for (int k = 0; k < size; ++k)
      put(k, s->pid_bundles[k].i1, s->pid_bundles[k].i2,
          s->pid_bundles[k].orig_dist, s->pid_bundles[k].type);
    break;
  case 5:
    size = (int)s->qa_free_pairs.size();
    for (int k = 0; k < size; ++k)
      put(k, s->qa_free_pairs[k].i1, s->qa_free_pairs[k].i2,
          s->qa_free_pairs[k].orig_dist, s->qa_free_pairs[k].vacant);
    break;
  case 6:
    size = (int)s->qa_contacts.size();
    for (int k = 0; k < size; ++k)
      put(k, s->qa_contacts[k].i1, s->qa_contacts[k].i2, s->qa_contacts[k].sat,
          s->qa_contacts[k].vacant ? 255 : s->qa_contacts[k].type % 255);
    break;
  }
  return size;
}
int32_t ora_static_list(const ora_sys *s, int32_t kind, int32_t cap,
                        int32_t *idx, double *v0, double *v1, int32_t *ity) {
  int32_t size = 0;
  auto put = [&](int k, int a, int b, int c, int d, double x, double y, int t) {
    if (k >= cap) return;
    if (idx) {
      idx[4 * k] = a;
      idx[4 * k + 1] = b;
      idx[4 * k + 2] = c;
      idx[4 * k + 3] = d;
    }
    if (v0) v0[k] = x;
    if (v1) v1[k] = y;
    if (ity) ity[k] = t;
  };
  switch (kind) {
  case 0:
    size = (int)s->tethers.size();
    for (int k = 0; k < size; ++k)
      put(k, s->tethers[k].i1, s->tethers[k].i2, -1, -1, s->tethers[k].nat, 0, 0);
    break;
  case 1:
    size = (int)s->lrep.size();
    for (int k = 0; k < size; ++k)
      put(k, s->lrep[k].first, s->lrep[k].second, -1, -1, 0, 0, 0);
    break;
  case 2:
    size = (int)s->chir.size();
    for (int k = 0; k < size; ++k) {
      auto const &q = s->chir[k];
      put(k, q.i1, q.i2, q.i3, q.i4, q.nat_factor, q.nat_chir, 0);
    }
    break;
  case 3:
  case 4: {
    auto const &L = kind == 3 ? s->nat_ang : s->heur_ang;
    size = (int)L.size();
    for (int k = 0; k < size; ++k)
      put(k, L[k].i1, L[k].i2, L[k].i3, -1, L[k].theta, 0, L[k].type);
    break;
  }
  case 5:
  case 6: {
    auto const &L = kind == 5 ? s->nat_dih : s->heur_dih;
    size = (int)L.size();
    for (int k = 0; k < size; ++k)
      put(k, L[k].i1, L[k].i2, L[k].i3, L[k].i4, L[k].phi, 0, L[k].type);
    break;
  }
  case 7:
    size = (int)s->all_nc.size();
    for (int k = 0; k < size; ++k)
      put(k, s->all_nc[k].i1, s->all_nc[k].i2, -1, -1, s->all_nc[k].nat_dist, 0,
          s->all_nc[k].type);
    break;
  }
  return size;
}
int32_t ora_qa_contacts(const ora_sys *s, int32_t cap, int32_t *ij,
                        double *orig, int32_t *type, double *sat, int8_t *sd,
                        uint8_t *vac) {
  int32_t size = (int32_t)s->qa_contacts.size();
  for (int k = 0; k < size && k < cap; ++k) {
    auto const &c = s->qa_contacts[k];
    ij[2 * k] = c.i1;
    ij[2 * k + 1] = c.i2;
    orig[k] = c.orig_dist;
    type[k] = c.type;
    sat[k] = c.sat;
    for (int a = 0; a < 4; ++a) {
      sd[8 * k + a] = c.d1.v[a];
      sd[8 * k + 4 + a] = c.d2.v[a];
    }
    vac[k] = c.vacant;
  }
  return size;
}
void ora_nat_cont_state(const ora_sys *s, uint8_t *active, double *change_t) {
  for (size_t k = 0; k < s->all_nc.size(); ++k) {
    active[k] = s->all_nc[k].active;
    change_t[k] = s->all_nc[k].change_t;
  }
}
int32_t ora_num_changed(const ora_sys *s) { return s->num_changed; }
int32_t ora_num_qa_contacts(const ora_sys *s) { return s->num_qa_contacts; }
void ora_sync(const ora_sys *s, int8_t *out) {
  for (int i = 0; i < s->n; ++i)
    for (int k = 0; k < 4; ++k)
      out[4 * i + k] = s->sync.empty() ? 0 : s->sync[i].v[k];
}
int32_t ora_nl_rebuilds(const ora_sys *s) { return s->nl_rebuilds; }
void ora_pid_counts(const ora_sys *s, int32_t *out) { s->pid_counts(out); }
void ora_set_tips(ora_sys *s, int32_t n_tips, const int32_t *tip_res, const double *tip_orig,
                  const double *tip_vel, int32_t n_ftips, const int32_t *ftip_res,
                  const double *ftip_force) {
  s->tips.clear();
  s->ftips.clear();
  for (int k = 0; k < n_tips; ++k) {
    tip_t tp;
    tp.res = tip_res[k];
    tp.orig = mk(tip_orig[3 * k], tip_orig[3 * k + 1], tip_orig[3 * k + 2]);
    tp.ref_t = s->t; /* afm::vel::tip ref_t = t at setup_afm (state.cpp:555-561) */
    tp.vel = mk(tip_vel[3 * k], tip_vel[3 * k + 1], tip_vel[3 * k + 2]);
    s->tips.push_back(tp);
  }
  for (int k = 0; k < n_ftips; ++k)
    s->ftips.emplace_back(ftip_res[k], mk(ftip_force[3 * k], ftip_force[3 * k + 1],
                                          ftip_force[3 * k + 2]));
  s->reset_tip_stats();
}
int32_t ora_afm_stats(const ora_sys *s, double *avg, double *perp, uint8_t *has) {
  const size_t nv = s->tips.size(), nf = s->ftips.size();
  for (size_t k = 0; k < nv + nf; ++k) {
    const bool vt = k < nv;
    const auto &a = vt ? s->tip_avg_F[k] : s->ftip_avg_v[k - nv];
    if (avg) {
      avg[3 * k] = a.avg.x;
      avg[3 * k + 1] = a.avg.y;
      avg[3 * k + 2] = a.avg.z;
    }
    if (perp) perp[k] = vt ? s->tip_avg_perp[k].avg : 0.0;
    if (has) has[k] = a.has ? 1 : 0;
  }
  return (int32_t)(nv + nf);
}
int64_t ora_saw_draws(const ora_sys *s) { return s->saw_draws; }

/* state::setup_walls' box part and reinit_wall_values (state.cpp:570-730) */
void ora_box_setup(ora_sys *s, const ora_box *b) {
  s->bx = *b;
  s->box_set = true;
  s->box_min = mk(b->box_min[0], b->box_min[1], b->box_min[2]);
  s->box_max = mk(b->box_max[0], b->box_max[1], b->box_max[2]);
  for (int side = 0; side < 6; ++side) s->side_wall[side] = -1;
  for (int w = 0; w < s->p.n_walls; ++w)
    if (s->p.wall_side[w] >= 0 && s->p.wall_side[w] < 6) s->side_wall[s->p.wall_side[w]] = w;
  for (int w = 0; w < 6; ++w) {
    s->avg_F[w] = {};
    s->avg_F[w].window = b->avg_forces_over;
    s->work[w] = 0.0;
  }
  s->avg_z = {};
  s->avg_z.window = b->avg_forces_over;
  s->log_on = s->side_wall[4] >= 0 && s->side_wall[5] >= 0;
  s->phase = 0;
  s->reinit_walls();
}
/* equil_step's hand-over (thread_main.cpp:130-135) */
void ora_box_start(ora_sys *s) {
  s->until = s->bx.total_time;
  s->since = s->t;
  s->phase = 1;
}
void ora_box_state(const ora_sys *s, double *out) {
  for (int k = 0; k < 120; ++k) out[k] = 0.0;
  out[0] = s->phase;
  out[1] = s->since;
  out[2] = s->until;
  out[3] = s->box_min.x, out[4] = s->box_min.y, out[5] = s->box_min.z;
  out[6] = s->box_max.x, out[7] = s->box_max.y, out[8] = s->box_max.z;
  out[9] = s->amplitude;
  out[10] = s->omega;
  out[11] = s->displacement;
  out[12] = s->shear ? 1.0 : 0.0;
  out[13] = s->avg_z.has ? 1.0 : 0.0;
  out[14] = s->avg_z.has ? s->avg_z.avg : 0.0;
  for (int w = 0; w < 6; ++w) {
    for (int c = 0; c < 3; ++c) {
      out[15 + 3 * w + c] = s->p.wall_origin[w][c];
      out[33 + 3 * w + c] = s->p.wall_normal[w][c];
    }
    const v3 f = s->wall_F[w], sh = s->shift[w], a = s->avg_F[w].avg;
    out[51 + 3 * w] = f.x, out[52 + 3 * w] = f.y, out[53 + 3 * w] = f.z;
    out[69 + 3 * w] = sh.x, out[70 + 3 * w] = sh.y, out[71 + 3 * w] = sh.z;
    out[87 + w] = s->work[w];
    out[93 + w] = s->avg_F[w].has ? 1.0 : 0.0;
    out[99 + 3 * w] = a.x, out[100 + 3 * w] = a.y, out[101 + 3 * w] = a.z;
  }
  out[117] = s->pbc.cell.x, out[118] = s->pbc.cell.y, out[119] = s->pbc.cell.z;
}
void ora_ljw_process(ora_sys *s) { s->ljw_process(); }
int32_t ora_ljw_slots(const ora_sys *s, int32_t cap, int32_t *res, int32_t *wall, double *off,
                      double *sat, uint8_t *vac) {
  int k = 0;
  for (auto const &c : s->ljw_conns) {
    if (k < cap) {
      res[k] = c.res, wall[k] = c.wall, sat[k] = c.sat, vac[k] = c.vacant ? 1 : 0;
      off[3 * k] = c.off.x, off[3 * k + 1] = c.off.y, off[3 * k + 2] = c.off.z;
    }
    ++k;
  }
  return k;
}
void ora_ljw_pending(const ora_sys *s, int32_t cap, int32_t *ncand, int32_t *cand_res,
                     int32_t *cand_wall, double *cand_d, int32_t *nrem, int32_t *rem,
                     uint8_t *is_conn, int32_t *num_conns) {
  *ncand = (int)s->ljw_cands.size();
  for (int a = 0; a < *ncand && a < cap; ++a)
    cand_res[a] = s->ljw_cands[a].res, cand_wall[a] = s->ljw_cands[a].wall,
    cand_d[a] = s->ljw_cands[a].dist;
  *nrem = (int)s->ljw_removed.size();
  for (int a = 0; a < *nrem && a < cap; ++a) rem[a] = s->ljw_removed[a];
  for (int i = 0; i < s->n; ++i) is_conn[i] = s->is_conn[i];
  for (int w = 0; w < 6; ++w) num_conns[w] = s->ljw_num_conns[w];
}
int32_t ora_wall_conns(const ora_sys *s, int32_t kind, int32_t cap, int32_t *res,
                       int32_t *wall, double *off, double *sat) {
  int k = 0;
  auto put = [&](int r_, int w_, v3 o, double x) {
    if (k < cap) {
      res[k] = r_, wall[k] = w_, sat[k] = x;
      off[3 * k] = o.x, off[3 * k + 1] = o.y, off[3 * k + 2] = o.z;
    }
    ++k;
  };
  if (kind == 0)
    for (auto const &c : s->hw_conns) put(c.res, c.wall, c.off, c.nat);
  else
    for (auto const &c : s->ljw_conns)
      if (!c.vacant) put(c.res, c.wall, c.off, c.sat);
  return k;
}

struct ora_rng {
  ran2_t g;
};
ora_rng *ora_rng_new(int32_t seed) {
  ora_rng *g = new ora_rng();
  g->g = ran2_t(seed);
  return g;
}
void ora_rng_free(ora_rng *g) { delete g; }
double ora_rng_uniform(ora_rng *g) { return g->g.uniform(); }
double ora_rng_normal(ora_rng *g) { return g->g.normal(); }

} // extern "C"
