/*
 * ref_harness.cpp -- drives the reference's own per-term kernels, compiled
 * from their sources where they lie under the reference tree (see
 * oracle/Makefile), on lists and positions handed in from the tests.
 *
 * TEST INFRASTRUCTURE ONLY (it pins the oracle): tests compare the
 * oracle with these calls. Nothing in the product loads it. Lists are built
 * the way cg/src/cg/simul/state.cpp and thread_setup.cpp build them, each
 * kernel object's parameters are bound as thread_setup.cpp binds them, and
 * the kernel runs through its own for_slice (runtime.h:17-21) over the whole
 * list, as one thread does.
 */
#include <cg/afm/force/eval_forces.h>
#include <cg/afm/vel/eval_forces.h>
#include <cg/angles/heur_ang/eval_forces.h>
#include <cg/angles/heur_dih/rel.h>
#include <cg/angles/nat_ang/eval_forces.h>
#include <cg/angles/nat_dih/complex.h>
#include <cg/angles/nat_dih/simple.h>
#include <cg/chir/eval_forces.h>
#include <cg/dh/const.h>
#include <cg/dh/rel.h>
#include <cg/dh/update_pairs.h>
#include <cg/langevin/step.h>
#include <cg/local_rep/eval_forces.h>
#include <cg/nat_cont/eval_forces.h>
#include <cg/nat_cont/update_contacts.h>
#include <cg/nl/legacy_update.h>
#include <cg/nl/verify.h>
#include <cg/pauli/eval_forces.h>
#include <cg/pauli/update_pairs.h>
#include <cg/pid/count_active.h>
#include <cg/pid/eval_forces.h>
#include <cg/pid/update_bundles.h>
#include <cg/qa/finish_processing.h>
#include <cg/qa/loop_over_candidates.h>
#include <cg/qa/prepare_nh.h>
#include <cg/qa/process_contacts.h>
#include <cg/qa/update_free_pairs.h>
#include <cg/tether/eval_forces.h>
#include <cg/wall/harmonic/eval_connected.h>
#include <cg/wall/harmonic/eval_free.h>
#include <cg/wall/lj/eval_connected.h>
#include <cg/wall/lj/process_candidates.h>
#include <cg/wall/lj/sift_free.h>
#include <cg/wall/log_forces.h>
#include <cg/wall/solid/eval_forces.h>

#include "oracle.h"

using namespace cg;

namespace {
vect::vector<vec3r> to_vec(int n, const double *r) {
  vect::vector<vec3r> v(n);
  for (int i = 0; i < n; ++i) v[i] = vec3r(r[3 * i], r[3 * i + 1], r[3 * i + 2]);
  return v;
}
void from_vec(vect::vector<vec3r> const &v, double *out) {
  for (int i = 0; i < v.size(); ++i) {
    vec3r x = v[i];
    out[3 * i] = x.x();
    out[3 * i + 1] = x.y();
    out[3 * i + 2] = x.z();
  }
}
sbox::pbc<real> make_pbc(const double *cell) {
  sbox::pbc<real> p;
  p.set_cell(vec3r(cell[0], cell[1], cell[2]));
  return p;
}
vect::vector<amino_acid> to_aa(int n, const uint8_t *atype) {
  vect::vector<amino_acid> v(n);
  for (int i = 0; i < n; ++i) v[i] = amino_acid((aa_code)atype[i]);
  return v;
}
cg::plane<real> plane_of(const ora_params *p, int w) {
  vec3r org(p->wall_origin[w][0], p->wall_origin[w][1], p->wall_origin[w][2]);
  vec3r nrm(p->wall_normal[w][0], p->wall_normal[w][1], p->wall_normal[w][2]);
  return cg::plane<real>(org, nrm);
}
// the planes of one kind in plane order (the reference keeps one vector per
// kind, state.h:174-183): plane index of each, and each plane's kind index
std::vector<int> planes_of_kind(const ora_params *p, int kind, int *kind_idx) {
  std::vector<int> v;
  for (int w = 0; w < p->n_walls; ++w) {
    kind_idx[w] = -1;
    if (p->wall_kind[w] == kind) {
      kind_idx[w] = (int)v.size();
      v.push_back(w);
    }
  }
  return v;
}
void put_wall_F(vect::vector<vec3r> const &wf, std::vector<int> const &planes, double *out) {
  for (int k = 0; k < (int)planes.size(); ++k) {
    vec3r f = wf[k];
    out[3 * planes[k]] = f.x(), out[3 * planes[k] + 1] = f.y(), out[3 * planes[k] + 2] = f.z();
  }
}
template <typename K> double run(K &k) {
  k.for_slice(0, k.total_size());
  return 0.0;
}
} // namespace

extern "C" {

double ref_chir(int n, const double *r, int k, const int *q,
                const double *nat_factor, const double *nat_chir, double e_chi,
                double *F) {
  auto R = to_vec(n, r);
  vect::vector<vec3r> Fv(n);
  vect::vector<chir::chiral_quad> quads;
  for (int a = 0; a < k; ++a) /* state.cpp:384 argument order */
    quads.emplace_back(q[4 * a], q[4 * a + 1], q[4 * a + 2], q[4 * a + 3],
                       nat_factor[a], nat_chir[a]);
  real V = 0;
  chir::eval_forces e;
  e.e_chi = e_chi;
  e.quads = quads;
  e.r = R;
  e.F = Fv;
  e.V = &V;
  run(e);
  from_vec(Fv, F);
  return V;
}

double ref_tether(int n, const double *r, int k, const int *ij,
                  const double *nat, double H1, double H2, double *F) {
  auto R = to_vec(n, r);
  vect::vector<vec3r> Fv(n);
  vect::vector<tether::pair> pairs;
  for (int a = 0; a < k; ++a) pairs.emplace_back(ij[2 * a], ij[2 * a + 1], nat[a]);
  real V = 0;
  tether::eval_forces e;
  e.H1 = H1;
  e.H2 = H2;
  e.r = R;
  e.tethers = pairs;
  e.F = Fv;
  e.V = &V;
  run(e);
  from_vec(Fv, F);
  return V;
}

double ref_lrep(int n, const double *r, int k, const int *ij, double depth,
                double cutoff, double *F) {
  auto R = to_vec(n, r);
  vect::vector<vec3r> Fv(n);
  vect::vector<local_rep::pair> pairs;
  for (int a = 0; a < k; ++a) pairs.emplace_back(ij[2 * a], ij[2 * a + 1]);
  real V = 0;
  local_rep::eval_forces e;
  e.depth = depth;
  e.cutoff = cutoff;
  e.r = R;
  e.pairs = pairs;
  e.F = Fv;
  e.V = &V;
  run(e);
  from_vec(Fv, F);
  return V;
}

double ref_pauli(int n, const double *r, const double *cell, int k,
                 const int *ij, double depth, double r_excl, double *F) {
  auto R = to_vec(n, r);
  vect::vector<vec3r> Fv(n);
  vect::vector<pauli::pair> pairs;
  for (int a = 0; a < k; ++a) pairs.emplace_back(ij[2 * a], ij[2 * a + 1]);
  auto box = make_pbc(cell);
  real V = 0;
  pauli::eval_forces e;
  e.depth = depth;
  e.r_excl = r_excl;
  e.r = R;
  e.simul_box = &box;
  e.pairs = &pairs;
  e.F = Fv;
  e.V = &V;
  run(e);
  from_vec(Fv, F);
  return V;
}

/* native contacts: evaluates the current list (all_cont_idx = position) and
 * returns updated active/change_t plus num_changed */
double ref_nat_cont(int n, const double *r, const double *cell, int k,
                    const int *ij, const double *nat, const int *type,
                    uint8_t *active, double *change_t, double t,
                    const ora_params *p, int *num_changed, double *F) {
  auto R = to_vec(n, r);
  vect::vector<vec3r> Fv(n);
  vect::vector<nat_cont::nat_cont> all, cur;
  for (int a = 0; a < k; ++a) {
    all.emplace_back(ij[2 * a], ij[2 * a + 1], nat[a], (nat_cont::type)type[a],
                     (bool)active[a], change_t[a], a);
    cur.emplace_back(ij[2 * a], ij[2 * a + 1], nat[a], (nat_cont::type)type[a],
                     (bool)active[a], change_t[a], a);
  }
  auto box = make_pbc(cell);
  real V = 0, tt = t;
  int nch = 0;
  nat_cont::eval_forces e;
  e.depth = p->nat_cont_depth;
  e.simul_box = &box;
  e.contacts = &cur;
  e.r = R;
  e.breaking_threshold = p->counting_factor;
  e.t = &tt;
  e.all_contacts = all;
  e.V = &V;
  e.F = Fv;
  e.cutoff = p->nl_cutoff;
  e.num_changed = &nch;
  if (p->nc_ss_kind == 1) {
    disulfide_force d;
    d.force = harmonic(p->nc_ss_H1, p->nc_ss_H2, p->nc_ss_nat_r);
    e.disulfide = d;
  } else if (p->nc_ss_kind == 2) {
    disulfide_force d;
    d.force = lj(p->nc_ss_depth, p->nc_ss_r_min);
    e.disulfide = d;
  } else {
    e.disulfide = std::nullopt;
  }
  run(e);
  for (int a = 0; a < k; ++a) {
    active[a] = all[a].active();
    change_t[a] = all[a].change_t();
  }
  *num_changed = nch;
  from_vec(Fv, F);
  return V;
}

double ref_dh(int variant, int n, const double *r, const double *cell, int k,
              const int *ij, const double *q, const ora_params *p, double *F) {
  auto R = to_vec(n, r);
  vect::vector<vec3r> Fv(n);
  vect::vector<dh::pair> pairs;
  for (int a = 0; a < k; ++a) pairs.emplace_back(ij[2 * a], ij[2 * a + 1], q[a]);
  auto box = make_pbc(cell);
  real V = 0;
  if (variant == 0) {
    const_dh::eval_forces e;
    e.set_V_factor(p->dh_perm);
    e.screen_dist_inv = 1.0 / p->dh_screen_dist;
    e.r = R;
    e.simul_box = &box;
    e.es_pairs = &pairs;
    e.V = &V;
    e.F = Fv;
    e.cutoff = p->dh_cutoff;
    run(e);
  } else {
    rel_dh::eval_forces e;
    e.set_V_factor(p->dh_perm);
    e.screen_dist_inv = 1.0 / p->dh_screen_dist;
    e.r = R;
    e.simul_box = &box;
    e.es_pairs = &pairs;
    e.V = &V;
    e.F = Fv;
    e.cutoff = p->dh_cutoff;
    run(e);
  }
  from_vec(Fv, F);
  return V;
}

double ref_nat_ang(int n, const double *r, int k, const int *ijk,
                   const double *theta, double CBA, double *F) {
  auto R = to_vec(n, r);
  vect::vector<vec3r> Fv(n);
  vect::vector<nat_ang::nat_ang> angles;
  for (int a = 0; a < k; ++a)
    angles.emplace_back(ijk[3 * a], ijk[3 * a + 1], ijk[3 * a + 2], theta[a]);
  real V = 0;
  nat_ang::eval_forces e;
  e.CBA = CBA;
  e.r = R;
  e.angles = angles;
  e.V = &V;
  e.F = Fv;
  run(e);
  from_vec(Fv, F);
  return V;
}

double ref_heur_ang(int n, const double *r, int k, const int *ijk,
                    const int *type, const ora_params *p, double *F) {
  auto R = to_vec(n, r);
  vect::vector<vec3r> Fv(n);
  vect::vector<heur_ang::heur_ang> angles;
  for (int a = 0; a < k; ++a)
    angles.emplace_back(ijk[3 * a], ijk[3 * a + 1], ijk[3 * a + 2],
                        aa_heur_pair((uint8_t)type[a]));
  real V = 0;
  heur_ang::eval_forces e;
  for (int d = 0; d < 7; ++d)
    for (int t = 0; t < 9; ++t) e.poly_coeffs[d][t] = p->heur_ang_poly[d][t];
  e.r = R;
  e.angles = angles;
  e.V = &V;
  e.F = Fv;
  run(e);
  from_vec(Fv, F);
  return V;
}

double ref_nat_dih(int variant, int n, const double *r, int k, const int *q,
                   const double *phi, const ora_params *p, double *F) {
  auto R = to_vec(n, r);
  vect::vector<vec3r> Fv(n);
  vect::vector<cg::nat_dih> dih;
  for (int a = 0; a < k; ++a)
    dih.emplace_back(q[4 * a], q[4 * a + 1], q[4 * a + 2], q[4 * a + 3], phi[a]);
  real V = 0;
  if (variant == 0) {
    cnd::eval_forces e;
    e.CDA = p->CDA;
    e.CDB = p->CDB;
    e.r = R;
    e.dihedrals = dih;
    e.V = &V;
    e.F = Fv;
    run(e);
  } else {
    snd::eval_forces e;
    e.CDH = p->CDH;
    e.r = R;
    e.dihedrals = dih;
    e.V = &V;
    e.F = Fv;
    run(e);
  }
  from_vec(Fv, F);
  return V;
}

double ref_heur_dih(int n, const double *r, int k, const int *q,
                    const int *type, const ora_params *p, double *F) {
  auto R = to_vec(n, r);
  vect::vector<vec3r> Fv(n);
  vect::vector<heur_dih::heur_dih> dih;
  for (int a = 0; a < k; ++a)
    dih.emplace_back(q[4 * a], q[4 * a + 1], q[4 * a + 2], q[4 * a + 3],
                     aa_heur_pair((uint8_t)type[a]));
  real V = 0;
  heur_dih::eval_forces e;
  for (int t = 0; t < 9; ++t) {
    e.coeffs.const_[t] = p->heur_dih_coeffs[0][t];
    e.coeffs.sin[t] = p->heur_dih_coeffs[1][t];
    e.coeffs.cos[t] = p->heur_dih_coeffs[2][t];
    e.coeffs.sin2[t] = p->heur_dih_coeffs[3][t];
    e.coeffs.cos2[t] = p->heur_dih_coeffs[4][t];
    e.coeffs.sin_cos[t] = p->heur_dih_coeffs[5][t];
  }
  e.r = R;
  e.dihedrals = dih;
  e.V = &V;
  e.F = Fv;
  run(e);
  from_vec(Fv, F);
  return V;
}

// the lambdas and sink LJs of pid::eval_forces (thread_setup.cpp:589-686)
static void pid_setup(pid::eval_forces &e, const ora_params *p) {
  auto ver = p->pid_lambda_variant == 0 ? pid::COSINE : pid::ALGEBRAIC;
  e.bb_plus_lam.version() = ver;
  e.bb_plus_lam.alpha() = p->pid_bbp_alpha;
  e.bb_plus_lam.psi_0() = p->pid_bbp_psi0;
  e.bb_minus_lam.version() = ver;
  e.bb_minus_lam.alpha() = p->pid_bbm_alpha;
  e.bb_minus_lam.psi_0() = p->pid_bbm_psi0;
  e.ss_lam.version() = ver;
  e.ss_lam.alpha() = p->pid_ss_alpha;
  e.ss_lam.psi_0() = p->pid_ss_psi0;
  e.bb_plus_lj = sink_lj(p->pid_bbp_lj[0], p->pid_bbp_lj[1], p->pid_bbp_lj[2]);
  e.bb_minus_lj = sink_lj(p->pid_bbm_lj[0], p->pid_bbm_lj[1], p->pid_bbm_lj[2]);
  e.ss_ljs = vect::vector<sink_lj>(441);
  for (int t = 0; t < 441; ++t)
    e.ss_ljs[t] = sink_lj(p->pid_ss_depth[t], p->pid_ss_r_low[t], p->pid_ss_r_high[t]);
  e.cutoff = p->nl_cutoff;
}

double ref_pid(int n, const double *r, const double *cell, const uint8_t *atype,
               int k, const int *ij, const double *orig_dist, const int *type,
               const ora_params *p, double *F) {
  auto R = to_vec(n, r);
  vect::vector<vec3r> Fv(n);
  auto at = to_aa(n, atype);
  vect::vector<pid::bundle> bundles;
  for (int a = 0; a < k; ++a)
    bundles.emplace_back(ij[2 * a], ij[2 * a + 1], orig_dist[a], (int16_t)type[a]);
  auto box = make_pbc(cell);
  vect::vector<int> prev(n, -1), next(n, -1);
  real V = 0, total_disp = 0;
  pid::eval_forces e;
  pid_setup(e, p);
  e.V = &V;
  e.F = Fv;
  e.total_disp = &total_disp;
  e.active_thr = p->counting_factor;
  e.atype = at;
  e.r = R;
  e.simul_box = &box;
  e.bundles = &bundles;
  e.prev = prev;
  e.next = next;
  run(e);
  from_vec(Fv, F);
  return V;
}

// pid::count_active (cg/src/cg/pid/count_active.cpp:4-164) over the given
// bundles: out = {intra, intra_bb, intra_ss, inter, inter_bb, inter_ss}
void ref_pid_count(int n, const double *r, const double *cell, const uint8_t *atype,
                   const int *chain, int k, const int *ij, const double *orig_dist,
                   const int *type, const ora_params *p, int *out) {
  auto R = to_vec(n, r);
  auto at = to_aa(n, atype);
  vect::vector<pid::bundle> bundles;
  for (int a = 0; a < k; ++a)
    bundles.emplace_back(ij[2 * a], ij[2 * a + 1], orig_dist[a], (int16_t)type[a]);
  auto box = make_pbc(cell);
  vect::vector<int> ch(n);
  for (int i = 0; i < n; ++i) ch[i] = chain[i];
  pid::eval_forces e;
  pid_setup(e, p);
  e.atype = at;
  e.r = R;
  e.simul_box = &box;
  e.bundles = &bundles;
  pid::count_active c;
  c.eval = &e;
  c.chain_idx = ch;
  c.count_factor = p->counting_factor;
  auto cnt = c();
  out[0] = cnt.intra;
  out[1] = cnt.intra_bb;
  out[2] = cnt.intra_ss;
  out[3] = cnt.inter;
  out[4] = cnt.inter_bb;
  out[5] = cnt.inter_ss;
}

// afm::force::eval_forces over the given tips (velocities only feed the
// tips' statistics)
double ref_force_afm(int n, double t, int k, const int *res, const double *force, double *F) {
  vect::vector<vec3r> Fv(n), v(n);
  vect::vector<afm::force::tip> tips;
  for (int a = 0; a < k; ++a)
    tips.emplace_back(res[a], vec3r(force[3 * a], force[3 * a + 1], force[3 * a + 2]),
                      moving_avg<vec3r, real>(1.0));
  real tt = t;
  afm::force::eval_forces e;
  e.F = Fv;
  e.afm_tips = tips;
  e.v = v;
  e.t = &tt;
  run(e);
  from_vec(Fv, F);
  return 0.0;
}

// The AFM tips' statistics after afm::vel::eval_forces and
// afm::force::eval_forces ran over a sequence of samples: sample j has the
// time t[j], positions r[j] and velocities v[j] ([ns][n][3] each). Outputs
// per tip (velocity tips, then force tips) the average (force on the tip,
// or residue velocity), the perpendicular force average (velocity tips) and
// whether each average has a value.
void ref_afm_stats(int n, int ns, const double *t, const double *r, const double *v, int kv,
                   const int *vres, const double *vorig, const double *vvel, int kf,
                   const int *fres, const double *fforce, double H1, double H2, double window,
                   double *avg, double *perp, uint8_t *has) {
  vect::vector<afm::vel::tip> vt;
  for (int a = 0; a < kv; ++a)
    vt.emplace_back(vres[a], vec3r(vorig[3 * a], vorig[3 * a + 1], vorig[3 * a + 2]),
                    (real)0.0, vec3r(vvel[3 * a], vvel[3 * a + 1], vvel[3 * a + 2]),
                    moving_avg<vec3r, real>(window), moving_avg<real, real>(window));
  vect::vector<afm::force::tip> ft;
  for (int a = 0; a < kf; ++a)
    ft.emplace_back(fres[a], vec3r(fforce[3 * a], fforce[3 * a + 1], fforce[3 * a + 2]),
                    moving_avg<vec3r, real>(window));
  for (int j = 0; j < ns; ++j) {
    auto R = to_vec(n, r + (size_t)3 * n * j), Vel = to_vec(n, v + (size_t)3 * n * j);
    vect::vector<vec3r> Fv(n);
    real V = 0, tt = t[j];
    afm::vel::eval_forces ev;
    ev.r = R;
    ev.F = Fv;
    ev.V = &V;
    ev.t = &tt;
    ev.afm_force.H1 = H1;
    ev.afm_force.H2 = H2;
    ev.afm_force.nat_r = 0;
    ev.afm_tips = vt;
    run(ev);
    afm::force::eval_forces ef;
    ef.F = Fv;
    ef.afm_tips = ft;
    ef.v = Vel;
    ef.t = &tt;
    run(ef);
  }
  for (int a = 0; a < kv + kf; ++a) {
    const bool isv = a < kv;
    auto const &opt = isv ? *vt[a].avg_force().operator->() : *ft[a - kv].avg_vel().operator->();
    has[a] = opt.has_value() ? 1 : 0;
    vec3r x = opt.has_value() ? opt.value() : vec3r::Zero();
    avg[3 * a] = x.x(), avg[3 * a + 1] = x.y(), avg[3 * a + 2] = x.z();
    perp[a] = 0.0;
    if (isv && vt[a].avg_perp_force()->has_value()) perp[a] = vt[a].avg_perp_force()->value();
  }
}

// wall::solid::eval_forces over the walls of the parameters
double ref_solid_wall(int n, const double *r, const ora_params *p, double *F) {
  auto R = to_vec(n, r);
  vect::vector<vec3r> Fv(n), wall_F(p->n_walls);
  vect::vector<wall::solid::wall> walls;
  for (int w = 0; w < p->n_walls; ++w) {
    if (p->wall_kind[w] != 0) continue;
    walls.emplace_back(plane_of(p, w), vect::iterator<vec3r>(wall_F[w]), (real)1.0);
  }
  wall::solid::eval_forces e;
  e.min_dist = p->wall_min_dist;
  e.walls = walls;
  e.depth = p->wall_depth;
  e.r = R;
  e.F = Fv;
  e.wall_F = wall_F;
  run(e);
  from_vec(Fv, F);
  return 0.0;
}

double ref_vel_afm(int n, const double *r, double t, int k, const int *res,
                   const double *orig, const double *vel, double H1, double H2,
                   double *F) {
  auto R = to_vec(n, r);
  vect::vector<vec3r> Fv(n);
  vect::vector<afm::vel::tip> tips;
  for (int a = 0; a < k; ++a)
    tips.emplace_back(res[a], vec3r(orig[3 * a], orig[3 * a + 1], orig[3 * a + 2]),
                      (real)0.0,
                      vec3r(vel[3 * a], vel[3 * a + 1], vel[3 * a + 2]),
                      moving_avg<vec3r, real>(1.0), moving_avg<real, real>(1.0));
  real V = 0, tt = t;
  afm::vel::eval_forces e;
  e.r = R;
  e.F = Fv;
  e.V = &V;
  e.t = &tt;
  e.afm_force.H1 = H1;
  e.afm_force.H2 = H2;
  e.afm_force.nat_r = 0;
  e.afm_tips = tips;
  run(e);
  from_vec(Fv, F);
  return V;
}

/* NL rebuild and the ordered list updates in the order of
 * thread_main.cpp:409-423. Outputs sizes; lists written up to cap. */
void ref_fix_nl(int n, const double *r, const double *cell, const uint8_t *atype,
                const int *chain, const int *seq, const ora_params *p,
                int n_cont, const int *cont_ij, const double *cont_nat,
                int cap, int *sizes /*[6]*/, int *nl_ij, double *nl_d,
                uint8_t *nl_taken, int *nc_ij, int *qa_ij, int *pid_ij,
                int *dh_ij, int *pauli_ij) {
  auto R = to_vec(n, r);
  auto box = make_pbc(cell);
  auto at = to_aa(n, atype);
  vect::vector<int> ch(n), sq(n), prev(n, -1), next(n, -1);
  for (int i = 0; i < n; ++i) {
    ch[i] = chain[i];
    sq[i] = seq[i];
  }
  nl::data nld;
  nld.orig_r = vect::vector<vec3r>(n);
  bool invalid = true;
  nl::legacy_update lu;
  lu.pad = p->nl_pad;
  lu.r = R;
  lu.simul_box = &box;
  lu.chain_idx = ch;
  lu.seq_idx = sq;
  lu.num_particles = n;
  lu.nl_data = &nld;
  lu.invalid = &invalid;
  lu.cutoff = std::max(p->nl_cutoff, p->dh_cutoff);
  lu.omp_async(); /* serial without an enclosing parallel region */

  bool qa = p->qa_on, pauli = p->pauli_on && !p->qa_on;
  vect::vector<nat_cont::nat_cont> all_nc, cur_nc;
  for (int a = 0; a < n_cont; ++a)
    all_nc.emplace_back(cont_ij[2 * a], cont_ij[2 * a + 1], cont_nat[a],
                        nat_cont::type::ANY, false, (real)-1.0, a);
  if (p->nat_cont_on && n_cont > 0) {
    nat_cont::update_contacts u;
    u.r = R;
    u.simul_box = &box;
    u.nl = &nld;
    u.all_contacts = all_nc;
    u.contacts = &cur_nc;
    u.cutoff = p->nl_cutoff;
    u();
  }
  vect::set<qa::free_pair> free_pairs;
  if (qa) {
    qa::update_free_pairs u;
    u.r = R;
    u.simul_box = &box;
    u.nl = &nld;
    u.pairs = &free_pairs;
    u.chain_idx = ch;
    u.seq_idx = sq;
    u.include4 = p->qa_include4;
    u.cutoff = p->nl_cutoff;
    u();
  }
  vect::vector<pid::bundle> bundles;
  if (p->pid_on) {
    pid::update_bundles u;
    u.r = R;
    u.prev = prev;
    u.next = next;
    u.atype = at;
    u.simul_box = &box;
    u.nl = &nld;
    u.bundles = &bundles;
    u.chain_idx = ch;
    u.seq_idx = sq;
    u.include4 = p->pid_include4;
    u.cutoff = p->nl_cutoff;
    u();
  }
  vect::vector<dh::pair> dh_pairs;
  if (p->dh_on) {
    dh::update_pairs u;
    u.r = R;
    u.simul_box = &box;
    u.nl = &nld;
    u.pairs = &dh_pairs;
    u.atype = at;
    u.cutoff = p->dh_cutoff;
    for (int a = 0; a < 21; ++a) u.q[a] = p->charge[a];
    u();
  }
  vect::vector<pauli::pair> pauli_pairs;
  if (pauli) {
    pauli::update_pairs u;
    u.r = R;
    u.simul_box = &box;
    u.nl = &nld;
    u.pairs = &pauli_pairs;
    u.r_excl = p->rep_cutoff;
    u();
  }
  sizes[0] = nld.pairs.size();
  sizes[1] = cur_nc.size();
  sizes[2] = free_pairs.size();
  sizes[3] = bundles.size();
  sizes[4] = dh_pairs.size();
  sizes[5] = pauli_pairs.size();
  for (int a = 0; a < sizes[0] && a < cap; ++a) {
    nl_ij[2 * a] = nld.pairs[a].i1();
    nl_ij[2 * a + 1] = nld.pairs[a].i2();
    nl_d[a] = nld.pairs[a].orig_dist();
    nl_taken[a] = nld.pairs[a].taken();
  }
  for (int a = 0; a < sizes[1] && a < cap; ++a) {
    nc_ij[2 * a] = cur_nc[a].i1();
    nc_ij[2 * a + 1] = cur_nc[a].i2();
  }
  for (int a = 0; a < sizes[2] && a < cap; ++a) {
    qa_ij[2 * a] = free_pairs[a].item().i1();
    qa_ij[2 * a + 1] = free_pairs[a].item().i2();
  }
  for (int a = 0; a < sizes[3] && a < cap; ++a) {
    pid_ij[2 * a] = bundles[a].i1();
    pid_ij[2 * a + 1] = bundles[a].i2();
  }
  for (int a = 0; a < sizes[4] && a < cap; ++a) {
    dh_ij[2 * a] = dh_pairs[a].i1();
    dh_ij[2 * a + 1] = dh_pairs[a].i2();
  }
  for (int a = 0; a < sizes[5] && a < cap; ++a) {
    pauli_ij[2 * a] = pauli_pairs[a].i1();
    pauli_ij[2 * a + 1] = pauli_pairs[a].i2();
  }
}

/* One QA round as the step runs it: prepare_nh, loop_over_candidates and
 * process_contacts through their for_slice, then finish_processing.
 * Free pairs / contacts are passed as (i1, i2, orig_dist, vacant) and
 * (i1, i2, orig_dist, type, saturation, sync_diff1[4], sync_diff2[4], vacant)
 * and returned updated; sizes in/out. */
double ref_qa_round(int n, const double *r, const double *cell,
                    const uint8_t *atype, const int *chain, const ora_params *p,
                    double dt, int8_t *sync_io /*[4n]*/, int *n_free,
                    int *free_ij, double *free_d, uint8_t *free_vac, int *n_cont,
                    int *cont_ij, double *cont_d, int *cont_type,
                    double *cont_sat, int8_t *cont_sd /*[8 per]*/,
                    uint8_t *cont_vac, int *num_contacts, int cap, double *F) {
  auto R = to_vec(n, r);
  vect::vector<vec3r> Fv(n), nv(n), hv(n);
  auto box = make_pbc(cell);
  auto at = to_aa(n, atype);
  vect::vector<int> prev(n, -1), next(n, -1);
  for (int i = 0; i < n; ++i) {
    if (i > 0 && chain[i - 1] == chain[i]) prev[i] = i - 1;
    if (i + 1 < n && chain[i + 1] == chain[i]) next[i] = i + 1;
  }
  vect::vector<sync_data> sync(n);
  for (int i = 0; i < n; ++i)
    sync[i] = sync_data(sync_io[4 * i], sync_io[4 * i + 1], sync_io[4 * i + 2],
                        sync_io[4 * i + 3]);
  vect::set<qa::free_pair> free_pairs;
  for (int a = 0; a < *n_free; ++a) {
    free_pairs.emplace(free_ij[2 * a], free_ij[2 * a + 1], free_d[a]);
    if (free_vac[a]) free_pairs.remove(a);
  }
  vect::set<qa::contact> contacts;
  for (int a = 0; a < *n_cont; ++a) {
    sync_data d1(cont_sd[8 * a], cont_sd[8 * a + 1], cont_sd[8 * a + 2],
                 cont_sd[8 * a + 3]);
    sync_data d2(cont_sd[8 * a + 4], cont_sd[8 * a + 5], cont_sd[8 * a + 6],
                 cont_sd[8 * a + 7]);
    contacts.emplace(cont_ij[2 * a], cont_ij[2 * a + 1], cont_d[a],
                     qa::contact_type::NONE(), cont_sat[a], d1, d2);
    if (cont_vac[a]) contacts.remove(a);
  }
  /* contact types are set through the item proxy (no public int16 ctor) */
  for (int a = 0; a < *n_cont; ++a) {
    int16_t tv = (int16_t)cont_type[a];
    std::memcpy((void *)&contacts[a].item().type(), &tv, sizeof(tv));
  }
  vect::vector<qa::candidate> cands;
  vect::vector<int> removed;
  real V = 0, total_disp = 0, tt = 0;

  qa::prepare_nh nh;
  nh.r = R;
  nh.num_particles = n;
  nh.n = nv;
  nh.h = hv;
  nh.prev = prev;
  nh.next = next;
  nh.simul_box = &box;
  nh.for_slice(0, n);

  vect::vector<sink_lj> ljs(qa::contact_type::NUM_TYPES);
  for (int t = 0; t < 445; ++t)
    ljs[t] = sink_lj(p->qa_lj_depth[t], p->qa_lj_r_low[t], p->qa_lj_r_high[t]);

  qa::loop_over_candidates loop;
  loop.r = R;
  loop.n = nv;
  loop.h = hv;
  loop.simul_box = &box;
  loop.atype = at;
  loop.sync = sync;
  loop.free_pairs = &free_pairs;
  loop.candidates = &cands;
  loop.total_disp = &total_disp;
  loop.min_abs_cos_hr = p->qa_min_cos_hr;
  loop.min_abs_cos_hh = p->qa_min_cos_hh;
  loop.max_cos_nr = p->qa_max_cos_nr;
  loop.formation_tolerance = p->qa_formation_tol;
  loop.disulfide_special_criteria = false;
  loop.rep_cutoff = p->rep_cutoff;
  loop.rep_depth = p->pauli_depth;
  loop.F = Fv;
  loop.V = &V;
  for (int a = 0; a < 21; ++a) loop.ptype[a] = (polarization_type)p->ptype[a];
  real max_req = 0;
  for (int t = 0; t < 445; ++t) {
    loop.req_min_dist[t] = ljs[t].r_high();
    max_req = std::max(max_req, loop.req_min_dist[t]);
  }
  loop.max_req_dist = max_req;

  qa::process_contacts proc;
  proc.saturation_speed = (real)1.0 / p->qa_phase_dur;
  proc.dt = dt;
  proc.set_factor(p->counting_factor);
  proc.t = &tt;
  proc.sync = sync;
  proc.contacts = &contacts;
  proc.simul_box = &box;
  proc.r = R;
  proc.free_pairs = &free_pairs;
  proc.removed = &removed;
  proc.V = &V;
  proc.F = Fv;
  proc.disulfide_special_criteria = false;
  proc.fixed_cutoff = p->nl_cutoff;
  proc.ljs = ljs;

  qa::finish_processing fin;
  fin.candidates = &cands;
  fin.sync = sync;
  fin.t = &tt;
  fin.contacts = &contacts;
  fin.free_pairs = &free_pairs;
  fin.removed = &removed;
  fin.num_contacts = num_contacts;
  fin.disulfide_special_criteria = false;

  loop.for_slice(0, loop.total_size());
  /* the defined part of process_contacts' free_pairs-sized loop */
  proc.for_slice(0, std::min(free_pairs.size(), contacts.size()));
  fin();

  *n_free = free_pairs.size();
  for (int a = 0; a < *n_free && a < cap; ++a) {
    free_ij[2 * a] = free_pairs[a].item().i1();
    free_ij[2 * a + 1] = free_pairs[a].item().i2();
    free_d[a] = free_pairs[a].item().orig_dist();
    free_vac[a] = free_pairs[a].is_vacant();
  }
  *n_cont = contacts.size();
  for (int a = 0; a < *n_cont && a < cap; ++a) {
    auto c = contacts[a].item();
    cont_ij[2 * a] = c.i1();
    cont_ij[2 * a + 1] = c.i2();
    cont_d[a] = c.orig_dist();
    cont_type[a] = (int16_t)(qa::contact_type)c.type();
    cont_sat[a] = c.saturation();
    cont_vac[a] = contacts[a].is_vacant();
  }
  for (int i = 0; i < n; ++i) {
    sync_data s = sync[i];
    sync_io[4 * i] = s.back();
    sync_io[4 * i + 1] = s.side_all();
    sync_io[4 * i + 2] = s.side_polar();
    sync_io[4 * i + 3] = s.side_hydrophobic();
  }
  from_vec(Fv, F);
  return V;
}


// wall::harmonic::eval_free over the harmonic planes of the parameters;
// wall_F [6][3] by plane index
double ref_hw_free(int n, const double *r, const ora_params *p, const uint8_t *is_conn,
                   double *F, double *wall_F) {
  auto R = to_vec(n, r);
  int kidx[6];
  auto planes = planes_of_kind(p, 1, kidx);
  vect::vector<vec3r> Fv(n), wf((int)planes.size());
  vect::vector<bool> conn(n);
  for (int i = 0; i < n; ++i) conn[i] = is_conn[i] != 0;
  vect::vector<wall::harmonic::wall> walls;
  for (int k = 0; k < (int)planes.size(); ++k)
    walls.emplace_back(plane_of(p, planes[k]), vect::iterator<vec3r>(wf[k]), (real)1.0,
                       p->wall_limit[planes[k]]);
  real V = 0.0;
  wall::harmonic::eval_free e;
  e.depth = p->wall_depth;
  e.min_dist = p->wall_min_dist;
  e.is_connected = conn;
  e.walls = walls;
  e.r = R;
  e.F = Fv;
  e.wall_F = wf;
  e.V = &V;
  run(e);
  from_vec(Fv, F);
  put_wall_F(wf, planes, wall_F);
  return V;
}

// wall::harmonic::eval_connected over k connections (plane index per connection)
double ref_hw_conn(int n, const double *r, const ora_params *p, int k, const int *res,
                   const int *wall_plane, const double *off, const double *nat, double *F,
                   double *wall_F) {
  auto R = to_vec(n, r);
  int kidx[6];
  auto planes = planes_of_kind(p, 1, kidx);
  vect::vector<vec3r> Fv(n), wf((int)planes.size());
  vect::vector<wall::harmonic::wall> walls;
  for (int a = 0; a < (int)planes.size(); ++a)
    walls.emplace_back(plane_of(p, planes[a]), vect::iterator<vec3r>(wf[a]), (real)1.0,
                       p->wall_limit[planes[a]]);
  vect::vector<wall::harmonic::connection> conns;
  for (int a = 0; a < k; ++a)
    conns.emplace_back(kidx[wall_plane[a]], res[a],
                       vec3r(off[3 * a], off[3 * a + 1], off[3 * a + 2]), nat[a]);
  real V = 0.0;
  wall::harmonic::eval_connected e;
  e.HH1 = p->wall_hh1;
  e.walls = walls;
  e.conns = conns;
  e.r = R;
  e.F = Fv;
  e.wall_F = wf;
  e.V = &V;
  run(e);
  from_vec(Fv, F);
  put_wall_F(wf, planes, wall_F);
  return V;
}

// wall::lj::sift_free: forces, wall_F and the candidates it emits
// (residue, plane index, distance; at most cap)
double ref_ljw_sift(int n, const double *r, const ora_params *p, const uint8_t *is_conn,
                    const int *num_conns, double *F, double *wall_F, int cap, int *cand_res,
                    int *cand_plane, double *cand_d, int *ncand) {
  auto R = to_vec(n, r);
  int kidx[6];
  auto planes = planes_of_kind(p, 2, kidx);
  vect::vector<vec3r> Fv(n), wf((int)planes.size());
  vect::vector<bool> conn(n);
  for (int i = 0; i < n; ++i) conn[i] = is_conn[i] != 0;
  vect::vector<wall::lj::wall> walls;
  for (int a = 0; a < (int)planes.size(); ++a) {
    auto &w = walls.emplace_back(plane_of(p, planes[a]), vect::iterator<vec3r>(wf[a]),
                                 (real)1.0, p->wall_limit[planes[a]]);
    w.num_conns = num_conns[planes[a]];
  }
  vect::vector<wall::lj::candidate> cands;
  vect::vector<int> removed;
  real V = 0.0;
  wall::lj::sift_free e;
  e.min_dist = p->wall_min_dist;
  e.force.depth = p->wall_lj_depth;
  e.force.r_min = p->wall_min_dist;
  e.walls = walls;
  e.is_connected = conn;
  e.candidates = &cands;
  e.removed = &removed;
  e.r = R;
  e.F = Fv;
  e.wall_F = wf;
  e.V = &V;
  run(e);
  from_vec(Fv, F);
  put_wall_F(wf, planes, wall_F);
  *ncand = cands.size();
  for (int a = 0; a < cands.size() && a < cap; ++a) {
    auto c = cands[a];
    cand_res[a] = c.res_idx();
    cand_plane[a] = planes[c.wall_idx()];
    cand_d[a] = c.dist();
  }
  return V;
}

// wall::lj::eval_connected over k connection slots (vacant ones included):
// forces, wall_F, the updated saturations (in place) and the released slots
double ref_ljw_conn(int n, const double *r, const ora_params *p, int k, const int *res,
                    const int *wall_plane, const double *off, double *sat, const uint8_t *vac,
                    double *F, double *wall_F, int cap, int *removed_out, int *nremoved) {
  auto R = to_vec(n, r);
  int kidx[6];
  auto planes = planes_of_kind(p, 2, kidx);
  vect::vector<vec3r> Fv(n), wf((int)planes.size());
  vect::vector<bool> conn(n, false);
  vect::vector<wall::lj::wall> walls;
  for (int a = 0; a < (int)planes.size(); ++a)
    walls.emplace_back(plane_of(p, planes[a]), vect::iterator<vec3r>(wf[a]), (real)1.0,
                       p->wall_limit[planes[a]]);
  vect::set<wall::lj::connection> conns;
  for (int a = 0; a < k; ++a) {
    conns.emplace(res[a], kidx[wall_plane[a]], vec3r(off[3 * a], off[3 * a + 1], off[3 * a + 2]),
                  sat[a]);
    if (vac[a]) conns.remove(a);
  }
  vect::vector<int> removed;
  real V = 0.0;
  wall::lj::eval_connected e;
  e.min_dist = p->wall_min_dist;
  e.force.depth() = p->wall_lj_depth;
  e.force.r_min() = p->wall_min_dist;
  e.breaking_dist = p->ljw_breaking_dist;
  e.saturation_diff = p->ljw_saturation_diff;
  e.walls = walls;
  e.conns = &conns;
  e.r = R;
  e.F = Fv;
  e.wall_F = wf;
  e.V = &V;
  e.is_connected = conn;
  e.removed = &removed;
  run(e);
  from_vec(Fv, F);
  put_wall_F(wf, planes, wall_F);
  for (int a = 0; a < k; ++a) sat[a] = conns[a].item().saturation();
  *nremoved = removed.size();
  for (int a = 0; a < removed.size() && a < cap; ++a) removed_out[a] = removed[a];
  return V;
}

// wall::lj::process_candidates on the given slots, releases and candidates;
// the slots after it go to the out arrays (at most cap), is_conn and
// num_conns (by plane) are updated in place
void ref_ljw_process(int n, const double *r, const ora_params *p, int k, const int *res,
                     const int *wall_plane, const double *off, const double *sat,
                     const uint8_t *vac, int nrem, const int *rem, int ncand,
                     const int *cand_res, const int *cand_plane, const double *cand_d,
                     uint8_t *is_conn, int *num_conns, int cap, int *nslots, int *res_out,
                     int *plane_out, double *off_out, double *sat_out, uint8_t *vac_out) {
  auto R = to_vec(n, r);
  int kidx[6];
  auto planes = planes_of_kind(p, 2, kidx);
  vect::vector<vec3r> wf((int)planes.size());
  vect::vector<bool> conn(n);
  for (int i = 0; i < n; ++i) conn[i] = is_conn[i] != 0;
  vect::vector<wall::lj::wall> walls;
  for (int a = 0; a < (int)planes.size(); ++a) {
    auto &w = walls.emplace_back(plane_of(p, planes[a]), vect::iterator<vec3r>(wf[a]),
                                 (real)1.0, p->wall_limit[planes[a]]);
    w.num_conns = num_conns[planes[a]];
  }
  vect::set<wall::lj::connection> conns;
  for (int a = 0; a < k; ++a) {
    conns.emplace(res[a], kidx[wall_plane[a]], vec3r(off[3 * a], off[3 * a + 1], off[3 * a + 2]),
                  sat[a]);
    if (vac[a]) conns.remove(a);
  }
  vect::vector<int> removed;
  for (int a = 0; a < nrem; ++a) removed.push_back(rem[a]);
  vect::vector<wall::lj::candidate> cands;
  for (int a = 0; a < ncand; ++a) cands.emplace_back(cand_res[a], kidx[cand_plane[a]], cand_d[a]);
  wall::lj::process_candidates e;
  e.r = R;
  e.walls = walls;
  e.conns = &conns;
  e.removed = &removed;
  e.candidates = &cands;
  e.is_connected = conn;
  e();
  for (int i = 0; i < n; ++i) is_conn[i] = conn[i] ? 1 : 0;
  for (int a = 0; a < (int)planes.size(); ++a) num_conns[planes[a]] = walls[a].num_conns;
  *nslots = conns.size();
  for (int a = 0; a < conns.size() && a < cap; ++a) {
    auto c = conns[a].item();
    res_out[a] = c.res_idx();
    plane_out[a] = planes[c.wall_idx()];
    vec3r o = c.bead_offset();
    off_out[3 * a] = o.x(), off_out[3 * a + 1] = o.y(), off_out[3 * a + 2] = o.z();
    sat_out[a] = c.saturation();
    vac_out[a] = conns[a].is_vacant() ? 1 : 0;
  }
}

// wall::log_forces over nsteps samples: walls on the sides present[6] (the
// plane list of the reference's state::walls, then the extra sides appended
// by setup_walls for LJ axes), with the given normals, shifts and per-step
// reaction forces F[nsteps][6][3]; returns avg_F [6][3], its has-flags, work
// [6] and the z-force average and flag
void ref_log_forces(int nsteps, const double *t, const double *F, const double *normal,
                    const double *shift, const int *present, int nextra, const int *extra,
                    double window, double *avg_F, double *avg_has, double *work,
                    double *avg_z, double *avg_z_has) {
  vect::vector<vec3r> wf(6);
  std::vector<wall::gen_wall> walls;
  walls.reserve(6);
  for (int w = 0; w < 6; ++w) {
    vec3r nrm(normal[3 * w], normal[3 * w + 1], normal[3 * w + 2]);
    walls.emplace_back(cg::plane<real>(vec3r::Zero(), nrm), vect::iterator<vec3r>(wf[w]),
                       (real)window);
    walls[w].shift = vec3r(shift[3 * w], shift[3 * w + 1], shift[3 * w + 2]);
  }
  vect::vector<wall::gen_wall *> ptrs(6, nullptr);
  for (int w = 0; w < 6; ++w)
    if (present[w]) ptrs[w] = &walls[w];
  for (int a = 0; a < nextra; ++a) ptrs.push_back(&walls[extra[a]]);
  moving_avg<real, real> avg_zf((real)window);
  real tt = 0.0;
  wall::log_forces e;
  e.t = &tt;
  e.walls = ptrs;
  e.neg_z = &walls[4];
  e.pos_z = &walls[5];
  e.avg_z_force = &avg_zf;
  for (int k = 0; k < nsteps; ++k) {
    tt = t[k];
    for (int w = 0; w < 6; ++w)
      wf[w] = vec3r(F[18 * k + 3 * w], F[18 * k + 3 * w + 1], F[18 * k + 3 * w + 2]);
    e();
  }
  for (int w = 0; w < 6; ++w) {
    auto const &a = walls[w].avg_F;
    avg_has[w] = a->has_value() ? 1.0 : 0.0;
    vec3r v = a->has_value() ? a->value() : vec3r::Zero();
    avg_F[3 * w] = v.x(), avg_F[3 * w + 1] = v.y(), avg_F[3 * w + 2] = v.z();
    work[w] = walls[w].work;
  }
  *avg_z_has = avg_zf->has_value() ? 1.0 : 0.0;
  *avg_z = avg_zf->has_value() ? avg_zf->value() : 0.0;
}

} // extern "C"
