/*
 * oracle.h -- C ABI of the FP64 CPU oracle for the pas-cg per-step loop.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only to check results.
 * The product (pas-cg_amd/) never links, loads or calls it.
 *
 * The oracle restates, in plain FP64 C++, the reference loop
 * thread::advance_by_step (cg/src/cg/simul/thread_main.cpp:373-536) with one
 * thread: NL verify/rebuild, the ordered list updates, every force term in the
 * reference term order (cg/src/cg/simul/thread_setup.cpp:27-48), the QA
 * finish, and the Gear-5 Langevin step with Numerical Recipes ran2 noise.
 * All quantities are in the reference's internal units (cg/include/cg/utils/units.h:
 * length 5 A, time tau, energy eps).
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Term ids, in the reference's evaluation order (thread_setup.cpp:27-48);
 * the wall terms: solid, harmonic free and connected (wall::harmonic::
 * eval_free, eval_connected), LJ free and connected (wall::lj::sift_free,
 * eval_connected). */
enum ora_term {
  ORA_CHIR = 0,
  ORA_TETHER,
  ORA_LREP,
  ORA_NAT_CONT,
  ORA_PAULI,
  ORA_CONST_DH,
  ORA_REL_DH,
  ORA_QA_CAND,
  ORA_QA_CONT,
  ORA_PID,
  ORA_NAT_ANG,
  ORA_HEUR_ANG,
  ORA_HEUR_DIH,
  ORA_CND,
  ORA_SND,
  ORA_VEL_AFM,
  ORA_FORCE_AFM,
  ORA_SOLID_WALL,
  ORA_HW_FREE,
  ORA_HW_CONN,
  ORA_LJW_FREE,
  ORA_LJW_CONN,
  ORA_NUM_TERMS
};

/* Force-field parameters in internal units. One flat struct so that ctypes
 * can fill it; the field meaning follows cg/default.yml. */
typedef struct ora_params {
  double rep_cutoff;      /* general.repulsive cutoff */
  double counting_factor; /* general.counting factor */
  double nl_cutoff;       /* neighbor list.cutoff (before max with DH cutoff) */
  double nl_pad;          /* neighbor list.pad */

  int32_t tether_on;
  double tether_H1, tether_H2;

  int32_t lrep_on;
  double lrep_depth;

  int32_t chir_on;
  double e_chi;

  int32_t nat_ang_on;
  double CBA;

  int32_t heur_ang_on;
  double heur_ang_poly[7][9]; /* [degree][aa_heur_pair] */

  int32_t nat_dih_on;
  int32_t nat_dih_variant; /* 0 complex, 1 simple */
  double CDA, CDB, CDH;

  int32_t heur_dih_on;
  double heur_dih_coeffs[6][9]; /* const, sin, cos, sin2, cos2, sin_cos */

  int32_t pauli_on;
  double pauli_depth;

  int32_t nat_cont_on;
  double nat_cont_depth;
  int32_t nc_ss_kind; /* 0 none, 1 harmonic, 2 lj */
  double nc_ss_H1, nc_ss_H2, nc_ss_nat_r, nc_ss_depth, nc_ss_r_min;

  int32_t dh_on;
  int32_t dh_variant; /* 0 constant, 1 relative */
  double dh_screen_dist, dh_cutoff, dh_perm; /* perm: permittivity or factor */

  int32_t qa_on;
  int32_t qa_include4;
  double qa_phase_dur, qa_formation_tol;
  double qa_min_cos_hr, qa_min_cos_hh, qa_max_cos_nr;
  /* sink-LJ per contact type: 0 NONE, 1 bb, 2 bs, 3 sb, 4 + a1 + 21*a2 ss */
  double qa_lj_depth[445], qa_lj_r_low[445], qa_lj_r_high[445];

  int32_t pid_on;
  int32_t pid_include4;
  int32_t pid_lambda_variant; /* 0 cosine, 1 algebraic */
  double pid_bbp_alpha, pid_bbp_psi0, pid_bbm_alpha, pid_bbm_psi0;
  double pid_ss_alpha, pid_ss_psi0;
  double pid_bbp_lj[3], pid_bbm_lj[3]; /* depth, r_low, r_high */
  double pid_ss_depth[441], pid_ss_r_low[441], pid_ss_r_high[441];

  int32_t afm_on;
  double afm_H1, afm_H2;

  int32_t lang_on;
  double gamma, temperature, dt;

  double mass[21];
  double charge[21];
  int32_t ptype[21];       /* 0 none, 1 polar, 2 hydrophobic */
  int32_t base_sync[21][4]; /* back, side_all, side_polar, side_hydrophobic */
  /* walls: planes (origin, normal), solid depth, threshold */
  int32_t n_walls;
  double wall_origin[6][3], wall_normal[6][3];
  double wall_depth, wall_min_dist;
  /* per plane: kind (0 solid, 1 harmonic, 2 lj), box side (0..5 = -x +x -y
   * +y -z +z, -1 none), connection limit */
  int32_t wall_kind[6], wall_side[6], wall_limit[6];
  int32_t hw_on, ljw_on; /* harmonic_walls_enabled, lj_walls_enabled */
  double wall_hh1, wall_lj_depth, ljw_breaking_dist, ljw_saturation_diff;
  double afm_stats_window; /* afm.average stats over */
} ora_params;

/* sbox parameters of the box protocol (the same fields as pcg_box). */
typedef struct ora_box {
  double box_min[3], box_max[3];
  int32_t pbc[3];
  double total_time;
  double rest_period, target_vel, accel_dist, avg_forces_over;
  int32_t squeezing;
  double target_density, vel_above_2V;
  int32_t force_min;
  double force_for_max_vel;
  int32_t oscillations;
  int32_t osc_shear;
  int32_t num_of_cycles;
  double angular_freq;
  int32_t amp_relative;
  double amp_abs, amp_rel;
} ora_box;

/* Input model, already compiled to residue indices (state::compile_model,
 * cg/src/cg/simul/state.cpp:144-249). Angles and dihedrals carry an optional
 * native value (has_* flag), as in cg/include/cg/input/model.h:313-323. */
typedef struct ora_model {
  int32_t n;
  const double *pos;      /* [3n] xyz, starting positions (model) */
  const double *orig_pos; /* [3n] xyz, orig_model positions (tether lengths) */
  const uint8_t *atype;   /* [n] */
  const int32_t *chain_idx, *seq_idx; /* [n] */
  int32_t n_teth;
  const int32_t *teth; /* [2 n_teth] */
  int32_t n_ang;
  const int32_t *ang; /* [3 n_ang] */
  const uint8_t *ang_has;
  const double *ang_theta;
  int32_t n_dih;
  const int32_t *dih; /* [4 n_dih] */
  const uint8_t *dih_has;
  const double *dih_phi;
  int32_t n_cont;
  const int32_t *cont; /* [2 n_cont] */
  const double *cont_len;
  const int32_t *cont_type; /* nat_cont::type: 5 == SSBOND */
  double cell[3];           /* PBC cell; 0 disables an axis */
  int32_t n_tips;           /* velocity AFM tips */
  const int32_t *tip_res;
  const double *tip_orig, *tip_vel; /* [3 n_tips] */
  int32_t n_ftips;                  /* constant-force AFM tips */
  const int32_t *ftip_res;
  const double *ftip_force;         /* [3 n_ftips] */
} ora_model;

typedef struct ora_saw {
  int32_t perform;
  double bond;          /* bond distance */
  double intersection;  /* intersection at */
  int32_t with_box;     /* start box given by size */
  double box_size;
  int32_t pbc;
  int32_t update_orig;  /* orig positions follow the SAW (seq-file input) */
} ora_saw;

typedef struct ora_sys ora_sys;

/* Builds the trajectory state exactly as state::traj_init does
 * (state.cpp:83-112): optional SAW morph, Langevin initial velocities from the
 * ran2 stream seeded with `seed`, then every per-term list. The SAW result is
 * written back to pos_out (if non-null). */
ora_sys *ora_create(const ora_params *p, const ora_model *m, int32_t seed,
                    const ora_saw *saw, double *pos_out);
void ora_destroy(ora_sys *s);

/* One full step (pre_eval, eval_forces, post_eval). V of the evaluation is
 * available from ora_last_V afterwards, as print_raw_data records it. */
void ora_step(ora_sys *s, int32_t nsteps, double *V_trace);
double ora_last_V(const ora_sys *s);
void ora_last_term_V(const ora_sys *s, double *out /* [ORA_NUM_TERMS] */);

/* Pieces of the step for per-call parity. */
int32_t ora_nl_verify(ora_sys *s);  /* returns nl_invalid after verify */
void ora_fix_nl(ora_sys *s);         /* rebuild + list updates */
void ora_eval_forces(ora_sys *s);    /* zero F,V then every enabled term */
void ora_eval_term(ora_sys *s, int32_t term, double *F /*[3n]*/, double *V);
void ora_qa_finish(ora_sys *s);
void ora_lang_step(ora_sys *s);
void ora_prepare_nh(ora_sys *s);

/* State access (xyz interleaved). */
void ora_get_r(const ora_sys *s, double *r);
void ora_set_r(ora_sys *s, const double *r); /* also sets y0 */
void ora_get_F(const ora_sys *s, double *F);
void ora_get_y(const ora_sys *s, int32_t k, double *y);
void ora_set_y(ora_sys *s, int32_t k, const double *y);
double ora_get_t(const ora_sys *s);
void ora_set_t(ora_sys *s, double t);
void ora_set_noise_scale(ora_sys *s, double scale); /* 0 turns noise off */

/* List access. kind: 0 NL, 1 nat_cont current, 2 pauli, 3 dh, 4 pid,
 * 5 qa free pairs (incl. vacant), 6 qa contacts (incl. vacant).
 * Writes i1,i2 (and a per-kind value) for up to cap items, returns size. */
int32_t ora_list(const ora_sys *s, int32_t kind, int32_t cap, int32_t *i1,
                 int32_t *i2, double *val, uint8_t *flag);
/* Static lists as built by the state setup. kind: 0 tethers (v0 nat_dist),
 * 1 local_rep pairs, 2 chiral quads (v0 nat_factor field, v1 nat_chir field,
 * as stored), 3 native angles (v0 theta), 4 heuristic angles (ity type),
 * 5 native dihedrals (v0 phi), 6 heuristic dihedrals (ity type),
 * 7 all native contacts (v0 nat_dist, ity type). idx holds 4 per item. */
int32_t ora_static_list(const ora_sys *s, int32_t kind, int32_t cap,
                        int32_t *idx, double *v0, double *v1, int32_t *ity);
/* QA contacts with every field (sync diffs as 8 int8 per contact). */
int32_t ora_qa_contacts(const ora_sys *s, int32_t cap, int32_t *ij,
                        double *orig, int32_t *type, double *sat, int8_t *sd,
                        uint8_t *vac);
/* Native contacts (all): active flags and change times. */
void ora_nat_cont_state(const ora_sys *s, uint8_t *active, double *change_t);
int32_t ora_num_changed(const ora_sys *s);
int32_t ora_num_qa_contacts(const ora_sys *s);
void ora_sync(const ora_sys *s, int8_t *sync /*[4n]*/);
int32_t ora_nl_rebuilds(const ora_sys *s);
/* pid::count_active over the current PID bundles and positions:
 * {intra, intra_bb, intra_ss, inter, inter_bb, inter_ss} */
void ora_pid_counts(const ora_sys *s, int32_t *out /*[6]*/);
/* replaces the AFM tips; their reference time becomes the current time */
void ora_set_tips(ora_sys *s, int32_t n_tips, const int32_t *tip_res, const double *tip_orig,
                  const double *tip_vel, int32_t n_ftips, const int32_t *ftip_res,
                  const double *ftip_force);
int64_t ora_saw_draws(const ora_sys *s); /* ran2 draws the SAW consumed */

/* The simulation box and its protocol (state::setup_walls and
 * reinit_wall_values, state.cpp:570-730; thread_main.cpp:144-330). After
 * ora_box_start every ora_step first runs the phase logic, then advances.
 * ora_box_state fills the layout of pcg_box_state (120 doubles). */
void ora_box_setup(ora_sys *s, const ora_box *b);
void ora_box_start(ora_sys *s);
void ora_box_state(const ora_sys *s, double *out);
/* wall::lj::process_candidates over the candidates and releases gathered
 * since the last call (then cleared) */
void ora_ljw_process(ora_sys *s);
/* every LJ connection slot (vacant ones too), and the LJ bookkeeping:
 * pending candidates and releases, the connected flags [n], the connection
 * counts by plane [6] */
int32_t ora_ljw_slots(const ora_sys *s, int32_t cap, int32_t *res, int32_t *wall, double *off,
                      double *sat, uint8_t *vac);
void ora_ljw_pending(const ora_sys *s, int32_t cap, int32_t *ncand, int32_t *cand_res,
                     int32_t *cand_wall, double *cand_d, int32_t *nrem, int32_t *rem,
                     uint8_t *is_conn, int32_t *num_conns);
/* AFM tip statistics: per velocity tip (first n_tips entries) the moving
 * average of the force on the tip and of its perpendicular part; per force
 * tip (after them) the moving average of the pulled residue's velocity;
 * has = the average has a value. avg [3 * tips], perp [tips], has [tips]. */
int32_t ora_afm_stats(const ora_sys *s, double *avg, double *perp, uint8_t *has);
/* wall connections, kind 0 harmonic, 1 lj (live ones in slot order) */
int32_t ora_wall_conns(const ora_sys *s, int32_t kind, int32_t cap, int32_t *res,
                       int32_t *wall, double *off, double *sat);

/* The ran2 generator by itself (cg/include/cg/utils/random/nr.h). */
typedef struct ora_rng ora_rng;
ora_rng *ora_rng_new(int32_t seed);
void ora_rng_free(ora_rng *g);
double ora_rng_uniform(ora_rng *g);
double ora_rng_normal(ora_rng *g);

#ifdef __cplusplus
}
#endif
