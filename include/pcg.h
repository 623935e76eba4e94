/*
 * pcg.h -- C ABI of the MI355X-native per-step loop of the PAS CG model.
 *
 * Drop-in boundary for the reference's per-step loop
 * (cg/src/cg/simul/thread_main.cpp:373-536). Every entry point returns an int
 * status (PCG_OK or a negative PCG_E* code); the message of the last failure
 * is available from pcg_last_error(). No exception crosses the ABI. Buffers
 * are caller-owned on the host side and context-owned on the device side.
 * A context (or batch, or server) is bound to one HIP stream of one device
 * and is not re-entrant; separate objects may be used from separate threads.
 *
 * Units are the reference's internal units (cg/include/cg/utils/units.h):
 * length 1 = 5 A, time 1 = tau, energy 1 = eps. Arrays of 3-vectors are
 * xyz-interleaved on the host side ([3n]).
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  PCG_OK = 0,
  PCG_E_ARG = -1,      /* invalid argument */
  PCG_E_HIP = -2,      /* HIP runtime failure */
  PCG_E_NOMEM = -3,    /* device allocation or budget refused */
  PCG_E_OVERFLOW = -4, /* a list outgrew its capacity */
  PCG_E_STATE = -5,    /* call not valid in the current state */
  PCG_E_NOTFOUND = -6
};

/* Force terms in the reference's evaluation order
 * (cg/src/cg/simul/thread_setup.cpp:27-48). */
enum pcg_term {
  PCG_CHIR = 0,
  PCG_TETHER,
  PCG_LREP,
  PCG_NAT_CONT,
  PCG_PAULI,
  PCG_CONST_DH,
  PCG_REL_DH,
  PCG_QA_CAND,
  PCG_QA_CONT,
  PCG_PID,
  PCG_NAT_ANG,
  PCG_HEUR_ANG,
  PCG_HEUR_DIH,
  PCG_CND,
  PCG_SND,
  PCG_VEL_AFM,
  PCG_FORCE_AFM, /* constant-force AFM tips (afm/force/eval_forces.cpp:10-13) */
  PCG_SOLID_WALL, /* solid walls (wall/solid/eval_forces.cpp:4-17) */
  PCG_HW_FREE,    /* harmonic walls, free residues (wall/harmonic/eval_free.cpp:15-31) */
  PCG_HW_CONN,    /* harmonic wall connections (wall/harmonic/eval_connected.cpp:17-29) */
  PCG_LJW_FREE,   /* LJ walls, free residues (wall/lj/sift_free.cpp:15-37) */
  PCG_LJW_CONN,   /* LJ wall connections (wall/lj/eval_connected.cpp:15-41) */
  PCG_NUM_TERMS
};

/* Force-field parameters (cg/default.yml sections, internal units). The
 * field meaning matches the parameters bound in thread::setup_<term>
 * (cg/src/cg/simul/thread_setup.cpp:122-720). */
typedef struct pcg_params {
  double rep_cutoff;      /* general.repulsive cutoff */
  double counting_factor; /* general.counting factor */
  double nl_cutoff;       /* neighbor list.cutoff; rebuild uses max(this, dh_cutoff) */
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
  double heur_ang_poly[7][9];
  int32_t nat_dih_on;
  int32_t nat_dih_variant; /* 0 complex, 1 simple */
  double CDA, CDB, CDH;
  int32_t heur_dih_on;
  double heur_dih_coeffs[6][9];
  int32_t pauli_on;
  double pauli_depth;
  int32_t nat_cont_on;
  double nat_cont_depth;
  int32_t nc_ss_kind; /* 0 none, 1 harmonic, 2 lj */
  double nc_ss_H1, nc_ss_H2, nc_ss_nat_r, nc_ss_depth, nc_ss_r_min;
  int32_t dh_on;
  int32_t dh_variant; /* 0 constant, 1 relative */
  double dh_screen_dist, dh_cutoff, dh_perm;
  int32_t qa_on;
  int32_t qa_include4;
  double qa_phase_dur, qa_formation_tol;
  double qa_min_cos_hr, qa_min_cos_hh, qa_max_cos_nr;
  double qa_lj_depth[445], qa_lj_r_low[445], qa_lj_r_high[445];
  int32_t pid_on;
  int32_t pid_include4;
  int32_t pid_lambda_variant; /* 0 cosine, 1 algebraic */
  double pid_bbp_alpha, pid_bbp_psi0, pid_bbm_alpha, pid_bbm_psi0;
  double pid_ss_alpha, pid_ss_psi0;
  double pid_bbp_lj[3], pid_bbm_lj[3];
  double pid_ss_depth[441], pid_ss_r_low[441], pid_ss_r_high[441];
  int32_t afm_on;
  double afm_H1, afm_H2;
  int32_t lang_on;
  double gamma, temperature, dt;
  double mass[21];
  double charge[21];
  int32_t ptype[21];
  int32_t base_sync[21][4];
  /* walls (state::setup_walls, state.cpp:570-700): n_walls planes through
   * origin with normal; n_walls 0 turns them off. wall_depth is
   * sbox.walls.solid wall.depth (the d^-10 force of solid walls and of
   * residues free of harmonic walls) and wall_min_dist sbox.walls.threshold.
   * Batches and servers take solid walls only, at fixed planes, shared by
   * every system. */
  int32_t n_walls;
  double wall_origin[6][3], wall_normal[6][3];
  double wall_depth, wall_min_dist;
  /* per plane: kind (0 solid, 1 harmonic, 2 Lennard-Jones), the side of the
   * simulation box it stands on (0..5: -x, +x, -y, +y, -z, +z; -1 none, a
   * plane the box protocol leaves in place) and its connection limit
   * (sbox.walls.<kind> wall.connection limit, "auto" resolved as
   * state.cpp:613-620 does). The planes of one axis come as the reference
   * creates them: the -axis side, then the +axis side. */
  int32_t wall_kind[6], wall_side[6], wall_limit[6];
  /* state::harmonic_walls_enabled and lj_walls_enabled: the reference's
   * trajectory set-up clears both right after setup_walls (state.cpp:111),
   * so its trajectories evaluate neither; 1 evaluates them */
  int32_t hw_on, ljw_on;
  double wall_hh1;          /* sbox.walls.harmonic wall.HH1 */
  double wall_lj_depth;     /* sbox.walls.lj wall.depth */
  double ljw_breaking_dist; /* breaking dist factor * 2^(-1/6) * threshold */
  double ljw_saturation_diff; /* dt / lj wall.cycle duration (thread_setup.cpp:787-790) */
  double afm_stats_window;    /* afm.average stats over: the tips' moving-average window */
} pcg_params;

/* One protein system after state::compile_model
 * (cg/src/cg/simul/state.cpp:144-249); see the reference model
 * cg/include/cg/input/model.h:268-329. */
typedef struct pcg_model {
  int32_t n;
  const double *pos;      /* [3n] starting positions */
  const double *orig_pos; /* [3n] positions the tether lengths come from */
  const uint8_t *atype;
  const int32_t *chain_idx, *seq_idx;
  int32_t n_teth;
  const int32_t *teth; /* [2k] */
  int32_t n_ang;
  const int32_t *ang; /* [3k] */
  const uint8_t *ang_has;
  const double *ang_theta;
  int32_t n_dih;
  const int32_t *dih; /* [4k] */
  const uint8_t *dih_has;
  const double *dih_phi;
  int32_t n_cont;
  const int32_t *cont; /* [2k] */
  const double *cont_len;
  const int32_t *cont_type; /* 5 = SSBOND */
  double cell[3];
  int32_t n_tips;
  const int32_t *tip_res;
  const double *tip_orig, *tip_vel;
  /* constant-force AFM tips: residue and pulling force [3k]
   * (afm::force::tip, state.cpp:563-566) */
  int32_t n_ftips;
  const int32_t *ftip_res;
  const double *ftip_force;
} pcg_model;

enum { PCG_MIXED = 0, PCG_FP64 = 1 };               /* force arithmetic */
enum { PCG_NOISE_PHILOX = 0, PCG_NOISE_RAN2 = 1 }; /* Langevin noise source */

typedef struct pcg_options {
  int32_t precision;  /* PCG_MIXED: FP32 forces, FP64 integration */
  int32_t noise;      /* PHILOX (device, counter-based) or RAN2 (reference stream) */
  int32_t seed;       /* ran2 seed (initial velocities; noise in RAN2 mode) */
  uint64_t philox_seed;
  double temperature; /* <0: take params.temperature */
  int32_t device;
  int64_t ran2_skip;  /* ran2 draws consumed by input preparation (a SAW) */
} pcg_options;

const char *pcg_last_error(void); /* thread-local message of the last failure */
int pcg_device_count(int *count);

/* ---- plain single-system context (the drop-in path) ------------------- */
typedef struct pcg_ctx pcg_ctx;

/* state::traj_init (state.cpp:83-112) + thread::init_kernels: initial
 * velocities from ran2(seed) (state.cpp:270-335), every per-term list. */
int pcg_create(const pcg_params *p, const pcg_model *m, const pcg_options *o,
               pcg_ctx **out);
int pcg_destroy(pcg_ctx *c);

/* thread::advance_by_step (thread_main.cpp:373-377) repeated nsteps times */
int pcg_step(pcg_ctx *c, int32_t nsteps);

/* The pieces of one step, for per-call parity (thread_main.cpp:379-536). */
int pcg_nl_verify(pcg_ctx *c, int32_t *invalid);     /* nl::verify */
int pcg_fix_nl(pcg_ctx *c);                          /* legacy_update + updates */
int pcg_prepare_nh(pcg_ctx *c);                      /* qa::prepare_nh */
int pcg_eval_forces(pcg_ctx *c);                     /* thread::eval_forces */
/* one term over items [from, to) of its list -- sliceable_task::for_slice
 * (cg/include/cg/simul/runtime.h:17-21). F (3n, may be NULL) receives the
 * term's forces alone, V its energy. to < 0 means the whole list. */
int pcg_eval_term(pcg_ctx *c, int32_t term, int32_t from, int32_t to,
                  double *F, double *V);
/* pcg_eval_term with the forces in three component arrays (see pcg_get3):
 * for_slice writing a reference dyn.F view directly. Fx, Fy, Fz NULL: energy
 * only. */
int pcg_eval_term3(pcg_ctx *c, int32_t term, int32_t from, int32_t to, double *Fx, double *Fy,
                   double *Fz, int64_t stride, double *V);
/* total_size(); for PCG_QA_CAND the free pairs whose NL distance is within
 * the QA maximum required distance, the only ones loop_over_candidates
 * evaluates past its first test (qa/loop_over_candidates.cpp:20) */
int pcg_term_size(pcg_ctx *c, int32_t term, int32_t *size);
int pcg_qa_finish(pcg_ctx *c);                       /* qa::finish_processing */
int pcg_lang_step(pcg_ctx *c);                       /* lang::step */

/* State access. which: 0..5 Gear y0..y5 (y0 = positions), 6 forces,
 * 7 velocities. */
int pcg_get(pcg_ctx *c, int32_t which, double *out);
int pcg_set(pcg_ctx *c, int32_t which, const double *in);
/* The same with the three components in separate arrays, element i of
 * component c at comp[c][i * stride]: stride 1 takes the reference's own
 * vect::vector<vec3r> storage, a structure of arrays of x, y, z
 * (cg/include/cg/types/vec3_def.h:53 with INST/FIELD), without a transposing
 * copy (the device keeps the same layout); x, x + 1, x + 2 with stride 3 is
 * the interleaved layout. */
int pcg_get3(pcg_ctx *c, int32_t which, double *x, double *y, double *z, int64_t stride);
int pcg_set3(pcg_ctx *c, int32_t which, const double *x, const double *y, const double *z,
             int64_t stride);
int pcg_get_energy(pcg_ctx *c, double *V, double *term_V /*[PCG_NUM_TERMS]*/);
/* The scalars make_report::add_cur_scalars writes for the trajectory
 * (cg/src/cg/output/make_report.cpp:196-247), of the current state, indexed
 * by pcg_report_field: time, potential, total and kinetic energy, radius of
 * gyration, end-to-end distance, RMSD to the original positions after the
 * optimal superposition, asphericity, active native contacts (all and
 * disulfide) and QA contacts between and within chains, then the PID
 * activity counts of pid::count_active (cg/src/cg/pid/count_active.cpp:4-164,
 * columns PID-INTRHC, PID-B1-B1, PID-B1-S1, PID-INTEHC, PID-B1-B2, PID-S1-S2
 * of make_report.cpp:253-261; zero when PID is off). */
typedef enum {
  PCG_REP_TIME = 0, PCG_REP_EPOT, PCG_REP_ETOT, PCG_REP_EKIN, PCG_REP_RG, PCG_REP_L,
  PCG_REP_RMSD, PCG_REP_W, PCG_REP_ICN, PCG_REP_ICNSS, PCG_REP_INTEHC, PCG_REP_INTRHC,
  PCG_REP_PID_INTRHC, PCG_REP_PID_B1B1, PCG_REP_PID_B1S1,
  PCG_REP_PID_INTEHC, PCG_REP_PID_B1B2, PCG_REP_PID_S1S2,
  PCG_REPORT_NUM
} pcg_report_field;
int pcg_report(pcg_ctx *c, double *out /*[PCG_REPORT_NUM]*/);
int pcg_get_time(pcg_ctx *c, double *t, int64_t *step);
int pcg_set_time(pcg_ctx *c, double t);
int pcg_set_noise_scale(pcg_ctx *c, double scale); /* 0: no noise */
/* Protocol phases (thread_main.cpp:84-142): replaces the AFM tips, as
 * state::setup_afm places them when a phase turns AFM on
 * (state.cpp:548-568). n_tips velocity tips (residue, position, velocity
 * [3k]) and n_ftips force tips (residue, force [3k]) fill the tip slots the
 * model reserved at creation (n_tips + n_ftips of pcg_model); their
 * reference time becomes the current time, so a velocity tip sits at
 * orig + (t - t_now) vel. Zero tips turn the AFM terms off (a phase without
 * AFM). The AFM terms must be enabled in pcg_params (afm_on). */
int pcg_set_tips(pcg_ctx *c, int32_t n_tips, const int32_t *tip_res, const double *tip_orig,
                 const double *tip_vel, int32_t n_ftips, const int32_t *ftip_res,
                 const double *ftip_force);
/* The AFM tips' statistics (moving_avg over afm.average stats over,
 * cg/include/cg/types/avg.h:14-29), as make_report::add_afm_stuff reads
 * them (make_report.cpp:264-310): entry k < n_tips of pcg_set_tips is a
 * velocity tip, avg [3k] the average force on the tip (afm/vel/
 * eval_forces.cpp:23) and perp [k] the average of its part perpendicular to
 * the pulling direction (:25-29); the force tips follow, avg the average
 * velocity of the pulled residue (afm/force/eval_forces.cpp:12) and perp 0.
 * has [k] = the average has a value (a window has closed). Every force
 * evaluation with the AFM terms on adds a sample; new tips start over.
 * Writes up to cap entries, *size = the number of tips. */
int pcg_afm_stats(pcg_ctx *c, int32_t cap, int32_t *size, double *avg, double *perp,
                  uint8_t *has);
/* Lists. kind: 0 NL, 1 current native contacts, 2 pauli, 3 dh, 4 pid,
 * 5 qa free pairs (flag = vacant), 6 qa contacts (val = saturation,
 * flag = vacant), 7 qa contacts (val = the qa::contact_type value: 1
 * back-back, 2 back-side, 3 side-back, 4 + a1 + 21 a2 side-side,
 * cg/include/cg/qa/contact_type.h:9-24; flag = vacant). Returns the size;
 * fills up to cap items. The persistent step form keeps no NL: kind 0 after
 * its steps is PCG_E_STATE until pcg_fix_nl rebuilds in the stage form. */
int pcg_get_list(pcg_ctx *c, int32_t kind, int32_t cap, int32_t *size,
                 int32_t *i1, int32_t *i2, double *val, uint8_t *flag);
int pcg_nat_cont_state(pcg_ctx *c, uint8_t *active, double *change_t,
                       int32_t *num_changed);
int pcg_qa_state(pcg_ctx *c, int8_t *sync /*[4n]*/, int32_t *num_contacts);
int pcg_nl_rebuilds(pcg_ctx *c, int32_t *count);
/* HIP stream the context launches on (hipStream_t as void*) */
int pcg_stream(pcg_ctx *c, void **stream);
/* Capture one full step as a hipGraph and replay it for later pcg_step
 * calls (steps without a NL rebuild); 0 turns it off. */
int pcg_use_graph(pcg_ctx *c, int32_t on);
/* How pcg_step advances: PCG_FORM_STAGES, one launch per stage of the
 * reference's step (what the per-call entry points above run); or
 * PCG_FORM_PERSISTENT, the batched form's persistent kernel for this one
 * system (one workgroup, many steps a launch, positions, forces and lists
 * in LDS), possible when the system fits LDS, the noise is Philox and there
 * are no walls or box (PCG_E_STATE otherwise). PCG_FORM_AUTO (the default)
 * takes the persistent form whenever it is possible. The first step in one
 * form after steps in the other rebuilds the neighbour list, and so does the
 * first nl_verify after persistent steps; per-term calls right after
 * persistent steps see the lists of the persistent kernel's last rebuild,
 * which keeps every per-term list but not the NL itself. *active receives
 * the form the next pcg_step takes. No counterpart in the reference. */
enum { PCG_FORM_AUTO = 0, PCG_FORM_STAGES = 1, PCG_FORM_PERSISTENT = 2 };
int pcg_step_form(pcg_ctx *c, int32_t form, int32_t *active);
/* Checkpoint and resume of the context (the fields of the reference's state
 * serialisation, cg/src/cg/simul/state_io.cpp:117-156: positions, time,
 * velocities, Gear y0..y5, QA contacts, sync budgets; here with the lists,
 * the native-contact state and, in the RAN2 noise mode, the host ran2
 * stream, which the reference does not save). The buffer is host memory of
 * pcg_checkpoint_size bytes; restore needs a context created from the same
 * model, parameters, precision and noise source. */
int pcg_checkpoint_size(pcg_ctx *c, int64_t *bytes);
int pcg_checkpoint(pcg_ctx *c, void *buf, int64_t bytes);
int pcg_restore(pcg_ctx *c, const void *buf, int64_t bytes);

/* ---- simulation box, walls and the box protocol (plain context) --------
 * sbox parameters (cg/include/cg/sbox/parameters.h, the "simulation box"
 * section of cg/default.yml) that the protocol phases of
 * thread_main.cpp:144-330 read, in internal units. */
typedef struct pcg_box {
  double box_min[3], box_max[3]; /* state::box at the trajectory start (state.cpp:173-225) */
  int32_t pbc[3];                /* axes of type "periodic": cell = box extent */
  double total_time;             /* general.total time */
  double rest_period, target_vel, accel_dist, avg_forces_over; /* sbox.common */
  int32_t squeezing;             /* sbox.squeezing */
  double target_density, vel_above_2V;
  int32_t force_min;             /* sbox.finding force minimum */
  double force_for_max_vel;
  int32_t oscillations;          /* sbox.oscillations */
  int32_t osc_shear;             /* type "shear" (else "squeeze") */
  int32_t num_of_cycles;
  double angular_freq;
  int32_t amp_relative;          /* amplitude variant "relative" (else "absolute") */
  double amp_abs, amp_rel;
} pcg_box;

/* phases of the box protocol (simul::phase, thread_main.cpp:32-54); NONE:
 * the walls stay where they are */
enum {
  PCG_PHASE_NONE = 0, PCG_PHASE_SQUEEZING, PCG_PHASE_REST_AFTER_SQUEEZING,
  PCG_PHASE_FIND_FORCE_MIN, PCG_PHASE_REST_AFTER_FORCE_MIN, PCG_PHASE_MAX_AMPLITUDE,
  PCG_PHASE_REST_AFTER_MAX_AMP, PCG_PHASE_OSCILLATIONS, PCG_PHASE_REST_AFTER_OSCILLATIONS,
  PCG_PHASE_FREEFORM
};

/* pcg_box_state layout (doubles): phase; since; until; box min [3], max [3];
 * amplitude; omega; displacement; shear; whether the averaged z force has a
 * value, and the value; per plane [6][3]: origin, normal, reaction force of
 * the last evaluation (the wall's dyn.*_wall_F), shift of the last move;
 * per plane [6]: work; whether avg_F has a value; avg_F [6][3]; the PBC
 * cell [3]. */
enum {
  PCG_BOX_PHASE = 0, PCG_BOX_SINCE = 1, PCG_BOX_UNTIL = 2, PCG_BOX_MIN = 3, PCG_BOX_MAX = 6,
  PCG_BOX_AMPLITUDE = 9, PCG_BOX_OMEGA = 10, PCG_BOX_DISPLACEMENT = 11, PCG_BOX_SHEAR = 12,
  PCG_BOX_AVG_Z_HAS = 13, PCG_BOX_AVG_Z = 14, PCG_BOX_ORIGIN = 15, PCG_BOX_NORMAL = 33,
  PCG_BOX_WALL_F = 51, PCG_BOX_SHIFT = 69, PCG_BOX_WORK = 87, PCG_BOX_AVG_F_HAS = 93,
  PCG_BOX_AVG_F = 99, PCG_BOX_CELL = 117, PCG_BOX_STATE_NUM = 120
};

/* The box of the trajectory (state::traj_init's box and setup_walls,
 * state.cpp:570-730): the planes with a box side move to its faces
 * (reinit_wall_values), periodic axes take the box extent as their PBC
 * cell, and, when both z sides have a plane, every step logs the wall forces
 * (wall/log_forces.cpp:4-17) averaged over avg_forces_over. The phase is
 * NONE until pcg_box_start. */
int pcg_box_setup(pcg_ctx *c, const pcg_box *box);
/* equil_step's hand-over to the box protocol (thread_main.cpp:130-135):
 * from the current time, SQUEEZING until total_time. Each later pcg_step
 * step first runs the phase logic of thread_main.cpp:144-330 on the device
 * (wall moves, phase changes), then advances. */
int pcg_box_start(pcg_ctx *c);
int pcg_box_state(pcg_ctx *c, double *out /*[PCG_BOX_STATE_NUM]*/);
/* wall::lj::process_candidates (wall/lj/process_candidates.cpp:5-31): removes
 * the connections eval_connected released, then connects the candidates
 * sift_free found, nearest first. The reference never clears its candidate
 * list (state.cpp:580 is its only clear); this call consumes the candidates
 * it processes. */
int pcg_ljw_process(pcg_ctx *c);
/* Wall connections: kind 0 harmonic (made at creation from the starting
 * positions, state.cpp:653-670), 1 Lennard-Jones (live slots in creation
 * order). Per connection: residue, wall (plane index), bead offset [3] and
 * saturation (harmonic: the natural distance). Returns the count in *size;
 * fills up to cap. */
int pcg_wall_conns(pcg_ctx *c, int32_t kind, int32_t cap, int32_t *size, int32_t *res,
                   int32_t *wall, double *offset, double *sat);

/* ---- batched form: many systems, one persistent launch ---------------- */
typedef struct pcg_batch pcg_batch;
typedef struct pcg_batch_options {
  int32_t precision;
  uint64_t philox_seed; /* per-system streams: (seed, system index) */
  int32_t block;        /* threads per workgroup, multiple of 64 in [64, 1024]
                           (0: from the largest system; pcg_batch_autotune
                           measures) */
  int32_t steps_per_launch; /* steps fused in one persistent launch (0: 100) */
  int32_t use_graph;    /* replay each run's launch sequence as a hipGraph */
  int32_t device;
} pcg_batch_options;

/* temps[i] < 0 takes params.temperature; seeds: ran2 seed of each system's
 * initial velocities. */
int pcg_batch_create(const pcg_params *p, int32_t nsys, const pcg_model *models,
                     const double *temps, const int32_t *seeds,
                     const pcg_batch_options *o, pcg_batch **out);
int pcg_batch_destroy(pcg_batch *b);
int pcg_batch_run(pcg_batch *b, int32_t nsteps);
int pcg_batch_sync(pcg_batch *b);
int pcg_batch_set_noise_scale(pcg_batch *b, double scale); /* 0: no noise */
/* Energies (default on) are evaluated on the last step a launch makes for a
 * system. Off: no step evaluates them (pcg_batch_energy keeps the last values)
 * and every step of the mixed precision takes its FP32 force paths. No
 * counterpart in the reference, which sums V on every step. */
int pcg_batch_energies(pcg_batch *b, int32_t on);
/* The report scalars of every system, [nsys][PCG_REPORT_NUM] (pcg_report). */
int pcg_batch_report(pcg_batch *b, double *out);
/* pcg_batch_report without any device or host allocation. No counterpart in
 * the reference: for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). The
 * per-system sums go to a device array of the library, one per GPU, held
 * until the process ends; batches of at most 4096 systems (PCG_E_ARG above
 * that). Calls that overlap on one GPU share that array and race. */
int pcg_batch_report_static(pcg_batch *b, double *out);
/* Checkpoint and resume of a whole batch (replaces the state serialisation of
 * cg/src/cg/simul/state_io.cpp:117-140: positions, time, velocities, Gear
 * y0..y5, QA contacts, sync budgets; here with the lists, the native-contact
 * state and the noise position as well, so a resumed run continues the same
 * trajectory). The buffer is host memory of pcg_batch_checkpoint_size bytes;
 * restore needs a batch built from the same models, parameters and
 * precision. */
int pcg_batch_checkpoint_size(pcg_batch *b, int64_t *bytes);
int pcg_batch_checkpoint(pcg_batch *b, void *buf, int64_t bytes);
int pcg_batch_restore(pcg_batch *b, const void *buf, int64_t bytes);
int pcg_batch_get(pcg_batch *b, int32_t sys, int32_t which, double *out);
/* which: 0..5 Gear y0..y5 of one system; setting positions (0) makes its
 * next step rebuild the neighbour list */
int pcg_batch_set(pcg_batch *b, int32_t sys, int32_t which, const double *in);
/* pcg_set_tips for system sys of a batch (residue indices within the system) */
int pcg_batch_set_tips(pcg_batch *b, int32_t sys, int32_t n_tips, const int32_t *tip_res,
                       const double *tip_orig, const double *tip_vel, int32_t n_ftips,
                       const int32_t *ftip_res, const double *ftip_force);
int pcg_batch_energy(pcg_batch *b, int32_t sys, double *V, double *term_V);
int pcg_batch_info(pcg_batch *b, int32_t *block, int32_t *steps_per_launch,
                   int64_t *device_bytes);
/* current list sizes of one system: [0, PCG_NUM_TERMS) per term (its
 * total_size()), then the NL size, the residue count and the number of NL
 * rebuilds so far */
int pcg_batch_sizes(pcg_batch *b, int32_t sys, int32_t *sizes /*[PCG_NUM_TERMS + 3]*/);
/* Autotuner: times `trial_steps` steps of the persistent kernel at each
 * workgroup size (128, 256, 512, 1024 threads) on the live state, restoring
 * the state after every trial, and keeps the fastest. Returns the choice and
 * its time per step (ms). */
int pcg_batch_autotune(pcg_batch *b, int32_t trial_steps, int32_t *block, double *ms_per_step);
/* HIP events around the last pcg_batch_run's launches, milliseconds */
int pcg_batch_last_ms(pcg_batch *b, double *ms, int32_t *launches);

/* ---- served form: continuous batching of simulation jobs -------------- */
typedef struct pcg_server pcg_server;
typedef struct pcg_server_options {
  int32_t precision;
  int64_t hbm_budget_bytes; /* admission limit for resident job state */
  int32_t max_resident;     /* slots */
  int32_t chunk_steps;      /* steps per scheduling round */
  uint64_t philox_seed;
  int32_t device;
} pcg_server_options;

enum { PCG_JOB_QUEUED = 0, PCG_JOB_RUNNING = 1, PCG_JOB_DONE = 2,
       PCG_JOB_CANCELLED = 3, PCG_JOB_FAILED = 4 };

int pcg_server_create(const pcg_params *p, const pcg_server_options *o,
                      pcg_server **out);
int pcg_server_destroy(pcg_server *s);
/* queue a job: model, temperature (<0: params), ran2 seed, step count;
 * pulling jobs carry their tips in the model */
int pcg_server_submit(pcg_server *s, const pcg_model *m, double temperature,
                      int32_t seed, int64_t nsteps, int64_t *job_id);
/* one scheduling round: admit by budget, run chunk_steps for every resident
 * job in one launch per size class (while no job waits for admission, up to
 * four chunks, never past the class's nearest job end), retire finished
 * jobs. Returns the number of jobs still queued or running in *active. */
int pcg_server_round(pcg_server *s, int32_t *active);
int pcg_server_status(pcg_server *s, int64_t job, int32_t *status,
                      int64_t *steps_done);
int pcg_server_result(pcg_server *s, int64_t job, double *pos /*[3n]*/,
                      double *V);
/* one array [3n] of a finished job at its end: which 0 positions, 6 the forces
 * of its last evaluation, 7 velocities (the reference's state::r, dyn.F, v) */
int pcg_server_get(pcg_server *s, int64_t job, int32_t which, double *out);
/* scales the Langevin noise of every resident and later job (0: no noise) */
int pcg_server_set_noise_scale(pcg_server *s, double scale);
int pcg_server_cancel(pcg_server *s, int64_t job);
/* wall-clock seconds (steady clock) of submission, admission and end */
int pcg_server_job_times(pcg_server *s, int64_t job, double *submit, double *start,
                         double *end);
int pcg_server_stats(pcg_server *s, int64_t *resident_bytes, int32_t *resident,
                     int32_t *queued);

/* ------------------------------------------------------ input preparation
 * model::morph_into_saw (cg/src/cg/input/model.cpp:108-234): every chain is
 * redrawn as a self-avoiding walk of fixed bond length from the ran2
 * generator of `seed`, up to max_tries times per chain until no residues
 * i, j >= i + 3 come closer than `intersection`. Chains are the runs of equal
 * chain_idx (contiguous, in index order, as state::compile_model lays them
 * out). The start point of each chain is drawn in the cube of side
 * box_size centred at the origin (box_size 0 and density > 0: side
 * cbrt(n / density); both 0: the origin); pbc wraps the intersection test in
 * that cube. pos receives 3n coordinates; draws the number of ran2 draws
 * used, which pcg_options.ran2_skip takes so that the initial velocities
 * and the noise continue the same stream, as in the reference. Host code;
 * no GPU is touched. */
typedef struct pcg_saw_options {
  double bond;          /* bond length (> 0) */
  double intersection;  /* minimum distance of residues three or more apart */
  double box_size;
  double density;
  int32_t pbc;
  int32_t max_tries;    /* <= 0: 9000, the reference's */
} pcg_saw_options;
int pcg_morph_saw(const pcg_saw_options *o, int32_t n, const int32_t *chain_idx, int32_t seed,
                  double *pos /*[3n]*/, int64_t *draws);

/* ------------------------------------------------------------ measurement
 * No counterpart in the reference. A streaming copy of `bytes` (read once,
 * written once) timed over `iters` launches: the achievable HBM bandwidth
 * (GB/s, read + write) and the calibration kernel for traffic counters. */
int pcg_stream_copy(int64_t bytes, int32_t iters, int32_t device, double *gbs,
                    double *ms_per_copy);

#ifdef __cplusplus
}
#endif
