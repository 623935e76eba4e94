// pcg_core.hpp -- device data layout shared by the plain, batched and served
// forms of the per-step loop.
//
// Every form works on a *batch*: systems concatenated residue-wise, each with
// a SysDesc naming its slices of the static lists and its capacity segments
// of the dynamic lists (NL, pauli, dh, pid, QA). The plain drop-in context is
// a batch of one. All arrays are structure-of-arrays; residue indices in
// lists are global (batch) indices.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pcg.h"

namespace pcg {

constexpr int NT = PCG_NUM_TERMS;
// partition classes of an NL pair in the stage kernels (per-chunk counts)
constexpr int NCL = 6;
// Static bonded lists whose items are chain-local: item of base residue r
// spans r, r+1 (tether), r, r+2 (local repulsion), r..r+2 (angles) or
// r..r+3 (dihedrals, chiral quads). The batched kernel evaluates them per
// base residue and sums the shares of neighbouring residues across lanes.
enum { BOND_TETHER, BOND_LREP, BOND_NAT_ANG, BOND_HEUR_ANG, BOND_NAT_DIH, BOND_HEUR_DIH,
       BOND_CHIR, NBOND };
constexpr double C216_INV = 0.89089871814033927; // cg/include/cg/utils/math.h:6

struct SysDesc {
  int res0, n;
  int teth0, nteth, lrep0, nlrep, chir0, nchir;
  int nang0, nnang, hang0, nhang, ndih0, nndih, hdih0, nhdih;
  int nc0, nnc, tip0, ntip;  // velocity tips [tip0, tip0 + ntip)
  int nftip;        // force tips after them: tip_vel holds the force
  int nl0, nlcap;   // NL and the NL-sized lists (pauli, dh, pid, QA candidates)
  int qf0, qfcap;   // QA free pairs
  int qc0, qccap;   // QA contacts and removed indices
  int bond_regular;  // every static bonded item is chain-local (see bond_item)
  // cell grid of the stage-form NL rebuild (cg[0] == 0: no grid): cg[a]
  // cells along axis a of the periodic cell, counts and starts at cell0
  int cell0, cg[3];
  double cell[3], cell_inv[3];
  double temperature, const2; // const2 = sqrt(2 kB T gamma dt) dt (step.cpp:7-17)
  uint64_t philox_key;
  int64_t steps_target;       // batched/served: steps this system still owes
  double afm_t0;              // time the tips were placed (afm::vel::tip ref_t)
};

// LDS layout of the batched kernel for a system of n residues: positions
// (3 n FP64), forces (3 n of fsz bytes) and, when staged, the per-residue
// bonded parameters (5 n FP64: tether length, native angle, native dihedral,
// the two chiral fields; 3 n bytes: flags, heuristic angle and dihedral
// types).
// With gear set, the Gear derivatives y1..y5 (15 n FP64) stay in LDS too.
// Dynamic LDS of one persistent workgroup, at most: 160 KiB per CU on
// gfx950, less the kernel's static LDS (about 5 KiB).
constexpr size_t LDS_DYN_MAX = 154 * 1024;
// the stage rebuild's cell grid: smallest system that takes one, most cells
// of one system, most 64-bit row-mask words (residues / 64) of one system
// (the NL row kernel holds one mask a wave in LDS: 4 x 8 x words bytes)
constexpr int CELL_MIN_N = 2048;
constexpr long long CELL_MAX = 1 << 21;
constexpr int CELL_WORDS_MAX = 2048;

struct LdsLayout {
  // 32-bit words of the per-wave stacks of QA pairs past the first test
  // (8 bytes a pair, 128 pairs for each of up to 16 waves)
  static constexpr size_t QA_QUEUE = 4 * 1024;
  size_t f, at, sy, bd, bu, g, q, total;
  __host__ __device__ LdsLayout(size_t n, size_t fsz, bool staged, bool gear, size_t nnc = 0) {
    f = 24 * n;
    at = f + 3 * n * fsz;
    sy = (at + n + 3) / 4 * 4;
    bd = (sy + 4 * n + 15) / 16 * 16;
    bu = bd + 48 * n;
    g = ((staged ? bu + 3 * n : bd) + 15) / 16 * 16;
    q = gear ? g + 120 * n : g;
    // the QA stacks' region also holds the list rebuild's tables (chain and
    // sequence index of every residue, native-contact row starts and the
    // contacts' second residues), which are never in use at the same time;
    // nnc bounds the native contacts of a system
    const size_t rb = rebuild_bytes(n, nnc);
    total = q + (rb > 4 * QA_QUEUE ? rb : 4 * QA_QUEUE) + 64;
  }
  __host__ __device__ static size_t rebuild_bytes(size_t n, size_t nnc) {
    return (8 * n + 4 + 2 * nnc + 15) / 16 * 16;
  }
};

// Served form: one slot per resident job of a round, uploaded and read back
// in one copy each (the step target in, the progress and health out).
struct JobSlot {
  int64_t target, step;
  double Vtot;
  int overflow, pad;
};

// Mutable per-system scalars.
struct SysState {
  int nl_n, nc_n, pauli_n, dh_n, pid_n, qf_n, qc_n, cand_n, rem_n, num_qa;
  int nl_invalid, verify_first, num_changed, overflow;
  int rebuilds, active;
  int qh_n;                   // QA free pairs that can become candidates (qh_slot)
  // free pairs held in the qf arrays: qf_n in the stage kernels; the
  // persistent kernel's rebuild stores only those within the QA maximum
  // required distance (the others cannot become candidates before the next
  // rebuild) and counts the rest in qf_n
  int qf_store;
  int n_hw, n_lj;             // harmonic wall connections, LJ wall connection slots
  int form;                   // plain context: the form (stages / persistent) whose lists these are
  // stage kernels' completion counts (rows, residues, blocks, residues):
  // the one that completes a system runs its follow-up (a scan, the tick)
  // and sets the count back to zero, so it is zero between kernels
  int done_rows, done_cells, done_part, done_lang;
  int64_t step;
  double t, true_t, orig_pad;
  double orig_cell[3];
  double V[NT];
  double Vtot;
  unsigned long long max_disp_bits; // FP64 max displacement since rebuild (as bits)
};

// Moving average of cg/include/cg/types/avg.h:5-38 (moving_avg<T, real>).
struct MAvg {
  double sum[3], avg[3], init_time, window;
  int num, has;
};
// An AFM tip's moving averages: a velocity tip's force on the tip (x, y, z)
// and its perpendicular part (entry 3), which the reference keeps in two
// moving_avg objects fed at the same times (so one count and start time
// serve both); a force tip's residue velocity (x, y, z)
struct TipStat {
  double sum[4], avg[4], init_time, window;
  int num, has;
};

// moving_avg::add (cg/include/cg/types/avg.h:14-29) over dim components
template <typename M>
__device__ __forceinline__ void mavg_add(M &a, double time, const double *sample, int dim) {
#pragma clang fp contract(off)
  if (a.num == 0) {
    a.init_time = time;
    for (int c = 0; c < dim; ++c) a.sum[c] = sample[c];
    a.num = 1;
  } else {
    for (int c = 0; c < dim; ++c) a.sum[c] += sample[c];
    ++a.num;
  }
  if (a.init_time + a.window <= time) {
    for (int c = 0; c < dim; ++c) a.avg[c] = a.sum[c] / (double)a.num;
    a.has = 1;
    a.num = 0;
  }
}

// Walls and the simulation box of a plain context (one system): the
// reaction force of every plane in the last evaluation, the wall
// connections, the LJ candidates and releases of the current step, and the
// state of the box protocol (thread_main.cpp:144-330). The planes themselves
// live in the parameters (wall_origin, wall_normal), which the box kernel
// moves between steps.
struct WallDev {
  double wall_F[6][3];
  // harmonic connections, fixed at creation (state.cpp:653-670)
  const int *hw_res, *hw_wall;
  const double *hw_off[3], *hw_nat;
  // LJ connection slots in creation order (a vect::set: releases leave
  // vacant slots, which the processing compacts in order)
  int *lj_res, *lj_wall;
  double *lj_off[3], *lj_sat;
  uint8_t *lj_gone;           // released by eval_connected this step
  int lj_cap;
  int *cand_res, *cand_wall;  // candidates of sift_free this step
  double *cand_d;
  int n_cand, cand_cap, n_gone;
  int num_conns[6];           // LJ connections per plane
  uint8_t *is_conn;           // state::is_connected_to_wall [n]
  int overflow;               // a candidate or slot beyond capacity
  // box protocol
  int box_on, log_on, phase, shear;
  int side_wall[6];           // plane on each box side, or -1
  double box_min[3], box_max[3];
  double since, until, amplitude, omega, displacement;
  double shift[6][3], work[6];
  MAvg avg_F[6], avg_z;
  pcg_box bx;
};

// Force-field parameters on the device (a copy of pcg_params plus derived
// constants).
struct DevParams {
  pcg_params p;
  double nl_eff_cutoff;      // max(nl.cutoff, dh.cutoff), thread_setup.cpp:368
  double V_factor, screen_inv;
  double gamma2, dt, dt_inv, deltsq;
  double qa_max_req, qa_factor, qa_sat_speed;
  int pauli_on;              // pauli.enabled && !qa.enabled (state.cpp:473)
  int nl_required;
  // wall terms on: a solid plane; harmonic planes with hw_on; LJ planes with ljw_on
  int solid_on, hw_terms, ljw_terms;
};

// Device code reads the parameters through the constant address space: they
// are never written while a kernel runs, so uniform fields become scalar
// loads (into SGPRs, hoisted freely) instead of per-lane loads at each use.
#if defined(__HIP_DEVICE_COMPILE__)
#define PCG_CONST_AS __attribute__((address_space(4)))
#else
#define PCG_CONST_AS
#endif
typedef const PCG_CONST_AS DevParams DevParamsC;
typedef const PCG_CONST_AS pcg_params ParamsC;
typedef const PCG_CONST_AS SysDesc SysDescC;  // written by the host between launches only

// What the report of one system needs, summed on the device
// (cg/src/cg/output/make_report.cpp:160-247).
struct ReportSums {
  double n, K, L;
  double S[6];       // sum of (r - c)(r - c)^T: xx, yy, zz, xy, xz, yz
  double H[9];       // sum of (q - q0)(r - c)^T, row-major (orig^T cur)
  double pp, qq;     // sums of |r - c|^2 and |q - q0|^2
  double icn, icnss, inter, intra;
  int pid[6];        // pid::count_active: intra, intra bb, intra ss, inter, inter bb, inter ss
  int pad_;
};

// Device arrays of a batch (pointers into one allocation).
struct DevBatch {
  int nsys, ntot;
  SysDescC *sys;
  SysState *st;
  DevParamsC *P;
  // residues
  double *y[6][3];            // Gear y0..y5 (y0 = positions)
  double *vel[3];
  void *F[3];                 // float or double forces (precision)
  double *orig_r[3];          // positions at the last NL rebuild
  double *nat_r[3];           // the model's original positions (the report's RMSD reference)
  double *nh[6];              // QA n and h vectors (FP64)
  const uint8_t *atype;
  const int *chain, *seq, *prev, *next, *sysid, *local;
  int8_t *sync;               // [4 ntot]
  double *noise[3];           // RAN2 mode: host-supplied noise
  // static lists
  const int *teth_i[2]; const double *teth_nat;
  const int *lrep_i[2];
  const int *chir_i[4]; const double *chir_f0, *chir_f1; // stored fields (nat_factor, nat_chir)
  const int *nang_i[3]; const double *nang_th;
  const int *hang_i[3]; const uint8_t *hang_t;
  const int *ndih_i[4]; const double *ndih_ph;
  const int *hdih_i[4]; const uint8_t *hdih_t;
  const int *nc_i[2]; const double *nc_nat; const uint8_t *nc_type;
  uint8_t *nc_active; double *nc_change_t;
  const int *tip_res; const double *tip_orig[3], *tip_vel[3];
  TipStat *tip_stat;          // per tip slot, as tip_res
  // dynamic lists
  int *nl_i[2]; double *nl_d; uint8_t *nl_cls;  // class bits of the partition
  int *ncur;                                      // current native contacts (index into all)
  int *ncur_i[2]; double *ncur_nat;  // their residues, length,
  uint8_t *ncur_t;                   // type | active << 7
  float *ncur_th2;  // squared transition distance (nat * 2^(-1/6) * counting factor), FP32
  // the same contact packed for the persistent kernel (systems under 4096
  // residues): i1 | i2 << 12 within the system, type << 24, active << 31
  uint32_t *ncur_ijt;
  // bond_item[L][r]: global index of the item of list L whose first residue
  // is r, or -1 (systems with bond_regular set)
  const int *bond_item[NBOND];
  int bond_lds;               // batched kernel: stage per-residue bonded data in LDS
  int gear_lds;               // batched kernel: keep y1..y5 in LDS for the launch
  int energy_off;             // batched kernel: no step evaluates energies
  int *pauli_i[2];
  int *dh_i[2]; double *dh_q;
  int *pid_i[2]; double *pid_d; int16_t *pid_t;
  int *qf_i[2]; double *qf_d; uint8_t *qf_vac;
  // the free pairs whose NL distance is within the QA maximum required
  // distance, in slot order, with their residues and a vacancy flag copied
  // next to the slot (dense, as the candidate search reads them every step):
  // loop_over_candidates skips every other free pair before reading a
  // position (orig_dist - total_disp > max_req_dist with total_disp == 0
  // after each verify, loop_over_candidates.cpp:20), so the search visits
  // only these
  int *qh_slot; int *qh_i[2]; uint8_t *qh_vac;
  // per near pair: the largest formation distance any contact type of its
  // residue pair can have (exact FP64), and the squared FP32 radius beyond
  // which the pair can have no effect at all (with a 1e-5 margin)
  double *qh_req; float *qh_thr2;  // thr2 < 0 once the pair is vacant
  uint32_t *qh_ij;  // the pair's residues within its system, i1 | i2 << 16
  int *cand_h;                // near-list index of each candidate
  int *qc_i[2]; double *qc_d; int16_t *qc_t; double *qc_sat; int8_t *qc_sd; uint8_t *qc_vac;
  int *cand_i[2]; double *cand_od, *cand_dist; int *cand_idx; int16_t *cand_t; int8_t *cand_sd;
  int *cand_perm;
  int *rem;
  int *row_cnt;               // NL rebuild scratch: per-residue row counts / offsets
  // cell grid of the NL rebuild: each residue's cell, the residues by cell
  // (system-local indices), each cell's first slot and its end (a count
  // between the rebuild's kernels, zero between rebuilds)
  int *cell_of, *cell_items, *cell_start, *cell_end;
  int *blk_cnt;               // partition scratch: per-chunk counts
  WallDev *wall;              // plain context: walls and the box protocol (else null)
};

// ------------------------------------------------------------- helpers ----
template <typename R> struct v3 { R x, y, z; };
template <typename R> __device__ __forceinline__ v3<R> mk3(R a, R b, R c) { return v3<R>{a, b, c}; }
template <typename R> __device__ __forceinline__ v3<R> operator+(v3<R> a, v3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename R> __device__ __forceinline__ v3<R> operator-(v3<R> a, v3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename R> __device__ __forceinline__ v3<R> operator-(v3<R> a) { return {-a.x, -a.y, -a.z}; }
template <typename R> __device__ __forceinline__ v3<R> operator*(v3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename R> __device__ __forceinline__ v3<R> operator*(R s, v3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename R> __device__ __forceinline__ v3<R> operator/(v3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <typename R> __device__ __forceinline__ R dot(v3<R> a, v3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename R> __device__ __forceinline__ R nsq(v3<R> a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
template <typename R> __device__ __forceinline__ v3<R> cross(v3<R> a, v3<R> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// 1/sqrt(x) for forces (never for a discrete decision): FP32 from the
// hardware's v_rsq_f32 (about 1 ulp); FP64 from v_rsq_f64 refined by two
// Newton steps y <- y + y (1 - x y^2) / 2 (an ulp or two from the correctly
// rounded value, against some 50 instructions for sqrt and a division).
// x = 0 and x = inf keep the hardware's inf and 0.
__device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double rsq(double x) {
  const double y0 = __builtin_amdgcn_rsq(x);
  double e = __builtin_fma(-(x * y0), y0, 1.0);
  const double y1 = __builtin_fma(0.5 * y0, e, y0);
  e = __builtin_fma(-(x * y1), y1, 1.0);
  const double y2 = __builtin_fma(0.5 * y1, e, y1);
  return (x == 0.0 || x == __builtin_inf()) ? y0 : y2;
}
// sqrt and 1/x for forces: FP32 on the hardware's v_sqrt_f32 / v_rcp_f32
// (about 1 ulp), FP64 correctly rounded
__device__ __forceinline__ float sq(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double sq(double x) { return sqrt(x); }
__device__ __forceinline__ float rcpR(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double rcpR(double x) { return 1.0 / x; }
// Makes a loaded value count as used here, so that the compiler issues its
// load where it stands (with the loads beside it) instead of sinking it past
// the branches that first read it. An empty statement: no instruction.
template <typename T> __device__ __forceinline__ void keep(T &x) { asm volatile("" : "+v"(x)); }
template <typename R> __device__ __forceinline__ R clampR(R v, R lo, R hi) { return v < lo ? lo : (v < hi ? v : hi); }
template <typename R> __device__ __forceinline__ R ipow6(R s) { R s2 = s * s; return s2 * s2 * s2; }
__device__ __forceinline__ float acosc(float x) { return acosf(fminf(fmaxf(x, -1.0f), 1.0f)); }
__device__ __forceinline__ double acosc(double x) { return acos(fmin(fmax(x, -1.0), 1.0)); }
__device__ __forceinline__ float sinR(float x) { return sinf(x); }
__device__ __forceinline__ double sinR(double x) { return sin(x); }
__device__ __forceinline__ float cosR(float x) { return cosf(x); }
__device__ __forceinline__ double cosR(double x) { return cos(x); }
__device__ __forceinline__ void sincosR(float x, float &s, float &c) { sincosf(x, &s, &c); }
__device__ __forceinline__ void sincosR(double x, double &s, double &c) { sincos(x, &s, &c); }
__device__ __forceinline__ float expR(float x) { return expf(x); }
__device__ __forceinline__ double expR(double x) { return exp(x); }

// Exact FP64 geometry for the decisions that must match the CPU path bit
// for bit (NL predicate, list filters, contact-state transitions, QA
// criteria, displacement check): IEEE sqrt and division, the reference's
// operand order (vec3_ops_expr.h:6-33, pbc.h:37-49), and no FMA contraction
// (the *_rn intrinsics are plain operators once inlined, so every function
// that must round each operation turns contraction off).
__device__ __forceinline__ double xwrap1(double v, double cell, double cinv) {
#pragma clang fp contract(off)
  return __dsub_rn(v, __dmul_rn(cell, round(__dmul_rn(v, cinv))));
}
__device__ __forceinline__ double xnorm(double x, double y, double z) {
#pragma clang fp contract(off)
  return __dsqrt_rn(__dadd_rn(__dadd_rn(__dmul_rn(x, x), __dmul_rn(y, y)), __dmul_rn(z, z)));
}

} // namespace pcg
