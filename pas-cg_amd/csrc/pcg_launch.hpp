// pcg_launch.hpp -- launchers of the grid-wide stage kernels (pcg_kernels.hip)
// and of the persistent batched kernel (pcg_batch.hip).
#pragma once
#include "pcg_kernels.hpp"

namespace pcg {
void launch_begin(const DevBatch &b, hipStream_t s);
void launch_zero_forces(const DevBatch &b, int prec, hipStream_t s);
void launch_prepare_nh(const DevBatch &b, hipStream_t s);
// begin, force clearing and (qa_on) prepare_nh of a full step in one launch
void launch_step_begin(const DevBatch &b, int prec, int qa_on, hipStream_t s);
// done: close the rebuild here (else the step's k_tick does, nl_done)
void launch_fix_nl(const DevBatch &b, int max_n, int max_nlcap, int cell_words, hipStream_t s,
                   bool done = true);
void launch_terms(const DevBatch &b, int prec, int max_items, int only_term, int from, int to,
                  hipStream_t s);
void launch_qa_finish(const DevBatch &b, int max_n, hipStream_t s);
// tick: the step's k_tick (advance, close a rebuild) by the residue that
// completes each system
void launch_lang(const DevBatch &b, int prec, int noise_mode, hipStream_t s, bool tick = false);
void launch_tick(const DevBatch &b, int advance, hipStream_t s, bool nl_done = false);
// walls and the box protocol of a plain context (b.wall set): LJ candidate
// processing; box set-up, the phase logic before a step, the force log after
// it. P and sd are the writable device parameters and system descriptor.
void launch_ljw_process(const DevBatch &b, hipStream_t s);
void launch_box_setup(const DevBatch &b, DevParams *P, SysDesc *sd, hipStream_t s);
void launch_box_phase(const DevBatch &b, DevParams *P, SysDesc *sd, hipStream_t s);
void launch_box_log(const DevBatch &b, hipStream_t s);
// persistent form: every system of the batch advances nsteps in one launch
void launch_persistent(const DevBatch &b, int prec, int block, int nsteps, size_t lds_bytes,
                       hipStream_t s);
void launch_persistent_jobs(const DevBatch *jobs, JobSlot *slots, int njobs, int prec, int block,
                            int nsteps, size_t lds_bytes, hipStream_t s);
// report sums of every system (pcg_report.hip)
void launch_report(const DevBatch &b, int pid_on, ReportSums *out, hipStream_t s);
// the same into the library's own device array (no allocation), for batches
// of at most REPORT_STATIC_MAX systems, with `device` current: 0, or -1 for
// too many systems, -2 for a device index beyond the library's table, -3 for
// a HIP failure
constexpr int REPORT_STATIC_MAX = 4096;
int report_static_sums(const DevBatch &b, int device, int pid_on, hipStream_t s,
                       ReportSums **sums);
} // namespace pcg
