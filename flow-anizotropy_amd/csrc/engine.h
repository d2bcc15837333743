// engine.h -- host side of the solver: structure geometry on the device,
// forced solves ("jobs"), and the batch that steps many jobs together.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>
#include "lbm_common.h"

namespace fp {

// kernels.hip
void kern_rest(int lattice, int prec, const InstDesc &D, hipStream_t s);
// ad: how the step kernels address populations (Addr: AD_B32 32-bit byte
// offsets, every instance spans less than 4 GiB; AD_E32 32-bit element
// indices; AD_E64 64-bit element indices, instances of 2^32 elements or more)
// defer: the deferred finish of the previous step (struct Defer), or null
// (the state st is current and a finish launch follows the step)
void kern_step(int lattice, int prec, bool local, const Tuning &tu, int ad, const InstDesc *insts,
               const InstState *st, const void *work, int nwork, Partial *part, hipStream_t s,
               const Defer *defer = nullptr);
// plain copies of n16 float4: kind 0 persistent grid, 1 one workgroup per
// 8 KiB, 2 the same with non-temporal loads and stores, 3 one workgroup per
// 4 KiB, non-temporal
constexpr int N_COPY_KINDS = 4;
void kern_copy(int kind, const void *a, void *b, int64_t n16, hipStream_t s);
// after a step: exact flux sums and the convergence update of every instance;
// adv > 1 also counts the adv - 1 quiet steps before it (quiet_run_step)
void kern_finish(int nwork, const RedWork *work, const InstDesc *insts, const InstCfg *cfgs,
                 InstState *st, const Partial *part, RPartial *red, unsigned int *counter,
                 bool local, hipStream_t s, int adv = 1);

// One voxel structure resident on the device: the segment table.
struct Geometry {
    int dim = 3, nx = 0, ny = 0, nz = 1, nseg = 0;
    int64_t nrows = 0, nunits = 0, n_fluid = 0, nn = 0;
    SegInfo *d_seg = nullptr;
    void *d_links = nullptr;         // LinkRec<Q-1> per segment (neighbour-step links)
    uint32_t *d_rowspec = nullptr;   // wrap-lane neighbour index per row and direction
    void *d_ctab = nullptr;          // CRec per fluid cell (built on first use)
    uint32_t *d_rowbase = nullptr;   // compact-step row table (start, last rank per slot)
    size_t links_bytes = 0, rowspec_bytes = 0, ctab_bytes = 0;
    std::vector<uint8_t> solid;      // host copy, x fastest
    std::vector<uint32_t> row_first; // compact index of each row's first fluid cell; [nrows] = n_fluid
    int refs = 0;                    // jobs using it
    ~Geometry();
    size_t device_bytes() const
    {
        return (size_t)nunits * sizeof(SegInfo) + links_bytes + rowspec_bytes + ctab_bytes;
    }
    // bytes the compact neighbour table of this structure takes
    size_t compact_bytes() const;
};

// The host half of geometry_build: the segment table, the link records
// (LinkRec<Q-1> per segment, as bytes) and the wrap table of a mask, plus the
// Geometry's sizes, solid copy and row starts; no device memory is touched
// (g's device pointers stay null).
std::string geometry_host_tables(Geometry &g, const uint8_t *solid, int dim, int nx, int ny, int nz,
                                 std::vector<SegInfo> &seg, std::vector<unsigned char> &links,
                                 std::vector<uint32_t> &rowspec);
// builds the segment table from a mask (x fastest; 1 = solid) and uploads it.
// Returns an empty string or the error.
std::string geometry_build(Geometry &g, const uint8_t *solid, int dim, int nx, int ny, int nz);
// host-only check of the neighbour link tables and of the compact neighbour
// table (no device memory); *bad = number of (cell, direction) pairs whose
// derived neighbour is wrong
std::string geometry_verify_links(const uint8_t *solid, int dim, int nx, int ny, int nz, int64_t *bad);
// build and upload the compact neighbour table (once; no-op when present)
std::string geometry_build_compact(Geometry &g);

// One forced solve.
struct Job {
    Geometry *geo = nullptr;
    int lattice = 3, prec = PREC_F64;
    double F[3] = {0, 0, 0};
    InstCfg cfg{};
    int64_t stride = 0;
    void *d_f = nullptr;
    float *d_vel[2] = {nullptr, nullptr};
    bool want_fields = false;        // keep the velocity fields (3D export, 2D fieldres)
    int nbr_mode = 0;                // neighbour-step form in use (set by the batch)
    int64_t items = 0;               // its work items in the batch's current tables
    InstState state{};
    bool admitted = false, done = false;
    int64_t id = -1;
    double t_submit = 0, t_admit = 0, t_done = 0;
    void *user = nullptr;
    size_t device_bytes() const;
};

// cells a job's population arrays hold (n_fluid rounded up to 1024)
int64_t pop_stride_of(int64_t n_fluid);
// the device descriptor of a job (its geometry's and its own device pointers)
InstDesc job_desc(const Job &j);

// bytes of device memory a job of this shape needs
size_t job_bytes(const Geometry &g, int lattice, int prec, bool fields);
size_t job_bytes_for(int64_t n_fluid, int dim, int lattice, int prec, bool fields);

struct BatchStats {
    int64_t steps = 0;               // lattice steps launched (chunks * chunk length)
    int64_t chunks = 0;
    double  gpu_ms = 0;              // time inside chunks
};

// A set of jobs of one (lattice, precision) stepped together.  Jobs join and
// leave only between chunks; a chunk is `chunk` steps (even) captured once in
// a hipGraph and replayed while the membership stays the same.
class Batch {
public:
    Batch(int lattice, int prec, hipStream_t stream, int chunk = 200, bool use_graph = true);
    ~Batch();
    Batch(const Batch &) = delete;
    Batch &operator=(const Batch &) = delete;

    // allocate the job's lattice and put it at rest; it joins at the next chunk
    std::string admit(Job *j);
    // run one chunk; returns an error string (device failure) or empty
    std::string run_chunk();
    // jobs that finished (or failed) since the last call; the caller reads
    // their fields and then calls release()
    std::vector<Job *> take_done();
    void release(Job *j);
    // forget every job without freeing its buffers (the caller owns them now)
    void detach_all();
    // jobs admitted and not yet retired
    int live() const
    {
        int n = (int)pending_.size();
        for(const Job *j : jobs_) n += (j && !j->done) ? 1 : 0;
        return n;
    }
    int lattice() const { return lattice_; }
    int prec() const { return prec_; }
    const BatchStats &stats() const { return stats_; }
    void set_chunk(int c) { chunk_ = c < 2 ? 2 : (c & ~1); dirty_ = true; }
    void set_tuning(const Tuning &t) { tune_ = t; dirty_ = true; }
    const Tuning &tuning() const { return tune_; }
    // neighbour-step forms in use: 0 row segments, 1 compact, 2 both
    int nbr_mode() const { return mode_; }
    // whether chunks run with the deferred finish (struct Defer)
    bool deferred() const { return defer_; }

    // Timing: launch `steps` steps (even) directly, with a HIP event pair on
    // the batch stream around every kernel; returns the summed kernel times
    // per kind and the wall time of the whole region (events too).
    struct KernelTimes {
        double total_ms = 0, local_ms = 0, nbr_ms = 0, monitor_ms = 0;
        int64_t n_local = 0, n_nbr = 0, n_monitor = 0;
    };
    std::string time_steps(int steps, KernelTimes &kt);
    // Timing as solves run: `steps` steps (even) as replays of one captured
    // chunk graph (plus one for the remainder), after one untimed replay;
    // `ms` is the event time of the timed replays.
    std::string graph_steps(int steps, float &ms);

private:
    std::string rebuild();
    void launch_step(bool local, const Defer *df = nullptr);
    void launch_finish(bool local, int adv = 1);
    // the kernels of n steps (even) from an even step count: each followed by
    // its finish, or by a finish only where adv[t] > 0 (plan_finishes), or
    // with the deferred finish and one finish at the end
    void enqueue_steps(int n, const std::vector<int> *adv = nullptr);
    // the finish plan of the next n steps, run `replays` times back to back
    // after `skip` steps, from the jobs' current states
    void plan_chunk(int64_t skip, int n, int replays, std::vector<int> &adv) const;
    // whether those steps run with that plan (true), or else with the deferred
    // finish (can_defer) or a finish after every step: a plan unless it has a
    // finish after every step, or the deferred finish applies and the plan has
    // more than one finish per PLAN_OVER_DEFER steps
    bool choose_plan(int64_t skip, int n, int replays, std::vector<int> &adv, bool &can_defer) const;
    bool wants_rebuild() const;
    bool choose_defer() const;
    int64_t work_items(const Job &j) const;
    void release_graph();
    void free_job(Job *j);

    int lattice_, prec_;
    hipStream_t stream_;
    int chunk_;
    bool use_graph_;
    bool dirty_ = true;
    std::vector<Job *> jobs_, pending_, finished_;
    // device tables
    InstDesc *d_inst_ = nullptr;
    InstCfg *d_cfg_ = nullptr;
    InstState *d_state_ = nullptr;
    InstState *d_state2_ = nullptr;          // the other state array of the deferred finish
    bool defer_ = false;                     // the mode of the last chunk (or of enqueue_steps)
    bool capped_ = false;                    // step launches apply Tuning::local_wgs / nbr_wgs
    bool defer_ok_ = false;                  // every instance within DEFER_MAX_PARTIALS
    bool defer_small_ = false;               // ... and within REDUCE_SPAN
    LocalWork *d_lwork_ = nullptr;
    NbrWork *d_nwork_ = nullptr;
    LocalWork *d_cwork_ = nullptr;           // compact neighbour-step items (mode 1)
    size_t cap_cwork_ = 0;
    int mode_ = 0;                           // forms in the tables: 0, 1 or 2 (both)
    int ad_ = AD_B32;                        // population addressing of the step kernels (Addr)
    int n_cwork_ = 0;
    Partial *d_lpart_ = nullptr, *d_npart_ = nullptr;
    RedWork *d_lredw_ = nullptr, *d_nredw_ = nullptr;     // second-level reduction items
    RPartial *d_lred_ = nullptr, *d_nred_ = nullptr;
    int n_lred_ = 0, n_nred_ = 0;
    unsigned int *d_count_ = nullptr;        // per-instance finished-item counters
    size_t cap_count_ = 0;
    size_t cap_lredw_ = 0, cap_nredw_ = 0, cap_lred_ = 0, cap_nred_ = 0;
    InstState *h_state_ = nullptr;       // pinned mirror
    int n_inst_ = 0, n_lwork_ = 0, n_nwork_ = 0;
    int64_t work_total_ = 0, work_dead_ = 0;     // launched items, items of retired jobs
    size_t cap_inst_ = 0, cap_lwork_ = 0, cap_nwork_ = 0, cap_lpart_ = 0, cap_npart_ = 0;
    Tuning tune_;
    hipGraph_t graph_[2] = {nullptr, nullptr};       // chunk graphs: finish per step, deferred
    hipGraphExec_t exec_[2] = {nullptr, nullptr};
    // chunk graphs of the first finish plans met since the last rebuild (the
    // plans of a stable job set repeat with the lags)
    struct PlanGraph {
        std::vector<int> adv;
        hipGraph_t g = nullptr;
        hipGraphExec_t x = nullptr;
    };
    std::vector<PlanGraph> plan_graphs_;
    bool sparse_ok_ = true;                  // finish plans in use (FLOWPERM_SPARSE_FINISH)
    BatchStats stats_;
};

// The work lists, flux-partial slots and reduction items of a batch (the
// host half of a table rebuild), and their check.
struct BatchTables {
    std::vector<InstDesc> desc;
    std::vector<LocalWork> lw, cw;       // local items; compact neighbour items
    std::vector<NbrWork> nw;             // row-segment neighbour items
    std::vector<RedWork> lr, nr;         // reduction items over local / neighbour partials
};
void build_tables(const std::vector<Job *> &jobs, int lattice, int prec, const Tuning &tu, BatchTables &t);
// x-adjacent row segments per row-segment work item for rows of nseg
// segments: the largest power of two within nbr_segs and nbr_waves that
// divides nseg
int segs_per_item(const Tuning &tu, int64_t nseg);
// the check (number of violations, the first in *first); nbr_waves: the
// wavefronts of a neighbour workgroup
int64_t check_tables(const BatchTables &t, std::string *first, int nbr_waves);
// host only: the tables of the structures' forcing runs as a batch would build
// them, checked (*bad violations); returns the first problem, or a mask error
std::string verify_batch_tables(const std::vector<const uint8_t *> &solids,
                                const std::vector<int> &dims, int lattice, int prec,
                                const Tuning &tu, int64_t *bad);

// Which steps of a chunk of n steps must end with a finish, when the chunk
// runs `replays` times back to back after `skip` steps from the states st
// (the same plan for every run: one captured graph).  adv[t] > 0: the finish
// after step t advances the state by adv[t] steps -- the quiet steps since the
// previous finish and step t (advance_state_n); adv[t] = 0: no finish.  A
// finish follows every step that is not quiet for some live instance
// (quiet_run_step, quiet_avg_step) and every step before one (its finish
// plans that step's outputs), the last step of the chunk, and every step from
// the first at which an instance's first phase may end (max_steps, or the
// check at which hold can first reach c.hold).  3D instances in their
// averaging phase (a flux sum every step) take a finish after every step;
// retired ones (phase >= PH_DONE) none.
void plan_finishes(const InstCfg *cfg, const InstState *st, int n_inst, int64_t skip, int n,
                   int replays, std::vector<int> &adv);
// host only: a batch of n instances with the flux sequences q (nsteps x 3 per
// instance), joining at the batch steps start[i], stepped both with a finish
// after every step and with the finish plans of chunks of `chunk` steps run
// `replays` times per plan; counts in *problems every step whose outputs
// (phase, flux_next, vel_next) and every finish whose state differ.  plant = 1
// replaces each plan by one finish at the chunk's end (a wrong plan the check
// must catch).
int64_t verify_finish_plan(const std::vector<InstCfg> &cfg, const double *q, int64_t nsteps,
                           const int64_t *start, int chunk, int replays, int plant,
                           int64_t *finishes);

// the D3Q19 FP32 per-direction force terms of InstDesc
void f32_force_terms(const double F[3], float g45[19], float h15[19]);

double now_seconds();

// Make the next n table rebuilds of any batch in the process fail as a device
// error would, before they touch the device (tests of the error paths).
void test_fail_rebuilds(int n);

} // namespace fp
