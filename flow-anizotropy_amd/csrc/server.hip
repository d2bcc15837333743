// server.hip -- the request server: whole-tensor requests admitted by device
// memory footprint into one continuously batched stepper (include/flowperm.h).
#include "solve_common.h"

#include <cmath>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <new>
#include <unordered_map>
#include <vector>

using namespace fp;

namespace {

struct Request {
    int64_t id = -1;
    int dim = 3, nx = 0, ny = 0, nz = 1;
    fp_config3d c3{};
    fp_config2d c2{};
    double F = 0;                                  // force magnitude of the runs
    int64_t n_pore = 0;
    std::vector<uint8_t> solid2d;                  // 2D: the mask, for the porosity column
    std::unique_ptr<Geometry> geo;
    std::vector<std::unique_ptr<Job>> jobs;        // forcing runs (+ the validate run, last)
    std::vector<fp_solve_out> outs;
    int nruns = 0, ndone = 0;
    int bypassed = 0;                              // later requests admitted ahead of it
    bool validate = false;
    size_t bytes = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    int64_t updates = 0;
    int status = FP_PENDING;
    fp_result3d r3{};
    fp_result2d r2{};
};

} // namespace

struct fp_server {
    fp_ctx *ctx = nullptr;
    int lattice = 3, prec = FP_F32, chunk = 200;
    int max_bypass = 16;
    int64_t budget = 0;
    std::unique_ptr<Batch> batch;                  // touched by the stepping thread only
    std::mutex mu;                                 // guards everything below
    std::deque<Request *> queue;                   // requests waiting for memory, oldest first
    std::unordered_map<int64_t, std::unique_ptr<Request>> reqs;   // every unpolled request
    std::unordered_map<const Job *, Request *> owner;
    int64_t next_id = 0;
    int64_t in_use = 0, peak = 0;
    fp_server_stats st{};
};

namespace {

// device memory a request holds while resident: its segment tables, the
// compact neighbour table when its structure will use that form, and the
// population (and, in 2D, velocity snapshot) arrays of each of its runs
size_t request_bytes(const Geometry &g, int lattice, int prec, int njobs, const Tuning &tu)
{
    const bool compact = tu.nbr_mode == 1 ||
                         (tu.nbr_mode == 2 && prefer_compact((double)g.n_fluid / (double)g.nn,
                                                             segs_per_item(tu, g.nseg), lattice, prec));
    return g.device_bytes() + (compact ? g.compact_bytes() : 0) +
           (size_t)njobs * job_bytes(g, lattice, prec, lattice == 2);
}

int64_t submit(fp_server *S, std::unique_ptr<Request> r, const uint8_t *solid)
{
    (void)hipSetDevice(S->ctx->device);
    r->geo.reset(new Geometry);
    std::string e = geometry_build(*r->geo, solid, r->dim, r->nx, r->ny, r->nz);
    if(!e.empty()) {
        std::lock_guard<std::mutex> lk(S->mu);
        S->st.rejected++;
        return -(int64_t)status_of_geometry_error(e);
    }
    const int njobs = r->nruns + (r->validate ? 1 : 0);
    r->bytes = request_bytes(*r->geo, S->lattice, S->prec, njobs, S->batch->tuning());
    if((int64_t)r->bytes > S->budget) {
        std::lock_guard<std::mutex> lk(S->mu);
        S->st.rejected++;
        return -(int64_t)FP_ERR_OOM;
    }
    for(int d = 0; d < njobs; d++) {
        std::unique_ptr<Job> j(new Job);
        j->geo = r->geo.get();
        j->lattice = S->lattice;
        j->prec = S->prec;
        if(r->dim == 3) {
            if(d < 3) j->F[d] = r->F;
            else { const double c = 1.0 / std::sqrt(3.0); j->F[0] = j->F[1] = j->F[2] = r->F * c; }
            j->cfg = cfg3(r->c3, r->geo->nn, false);
            j->want_fields = false;
        } else {
            if(d < 2) j->F[d] = r->F;
            else validate2d_force(r->F, r->c2.validate_deg, j->F);
            j->cfg = cfg2(r->c2, r->nx, r->ny);
            j->want_fields = true;                 // fieldres needs both snapshots
        }
        r->jobs.push_back(std::move(j));
    }
    r->outs.assign((size_t)njobs, fp_solve_out{});
    r->t_submit = now_seconds();
    std::lock_guard<std::mutex> lk(S->mu);
    r->id = S->next_id++;
    const int64_t id = r->id;
    for(auto &j : r->jobs) S->owner[j.get()] = r.get();
    S->queue.push_back(r.get());
    S->reqs[id] = std::move(r);
    S->st.submitted++;
    return id;
}

// assemble a request's row from its runs as fp_solve3d / fp_solve2d do;
// returns its final status
int finish_request(Request *r)
{
    int status = FP_OK;
    bool conv = true;
    for(int d = 0; d < (int)r->outs.size(); d++) {
        if(r->outs[d].status == FP_ERR_DEVICE) status = FP_ERR_DEVICE;
        else if(r->outs[d].status != FP_OK && status == FP_OK) status = FP_ERR_NONFINITE;
        if(d < r->nruns) conv = conv && r->outs[d].converged;
    }
    if(r->dim == 3) {
        for(int d = 0; d < 3; d++) r->r3.run[d] = r->outs[d];
        if(status == FP_OK) {
            assemble3d(&r->r3, r->nx, r->ny, r->nz, r->n_pore, r->F);
            if(r->validate) { r->r3.val = r->outs[3]; assemble3d_val(&r->r3, r->F); }
        }
        r->r3.seconds = r->t_done - r->t_admit;
        r->r3.fluid_updates = r->updates;
    } else {
        for(int d = 0; d < 2; d++) r->r2.run[d] = r->outs[d];
        if(status == FP_OK) {
            assemble2d(&r->r2, r->nx, r->ny, r->solid2d.data(), r->F);
            if(r->validate) { r->r2.val = r->outs[2]; assemble2d_val(&r->r2, r->F, r->c2.validate_deg); }
        }
        r->r2.seconds = r->t_done - r->t_admit;
        r->r2.fluid_updates = r->updates;
    }
    if(status == FP_OK && !conv) status = FP_NOT_CONVERGED;
    return status;
}

// Admit queued requests that fit under the budget.  The queue is walked in
// arrival order; a request that does not fit is passed over and later ones that
// do fit are admitted ahead of it (a small request does not wait behind a large
// one for memory it does not need), until it has been passed over max_bypass
// times: from then on nothing behind it is admitted before it.  max_bypass < 0
// keeps strict arrival order.
void admit_queued(fp_server *S)
{
    for(;;) {
        Request *r = nullptr;
        size_t pos = 0;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            for(size_t i = 0; i < S->queue.size(); i++) {
                Request *q = S->queue[i];
                if(S->in_use + (int64_t)q->bytes <= S->budget) { r = q; pos = i; break; }
                if(S->max_bypass < 0 || q->bypassed >= S->max_bypass) break;
            }
            if(!r) return;
            for(size_t i = 0; i < pos; i++) S->queue[i]->bypassed++;
            S->queue.erase(S->queue.begin() + (std::ptrdiff_t)pos);
            S->in_use += (int64_t)r->bytes;
            S->peak = std::max(S->peak, S->in_use);
        }
        std::string e;
        size_t k = 0;
        for(; k < r->jobs.size() && e.empty(); k++) e = S->batch->admit(r->jobs[k].get());
        if(e.empty()) {
            std::lock_guard<std::mutex> lk(S->mu);
            r->t_admit = now_seconds();
            S->st.admitted++;
            continue;
        }
        // the device could not hold it after all: undo, and retry once memory
        // is returned, or fail it when nothing else holds memory
        for(auto &j : r->jobs) if(j->admitted) S->batch->release(j.get());
        for(auto &j : r->jobs) { j->done = false; j->admitted = false; }
        std::lock_guard<std::mutex> lk(S->mu);
        S->in_use -= (int64_t)r->bytes;
        if(S->batch->live() == 0) {
            r->status = e.find("memory") != std::string::npos ? FP_ERR_OOM : FP_ERR_DEVICE;
            r->t_done = now_seconds();
            for(auto &jj : r->jobs) S->owner.erase(jj.get());
            r->geo.reset();
            S->st.rejected++;
            continue;
        }
        // back to its place (other threads only append to the queue)
        for(size_t i = 0; i < pos; i++) S->queue[i]->bypassed--;
        S->queue.insert(S->queue.begin() + (std::ptrdiff_t)pos, r);
        return;
    }
}

// retire finished runs: read their results (2D: the field residual from the
// device snapshots), free their lattices, and finish requests whose last run
// this was
void collect(fp_server *S)
{
    for(Job *j : S->batch->take_done()) {
        Request *r;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            r = S->owner[j];
        }
        size_t d = 0;
        while(r->jobs[d].get() != j) d++;
        fill_out(*j, &r->outs[d]);                 // outs are not read by poll
        if(r->dim == 2 && r->outs[d].status == FP_OK) {
            std::string e = finish2d(*j, r->nx, r->ny, r->c2.avg_steps, &r->outs[d], nullptr);
            if(!e.empty()) r->outs[d].status = FP_ERR_DEVICE;
        }
        const int64_t upd = j->state.steps * r->geo->n_fluid;
        S->batch->release(j);
        std::lock_guard<std::mutex> lk(S->mu);
        r->updates += upd;
        if(++r->ndone < (int)r->jobs.size()) continue;
        r->t_done = now_seconds();
        r->status = finish_request(r);
        for(auto &jj : r->jobs) S->owner.erase(jj.get());
        r->geo.reset();                            // no live instance reads it any more
        S->in_use -= (int64_t)r->bytes;
        S->st.completed++;
        S->st.fluid_updates += r->updates;
    }
}

// A failed chunk (a device error, or a table rebuild that could not get its
// memory) ends every resident request: its runs leave the batch, its lattices
// and footprint are returned, and it finishes with FP_ERR_DEVICE.  Nothing of
// it stays behind for a later step or poll to touch; queued requests are
// untouched and are admitted by later steps.
void fail_resident(fp_server *S, const std::string &err)
{
    std::vector<Request *> failed;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->ctx->err = err;
        for(auto &kv : S->reqs) {
            Request *r = kv.second.get();
            if(r->status == FP_PENDING && r->t_admit > 0) failed.push_back(r);
        }
    }
    for(Request *r : failed)
        for(auto &j : r->jobs)
            if(j->admitted) S->batch->release(j.get());
    std::lock_guard<std::mutex> lk(S->mu);
    const double t = now_seconds();
    for(Request *r : failed) {
        for(auto &j : r->jobs) S->owner.erase(j.get());
        r->status = FP_ERR_DEVICE;
        r->t_done = t;
        r->geo.reset();
        S->in_use -= (int64_t)r->bytes;
        S->st.completed++;
    }
}

} // namespace

extern "C" {

double fp_now(void) { return now_seconds(); }

fp_server *fp_server_create(fp_ctx *ctx, const fp_server_config *cfg, int *status)
{
    auto bad = [&](int st) -> fp_server * { if(status) *status = st; return nullptr; };
    if(!ctx || !cfg) return bad(FP_ERR_ARG);
    const int L = cfg->lattice, P = cfg->precision;
    if(!((L == 3 && (P == FP_F64 || P == FP_F32)) ||
         (L == 2 && (P == FP_F64 || P == FP_F32 || P == FP_F32_REF2D))))
        return bad(FP_ERR_ARG);
    if(cfg->budget_bytes < 0 || cfg->chunk_steps < 0) return bad(FP_ERR_ARG);
    if(hipSetDevice(ctx->device) != hipSuccess) return bad(FP_ERR_DEVICE);
    fp_server *S = new(std::nothrow) fp_server;
    if(!S) return bad(FP_ERR_OOM);
    S->ctx = ctx;
    S->lattice = L;
    S->prec = P;
    S->chunk = cfg->chunk_steps > 0 ? cfg->chunk_steps : ctx->chunk;
    S->max_bypass = cfg->max_bypass == 0 ? 16 : (cfg->max_bypass < 0 ? -1 : cfg->max_bypass);
    if(cfg->budget_bytes > 0) S->budget = cfg->budget_bytes;
    else {
        size_t fr = 0, tot = 0;
        if(hipMemGetInfo(&fr, &tot) != hipSuccess) { delete S; return bad(FP_ERR_DEVICE); }
        S->budget = (int64_t)(0.9 * (double)fr);
    }
    S->batch.reset(new Batch(L, P, ctx->stream, S->chunk, true));
    S->batch->set_tuning(ctx->tune);
    if(status) *status = FP_OK;
    return S;
}

void fp_server_destroy(fp_server *S)
{
    if(!S) return;
    (void)hipSetDevice(S->ctx->device);
    S->batch.reset();                              // frees resident lattices first
    delete S;
}

int64_t fp_server_submit3d(fp_server *S, const uint8_t *solid, int nx, int ny, int nz,
                           const fp_config3d *cfg)
{
    if(!S || !solid || !cfg3_ok(cfg) || S->lattice != 3 || cfg->precision != S->prec)
        return -FP_ERR_ARG;
    if(nx <= 0 || ny <= 0 || nz <= 0) return -FP_ERR_BAD_DIMS;
    int64_t np = 0;
    const int st = check_mask(solid, (int64_t)nx * ny * nz, &np);
    if(st != FP_OK) return -st;
    std::unique_ptr<Request> r(new Request);
    r->dim = 3; r->nx = nx; r->ny = ny; r->nz = nz;
    r->c3 = *cfg;
    r->F = FX_BASE * cfg->ff;                      // LMB3d/lbm3d-perm.cpp:512
    r->n_pore = np;
    r->nruns = 3;
    r->validate = cfg->validate != 0;
    return submit(S, std::move(r), solid);
}

int64_t fp_server_submit2d(fp_server *S, const uint8_t *solid, int lx, int ly, const fp_config2d *cfg)
{
    if(!S || !solid || !cfg2_ok(cfg) || S->lattice != 2 || cfg->precision != S->prec)
        return -FP_ERR_ARG;
    if(lx <= 0 || ly <= 0) return -FP_ERR_BAD_DIMS;
    const int st = check_mask(solid, (int64_t)lx * ly, nullptr);
    if(st != FP_OK) return -st;
    std::unique_ptr<Request> r(new Request);
    r->dim = 2; r->nx = lx; r->ny = ly; r->nz = 1;
    r->c2 = *cfg;
    r->F = FX_BASE * cfg->ff;                      // LMB2d/main.cpp:314
    r->nruns = 2;
    r->validate = cfg->validate_deg >= 0.0;
    r->solid2d.assign(solid, solid + (size_t)lx * ly);
    return submit(S, std::move(r), solid);
}

int fp_server_step(fp_server *S, int64_t *pending)
{
    if(!S) return FP_ERR_ARG;
    (void)hipSetDevice(S->ctx->device);
    admit_queued(S);
    collect(S);                                    // runs with nothing to step finish at once
    int rc = FP_OK;
    if(S->batch->live() > 0) {
        const double t0 = now_seconds();
        const int64_t steps0 = S->batch->stats().steps;
        std::string e = S->batch->run_chunk();
        const double dt = now_seconds() - t0;
        if(!e.empty()) {
            fail_resident(S, e);
            rc = FP_ERR_DEVICE;
        } else {
            collect(S);
            std::lock_guard<std::mutex> lk(S->mu);
            S->st.busy_seconds += dt;
            S->st.chunks++;
            S->st.steps += S->batch->stats().steps - steps0;
        }
    }
    std::lock_guard<std::mutex> lk(S->mu);
    int64_t resident = 0, live = 0;
    for(auto &kv : S->reqs) {
        const Request *r = kv.second.get();
        if(r->status != FP_PENDING || r->t_admit <= 0) continue;
        resident++;
        live += (int64_t)r->jobs.size() - r->ndone;
    }
    S->st.resident = resident;
    S->st.live_instances = live;
    S->st.queued = (int64_t)S->queue.size();
    if(pending) *pending = resident + S->st.queued;
    return rc;
}

static int poll(fp_server *S, int64_t id, int dim, void *out, fp_request_info *info)
{
    if(!S || !out) return FP_ERR_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if(it == S->reqs.end() || it->second->dim != dim) return FP_ERR_NOT_FOUND;
    Request *r = it->second.get();
    if(info) {
        info->t_submit = r->t_submit;
        info->t_admit = r->t_admit;
        info->t_done = r->t_done;
        info->device_bytes = (int64_t)r->bytes;
        info->fluid_updates = r->updates;
    }
    if(r->status == FP_PENDING) return FP_PENDING;
    if(dim == 3) *(fp_result3d *)out = r->r3;
    else *(fp_result2d *)out = r->r2;
    const int st = r->status;
    S->reqs.erase(it);
    return st;
}

int fp_server_poll3d(fp_server *S, int64_t id, fp_result3d *out, fp_request_info *info)
{
    return poll(S, id, 3, out, info);
}

int fp_server_poll2d(fp_server *S, int64_t id, fp_result2d *out, fp_request_info *info)
{
    return poll(S, id, 2, out, info);
}

void fp_server_get_stats(fp_server *S, fp_server_stats *out)
{
    if(!S || !out) return;
    std::lock_guard<std::mutex> lk(S->mu);
    *out = S->st;
    out->bytes_in_use = S->in_use;
    out->bytes_peak = S->peak;
    out->budget_bytes = S->budget;
}

} // extern "C"
