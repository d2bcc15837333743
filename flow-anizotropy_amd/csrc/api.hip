// api.hip -- the extern "C" surface of libflowperm.so (include/flowperm.h).
#include "solve_common.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

using namespace fp;

namespace fp {

float mu2d()                                  // LMB2d/lbm.cpp:44-45
{
    float tau = 1;
    float mu = 0.33 * (tau - 0.5);
    return mu;
}

int fail(fp_ctx *ctx, int status, const std::string &msg)
{
    if(ctx) ctx->err = msg;
    return status;
}

int status_of_geometry_error(const std::string &e)
{
    if(e == "bad dimensions") return FP_ERR_BAD_DIMS;
    if(e == "non-binary voxel") return FP_ERR_NON_BINARY;
    if(e == "structure has no pore space") return FP_ERR_NO_PORE;
    if(e.find("hipMalloc") != std::string::npos) return FP_ERR_OOM;
    if(e.find("too many") != std::string::npos || e.find("too large") != std::string::npos)
        return FP_ERR_ARG;
    return FP_ERR_DEVICE;
}

InstCfg cfg3(const fp_config3d &c, int64_t nn, bool want_vel)
{
    InstCfg k;
    std::memset(&k, 0, sizeof(k));
    k.eps = c.eps;
    k.lag = c.lag;
    k.min_steps = c.min_steps;
    k.max_steps = c.max_steps;
    k.avg_steps = c.avg_steps;
    k.avg_sample = 1;
    k.hold = c.hold;
    k.kind = 3;
    k.want_vel = want_vel ? 1 : 0;
    k.denom[0] = k.denom[1] = k.denom[2] = (double)nn;   // q = sum / NN  (:302)
    return k;
}

InstCfg cfg2(const fp_config2d &c, int lx, int ly)
{
    InstCfg k;
    std::memset(&k, 0, sizeof(k));
    k.eps = c.eps;
    k.qfloor = c.qfloor;
    k.ksigma = c.ksigma;
    k.lag = c.lag;
    k.min_steps = c.min_steps;
    k.max_steps = c.max_steps;
    k.avg_steps = c.avg_steps;
    k.avg_sample = c.avg_sample;
    k.hold = c.hold;
    k.kind = 2;
    k.want_vel = 1;
    // q_x = su / (LX * L0 * L0),  q_y = sv / (LY * L0 * L0)   (LMB2d/lbm.cpp:150-151)
    k.denom[0] = (double)lx * c.L0 * c.L0;
    k.denom[1] = (double)ly * c.L0 * c.L0;
    k.denom[2] = 1.0;
    return k;
}

// compact field (dim floats per fluid cell, .raw order) -> dense, zero at solid
void scatter_field(const Geometry &g, const std::vector<float> &compact, float *dense)
{
    const int dim = g.dim;
    int64_t c = 0;
    for(int64_t n = 0; n < g.nn; n++) {
        if(g.solid[n] == 0) {
            for(int k = 0; k < dim; k++) dense[n * dim + k] = compact[c * dim + k];
            c++;
        } else {
            for(int k = 0; k < dim; k++) dense[n * dim + k] = 0.0f;
        }
    }
}

std::string copy_field(const Job &j, int slot, std::vector<float> &out)
{
    out.resize((size_t)j.geo->dim * j.geo->n_fluid);
    hipError_t e = hipMemcpy(out.data(), j.d_vel[slot], out.size() * sizeof(float), hipMemcpyDeviceToHost);
    return e == hipSuccess ? "" : std::string("hipMemcpy(field): ") + hipGetErrorString(e);
}

// field_residual of LMB2d/lbm.cpp:169-182 over dense [y][x] pairs, i-outer order
double field_residual(int lx, int ly, const float *cur, const float *snap)
{
    double num = 0.0, den = 0.0;
    for(int i = 0; i < lx; i++)
        for(int j = 0; j < ly; j++) {
            const size_t n = ((size_t)j * lx + i) * 2;
            double du = cur[n] - snap[n];
            double dv = cur[n + 1] - snap[n + 1];
            num += du * du + dv * dv;
            den += (double)cur[n] * cur[n] + (double)cur[n + 1] * cur[n + 1];
        }
    return den > 0.0 ? (double)(float)std::sqrt(num / den) : 10.0;
}

static void fill_check(const InstState &S, fp_solve_out *o)
{
    o->check_step = S.chk_s;
    o->checks = S.chk_n;
    for(int k = 0; k < 3; k++) o->check_q[k] = S.chk_q[k];
    for(int k = 0; k < 2; k++) o->check_ratio[k] = S.chk_r[k];
    o->check_hold = S.chk_hold;
}

void fill_out(const Job &j, fp_solve_out *o)
{
    std::memset(o, 0, sizeof(*o));
    for(int k = 0; k < 3; k++) { o->q[k] = j.state.q[k]; o->se[k] = j.state.se[k]; }
    o->steps = j.state.steps;
    o->converged = j.state.converged;
    o->status = j.state.status == 2 ? FP_ERR_NONFINITE : FP_OK;
    fill_check(j.state, o);
}

// Run a structure's forced solves to completion: together on one batch, or one
// after another.  The jobs keep their device buffers (the caller reads the
// velocity fields, then frees them with free_job_memory).
std::string run_jobs(fp_ctx *ctx, std::vector<Job *> &jobs, int lattice, int prec, bool together,
                     int run0)
{
    int64_t maxsteps = 0;
    for(Job *j : jobs)
        maxsteps = std::max<int64_t>(maxsteps, std::max<int64_t>(0, j->cfg.max_steps) +
                                                   std::max<int64_t>(0, j->cfg.avg_steps));
    int chunk = (int)std::min<int64_t>(200, std::max<int64_t>(2, (maxsteps + 1) & ~1ll));
    // with a progress callback a chunk is no longer than the check interval,
    // so that every check is reported (a chunk has an even length: with
    // lag 1 only every other check is)
    std::vector<int64_t> reported(jobs.size(), 0);
    if(ctx->progress) {
        int64_t lag = INT64_MAX;
        for(Job *j : jobs) lag = std::min<int64_t>(lag, j->cfg.lag);
        chunk = (int)std::max<int64_t>(2, std::min<int64_t>(chunk, lag) & ~1ll);
    }
    auto report = [&]() {
        if(!ctx->progress) return;
        for(size_t r = 0; r < jobs.size(); r++) {
            const InstState &S = jobs[r]->state;
            if(S.chk_n <= reported[r]) continue;
            reported[r] = S.chk_n;
            fp_solve_out o;
            std::memset(&o, 0, sizeof(o));
            o.steps = S.s;
            o.converged = S.converged;
            fill_check(S, &o);
            ctx->progress(ctx->progress_user, run0 + (int)r, &o);
        }
    };
    auto step_until_done = [&](const std::vector<Job *> &set) -> std::string {
        Batch b(lattice, prec, ctx->stream, std::min(chunk, ctx->chunk), true);
        b.set_tuning(ctx->tune);
        std::string err;
        for(Job *j : set) {
            err = b.admit(j);
            if(!err.empty()) break;
        }
        while(err.empty()) {
            bool all = true;
            for(Job *j : set) all = all && j->done;
            if(all) break;
            err = b.run_chunk();
            (void)b.take_done();
            report();
        }
        b.detach_all();
        return err;
    };
    std::string err;
    if(together) err = step_until_done(jobs);
    else
        for(Job *j : jobs) {
            err = step_until_done(std::vector<Job *>{j});
            if(!err.empty()) break;
        }
    return err;
}

void free_job_memory(Job &j)
{
    if(j.d_f) { (void)hipFree(j.d_f); j.d_f = nullptr; }
    for(int s = 0; s < 2; s++)
        if(j.d_vel[s]) { (void)hipFree(j.d_vel[s]); j.d_vel[s] = nullptr; }
}

// symmetric 3x3 cyclic Jacobi, LMB3d/lbm3d-perm.cpp:401-443
void jacobi3(const double Ain[3][3], double val[3], double vec[3][3])
{
    double a[3][3];
    std::memcpy(a, Ain, sizeof(a));
    for(int i = 0; i < 3; i++) for(int j = 0; j < 3; j++) vec[i][j] = (i == j) ? 1.0 : 0.0;
    for(int sweep = 0; sweep < 50; sweep++) {
        double off = std::fabs(a[0][1]) + std::fabs(a[0][2]) + std::fabs(a[1][2]);
        if(off < 1e-18) break;
        for(int p = 0; p < 2; p++)
            for(int q = p + 1; q < 3; q++) {
                if(std::fabs(a[p][q]) < 1e-300) continue;
                const double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
                const double t = (theta >= 0 ? 1.0 : -1.0) /
                                 (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for(int k = 0; k < 3; k++) {
                    const double akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - s * akq;  a[k][q] = s * akp + c * akq;
                }
                for(int k = 0; k < 3; k++) {
                    const double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - s * aqk;  a[q][k] = s * apk + c * aqk;
                    const double vkp = vec[k][p], vkq = vec[k][q];
                    vec[k][p] = c * vkp - s * vkq; vec[k][q] = s * vkp + c * vkq;
                }
            }
    }
    for(int i = 0; i < 3; i++) val[i] = a[i][i];
    for(int i = 0; i < 2; i++)
        for(int j = i + 1; j < 3; j++)
            if(val[j] > val[i]) {
                std::swap(val[i], val[j]);
                for(int k = 0; k < 3; k++) std::swap(vec[k][i], vec[k][j]);
            }
}

int check_mask(const uint8_t *solid, int64_t nn, int64_t *n_pore)
{
    int64_t np = 0;
    for(int64_t i = 0; i < nn; i++) {
        if(solid[i] > 1) return FP_ERR_NON_BINARY;
        if(!solid[i]) np++;
    }
    if(n_pore) *n_pore = np;
    return np == 0 ? FP_ERR_NO_PORE : FP_OK;
}

const char *msg_of(int st)
{
    switch(st) {
    case FP_ERR_BAD_DIMS: return "bad dimensions";
    case FP_ERR_NON_BINARY: return "non-binary voxel";
    case FP_ERR_NO_PORE: return "structure has no pore space";
    default: return "error";
    }
}

bool cfg3_ok(const fp_config3d *c) { return c && c->lag > 0; }
bool cfg2_ok(const fp_config2d *c) { return c && c->lag > 0 && c->avg_sample > 0 && c->L0 > 0; }

// K and everything derived from it, LMB3d/lbm3d-perm.cpp:523-553; out->run[]
// holds the three solves
void assemble3d(fp_result3d *out, int nx, int ny, int nz, int64_t n_pore, double F)
{
    out->nx = nx; out->ny = ny; out->nz = nz;
    out->n_pore = n_pore;
    const double NN = (double)((int64_t)nx * ny * nz);
    out->porosity = (double)n_pore / NN;
    out->force = F;
    out->mu = MU3;
    double (*K)[3] = out->K;
    for(int c = 0; c < 3; c++) {
        K[c][0] = MU3 * out->run[0].q[c] / F;
        K[c][1] = MU3 * out->run[1].q[c] / F;
        K[c][2] = MU3 * out->run[2].q[c] / F;
    }
    const double trace = K[0][0] + K[1][1] + K[2][2];
    out->k_xy = 0.5 * (K[0][1] + K[1][0]);
    out->k_xz = 0.5 * (K[0][2] + K[2][0]);
    out->k_yz = 0.5 * (K[1][2] + K[2][1]);
    out->recip_xy = std::fabs(K[0][1] - K[1][0]) / std::fabs(trace);
    out->recip_xz = std::fabs(K[0][2] - K[2][0]) / std::fabs(trace);
    out->recip_yz = std::fabs(K[1][2] - K[2][1]) / std::fabs(trace);
    out->recip_resid = dmax_ref(out->recip_xy, dmax_ref(out->recip_xz, out->recip_yz));
    double Ks[3][3] = {{K[0][0], out->k_xy, out->k_xz},
                       {out->k_xy, K[1][1], out->k_yz},
                       {out->k_xz, out->k_yz, K[2][2]}};
    double *val = out->k;
    jacobi3(Ks, val, out->v);
    out->k_mean = trace / 3.0;
    const double dv = (val[0]-val[1])*(val[0]-val[1]) + (val[1]-val[2])*(val[1]-val[2])
                    + (val[2]-val[0])*(val[2]-val[0]);
    const double nrm = val[0]*val[0] + val[1]*val[1] + val[2]*val[2];
    out->fa = nrm > 0 ? std::sqrt(0.5 * dv / nrm) : NAN;
}

// the tensor test of LMB3d/lbm3d-perm.cpp:555-572 once out->val holds the
// solve under F (1,1,1)/sqrt(3)
void assemble3d_val(fp_result3d *out, double F)
{
    const double c = 1.0 / std::sqrt(3.0);
    const double Ks[3][3] = {{out->K[0][0], out->k_xy, out->k_xz},
                             {out->k_xy, out->K[1][1], out->k_yz},
                             {out->k_xz, out->k_yz, out->K[2][2]}};
    double err = 0, qn = 0;
    for(int r = 0; r < 3; r++) {
        out->val_pred[r] = (Ks[r][0]*F*c + Ks[r][1]*F*c + Ks[r][2]*F*c) / MU3;
        err += (out->val.q[r] - out->val_pred[r]) * (out->val.q[r] - out->val_pred[r]);
        qn += out->val.q[r] * out->val.q[r];
    }
    out->val_err = std::sqrt(err) / (qn > 0 ? std::sqrt(qn) : 1);
    out->validated = 1;
}

// the 2x2 tensor and derived values of LMB2d/main.cpp:314-353
void assemble2d(fp_result2d *out, int lx, int ly, const uint8_t *solid, double F)
{
    int nfluid = 0;                                      // getporosity_full, lbm.cpp:188-195
    for(int64_t i = 0; i < (int64_t)lx * ly; i++) if(solid[i] == 0) nfluid++;
    const float por = nfluid / (float)(lx * ly);
    out->lx = lx; out->ly = ly;
    out->porosity = por;
    out->force = F;
    const float mu = mu2d();
    out->mu = mu;
    const fp_solve_out &rx = out->run[0], &ry = out->run[1];
    const double k_xx = mu * rx.q[0] / F;
    const double k_yx = mu * rx.q[1] / F;
    const double k_xy = mu * ry.q[0] / F;
    const double k_yy = mu * ry.q[1] / F;
    const double k_off = 0.5 * (k_xy + k_yx);
    const double trace = k_xx + k_yy;
    const double recip = trace != 0.0 ? std::fabs(k_xy - k_yx) / std::fabs(trace) : NAN;
    const double km = 0.5 * (k_xx + k_yy);
    const double d1 = 0.5 * (k_xx - k_yy);
    const double d2 = k_off;
    const double r = std::sqrt(d1 * d1 + d2 * d2);
    out->k_xx = k_xx; out->k_yx = k_yx; out->k_xy = k_xy; out->k_yy = k_yy;
    out->k_off = k_off; out->recip_resid = recip;
    out->k1 = km + r; out->k2 = km - r; out->k_mean = km;
    out->anisotropy = km != 0.0 ? r / km : NAN;
    out->theta_deg = 0.5 * std::atan2(d2, d1) * 180.0 / M_PI;
    out->alpha_x_deg = std::atan2(rx.q[1], rx.q[0]) * 180.0 / M_PI;
    out->alpha_y_deg = std::atan2(ry.q[1], ry.q[0]) * 180.0 / M_PI - 90.0;
}

// force of the 2D --validate DEG solve (LMB2d/main.cpp:360-364)
void validate2d_force(double F, double deg, double Fv[2])
{
    const double th = deg * M_PI / 180.0;
    Fv[0] = F * std::cos(th);
    Fv[1] = F * std::sin(th);
}

// the 2D tensor test (LMB2d/main.cpp:365-380) once out->val holds that solve
void assemble2d_val(fp_result2d *out, double F, double deg)
{
    double Fv[2];
    validate2d_force(F, deg, Fv);
    const double mu = out->mu;
    const double pred_qx = (out->k_xx * Fv[0] + out->k_off * Fv[1]) / mu;
    const double pred_qy = (out->k_off * Fv[0] + out->k_yy * Fv[1]) / mu;
    const double qn = std::sqrt(out->val.q[0] * out->val.q[0] + out->val.q[1] * out->val.q[1]);
    out->val_pred[0] = pred_qx; out->val_pred[1] = pred_qy;
    out->val_err = std::sqrt((out->val.q[0] - pred_qx) * (out->val.q[0] - pred_qx) +
                             (out->val.q[1] - pred_qy) * (out->val.q[1] - pred_qy)) / (qn > 0 ? qn : 1);
    out->validated = 1;
}

// fieldres of a finished 2D solve from its two velocity snapshots (before
// averaging, final), and the final field as dense [y][x] (U, V) pairs into
// vel when not NULL
std::string finish2d(const Job &j, int lx, int ly, int64_t avg_steps, fp_solve_out *o, float *vel)
{
    const Geometry &g = *j.geo;
    const size_t nn = (size_t)lx * ly;
    std::vector<float> snap(nn * 2), cur, c;
    std::string ee = copy_field(j, 0, c);
    if(!ee.empty()) return ee;
    scatter_field(g, c, snap.data());
    if(avg_steps > 0) {
        cur.resize(nn * 2);
        ee = copy_field(j, 1, c);
        if(!ee.empty()) return ee;
        scatter_field(g, c, cur.data());
    } else cur = snap;
    o->fieldres = field_residual(lx, ly, cur.data(), snap.data());
    if(vel) std::memcpy(vel, cur.data(), nn * 2 * sizeof(float));
    return "";
}

} // namespace fp

// ===========================================================================
extern "C" {

int fp_exit_code(int status)
{
    if(status == FP_OK) return 0;
    if(status == FP_NOT_CONVERGED) return 3;
    return 2;
}

int fp_abi_version(void) { return FP_ABI_VERSION; }

int fp_set_progress(fp_ctx *ctx, fp_progress_fn fn, void *user)
{
    if(!ctx) return FP_ERR_ARG;
    ctx->progress = fn;
    ctx->progress_user = fn ? user : nullptr;
    return FP_OK;
}

fp_ctx *fp_create(int device, int *status)
{
    fp_ctx *c = new(std::nothrow) fp_ctx;
    if(!c) { if(status) *status = FP_ERR_OOM; return nullptr; }
    c->device = device;
    hipError_t e = hipSetDevice(device);
    if(e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if(e != hipSuccess) {
        if(status) *status = FP_ERR_DEVICE;
        delete c;
        return nullptr;
    }
    // the address of fp_solve_dir3d_noalloc's scratch array on this device
    // (a failure is reported by that entry point)
    (void)noalloc_setup(device);
    if(status) *status = FP_OK;
    return c;
}

void fp_destroy(fp_ctx *ctx)
{
    if(!ctx) return;
    (void)hipSetDevice(ctx->device);
    if(ctx->stream) { (void)hipStreamSynchronize(ctx->stream); (void)hipStreamDestroy(ctx->stream); }
    delete ctx;
}

const char *fp_last_error(fp_ctx *ctx) { return ctx ? ctx->err.c_str() : ""; }

void fp_config3d_default(fp_config3d *c)
{
    std::memset(c, 0, sizeof(*c));
    c->ff = 1.0; c->eps = 1e-5; c->lag = 2000; c->hold = 2;
    c->min_steps = 4000; c->max_steps = 200000; c->avg_steps = 2000;
    c->validate = 0; c->precision = FP_F64; c->concurrent = 0;
}

void fp_config2d_default(fp_config2d *c)
{
    std::memset(c, 0, sizeof(*c));
    c->ff = 1.0; c->eps = 1e-3; c->qfloor = 1e-2; c->ksigma = 3.0; c->L0 = 4.0;
    c->validate_deg = -1.0;
    c->lag = 5000; c->min_steps = 10000; c->max_steps = 500000; c->avg_steps = 20000;
    c->avg_sample = 50; c->hold = 3; c->precision = FP_F32_REF2D; c->concurrent = 0;
}

void fp_jacobi3(const double A9[9], double val[3], double vec9[9])
{
    double A[3][3], vec[3][3];
    std::memcpy(A, A9, sizeof(A));
    jacobi3(A, val, vec);
    std::memcpy(vec9, vec, sizeof(vec));
}

int fp_verify_links(const uint8_t *solid, int dim, int nx, int ny, int nz, int64_t *mismatches)
{
    if(!solid || !mismatches || (dim != 2 && dim != 3)) return FP_ERR_ARG;
    std::string e = geometry_verify_links(solid, dim, nx, ny, dim == 3 ? nz : 1, mismatches);
    return e.empty() ? FP_OK : status_of_geometry_error(e);
}

// ---------------------------------------------------------------------------
// host-only checks of the convergence rules and the field residual
// ---------------------------------------------------------------------------
// The device state machine (advance_state / plan_next, the code k_finish runs
// after every step) driven by a given flux sequence: step t (1-based) sees
// q[(t - 1) * dim ...] when the machine asked that step for a flux, and a zero
// flux otherwise, as on the device.
static int replay_convergence(const InstCfg &c, int dim, const double *q, int64_t n, fp_solve_out *out)
{
    InstState S;
    init_state(c, S);
    for(int64_t t = 0; S.phase < PH_DONE; t++) {
        if(t >= n) return FP_ERR_ARG;
        const int flux = S.flux_next;
        double qt[3] = {0.0, 0.0, 0.0};
        if(flux)
            for(int k = 0; k < dim; k++) qt[k] = q[t * dim + k];
        advance_state(c, S, qt, flux);
        plan_next(c, S);
    }
    Job j;
    j.state = S;
    fill_out(j, out);
    if(S.status == 2) return FP_ERR_NONFINITE;
    return S.converged ? FP_OK : FP_NOT_CONVERGED;
}

int fp_replay_convergence3d(const fp_config3d *cfg, const double *q, int64_t nsteps, fp_solve_out *out)
{
    if(!cfg3_ok(cfg) || !q || !out || nsteps < 0) return FP_ERR_ARG;
    std::memset(out, 0, sizeof(*out));
    return replay_convergence(cfg3(*cfg, 1, false), 3, q, nsteps, out);
}

int fp_replay_convergence2d(const fp_config2d *cfg, const double *q, int64_t nsteps, fp_solve_out *out)
{
    if(!cfg2_ok(cfg) || !q || !out || nsteps < 0) return FP_ERR_ARG;
    std::memset(out, 0, sizeof(*out));
    return replay_convergence(cfg2(*cfg, 1, 1), 2, q, nsteps, out);
}

double fp_field_residual2d(int lx, int ly, const float *cur, const float *snap)
{
    if(lx <= 0 || ly <= 0 || !cur || !snap) return NAN;
    return field_residual(lx, ly, cur, snap);
}

void fp_test_fail_rebuilds(int n) { test_fail_rebuilds(n); }

// the launch shapes of an fp_tuning (nbr_segs 0: the default, 4)
static void copy_tuning(const fp_tuning *t, Tuning &tu)
{
    tu.local_cpt = t->local_cpt; tu.nbr_waves = t->nbr_waves;
    tu.nontemporal = t->nontemporal; tu.local_bpc = t->local_bpc;
    tu.nbr_bpc = t->nbr_bpc; tu.nbr_mode = t->nbr_mode;
    tu.nbr_segs = t->nbr_segs > 0 ? t->nbr_segs : Tuning().nbr_segs;
    tu.local_wgs = t->local_wgs; tu.nbr_wgs = t->nbr_wgs;
}

int fp_verify_batch_tables(const uint8_t *const *solids, const int32_t *dims, int n, int lattice,
                           int precision, const fp_tuning *tune, int64_t *problems)
{
    if(!solids || !dims || n <= 0 || !problems || (lattice != 2 && lattice != 3) ||
       precision < FP_F64 || precision > FP_F32_REF2D)
        return FP_ERR_ARG;
    Tuning tu;
    if(tune) copy_tuning(tune, tu);
    std::vector<const uint8_t *> s(solids, solids + n);
    std::vector<int> d(dims, dims + 3 * n);
    for(const uint8_t *p : s) if(!p) return FP_ERR_ARG;
    *problems = 0;
    std::string e = verify_batch_tables(s, d, lattice, precision, tu, problems);
    if(!e.empty() && *problems == 0) return status_of_geometry_error(e);
    return FP_OK;
}

int fp_verify_finish_plan(const fp_config3d *c3, const fp_config2d *c2, int n, const double *q,
                          int64_t nsteps, const int64_t *start, int chunk, int replays, int plant,
                          int64_t *problems, int64_t *finishes)
{
    if((!c3 == !c2) || n <= 0 || !q || nsteps < 0 || !start || chunk < 2 || (chunk & 1) ||
       replays < 1 || !problems)
        return FP_ERR_ARG;
    std::vector<InstCfg> cfg(n);
    for(int i = 0; i < n; i++) {
        if(c3) {
            if(!cfg3_ok(&c3[i])) return FP_ERR_ARG;
            cfg[i] = cfg3(c3[i], 1, (i & 1) != 0);          // every other one writes its velocity
        } else {
            if(!cfg2_ok(&c2[i])) return FP_ERR_ARG;
            cfg[i] = cfg2(c2[i], 1, 1);
        }
    }
    *problems = verify_finish_plan(cfg, q, nsteps, start, chunk, replays, plant, finishes);
    return FP_OK;
}

int fp_check_mask3d(const uint8_t *solid, int nx, int ny, int nz, int64_t *n_pore)
{
    if(nx <= 0 || ny <= 0 || nz <= 0) return FP_ERR_BAD_DIMS;
    if(!solid) return FP_ERR_ARG;
    return check_mask(solid, (int64_t)nx * ny * nz, n_pore);
}

// ---------------------------------------------------------------------------
// 3D
// ---------------------------------------------------------------------------
static int solve3d_dirs(fp_ctx *ctx, const uint8_t *solid, int nx, int ny, int nz,
                        const fp_config3d *cfg, const double (*F)[3], int ndir,
                        fp_solve_out *outs, float *vel, int64_t *updates, int run0 = 0)
{
    if(!ctx) return FP_ERR_ARG;
    if(!cfg3_ok(cfg)) return fail(ctx, FP_ERR_ARG, "lag must be positive");
    if(cfg->precision != FP_F64 && cfg->precision != FP_F32)
        return fail(ctx, FP_ERR_ARG, "3D precision must be FP_F64 or FP_F32");
    if(nx <= 0 || ny <= 0 || nz <= 0) return fail(ctx, FP_ERR_BAD_DIMS, "bad dimensions");
    int64_t np = 0;
    int st = check_mask(solid, (int64_t)nx * ny * nz, &np);
    if(st != FP_OK) return fail(ctx, st, msg_of(st));
    (void)hipSetDevice(ctx->device);
    Geometry g;
    std::string e = geometry_build(g, solid, 3, nx, ny, nz);
    if(!e.empty()) return fail(ctx, status_of_geometry_error(e), e);
    std::vector<Job> jobs(ndir);
    std::vector<Job *> ptrs;
    for(int d = 0; d < ndir; d++) {
        Job &j = jobs[d];
        j.geo = &g;
        j.lattice = 3;
        j.prec = cfg->precision;
        for(int k = 0; k < 3; k++) j.F[k] = F[d][k];
        j.cfg = cfg3(*cfg, g.nn, vel != nullptr);
        j.want_fields = vel != nullptr;
        ptrs.push_back(&j);
    }
    e = run_jobs(ctx, ptrs, 3, cfg->precision, cfg->concurrent != 0, run0);
    int rc = FP_OK;
    if(!e.empty()) rc = fail(ctx, e.find("memory") != std::string::npos ? FP_ERR_OOM : FP_ERR_DEVICE, e);
    int64_t upd = 0;
    for(int d = 0; d < ndir && rc == FP_OK; d++) {
        fill_out(jobs[d], &outs[d]);
        upd += jobs[d].state.steps * g.n_fluid;
        if(vel) {
            std::vector<float> c;
            std::string ee = copy_field(jobs[d], 1, c);
            if(!ee.empty()) { rc = fail(ctx, FP_ERR_DEVICE, ee); break; }
            scatter_field(g, c, vel + (size_t)d * g.nn * 3);
        }
    }
    for(Job &j : jobs) free_job_memory(j);
    if(updates) *updates = upd;
    return rc;
}

int fp_solve_dir3d(fp_ctx *ctx, const uint8_t *solid, int nx, int ny, int nz,
                   const fp_config3d *cfg, const double F[3], fp_solve_out *out, float *vel)
{
    const double Fd[1][3] = {{F[0], F[1], F[2]}};
    int rc = solve3d_dirs(ctx, solid, nx, ny, nz, cfg, Fd, 1, out, vel, nullptr);
    if(rc != FP_OK) return rc;
    if(out->status != FP_OK)
        return fail(ctx, FP_ERR_NONFINITE, "non-finite flux in run at step " +
                                            std::to_string(out->steps));
    return out->converged ? FP_OK : FP_NOT_CONVERGED;
}

int fp_solve3d(fp_ctx *ctx, const uint8_t *solid, int nx, int ny, int nz,
               const fp_config3d *cfg, fp_result3d *out, float *vel)
{
    if(!ctx || !out || !cfg) return FP_ERR_ARG;
    std::memset(out, 0, sizeof(*out));
    const double t0 = now_seconds();
    const double F = FX_BASE * cfg->ff;
    const double Fd[3][3] = {{F, 0, 0}, {0, F, 0}, {0, 0, F}};
    int64_t upd = 0;
    int rc = solve3d_dirs(ctx, solid, nx, ny, nz, cfg, Fd, 3, out->run, vel, &upd);
    if(rc != FP_OK) return rc;
    const char *tags[3] = {"fx", "fy", "fz"};
    for(int d = 0; d < 3; d++)
        if(out->run[d].status != FP_OK)
            return fail(ctx, FP_ERR_NONFINITE, std::string("non-finite flux in run ") + tags[d] +
                                                   " at step " + std::to_string(out->run[d].steps));
    int64_t np = 0;
    (void)check_mask(solid, (int64_t)nx * ny * nz, &np);
    assemble3d(out, nx, ny, nz, np, F);
    if(cfg->validate) {                                  // :555-572
        const double c = 1.0 / std::sqrt(3.0);
        const double Fv[1][3] = {{F * c, F * c, F * c}};
        int64_t u2 = 0;
        rc = solve3d_dirs(ctx, solid, nx, ny, nz, cfg, Fv, 1, &out->val, nullptr, &u2, 3);
        if(rc != FP_OK) return rc;
        if(out->val.status != FP_OK)
            return fail(ctx, FP_ERR_NONFINITE, "non-finite flux in run val at step " +
                                                   std::to_string(out->val.steps));
        upd += u2;
        assemble3d_val(out, F);
    }
    out->seconds = now_seconds() - t0;
    out->fluid_updates = upd;
    ctx->err.clear();
    return (out->run[0].converged && out->run[1].converged && out->run[2].converged)
               ? FP_OK : FP_NOT_CONVERGED;
}

// ---------------------------------------------------------------------------
// 2D
// ---------------------------------------------------------------------------
static int solve2d_dirs(fp_ctx *ctx, const uint8_t *solid, int lx, int ly, const fp_config2d *cfg,
                        const double (*F)[2], int ndir, fp_solve_out *outs, float *vel, int64_t *updates,
                        int run0 = 0)
{
    if(!ctx) return FP_ERR_ARG;
    if(!cfg2_ok(cfg)) return fail(ctx, FP_ERR_ARG, "--lag and --avg-sample must be positive");
    if(cfg->precision != FP_F64 && cfg->precision != FP_F32 && cfg->precision != FP_F32_REF2D)
        return fail(ctx, FP_ERR_ARG, "unknown precision");
    if(lx <= 0 || ly <= 0) return fail(ctx, FP_ERR_BAD_DIMS, "bad dimensions");
    int64_t np = 0;
    int st = check_mask(solid, (int64_t)lx * ly, &np);
    if(st == FP_ERR_NON_BINARY) return fail(ctx, st, msg_of(st));
    if(st == FP_ERR_NO_PORE) return fail(ctx, st, msg_of(st));
    (void)hipSetDevice(ctx->device);
    Geometry g;
    std::string e = geometry_build(g, solid, 2, lx, ly, 1);
    if(!e.empty()) return fail(ctx, status_of_geometry_error(e), e);
    std::vector<Job> jobs(ndir);
    std::vector<Job *> ptrs;
    for(int d = 0; d < ndir; d++) {
        Job &j = jobs[d];
        j.geo = &g;
        j.lattice = 2;
        j.prec = cfg->precision;
        j.F[0] = F[d][0]; j.F[1] = F[d][1]; j.F[2] = 0.0;
        j.cfg = cfg2(*cfg, lx, ly);
        j.want_fields = true;
        ptrs.push_back(&j);
    }
    e = run_jobs(ctx, ptrs, 2, cfg->precision, cfg->concurrent != 0, run0);
    int rc = FP_OK;
    if(!e.empty()) rc = fail(ctx, e.find("memory") != std::string::npos ? FP_ERR_OOM : FP_ERR_DEVICE, e);
    int64_t upd = 0;
    const size_t nn = (size_t)lx * ly;
    for(int d = 0; d < ndir && rc == FP_OK; d++) {
        fill_out(jobs[d], &outs[d]);
        upd += jobs[d].state.steps * g.n_fluid;
        if(jobs[d].state.status != 0) continue;
        std::string ee = finish2d(jobs[d], lx, ly, cfg->avg_steps, &outs[d],
                                  vel ? vel + (size_t)d * nn * 2 : nullptr);
        if(!ee.empty()) { rc = fail(ctx, FP_ERR_DEVICE, ee); break; }
    }
    for(Job &j : jobs) free_job_memory(j);
    if(updates) *updates = upd;
    return rc;
}

int fp_solve_dir2d(fp_ctx *ctx, const uint8_t *solid, int lx, int ly, const fp_config2d *cfg,
                   const double F[2], fp_solve_out *out, float *vel)
{
    const double Fd[1][2] = {{F[0], F[1]}};
    int rc = solve2d_dirs(ctx, solid, lx, ly, cfg, Fd, 1, out, vel, nullptr);
    if(rc != FP_OK) return rc;
    if(out->status != FP_OK)
        return fail(ctx, FP_ERR_NONFINITE, "non-finite flux in run at step " + std::to_string(out->steps));
    return out->converged ? FP_OK : FP_NOT_CONVERGED;
}

int fp_solve2d(fp_ctx *ctx, const uint8_t *solid, int lx, int ly, const fp_config2d *cfg,
               fp_result2d *out, float *vel)
{
    if(!ctx || !out || !cfg) return FP_ERR_ARG;
    std::memset(out, 0, sizeof(*out));
    const double t0 = now_seconds();
    const double F = FX_BASE * cfg->ff;                 // main.cpp:314
    const double Fd[2][2] = {{F, 0.0}, {0.0, F}};
    int64_t upd = 0;
    int rc = solve2d_dirs(ctx, solid, lx, ly, cfg, Fd, 2, out->run, vel, &upd);
    if(rc != FP_OK) return rc;
    const char *tags[2] = {"fx", "fy"};
    for(int d = 0; d < 2; d++)
        if(out->run[d].status != FP_OK)
            return fail(ctx, FP_ERR_NONFINITE, std::string("non-finite flux in run ") + tags[d] +
                                                   " at step " + std::to_string(out->run[d].steps));
    assemble2d(out, lx, ly, solid, F);                   // main.cpp:314-353
    if(cfg->validate_deg >= 0.0) {                       // main.cpp:360-380
        double Fv2[2];
        validate2d_force(F, cfg->validate_deg, Fv2);
        const double Fv[1][2] = {{Fv2[0], Fv2[1]}};
        int64_t u2 = 0;
        rc = solve2d_dirs(ctx, solid, lx, ly, cfg, Fv, 1, &out->val, nullptr, &u2, 2);
        if(rc != FP_OK) return rc;
        if(out->val.status != FP_OK)
            return fail(ctx, FP_ERR_NONFINITE, "non-finite flux while averaging run val");
        upd += u2;
        assemble2d_val(out, F, cfg->validate_deg);
    }
    out->seconds = now_seconds() - t0;
    out->fluid_updates = upd;
    ctx->err.clear();
    return (out->run[0].converged && out->run[1].converged) ? FP_OK : FP_NOT_CONVERGED;
}

} // extern "C"

// ---------------------------------------------------------------------------
// batches
// ---------------------------------------------------------------------------
struct fp_batch {
    fp_ctx *ctx = nullptr;
    int lattice = 3, prec = FP_F64;
    std::unique_ptr<Batch> b;
    std::vector<std::unique_ptr<Geometry>> geos;
    std::unordered_multimap<uint64_t, Geometry *> geo_index;
    std::vector<std::unique_ptr<Job>> jobs;          // index = job id
    std::vector<fp_solve_out> results;
    std::vector<char> have;
    int64_t updates = 0;
    double seconds = 0;
};

static uint64_t fnv1a(const uint8_t *p, size_t n, uint64_t h)
{
    for(size_t i = 0; i < n; i++) { h ^= p[i]; h *= 1099511628211ull; }
    return h;
}

// one device geometry per distinct mask
static Geometry *batch_geometry(fp_batch *B, const uint8_t *solid, int dim, int nx, int ny, int nz,
                                std::string &err)
{
    const int64_t nn = (int64_t)nx * ny * (dim == 3 ? nz : 1);
    uint64_t h = fnv1a(solid, (size_t)nn, 1469598103934665603ull);
    h = fnv1a((const uint8_t *)&nx, sizeof(int), h);
    h = fnv1a((const uint8_t *)&ny, sizeof(int), h);
    h = fnv1a((const uint8_t *)&nz, sizeof(int), h);
    auto range = B->geo_index.equal_range(h);
    for(auto it = range.first; it != range.second; ++it) {
        Geometry *g = it->second;
        if(g->dim == dim && g->nx == nx && g->ny == ny && g->nz == (dim == 3 ? nz : 1) &&
           std::memcmp(g->solid.data(), solid, (size_t)nn) == 0)
            return g;
    }
    std::unique_ptr<Geometry> g(new Geometry);
    err = geometry_build(*g, solid, dim, nx, ny, nz);
    if(!err.empty()) return nullptr;
    Geometry *raw = g.get();
    B->geos.push_back(std::move(g));
    B->geo_index.emplace(h, raw);
    return raw;
}

extern "C" {

fp_batch *fp_batch_create(fp_ctx *ctx, int lattice, int precision, int *status)
{
    if(!ctx || (lattice != 2 && lattice != 3) ||
       (lattice == 3 && precision != FP_F64 && precision != FP_F32) ||
       (lattice == 2 && precision != FP_F64 && precision != FP_F32 && precision != FP_F32_REF2D)) {
        if(status) *status = FP_ERR_ARG;
        return nullptr;
    }
    fp_batch *B = new(std::nothrow) fp_batch;
    if(!B) { if(status) *status = FP_ERR_OOM; return nullptr; }
    B->ctx = ctx;
    B->lattice = lattice;
    B->prec = precision;
    B->b.reset(new Batch(lattice, precision, ctx->stream, ctx->chunk, true));
    B->b->set_tuning(ctx->tune);
    if(status) *status = FP_OK;
    return B;
}

void fp_batch_destroy(fp_batch *B)
{
    if(!B) return;
    B->b.reset();
    delete B;
}

static int64_t batch_add(fp_batch *B, Geometry *g, const InstCfg &cfg, const double F[3])
{
    std::unique_ptr<Job> j(new Job);
    j->geo = g;
    j->lattice = B->lattice;
    j->prec = B->prec;
    for(int k = 0; k < 3; k++) j->F[k] = F[k];
    j->cfg = cfg;
    j->want_fields = (B->lattice == 2);
    j->id = (int64_t)B->jobs.size();
    std::string e = B->b->admit(j.get());
    if(!e.empty()) {
        fail(B->ctx, e.find("memory") != std::string::npos ? FP_ERR_OOM : FP_ERR_DEVICE, e);
        return e.find("memory") != std::string::npos ? -FP_ERR_OOM : -FP_ERR_DEVICE;
    }
    B->jobs.push_back(std::move(j));
    B->results.emplace_back();
    B->have.push_back(0);
    return (int64_t)B->jobs.size() - 1;
}

int64_t fp_batch_add3d(fp_batch *B, const uint8_t *solid, int nx, int ny, int nz,
                       const fp_config3d *cfg, const double F[3])
{
    if(!B || !solid || !cfg3_ok(cfg) || B->lattice != 3) return -FP_ERR_ARG;
    if(nx <= 0 || ny <= 0 || nz <= 0) return -FP_ERR_BAD_DIMS;
    int st = check_mask(solid, (int64_t)nx * ny * nz, nullptr);
    if(st != FP_OK) return -st;
    (void)hipSetDevice(B->ctx->device);
    std::string err;
    Geometry *g = batch_geometry(B, solid, 3, nx, ny, nz, err);
    if(!g) return -(int64_t)status_of_geometry_error(err);
    return batch_add(B, g, cfg3(*cfg, g->nn, false), F);
}

int64_t fp_batch_add2d(fp_batch *B, const uint8_t *solid, int lx, int ly,
                       const fp_config2d *cfg, const double F[2])
{
    if(!B || !solid || !cfg2_ok(cfg) || B->lattice != 2) return -FP_ERR_ARG;
    if(lx <= 0 || ly <= 0) return -FP_ERR_BAD_DIMS;
    int st = check_mask(solid, (int64_t)lx * ly, nullptr);
    if(st != FP_OK) return -st;
    (void)hipSetDevice(B->ctx->device);
    std::string err;
    Geometry *g = batch_geometry(B, solid, 2, lx, ly, 1, err);
    if(!g) return -(int64_t)status_of_geometry_error(err);
    const double F3[3] = {F[0], F[1], 0.0};
    return batch_add(B, g, cfg2(*cfg, lx, ly), F3);
}

int fp_batch_run(fp_batch *B)
{
    if(!B) return FP_ERR_ARG;
    (void)hipSetDevice(B->ctx->device);
    const double t0 = now_seconds();
    auto collect = [&]() {
        for(Job *j : B->b->take_done()) {
            fill_out(*j, &B->results[(size_t)j->id]);
            B->have[(size_t)j->id] = 1;
            B->updates += j->state.steps * j->geo->n_fluid;
            B->b->release(j);
        }
    };
    collect();
    while(B->b->live() > 0) {
        std::string e = B->b->run_chunk();
        if(!e.empty()) return fail(B->ctx, FP_ERR_DEVICE, e);
        collect();
    }
    B->seconds += now_seconds() - t0;
    return FP_OK;
}

int fp_batch_result(fp_batch *B, int64_t id, fp_solve_out *out)
{
    if(!B || !out) return FP_ERR_ARG;
    if(id < 0 || id >= (int64_t)B->jobs.size() || !B->have[(size_t)id]) return FP_ERR_NOT_FOUND;
    *out = B->results[(size_t)id];
    return FP_OK;
}

void fp_batch_stats(fp_batch *B, int64_t *steps, int64_t *fluid_updates, double *seconds)
{
    if(!B) return;
    if(steps) *steps = B->b->stats().steps;
    if(fluid_updates) *fluid_updates = B->updates;
    if(seconds) *seconds = B->seconds;
}

void fp_batch_set_chunk(fp_batch *B, int steps)
{
    if(B) B->b->set_chunk(steps);
}

} // extern "C"

// ---------------------------------------------------------------------------
// benchmark entry: time stepping of one resident lattice
// ---------------------------------------------------------------------------
static bool tuning_ok(const fp_tuning *t)
{
    return t && (t->local_cpt == 1 || t->local_cpt == 2 || t->local_cpt == 4) &&
           t->nbr_waves >= 1 && t->nbr_waves <= 8 && t->chunk_steps >= 2 &&
           (t->nontemporal == 0 || t->nontemporal == 1) &&
           t->local_bpc >= 0 && t->local_bpc <= 32 && t->nbr_bpc >= 0 && t->nbr_bpc <= 32 &&
           (t->nbr_mode >= 0 && t->nbr_mode <= 2) &&
           (t->nbr_segs == 0 || t->nbr_segs == 1 || t->nbr_segs == 2 || t->nbr_segs == 4 ||
            t->nbr_segs == 8) &&
           t->local_wgs >= 0 && t->local_wgs <= 32 && t->nbr_wgs >= 0 && t->nbr_wgs <= 32;
}

extern "C" int fp_set_tuning(fp_ctx *ctx, const fp_tuning *t)
{
    if(!ctx || !tuning_ok(t)) return FP_ERR_ARG;
    copy_tuning(t, ctx->tune);
    ctx->chunk = t->chunk_steps & ~1;
    return FP_OK;
}

extern "C" void fp_get_tuning(fp_ctx *ctx, fp_tuning *t)
{
    if(!ctx || !t) return;
    t->local_cpt = ctx->tune.local_cpt;
    t->nbr_waves = ctx->tune.nbr_waves;
    t->chunk_steps = ctx->chunk;
    t->nontemporal = ctx->tune.nontemporal;
    t->local_bpc = ctx->tune.local_bpc;
    t->nbr_bpc = ctx->tune.nbr_bpc;
    t->nbr_mode = ctx->tune.nbr_mode;
    t->nbr_segs = ctx->tune.nbr_segs;
    t->local_wgs = ctx->tune.local_wgs;
    t->nbr_wgs = ctx->tune.nbr_wgs;
}

extern "C" int fp_bench_steps(fp_ctx *ctx, const uint8_t *solid, int nx, int ny, int nz, int lattice,
                              int precision, const double F[3], int warmup, int steps,
                              const fp_tuning *tune, int flux_every_step, fp_bench_out *out)
{
    if(tune && !tuning_ok(tune)) return FP_ERR_ARG;
    if(!ctx || !solid || !out || (lattice != 2 && lattice != 3) || steps < 2 || warmup < 0)
        return FP_ERR_ARG;
    std::memset(out, 0, sizeof(*out));
    (void)hipSetDevice(ctx->device);
    Geometry g;
    std::string e = geometry_build(g, solid, lattice, nx, ny, lattice == 3 ? nz : 1);
    if(!e.empty()) return fail(ctx, status_of_geometry_error(e), e);
    Job j;
    j.geo = &g;
    j.lattice = lattice;
    j.prec = precision;
    for(int k = 0; k < 3; k++) j.F[k] = F[k];
    if(lattice == 3) {
        fp_config3d c;
        fp_config3d_default(&c);
        c.eps = 0.0;                                  // never converges: pure stepping
        c.max_steps = INT64_MAX / 4;
        if(flux_every_step) c.lag = 1;                // a convergence check every step
        j.cfg = cfg3(c, g.nn, false);
    } else {
        fp_config2d c;
        fp_config2d_default(&c);
        c.eps = 0.0; c.ksigma = 0.0; c.qfloor = 0.0;
        c.max_steps = INT64_MAX / 4;
        if(flux_every_step) { c.lag = 1; c.avg_sample = 1; }
        j.cfg = cfg2(c, nx, ny);
    }
    Batch b(lattice, precision, ctx->stream, 2, false);
    Tuning tu = ctx->tune;
    if(tune) copy_tuning(tune, tu);
    b.set_tuning(tu);
    e = b.admit(&j);
    if(!e.empty()) return fail(ctx, e.find("memory") != std::string::npos ? FP_ERR_OOM : FP_ERR_DEVICE, e);
    Batch::KernelTimes kw, kt;
    if(warmup >= 2) {
        e = b.time_steps(warmup, kw);
        if(!e.empty()) { b.detach_all(); free_job_memory(j); return fail(ctx, FP_ERR_DEVICE, e); }
    }
    e = b.time_steps(steps, kt);
    float graph_ms = 0.0f;
    if(e.empty()) e = b.graph_steps(steps, graph_ms);
    const bool deferred = b.deferred();
    b.detach_all();
    free_job_memory(j);
    if(!e.empty()) return fail(ctx, FP_ERR_DEVICE, e);
    out->n_fluid = g.n_fluid;
    out->n_cells = g.nn;
    out->steps = steps & ~1;
    out->total_ms = graph_ms;
    out->eager_total_ms = kt.total_ms;
    out->local_ms = kt.n_local ? kt.local_ms / kt.n_local : 0;
    out->nbr_ms = kt.n_nbr ? kt.nbr_ms / kt.n_nbr : 0;
    out->monitor_ms = kt.n_monitor ? kt.monitor_ms / kt.n_monitor : 0;
    out->bytes_per_update = (double)(lattice == 3 ? 19 : 9) * 2 * (precision == FP_F64 ? 8 : 4);
    out->deferred = deferred ? 1 : 0;
    out->pad = 0;
    return FP_OK;
}

extern "C" int fp_bench_copy(fp_ctx *ctx, int64_t bytes, int iters, double *ms_per_copy)
{
    if(!ctx || bytes < 16 || iters < 1 || !ms_per_copy) return FP_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    const int64_t n16 = bytes / 16;
    void *a = nullptr, *b = nullptr;
    if(hipMalloc(&a, n16 * 16) != hipSuccess) { (void)hipGetLastError(); return fail(ctx, FP_ERR_OOM, "copy buffers"); }
    if(hipMalloc(&b, n16 * 16) != hipSuccess) { (void)hipFree(a); (void)hipGetLastError(); return fail(ctx, FP_ERR_OOM, "copy buffers"); }
    (void)hipMemsetAsync(a, 0, n16 * 16, ctx->stream);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    double best = 1e30;
    hipError_t e = hipSuccess;
    for(int kind = 0; kind < N_COPY_KINDS && e == hipSuccess; kind++) {
        kern_copy(kind, a, b, n16, ctx->stream);           // warm
        (void)hipEventRecord(e0, ctx->stream);
        for(int i = 0; i < iters; i++) kern_copy(kind, a, b, n16, ctx->stream);
        (void)hipEventRecord(e1, ctx->stream);
        e = hipEventSynchronize(e1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if(e == hipSuccess) best = std::min(best, (double)ms / iters);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if(e != hipSuccess) return fail(ctx, FP_ERR_DEVICE, hipGetErrorString(e));
    *ms_per_copy = best;
    return FP_OK;
}

// ---------------------------------------------------------------------------
// autotuner: time every launch shape of the two step kernels on one resident
// lattice and keep the fastest of each.  Run i pairs local shape i with
// neighbour shape i mod FP_TUNE_NNBR (the kernels are timed independently by
// events).
// ---------------------------------------------------------------------------
// local step: cells/thread, blocks/CU (0: one per item), non-temporal,
// resident workgroups per CU (local_wgs, 0: no cap)
static const int TUNE_LOCAL[FP_TUNE_NLOCAL][4] = {
    {1, 0, 0, 0}, {2, 0, 0, 0}, {4, 0, 0, 0},
    {1, 0, 1, 0}, {1, 2, 1, 0}, {1, 8, 1, 0}, {1, 16, 1, 0},
    {2, 0, 1, 0}, {2, 2, 1, 0}, {2, 8, 1, 0}, {2, 16, 1, 0},
    {4, 0, 1, 0}, {4, 2, 1, 0}, {4, 8, 1, 0}, {4, 16, 1, 0},
    {1, 0, 1, 2}, {1, 0, 1, 3}, {1, 0, 1, 4}, {1, 0, 1, 5}, {1, 0, 1, 6},
    {2, 0, 1, 2}, {2, 0, 1, 3}, {2, 0, 1, 4}, {4, 0, 1, 2}};
// neighbour step: segments per workgroup (row-segment form), blocks/CU (0:
// one per item), form (0 row segments, 1 compact), x-adjacent segments of a
// row among them (nbr_segs), resident workgroups per CU (nbr_wgs)
static const int TUNE_NBR[FP_TUNE_NNBR][5] = {
    {4, 0, 0, 4, 0}, {8, 0, 0, 4, 0}, {8, 0, 0, 8, 0}, {2, 0, 0, 2, 0}, {4, 0, 0, 2, 0},
    {8, 0, 0, 2, 0}, {2, 0, 0, 1, 0}, {4, 0, 0, 1, 0}, {8, 0, 0, 1, 0}, {4, 0, 0, 4, 5},
    {4, 0, 0, 4, 6}, {1, 0, 1, 1, 0}, {1, 2, 1, 1, 0}, {1, 0, 1, 1, 5}, {1, 8, 1, 1, 0},
    {1, 16, 1, 1, 0}};

extern "C" void fp_tune_shape(int kind, int i, int *a, int *b, int *c)
{
    *a = *b = *c = 0;
    if(kind == 0 && i >= 0 && i < FP_TUNE_NLOCAL) { *a = TUNE_LOCAL[i][0]; *b = TUNE_LOCAL[i][1]; *c = TUNE_LOCAL[i][2]; }
    else if(kind == 1 && i >= 0 && i < FP_TUNE_NNBR) { *a = TUNE_NBR[i][0]; *b = TUNE_NBR[i][1]; *c = TUNE_NBR[i][2]; }
    else if(kind == 2 && i >= 0 && i < FP_TUNE_NNBR) *a = TUNE_NBR[i][3];
    else if(kind == 3 && i >= 0 && i < FP_TUNE_NLOCAL) *a = TUNE_LOCAL[i][3];
    else if(kind == 4 && i >= 0 && i < FP_TUNE_NNBR) *a = TUNE_NBR[i][4];
}

extern "C" int fp_autotune(fp_ctx *ctx, const uint8_t *solid, int nx, int ny, int nz, int lattice,
                           int precision, int steps, fp_tuning *best, double *local_ms, double *nbr_ms)
{
    if(!ctx || !solid || !best || steps < 2 || (lattice != 2 && lattice != 3)) return FP_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    Geometry g;
    std::string e = geometry_build(g, solid, lattice, nx, ny, lattice == 3 ? nz : 1);
    if(!e.empty()) return fail(ctx, status_of_geometry_error(e), e);
    Job j;
    j.geo = &g;
    j.lattice = lattice;
    j.prec = precision;
    j.F[0] = 2.5e-7 * 40;
    if(lattice == 3) {
        fp_config3d c;
        fp_config3d_default(&c);
        c.eps = 0.0;
        c.max_steps = INT64_MAX / 4;
        j.cfg = cfg3(c, g.nn, false);
    } else {
        fp_config2d c;
        fp_config2d_default(&c);
        c.eps = 0.0; c.ksigma = 0.0; c.qfloor = 0.0;
        c.max_steps = INT64_MAX / 4;
        j.cfg = cfg2(c, nx, ny);
    }
    Batch b(lattice, precision, ctx->stream, 2, false);
    e = b.admit(&j);
    if(!e.empty()) { b.detach_all(); free_job_memory(j); return fail(ctx, FP_ERR_OOM, e); }
    double lt[FP_TUNE_NLOCAL], nt[FP_TUNE_NNBR];
    for(auto &v : lt) v = 1e30;
    for(auto &v : nt) v = 1e30;
    const int runs = std::max(FP_TUNE_NLOCAL, FP_TUNE_NNBR);
    for(int r = 0; r < runs && e.empty(); r++) {
        const int i = r % FP_TUNE_NLOCAL, n = r % FP_TUNE_NNBR;
        Tuning t = ctx->tune;
        t.local_cpt = TUNE_LOCAL[i][0]; t.local_bpc = TUNE_LOCAL[i][1]; t.nontemporal = TUNE_LOCAL[i][2];
        t.local_wgs = TUNE_LOCAL[i][3];
        t.nbr_waves = TUNE_NBR[n][0]; t.nbr_bpc = TUNE_NBR[n][1]; t.nbr_mode = TUNE_NBR[n][2];
        t.nbr_segs = TUNE_NBR[n][3]; t.nbr_wgs = TUNE_NBR[n][4];
        b.set_tuning(t);
        Batch::KernelTimes kw, kt;
        e = b.time_steps(2, kw);
        if(e.empty()) e = b.time_steps(steps, kt);
        if(!e.empty()) break;
        lt[i] = std::min(lt[i], kt.n_local ? kt.local_ms / kt.n_local : 1e30);
        // a form the structure cannot use (the batch fell back) is not a result
        if(b.nbr_mode() == t.nbr_mode)
            nt[n] = std::min(nt[n], kt.n_nbr ? kt.nbr_ms / kt.n_nbr : 1e30);
    }
    b.detach_all();
    free_job_memory(j);
    if(!e.empty()) return fail(ctx, FP_ERR_DEVICE, e);
    int bl = 0, bn = 0;
    for(int i = 1; i < FP_TUNE_NLOCAL; i++) if(lt[i] < lt[bl]) bl = i;
    for(int i = 1; i < FP_TUNE_NNBR; i++) if(nt[i] < nt[bn]) bn = i;
    best->local_cpt = TUNE_LOCAL[bl][0];
    best->local_bpc = TUNE_LOCAL[bl][1];
    best->nontemporal = TUNE_LOCAL[bl][2];
    best->nbr_waves = TUNE_NBR[bn][0];
    best->nbr_bpc = TUNE_NBR[bn][1];
    best->nbr_mode = TUNE_NBR[bn][2];
    best->nbr_segs = TUNE_NBR[bn][3];
    best->local_wgs = TUNE_LOCAL[bl][3];
    best->nbr_wgs = TUNE_NBR[bn][4];
    best->chunk_steps = ctx->chunk;
    if(local_ms) for(int i = 0; i < FP_TUNE_NLOCAL; i++) local_ms[i] = lt[i];
    if(nbr_ms) for(int i = 0; i < FP_TUNE_NNBR; i++) nbr_ms[i] = nt[i];
    return fp_set_tuning(ctx, best);
}
