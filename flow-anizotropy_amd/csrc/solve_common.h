// solve_common.h -- host helpers shared by the one-shot solves (api.hip) and
// the request server (server.hip): the context, config translation, result
// assembly (K tensor and derived values) and mask checks.
#pragma once
#include "../../include/flowperm.h"
#include "engine.h"

#include <string>
#include <vector>

struct fp_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    fp::Tuning tune;
    int chunk = 200;
    fp_progress_fn progress = nullptr;   // called after every convergence check (fp_set_progress)
    void *progress_user = nullptr;
};

namespace fp {

constexpr double FX_BASE = 2.5e-07;             // LMB3d/lbm3d-perm.cpp:115, LMB2d/main.cpp:62
constexpr double TAU = 1.0, CS2 = 1.0 / 3.0;
constexpr double MU3 = CS2 * (TAU - 0.5);       // LMB3d/lbm3d-perm.cpp:114
float mu2d();                                   // 0.33 * (tau - 1/2) in float, LMB2d/lbm.cpp:44-45

int fail(fp_ctx *ctx, int status, const std::string &msg);
int status_of_geometry_error(const std::string &e);
int check_mask(const uint8_t *solid, int64_t nn, int64_t *n_pore);
const char *msg_of(int st);
bool cfg3_ok(const fp_config3d *c);
bool cfg2_ok(const fp_config2d *c);

InstCfg cfg3(const fp_config3d &c, int64_t nn, bool want_vel);
InstCfg cfg2(const fp_config2d &c, int lx, int ly);
void fill_out(const Job &j, fp_solve_out *o);
void jacobi3(const double Ain[3][3], double val[3], double vec[3][3]);

void assemble3d(fp_result3d *out, int nx, int ny, int nz, int64_t n_pore, double F);
void assemble3d_val(fp_result3d *out, double F);
void assemble2d(fp_result2d *out, int lx, int ly, const uint8_t *solid, double F);
void validate2d_force(double F, double deg, double Fv[2]);
void assemble2d_val(fp_result2d *out, double F, double deg);
std::string finish2d(const Job &j, int lx, int ly, int64_t avg_steps, fp_solve_out *o, float *vel);
// the allocation-free solve's scratch array: look up its address on `device`
// (the current device), once (noalloc.hip)
hipError_t noalloc_setup(int device);

} // namespace fp
