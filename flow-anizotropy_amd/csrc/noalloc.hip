// noalloc.hip -- fp_solve_dir3d_noalloc: one forced D3Q19 solve whose device
// memory all lies in a fixed array of the library, for callers that cannot
// allocate device memory (include/flowperm.h).
//
// Every buffer the step and finish kernels touch -- the segment table, link
// records and wrap table, the populations, the instance descriptor, config and
// state, the work lists, flux partials, reduction items and the finish
// counter -- is carved out of `scratch`, a file-scope __device__ array of
// FP_NOALLOC_SCRATCH_BYTES.  Its device address is looked up once per device
// (hipGetSymbolAddress, from fp_create) and only that address is used by host
// code: kernel arguments and copies.  The kernels are the ones fp_solve_dir3d
// runs (row-segment neighbour form), launched one by one on the context's
// stream; the results are bit-identical to fp_solve_dir3d's.
#include "solve_common.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace fp;

static __device__ __align__(256) unsigned char scratch[FP_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_setup_mu;
void *g_scratch[MAX_DEVICES] = {};            // device address of scratch, per device
hipError_t g_scratch_err[MAX_DEVICES] = {};
std::mutex g_call_mu[MAX_DEVICES];            // one call at a time per device

size_t align256(size_t v) { return (v + 255) / 256 * 256; }

// offsets of the buffers inside scratch
struct Layout {
    size_t seg, links, rowspec, f, inst, cfg, state, lwork, nwork, lpart, npart;
    size_t lredw, nredw, lred, nred, count, total;
};

Layout plan(const Geometry &g, size_t links_bytes, size_t rowspec_bytes, size_t f_bytes,
            const BatchTables &t)
{
    Layout L;
    size_t at = 0;
    auto take = [&](size_t bytes) { const size_t o = align256(at); at = o + bytes; return o; };
    L.seg = take((size_t)g.nunits * sizeof(SegInfo));
    L.links = take(links_bytes);
    L.rowspec = take(rowspec_bytes);
    L.f = take(f_bytes);
    L.inst = take(sizeof(InstDesc));
    L.cfg = take(sizeof(InstCfg));
    L.state = take(sizeof(InstState));
    L.lwork = take(t.lw.size() * sizeof(LocalWork));
    L.nwork = take(t.nw.size() * sizeof(NbrWork));
    L.lpart = take(t.lw.size() * sizeof(Partial));
    L.npart = take(t.nw.size() * sizeof(Partial));
    L.lredw = take(t.lr.size() * sizeof(RedWork));
    L.nredw = take(t.nr.size() * sizeof(RedWork));
    L.lred = take(t.lr.size() * sizeof(RPartial));
    L.nred = take(t.nr.size() * sizeof(RPartial));
    L.count = take(sizeof(unsigned int));
    L.total = at;
    return L;
}

} // namespace

namespace fp {

// the device address of scratch on `device` (the current device), looked up
// once; the library's context creation calls it
hipError_t noalloc_setup(int device)
{
    if(device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_setup_mu);
    if(g_scratch[device]) return hipSuccess;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_err[device] = e;
    if(e == hipSuccess) g_scratch[device] = p;
    return e;
}

} // namespace fp

extern "C" int fp_solve_dir3d_noalloc(fp_ctx *ctx, const uint8_t *solid, int nx, int ny, int nz,
                                      const fp_config3d *cfg, const double F[3], fp_solve_out *out)
{
    if(!ctx || !solid || !F || !out) return FP_ERR_ARG;
    std::memset(out, 0, sizeof(*out));
    if(!cfg3_ok(cfg)) return fail(ctx, FP_ERR_ARG, "lag must be positive");
    if(cfg->precision != FP_F64 && cfg->precision != FP_F32)
        return fail(ctx, FP_ERR_ARG, "3D precision must be FP_F64 or FP_F32");
    if(nx <= 0 || ny <= 0 || nz <= 0) return fail(ctx, FP_ERR_BAD_DIMS, "bad dimensions");
    const int st = check_mask(solid, (int64_t)nx * ny * nz, nullptr);
    if(st != FP_OK) return fail(ctx, st, msg_of(st));
    const int dev = ctx->device;
    if(dev < 0 || dev >= MAX_DEVICES) return fail(ctx, FP_ERR_DEVICE, "device index out of range");
    if(hipSetDevice(dev) != hipSuccess) return fail(ctx, FP_ERR_DEVICE, "hipSetDevice failed");
    unsigned char *base;
    {
        std::lock_guard<std::mutex> lk(g_setup_mu);
        base = (unsigned char *)g_scratch[dev];
        if(!base)
            return fail(ctx, FP_ERR_DEVICE, std::string("scratch array not available: ") +
                                                hipGetErrorString(g_scratch_err[dev]));
    }
    std::lock_guard<std::mutex> call(g_call_mu[dev]);

    // host tables (no device memory)
    Geometry g;
    std::vector<SegInfo> seg;
    std::vector<unsigned char> links;
    std::vector<uint32_t> rowspec;
    std::string e = geometry_host_tables(g, solid, 3, nx, ny, nz, seg, links, rowspec);
    if(!e.empty()) return fail(ctx, status_of_geometry_error(e), e);
    const int prec = cfg->precision;
    const size_t elem = prec == FP_F64 ? 8 : 4;
    Job j;
    j.geo = &g;
    j.lattice = 3;
    j.prec = prec;
    for(int k = 0; k < 3; k++) j.F[k] = F[k];
    j.cfg = cfg3(*cfg, g.nn, false);
    j.stride = pop_stride_of(g.n_fluid);
    j.nbr_mode = 0;
    Tuning tu = ctx->tune;
    tu.nbr_mode = 0;
    std::vector<Job *> jobs{&j};
    BatchTables t;
    build_tables(jobs, 3, prec, tu, t);
    const size_t f_bytes = (size_t)19 * (size_t)j.stride * elem;
    const Layout L = plan(g, links.size(), rowspec.size() * sizeof(uint32_t), f_bytes, t);
    if(L.total > (size_t)FP_NOALLOC_SCRATCH_BYTES)
        return fail(ctx, FP_ERR_ARG, "lattice too large for fp_solve_dir3d_noalloc: it needs " +
                                         std::to_string(L.total) + " bytes of device memory, the " +
                                         "scratch array holds " +
                                         std::to_string((size_t)FP_NOALLOC_SCRATCH_BYTES));
    // the descriptor points into scratch (the Geometry never owns these)
    g.d_seg = (SegInfo *)(base + L.seg);
    g.d_links = base + L.links;
    g.d_rowspec = (uint32_t *)(base + L.rowspec);
    j.d_f = base + L.f;
    build_tables(jobs, 3, prec, tu, t);
    g.d_seg = nullptr;
    g.d_links = nullptr;
    g.d_rowspec = nullptr;

    InstState S;
    init_state(j.cfg, S);
    const unsigned int zero = 0u;
    hipStream_t s = ctx->stream;
    auto put = [&](size_t off, const void *src, size_t bytes) {
        return bytes ? hipMemcpyAsync(base + off, src, bytes, hipMemcpyHostToDevice, s) : hipSuccess;
    };
    hipError_t he = put(L.seg, seg.data(), seg.size() * sizeof(SegInfo));
    if(he == hipSuccess) he = put(L.links, links.data(), links.size());
    if(he == hipSuccess) he = put(L.rowspec, rowspec.data(), rowspec.size() * sizeof(uint32_t));
    if(he == hipSuccess) he = put(L.inst, t.desc.data(), sizeof(InstDesc));
    if(he == hipSuccess) he = put(L.cfg, &j.cfg, sizeof(InstCfg));
    if(he == hipSuccess) he = put(L.state, &S, sizeof(InstState));
    if(he == hipSuccess) he = put(L.lwork, t.lw.data(), t.lw.size() * sizeof(LocalWork));
    if(he == hipSuccess) he = put(L.nwork, t.nw.data(), t.nw.size() * sizeof(NbrWork));
    if(he == hipSuccess) he = put(L.lredw, t.lr.data(), t.lr.size() * sizeof(RedWork));
    if(he == hipSuccess) he = put(L.nredw, t.nr.data(), t.nr.size() * sizeof(RedWork));
    if(he == hipSuccess) he = put(L.count, &zero, sizeof(zero));      // the finish counter starts at 0
    if(he != hipSuccess) return fail(ctx, FP_ERR_DEVICE, std::string("upload: ") + hipGetErrorString(he));
    kern_rest(3, prec, t.desc[0], s);

    const InstDesc *d_inst = (const InstDesc *)(base + L.inst);
    const InstCfg *d_cfg = (const InstCfg *)(base + L.cfg);
    InstState *d_state = (InstState *)(base + L.state);
    const int n_l = (int)t.lw.size(), n_n = (int)t.nw.size();
    const int n_lr = (int)t.lr.size(), n_nr = (int)t.nr.size();
    const int chunk = std::max(2, ctx->chunk & ~1);
    // a finish only after the steps on which the convergence rules can act
    // (plan_finishes), as in a batch; FLOWPERM_SPARSE_FINISH=0: after every step
    const char *sv = getenv("FLOWPERM_SPARSE_FINISH");
    const bool sparse = !(sv && sv[0] == '0');
    std::vector<int> adv(chunk, 1);
    while(S.phase < PH_DONE) {
        if(sparse) plan_finishes(&j.cfg, &S, 1, 0, chunk, 1, adv);
        // the workgroup caps only in chunks that skip finishes, as a batch
        // applies them (a flux sum every step runs faster uncapped)
        Tuning tc = tu;
        if(!sparse || std::count(adv.begin(), adv.end(), 0) == 0) tc.local_wgs = tc.nbr_wgs = 0;
        // every chunk starts at an even step count: local (odd) step first
        for(int step = 0; step < chunk; step++) {
            const bool local = (step & 1) == 0;
            if(local) {
                kern_step(3, prec, true, tc, AD_B32, d_inst, d_state, base + L.lwork, n_l,
                          (Partial *)(base + L.lpart), s);
                if(adv[step] > 0)
                    kern_finish(n_lr, (const RedWork *)(base + L.lredw), d_inst, d_cfg, d_state,
                                (const Partial *)(base + L.lpart), (RPartial *)(base + L.lred),
                                (unsigned int *)(base + L.count), true, s, adv[step]);
            } else {
                kern_step(3, prec, false, tc, AD_B32, d_inst, d_state, base + L.nwork, n_n,
                          (Partial *)(base + L.npart), s);
                if(adv[step] > 0)
                    kern_finish(n_nr, (const RedWork *)(base + L.nredw), d_inst, d_cfg, d_state,
                                (const Partial *)(base + L.npart), (RPartial *)(base + L.nred),
                                (unsigned int *)(base + L.count), false, s, adv[step]);
            }
        }
        he = hipGetLastError();
        if(he == hipSuccess) he = hipMemcpyAsync(&S, d_state, sizeof(InstState), hipMemcpyDeviceToHost, s);
        if(he == hipSuccess) he = hipStreamSynchronize(s);
        if(he != hipSuccess) return fail(ctx, FP_ERR_DEVICE, std::string("device: ") + hipGetErrorString(he));
    }
    j.state = S;
    fill_out(j, out);
    if(out->status != FP_OK)
        return fail(ctx, FP_ERR_NONFINITE, "non-finite flux in run at step " + std::to_string(out->steps));
    ctx->err.clear();
    return out->converged ? FP_OK : FP_NOT_CONVERGED;
}
