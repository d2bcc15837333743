// kernels.hip -- A-A in-place stream-collide kernels on fluid-only compact
// storage, exact flux partials, and the device-side convergence monitor.
//
// Storage.  Fluid cells are numbered in the .raw order (z, y, x fastest)
// skipping solid ones; population k of cell i lives at f[k*stride + i].
// Solid cells are never stored.  Each lattice row is cut into 64-cell
// segments; SegInfo holds the segment's fluid bits and the compact index of
// its first fluid cell, from which any neighbour's index and solidity follow
// (start + popcount of the bits below it) -- no per-cell neighbour table.
//
// A-A streaming (one array, no second buffer).  The reference alternates a
// priming collide with pull stream+collide steps over two buffers
// (LMB3d/lbm3d-perm.cpp:249-295).  The same sequence of collisions is:
//   odd step  t (1, 3, ...): "local"  -- read f[k][i], collide, write f[opp k][i]
//   even step t (2, 4, ...): "neighbour" -- read f[opp k][i - e_k] (or f[k][i]
//       when i - e_k is solid: halfway bounce-back), collide, write to
//       f[k][i + e_k] (or f[opp k][i] when i + e_k is solid).
// Step 1 is the reference's collide_only on the rest state; every later step
// streams the previous step's post-collision values, exactly as the pull
// stream_collide does.  Each location is read and written by one thread only.
#include "collide.h"

namespace fp {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

// Workgroups are dealt round-robin to the 8 XCDs; give each XCD one
// contiguous range of work so neighbouring segments share its L2.
__device__ __forceinline__ int xcd_remap(int b, int G)
{
    const int q = G >> 3, r = G & 7;
    const int x = b & 7, j = b >> 3;
    return x < r ? x * (q + 1) + j : r * (q + 1) + (x - r) * q + j;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for(int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// exact block reduction of the three chunk sums of up to three components
template<int NC>
__device__ __forceinline__ void block_flux(const double *u, int bad, Partial *out)
{
    __shared__ double sh[16][9];
    __shared__ int shbad[16];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double c[NC][3];
#pragma unroll
    for(int k = 0; k < NC; k++) fix_split(u[k], c[k][0], c[k][1], c[k][2], bad);
#pragma unroll
    for(int k = 0; k < NC; k++)
#pragma unroll
        for(int j = 0; j < 3; j++) c[k][j] = wave_sum(c[k][j]);
    const int anybad = __any(bad);
    if(lane == 0) {
#pragma unroll
        for(int k = 0; k < NC; k++)
#pragma unroll
            for(int j = 0; j < 3; j++) sh[wave][k*3 + j] = c[k][j];
        shbad[wave] = anybad;
    }
    __syncthreads();
    if(threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        Partial p;
        p.bad = 0; p.pad = 0;
        for(int k = 0; k < 3; k++)
            for(int j = 0; j < 3; j++) {
                int64_t acc = 0;
                if(k < NC)
                    for(int w = 0; w < nw; w++) acc += (int64_t)sh[w][k*3 + j];
                p.c[k][j] = acc;
            }
        for(int w = 0; w < nw; w++) p.bad |= shbad[w];
        *out = p;
    }
}

template<class L, int P> struct Store;
template<> struct Store<D3Q19, PREC_F64> { using T = double; };
template<> struct Store<D3Q19, PREC_F32> { using T = float; };
template<> struct Store<D2Q9, PREC_F64> { using T = double; };
template<> struct Store<D2Q9, PREC_F32> { using T = float; };
template<> struct Store<D2Q9, PREC_F32_REF2D> { using T = float; };

// moments + collision of one node in place; u receives the velocity that
// the flux sums use (Guo physical velocity in 3D, raw moment in 2D)
template<class L, int P>
__device__ __forceinline__ void node(typename Store<L, P>::T *v, const InstDesc &D, double *u)
{
    if constexpr(L::DIM == 3) {
        if constexpr(P == PREC_F64) {
            const double F[3] = {D.F[0], D.F[1], D.F[2]};
            collide3_f64(v, F, u[0], u[1], u[2]);
        } else {
            const float F[3] = {D.Ff[0], D.Ff[1], D.Ff[2]};
            float a, b, c;
            collide3_f32(v, F, D.g45, D.h15, a, b, c);
            u[0] = a; u[1] = b; u[2] = c;
        }
    } else {
        if constexpr(P == PREC_F64) {
            double a, b;
            collide2_ref<double>(v, D.F[0], D.F[1], a, b);
            u[0] = a; u[1] = b; u[2] = 0.0;
        } else if constexpr(P == PREC_F32_REF2D) {
            float a, b;
            collide2_ref<float>(v, D.Ff[0], D.Ff[1], a, b);
            u[0] = a; u[1] = b; u[2] = 0.0;
        } else {
            float a, b;
            collide2_f32(v, D.Ff[0], D.Ff[1], a, b);
            u[0] = a; u[1] = b; u[2] = 0.0;
        }
    }
}

template<class L>
__device__ __forceinline__ void write_vel(const InstDesc &D, int slot, int64_t i, const double *u)
{
    float *v = D.vel[slot - 1];
    if(!v) return;
#pragma unroll
    for(int c = 0; c < L::DIM; c++) v[i * L::DIM + c] = (float)u[c];
}

// ---------------------------------------------------------------------------
// rest state: f = w (unshifted) or h = 0 (shifted)
// ---------------------------------------------------------------------------
template<class L, int P>
__global__ void __launch_bounds__(256)
k_rest(InstDesc D)
{
    using T = typename Store<L, P>::T;
    T *f = (T *)D.f;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= D.stride) return;
#pragma unroll
    for(int k = 0; k < L::Q; k++) {
        T w;
        if constexpr(P == PREC_F32) w = 0;
        else if constexpr(L::DIM == 3) w = (T)L::w(k);
        else if constexpr(P == PREC_F64) w = (T)(k == 0 ? 4.0/9.0 : (k < 5 ? 1.0/9.0 : 1.0/36.0));
        else w = (T)L::wf(k);
        f[k * D.stride + i] = w;
    }
}

// ---------------------------------------------------------------------------
// local step (odd t): one thread per compact cell
// ---------------------------------------------------------------------------
template<class L, int P>
__global__ void __launch_bounds__(256)
k_local(const InstDesc *__restrict__ insts, const InstState *__restrict__ st,
        const LocalWork *__restrict__ work, Partial *__restrict__ part)
{
    using T = typename Store<L, P>::T;
    const int wb = xcd_remap(blockIdx.x, gridDim.x);
    const LocalWork w = work[wb];
    const int phase = st[w.inst].phase;
    if(phase >= PH_DONE) return;
    const int flux = st[w.inst].flux_next;
    const int vslot = st[w.inst].vel_next;
    const InstDesc &D = insts[w.inst];
    T *f = (T *)D.f;
    const int64_t stride = D.stride;

    const bool act = (int)threadIdx.x < w.ncell;
    const int64_t i = w.first + threadIdx.x;
    double u[3] = {0.0, 0.0, 0.0};
    if(act) {
        T v[L::Q];
#pragma unroll
        for(int k = 0; k < L::Q; k++) v[k] = f[k * stride + i];
        node<L, P>(v, D, u);
#pragma unroll
        for(int k = 0; k < L::Q; k++) f[L::opp(k) * stride + i] = v[k];
        if(vslot) write_vel<L>(D, vslot, i, u);
    }
    if(flux) block_flux<L::DIM>(u, 0, &part[wb]);
}

// ---------------------------------------------------------------------------
// neighbour step (even t): one wavefront per (row, 64-cell segment), lane = x
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lane_rank(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

template<class L, int P>
__global__ void __launch_bounds__(256)
k_nbr(const InstDesc *__restrict__ insts, const InstState *__restrict__ st,
      const NbrWork *__restrict__ work, Partial *__restrict__ part)
{
    using T = typename Store<L, P>::T;
    constexpr int Q = L::Q;
    const int wb = xcd_remap(blockIdx.x, gridDim.x);
    const NbrWork w = work[wb];
    const int phase = st[w.inst].phase;
    if(phase >= PH_DONE) return;
    const int flux = st[w.inst].flux_next;
    const int vslot = st[w.inst].vel_next;
    const InstDesc &D = insts[w.inst];
    T *f = (T *)D.f;
    const int64_t stride = D.stride;
    const int nx = D.nx, ny = D.ny, nz = D.nz, nseg = D.nseg;
    const SegInfo *seg = D.seg;

    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const bool wave_act = wave < w.nunit;
    const int unit = (int)w.first + wave;
    const int row = unit / nseg;
    const int sg = unit - row * nseg;
    const int y = (L::DIM == 3) ? row % ny : row;
    const int z = (L::DIM == 3) ? row / ny : 0;
    const int x = sg * 64 + lane;

    uint64_t m0 = 0;
    uint32_t st0 = 0;
    if(wave_act) { m0 = seg[unit].mask; st0 = seg[unit].start; }
    const bool act = wave_act && x < nx && ((m0 >> lane) & 1ull);
    const uint32_t i = st0 + lane_rank(m0);

    double u[3] = {0.0, 0.0, 0.0};
    if(act) {
        // neighbour n_k = cell - e_k for k >= 1: solidity and compact index
        bool fl[Q];
        uint32_t nb[Q];
        fl[0] = true; nb[0] = i;
#pragma unroll
        for(int k = 1; k < Q; k++) {
            const int dx = -L::ex(k), dy = -L::ey(k), dz = -L::ez(k);
            int yy = y + dy; yy = yy < 0 ? yy + ny : (yy >= ny ? yy - ny : yy);
            int zz = z;
            if constexpr(L::DIM == 3) { zz = z + dz; zz = zz < 0 ? zz + nz : (zz >= nz ? zz - nz : zz); }
            const int rb = (zz * ny + yy) * nseg;
            if(dx == 0) {
                const SegInfo s = seg[rb + sg];
                fl[k] = (s.mask >> lane) & 1ull;
                nb[k] = s.start + lane_rank(s.mask);
            } else {
                int xx = x + dx; xx = xx < 0 ? xx + nx : (xx >= nx ? xx - nx : xx);
                const int b = xx & 63;
                const SegInfo s = seg[rb + (xx >> 6)];
                fl[k] = (s.mask >> b) & 1ull;
                nb[k] = s.start + (uint32_t)__popcll(s.mask & ((1ull << b) - 1ull));
            }
        }
        T v[Q];
        v[0] = f[i];
#pragma unroll
        for(int k = 1; k < Q; k++) {
            const int64_t off = fl[k] ? (int64_t)L::opp(k) * stride + nb[k]
                                      : (int64_t)k * stride + i;
            v[k] = f[off];
        }
        node<L, P>(v, D, u);
        f[i] = v[0];
#pragma unroll
        for(int k = 1; k < Q; k++) {
            const int ko = L::opp(k);
            const int64_t off = fl[ko] ? (int64_t)k * stride + nb[ko]
                                       : (int64_t)ko * stride + i;
            f[off] = v[k];
        }
        if(vslot) write_vel<L>(D, vslot, (int64_t)i, u);
    }
    if(flux) block_flux<L::DIM>(u, 0, &part[wb]);
}

// ---------------------------------------------------------------------------
// monitor: exact flux of every instance that produced one this step, then the
// convergence state machine.  One workgroup per instance.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_monitor(const InstDesc *__restrict__ insts, const InstCfg *__restrict__ cfgs,
          InstState *__restrict__ st, const Partial *__restrict__ lpart,
          const Partial *__restrict__ npart)
{
    const int inst = blockIdx.x;
    // thread 0 rewrites st[inst] at the end: read it once, through LDS
    __shared__ int sh_phase, sh_flux;
    __shared__ int64_t sh_t;
    if(threadIdx.x == 0) {
        sh_phase = st[inst].phase;
        sh_flux = st[inst].flux_next;
        sh_t = st[inst].s + 1;
    }
    __syncthreads();
    const int phase = sh_phase;
    if(phase >= PH_DONE) return;
    const int flux = sh_flux;
    const int64_t t = sh_t;
    const InstDesc &D = insts[inst];

    __shared__ I128 acc[3][256];
    __shared__ int accbad[256];
    double q[3] = {0.0, 0.0, 0.0};
    if(flux) {
        const bool local = (t & 1) != 0;
        const Partial *P = local ? lpart : npart;
        const int p0 = local ? D.lp0 : D.np0, p1 = local ? D.lp1 : D.np1;
        I128 s[3] = {i128_zero(), i128_zero(), i128_zero()};
        int bad = 0;
        for(int p = p0 + (int)threadIdx.x; p < p1; p += blockDim.x) {
            const Partial pp = P[p];
            for(int k = 0; k < 3; k++)
                s[k] = i128_add(s[k], i128_from_chunks(pp.c[k][0], pp.c[k][1], pp.c[k][2]));
            bad |= pp.bad;
        }
        for(int k = 0; k < 3; k++) acc[k][threadIdx.x] = s[k];
        accbad[threadIdx.x] = bad;
        __syncthreads();
        for(int h = blockDim.x >> 1; h > 0; h >>= 1) {
            if((int)threadIdx.x < h) {
                for(int k = 0; k < 3; k++)
                    acc[k][threadIdx.x] = i128_add(acc[k][threadIdx.x], acc[k][threadIdx.x + h]);
                accbad[threadIdx.x] |= accbad[threadIdx.x + h];
            }
            __syncthreads();
        }
        if(threadIdx.x == 0) {
            const InstCfg &c = cfgs[inst];
            for(int k = 0; k < 3; k++) q[k] = i128_to_double_scaled(acc[k][0]) / c.denom[k];
            if(accbad[0]) q[0] = q[1] = q[2] = __builtin_nan("");
        }
    }
    if(threadIdx.x == 0) {
        const InstCfg &c = cfgs[inst];
        InstState S = st[inst];
        advance_state(c, S, q, flux);
        plan_next(c, S);
        st[inst] = S;
    }
}

// ---------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------
template<class L, int P>
static void launch_rest(const InstDesc &D, hipStream_t s)
{
    const int64_t n = D.stride;
    hipLaunchKernelGGL((k_rest<L, P>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, D);
}

template<class L, int P>
static void launch_step(bool local, const InstDesc *insts, const InstState *st,
                        const void *work, int nwork, Partial *part, hipStream_t s)
{
    if(nwork <= 0) return;
    if(local)
        hipLaunchKernelGGL((k_local<L, P>), dim3(nwork), dim3(256), 0, s, insts, st,
                           (const LocalWork *)work, part);
    else
        hipLaunchKernelGGL((k_nbr<L, P>), dim3(nwork), dim3(256), 0, s, insts, st,
                           (const NbrWork *)work, part);
}

void kern_rest(int lattice, int prec, const InstDesc &D, hipStream_t s)
{
    if(lattice == 3) {
        if(prec == PREC_F64) launch_rest<D3Q19, PREC_F64>(D, s);
        else                 launch_rest<D3Q19, PREC_F32>(D, s);
    } else {
        if(prec == PREC_F64)      launch_rest<D2Q9, PREC_F64>(D, s);
        else if(prec == PREC_F32) launch_rest<D2Q9, PREC_F32>(D, s);
        else                      launch_rest<D2Q9, PREC_F32_REF2D>(D, s);
    }
}

void kern_step(int lattice, int prec, bool local, const InstDesc *insts, const InstState *st,
               const void *work, int nwork, Partial *part, hipStream_t s)
{
    if(lattice == 3) {
        if(prec == PREC_F64) launch_step<D3Q19, PREC_F64>(local, insts, st, work, nwork, part, s);
        else                 launch_step<D3Q19, PREC_F32>(local, insts, st, work, nwork, part, s);
    } else {
        if(prec == PREC_F64)      launch_step<D2Q9, PREC_F64>(local, insts, st, work, nwork, part, s);
        else if(prec == PREC_F32) launch_step<D2Q9, PREC_F32>(local, insts, st, work, nwork, part, s);
        else                      launch_step<D2Q9, PREC_F32_REF2D>(local, insts, st, work, nwork, part, s);
    }
}

void kern_monitor(int ninst, const InstDesc *insts, const InstCfg *cfgs, InstState *st,
                  const Partial *lpart, const Partial *npart, hipStream_t s)
{
    if(ninst <= 0) return;
    hipLaunchKernelGGL(k_monitor, dim3(ninst), dim3(256), 0, s, insts, cfgs, st, lpart, npart);
}

} // namespace fp
