// kernels.hip -- A-A in-place stream-collide kernels on fluid-only compact
// storage, exact flux partials, and the device-side convergence monitor.
//
// Storage.  Fluid cells are numbered in the .raw order (z, y, x fastest)
// skipping solid ones; population k of cell i lives at f[pop_at<Q>(k, i)]
// (blocks of 128 cells, each holding its Q populations one after another).
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

#include <algorithm>

namespace fp {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

// Pointers read out of InstDesc are generic to the compiler, which then emits
// flat memory instructions.  These casts state the address space: global for
// lattice data, constant for the read-only segment table (uniform reads of it
// become scalar loads).
template<class T> using gptr = __attribute__((address_space(1))) T *;
template<class T> using cptr = const __attribute__((address_space(4))) T *;
template<class T> __device__ __forceinline__ gptr<T> as_global(void *p) { return (gptr<T>)p; }

// population traffic is streamed (each value is read and written once a
// step); NT selects loads/stores that do not allocate in the caches
template<bool NT, class V> __device__ __forceinline__ V ld(gptr<V> p)
{
    if constexpr(NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template<bool NT, class V> __device__ __forceinline__ void stv(gptr<V> p, V v)
{
    if constexpr(NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}
template<class T> __device__ __forceinline__ cptr<T> as_const(const void *p) { return (cptr<T>)p; }

// population element `off` of an instance, addressed in mode AD (Addr).
// AD_B32: the instance's arrays span less than 4 GiB, so the byte offset is a
// 32-bit lane value added to the wave-uniform base (the `global_load v, v_off,
// s[base]` form: one VGPR per address instead of two).  The neighbour kernels
// hold 19 addresses at once; at 60 instead of 68 VGPRs eight wavefronts fit per
// SIMD instead of seven and the FP32 row-segment step takes 0.400 instead of
// 0.440 ms at 256^3.  AD_E32: a 32-bit element index; AD_E64: a 64-bit one (an
// instance of 2^32 population elements or more, 226 M fluid cells in D3Q19).
template<int AD> struct OffOf { using T = uint32_t; };
template<> struct OffOf<AD_E64> { using T = uint64_t; };
template<int AD, class T> __device__ __forceinline__ gptr<T> pop_ptr(gptr<T> f, typename OffOf<AD>::T off)
{
    if constexpr(AD == AD_B32) return (gptr<T>)((gptr<char>)f + off * (uint32_t)sizeof(T));
    else return f + off;
}

// Workgroups are dealt round-robin to the 8 XCDs; give each XCD one
// contiguous range of work so neighbouring segments share its L2.
__device__ __forceinline__ int xcd_remap(int b, int G)
{
    const int q = G >> 3, r = G & 7;
    const int x = b & 7, j = b >> 3;
    return x < r ? x * (q + 1) + j : r * (q + 1) + (x - r) * q + j;
}

// v from the lane that the DPP control CTRL names (all 64 lanes active)
template<int CTRL> __device__ __forceinline__ double dpp_f64(double v)
{
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

// sum over the wavefront, in every lane.  Inside each row of 16 lanes the
// partner comes through DPP (quad_perm [1,0,3,2], [2,3,0,1], row_half_mirror,
// row_mirror), across rows through ds_bpermute: 4 LDS-crossbar operations
// per double instead of 12.  The callers add integer-valued doubles whose
// sums stay below 2^53, so the order of the additions does not matter.
__device__ __forceinline__ double wave_sum(double v)
{
    v += dpp_f64<0xB1>(v);
    v += dpp_f64<0x4E>(v);
    v += dpp_f64<0x141>(v);
    v += dpp_f64<0x140>(v);
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    return v;
}

// Exact flux partials.  Each thread adds its cells' chunks (fix_split) into
// three integer-valued doubles per component; with at most 4 cells a thread
// and 64 lanes a wavefront, every sum stays below 2^53 and is exact.  The
// wavefront sums are widened to 64-bit integers and added across the block.
template<int NC>
__device__ __forceinline__ void fix_add(const double *u, double (&c)[3][3], int &bad)
{
#pragma unroll
    for(int k = 0; k < NC; k++) {
        double a, b, d;
        fix_split(u[k], a, b, d, bad);
        c[k][0] += a; c[k][1] += b; c[k][2] += d;
    }
}

template<int NC>
__device__ __forceinline__ void block_flux(double (&c)[3][3], int bad, Partial *out)
{
    __shared__ double sh[16][9];
    __shared__ int shbad[16];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
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
    // lanes 0-8 of the first wavefront add one (component, part) each across
    // the wavefronts, lane 9 the flags (a single lane doing all of it kept
    // the workgroup's other wavefronts waiting: 0.49 against 0.44 ms for the
    // FP32 neighbour step at 256^3 with the flux on)
    const int nw = (blockDim.x + 63) >> 6;
    const int q = threadIdx.x;
    if(q < 9) {
        int64_t acc = 0;
        if(q / 3 < NC)
            for(int w = 0; w < nw; w++) acc += (int64_t)sh[w][q];
        out->c[q / 3][q % 3] = acc;
    } else if(q == 9) {
        int b = 0;
        for(int w = 0; w < nw; w++) b |= shbad[w];
        out->bad = b;
        out->pad = 0;
    }
}

// ---------------------------------------------------------------------------
// deferred finish (struct Defer): the state a step runs with
// ---------------------------------------------------------------------------
// sum of 128-bit values over a workgroup of up to 1024 threads (butterflies
// inside each wavefront, then the wavefronts' sums in LDS); exact, result in
// thread 0
__device__ __forceinline__ void wg_sum_i128(I128 (&s)[3], int &bad)
{
    for(int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for(int k = 0; k < 3; k++) {
            I128 o;
            o.lo = __shfl_down(s[k].lo, off);
            o.hi = __shfl_down(s[k].hi, off);
            s[k] = i128_add(s[k], o);
        }
        bad |= __shfl_down(bad, off);
    }
    __shared__ I128 ws[16][3];
    __shared__ int wb[16];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if(lane == 0) {
#pragma unroll
        for(int k = 0; k < 3; k++) ws[wv][k] = s[k];
        wb[wv] = bad;
    }
    __syncthreads();
    if(threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for(int w = 1; w < nw; w++) {
#pragma unroll
            for(int k = 0; k < 3; k++) s[k] = i128_add(s[k], ws[w][k]);
            bad |= wb[w];
        }
    }
}

// copy of one InstState by the whole workgroup, a 32-bit word per thread
__device__ __forceinline__ void wg_copy_state(InstState *dst, const InstState *src)
{
    constexpr int NW = (int)(sizeof(InstState) / 4);
    static_assert(sizeof(InstState) % 4 == 0, "InstState is copied in 32-bit words");
    for(int w = (int)threadIdx.x; w < NW; w += (int)blockDim.x)
        reinterpret_cast<uint32_t *>(dst)[w] = reinterpret_cast<const uint32_t *>(src)[w];
}

// The state this step runs with for instance `inst`, into shS (LDS): st[inst]
// itself when it is current (df.pending == 0), else with the previous step's
// finish applied -- the reduction and update k_finish would have done, in the
// same integer arithmetic (the sum is exact, so the order of the partials
// does not matter) and the same floating-point operations.  PREV_LOCAL: the
// previous step was a local step (partials [lp0, lp1)), else a neighbour step
// ([np0, np1)).  Every thread of the workgroup calls it.
template<bool PREV_LOCAL>
__device__ void deferred_state(const InstDesc &D, const InstState *__restrict__ st, int inst,
                               const Defer &df, InstState &shS)
{
    __syncthreads();                               // every thread has read the previous shS
    wg_copy_state(&shS, &st[inst]);
    __syncthreads();
    if(!df.pending || shS.phase >= PH_DONE) return;             // the same for every thread
    const int flux = shS.flux_next;                // whether the previous step summed the flux
    double q[3] = {0.0, 0.0, 0.0};
    if(flux) {
        I128 s[3] = {i128_zero(), i128_zero(), i128_zero()};
        int bad = 0;
        const int p0 = PREV_LOCAL ? D.lp0 : D.np0, p1 = PREV_LOCAL ? D.lp1 : D.np1;
        for(int p = p0 + (int)threadIdx.x; p < p1; p += (int)blockDim.x) {
            const Partial pp = df.prev[p];
#pragma unroll
            for(int k = 0; k < 3; k++)
                s[k] = i128_add(s[k], i128_from_chunks(pp.c[k][0], pp.c[k][1], pp.c[k][2]));
            bad |= pp.bad;
        }
        wg_sum_i128(s, bad);
        if(threadIdx.x == 0) {
            const InstCfg &c = df.cfg[inst];
            for(int k = 0; k < 3; k++) q[k] = i128_to_double_scaled(s[k]) / c.denom[k];
            if(bad) q[0] = q[1] = q[2] = __builtin_nan("");
        }
    }
    if(threadIdx.x == 0) {
        const InstCfg &c = df.cfg[inst];
        advance_state(c, shS, q, flux);
        plan_next(c, shS);
    }
    __syncthreads();
}

// phase, flux and velocity flags of the step for the item's instance, and
// the deferred state handed on by the instance's first item
struct StepFlags { int phase, flux, vslot; };

// DF: the kernel variant with the deferred finish; the convergence update
// it carries needs about twice the registers of the step itself, so the
// kernels without it (every lattice too large for it) are separate variants
template<bool DF, bool PREV_LOCAL>
__device__ __forceinline__ StepFlags step_flags(const InstDesc &D, const InstState *__restrict__ st, int inst,
                                               int it, int first_slot, const Defer &df, InstState &shS,
                                               int &cur)
{
    static_assert(DF, "only the deferred-finish kernels derive their flags here");
    StepFlags r;
    if(inst != cur) {
        deferred_state<PREV_LOCAL>(D, st, inst, df, shS);
        cur = inst;
    }
    r.phase = (int)__builtin_amdgcn_readfirstlane((uint32_t)shS.phase);
    r.flux = (int)__builtin_amdgcn_readfirstlane((uint32_t)shS.flux_next);
    r.vslot = (int)__builtin_amdgcn_readfirstlane((uint32_t)shS.vel_next);
    if(df.slot0 + it == first_slot) wg_copy_state(&df.sout[inst], &shS);
    return r;
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
    if(slot != 1 && slot != 2) return;             // the two snapshot slots only
    if(!D.vel[slot - 1]) return;
    gptr<float> v = as_global<float>(D.vel[slot - 1]);
#pragma unroll
    for(int c = 0; c < L::DIM; c++) v[i * L::DIM + c] = (float)u[c];
}

// ---------------------------------------------------------------------------
// rest state: f = w (unshifted) or h = 0 (shifted), and zero velocity fields
// (resetdistributions sets U = V = 0, LMB2d/lbm.cpp:200-211: a 2D solve
// whose first phase runs no step snapshots that rest field)
// ---------------------------------------------------------------------------
template<class L, int P>
__global__ void __launch_bounds__(256)
k_rest(InstDesc D)
{
    using T = typename Store<L, P>::T;
    gptr<T> f = as_global<T>(D.f);
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= D.stride) return;
#pragma unroll
    for(int k = 0; k < L::Q; k++) {
        T w;
        if constexpr(P == PREC_F32) w = 0;
        else if constexpr(L::DIM == 3) w = (T)L::w(k);
        else if constexpr(P == PREC_F64) w = (T)(k == 0 ? 4.0/9.0 : (k < 5 ? 1.0/9.0 : 1.0/36.0));
        else w = (T)L::wf(k);
        f[pop_at64<L::Q>(k, (uint64_t)i)] = w;
    }
    if(i < D.n_fluid) {
#pragma unroll
        for(int s = 0; s < 2; s++) {
            if(!D.vel[s]) continue;
            gptr<float> v = as_global<float>(D.vel[s]);
#pragma unroll
            for(int c = 0; c < L::DIM; c++) v[i * L::DIM + c] = 0.0f;
        }
    }
}

// Work items of a launch are dealt to its workgroups so that XCD x (which the
// hardware gives every 8th workgroup) owns items [x n/8, (x+1) n/8) and its
// workgroups walk that range together: those running at the same time on an
// XCD work on neighbouring items and share its L2.
struct ItemRange { int begin, end, step; };
__device__ __forceinline__ ItemRange my_items(int nitems)
{
    const int G = gridDim.x, b = blockIdx.x;
    if(G < 8) return {b, nitems, G};
    const int x = b & 7, j = b >> 3;
    const int gx = (G - x + 7) >> 3;
    const int lo = (int)((int64_t)nitems * x / 8), hi = (int)((int64_t)nitems * (x + 1) / 8);
    return {lo + j, hi, gx};
}

// ---------------------------------------------------------------------------
// local step (odd t): each thread updates CPT consecutive compact cells with
// vector loads/stores; a workgroup covers 256*CPT cells of one instance.
// The populations are padded to a multiple of 1024 cells and stored in 128-cell
// blocks, so a vector never leaves its block; padding cells keep their rest
// values.
// ---------------------------------------------------------------------------
template<class T, int N> struct VecT { typedef T type __attribute__((ext_vector_type(N))); };
template<class T> struct VecT<T, 1> { typedef T type; };
template<class T, int N> __device__ __forceinline__ T &lane_of(typename VecT<T, N>::type &v, int c)
{
    if constexpr(N == 1) return v;
    else return reinterpret_cast<T *>(&v)[c];
}

template<class L, int P, int CPT, bool NT, bool W, bool DF>
__global__ void __launch_bounds__(256)
k_local(const InstDesc *__restrict__ insts, const InstState *__restrict__ st,
        const LocalWork *__restrict__ work, int nwork, Partial *__restrict__ part, Defer df)
{
    __shared__ InstState shS;
    int cur = -1;
    // W: 64-bit element indices (arrays of 2^32 elements or more)
    auto at = [](int k, int64_t i) -> uint64_t {
        if constexpr(W) return pop_at64<L::Q>(k, (uint64_t)i);
        else return pop_at<L::Q>(k, (uint32_t)i);
    };
    using T = typename Store<L, P>::T;
    using V = typename VecT<T, CPT>::type;
    const ItemRange R = my_items(nwork);
    for(int it = R.begin; it < R.end; it += R.step) {
        const LocalWork w = work[it];
        int flux, vslot;
        if constexpr(DF) {
            const InstDesc &D0 = insts[w.inst];
            const StepFlags sf = step_flags<DF, false>(D0, st, w.inst, it, D0.lp0, df, shS, cur);
            if(sf.phase >= PH_DONE) continue;
            flux = sf.flux;
            vslot = sf.vslot;
        } else {
            if(st[w.inst].phase >= PH_DONE) continue;
            flux = st[w.inst].flux_next;
            vslot = st[w.inst].vel_next;
        }
        const InstDesc &D = insts[w.inst];
        gptr<T> f = as_global<T>(D.f);
        const int c0 = (int)threadIdx.x * CPT;
        const int nreal = w.ncell - c0;             // real cells of this thread (may be <= 0)
        const int64_t i = w.first + c0;
        double uc[CPT][3];
        if(nreal > 0) {
            V in[L::Q];
#pragma unroll
            for(int k = 0; k < L::Q; k++) in[k] = ld<NT>((gptr<V>)(f + at(k, i)));
            V out[L::Q];
#pragma unroll
            for(int c = 0; c < CPT; c++) {
                T v[L::Q];
#pragma unroll
                for(int k = 0; k < L::Q; k++) v[k] = lane_of<T, CPT>(in[k], c);
                node<L, P>(v, D, uc[c]);
                const bool real = c < nreal;
#pragma unroll
                for(int k = 0; k < L::Q; k++)
                    lane_of<T, CPT>(out[L::opp(k)], c) = real ? v[k] : lane_of<T, CPT>(in[L::opp(k)], c);
            }
#pragma unroll
            for(int k = 0; k < L::Q; k++) stv<NT>((gptr<V>)(f + at(k, i)), out[k]);
        }
        if(vslot && nreal > 0) {
#pragma unroll
            for(int c = 0; c < CPT; c++)
                if(c < nreal) write_vel<L>(D, vslot, i + c, uc[c]);
        }
        if(flux) {
            double ch[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
            int bad = 0;
#pragma unroll
            for(int c = 0; c < CPT; c++)
                if(c < nreal) fix_add<L::DIM>(uc[c], ch, bad);
            block_flux<L::DIM>(ch, bad, &part[it]);
            if(it + R.step < R.end) __syncthreads();   // LDS reused by the next item
        }
    }
}

// ---------------------------------------------------------------------------
// neighbour step (even t): one wavefront per (row, 64-cell segment), lane = x
//
// Every lookup of a neighbour row goes through scalar loads of its SegInfo
// (the row and segment are the same for the whole wavefront); a lane then
// needs one mbcnt for its neighbour's compact index.  A neighbour at x-1 or
// x+1 is the same segment's mask shifted by one, with the single lane whose
// neighbour sits in the next/previous segment (or across the periodic wrap)
// patched from uniform values.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lane_rank(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint32_t lane_rank_plus(uint64_t m, uint32_t base)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, base));
}

__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t uni64(uint64_t v)
{
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// a when bit `lane` of the wave-uniform mask M is set, else b: one
// v_cndmask with M as the lane-select operand
__device__ __forceinline__ uint32_t lane_select(uint64_t M, uint32_t a, uint32_t b)
{
    uint32_t r;
    asm("v_cndmask_b32_e64 %0, %1, %2, %3" : "=v"(r) : "v"(b), "v"(a), "s"(M));
    return r;
}

template<class L, int P, int AD, bool DF>
__global__ void __launch_bounds__(512)
k_nbr(const InstDesc *__restrict__ insts, const InstState *__restrict__ st,
      const NbrWork *__restrict__ work, int nwork, Partial *__restrict__ part, Defer df)
{
    __shared__ InstState shS;
    int cur = -1;
    using T = typename Store<L, P>::T;
    constexpr int Q = L::Q;
    constexpr int NK = Q - 1;
    const int wave = uni(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const ItemRange R = my_items(nwork);
    for(int it = R.begin; it < R.end; it += R.step) {
        const NbrWork w = work[it];
        int flux, vslot;
        if constexpr(DF) {
            const InstDesc &D0 = insts[w.inst];
            const StepFlags sf = step_flags<DF, true>(D0, st, w.inst, it, D0.np0, df, shS, cur);
            if(sf.phase >= PH_DONE) continue;
            flux = sf.flux;
            vslot = sf.vslot;
        } else {
            // this order keeps the work item's fields in scalar loads (with the
            // flags read first, the compiler turned two of them into vector
            // loads: 0.449 against 0.379 ms at 256^3)
            if(st[w.inst].phase >= PH_DONE) continue;
            flux = st[w.inst].flux_next;
            vslot = st[w.inst].vel_next;
        }
        const InstDesc &D = insts[w.inst];
        gptr<T> f = as_global<T>(D.f);
        const int ny = D.ny, nseg = D.nseg;
        cptr<SegInfo> seg = as_const<SegInfo>(D.seg);

        // wave -> (segment seg + dsg, row p0 + dz) of the item's S x nunit units
        const int sh = w.sshift(), seg0 = w.seg();
        const int dsg = wave & ((1 << sh) - 1), dz = wave >> sh;
        const bool wave_act = dz < w.nunit() && seg0 + dsg < nseg;
        const int p = uni(w.p0 + (wave_act ? dz : 0));
        const int sg = uni(seg0 + (wave_act ? dsg : 0));
        const int row = (L::DIM == 3) ? p * ny + w.y : p;
        const int unit = row * nseg + sg;

        const uint64_t m0 = uni64(seg[unit].mask);
        const uint32_t i = lane_rank_plus(m0, uni(seg[unit].start));
        const bool act = wave_act && ((m0 >> lane) & 1ull);

        double u[3] = {0.0, 0.0, 0.0};
        if(act) {
            cptr<LinkRec<NK>> lr = as_const<LinkRec<NK>>(D.links) + unit;
            cptr<uint32_t> rs = as_const<uint32_t>(D.rowspec) + (int64_t)row * NK;
            // the lanes whose x-neighbour wraps around the row
            const bool wrapL = sg == 0 && lane == 0;
            const bool wrapR = sg == nseg - 1 && lane == ((D.nx - 1) & 63);
            // off[k]: the location population k is read from this step and the
            // location population opp(k) is written to: f[opp k][cell - e_k]
            // when that neighbour is fluid, else (halfway bounce-back) f[k][cell]
            // the segment's whole link record and the row's wrap indices, in
            // scalar registers before any use (the loads overlap)
            uint64_t Mk[NK];
            uint32_t Bk[NK], Sk[NK];
#pragma unroll
            for(int k = 0; k < NK; k++) {
                Mk[k] = lr->M[k];
                Bk[k] = lr->base[k];
                Sk[k] = L::ex(k + 1) != 0 ? rs[k] : 0u;
            }
            using O = typename OffOf<AD>::T;
            O off[Q];
            off[0] = pop_at_as<Q, O>(0, i);
#pragma unroll
            for(int k = 1; k < Q; k++) {
                uint32_t nbi = lane_rank_plus(Mk[k - 1], Bk[k - 1]);
                if(L::ex(k) > 0) nbi = wrapL ? Sk[k - 1] : nbi;
                if(L::ex(k) < 0) nbi = wrapR ? Sk[k - 1] : nbi;
                if constexpr(AD == AD_E64)
                    off[k] = ((Mk[k - 1] >> lane) & 1ull) ? pop_at64<Q>(L::opp(k), nbi) : pop_at64<Q>(k, i);
                else
                    off[k] = lane_select(Mk[k - 1], pop_at<Q>(L::opp(k), nbi), pop_at<Q>(k, i));
            }
            T v[Q];
#pragma unroll
            for(int k = 0; k < Q; k++) v[k] = *pop_ptr<AD>(f, off[k]);
            node<L, P>(v, D, u);
#pragma unroll
            for(int k = 0; k < Q; k++) *pop_ptr<AD>(f, off[L::opp(k)]) = v[k];
            if(vslot) write_vel<L>(D, vslot, (int64_t)i, u);
        }
        if(flux) {
            double ch[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
            int bad = 0;
            if(act) fix_add<L::DIM>(u, ch, bad);
            block_flux<L::DIM>(ch, bad, &part[it]);
            if(it + R.step < R.end) __syncthreads();   // LDS reused by the next item
        }
    }
}

// ---------------------------------------------------------------------------
// neighbour step, compact form (even t): 64 consecutive fluid cells per
// wavefront, whatever rows they lie in.  Each cell reads its 24-byte CRec
// (fluid bits at x - 1, x, x + 1 and the rank of x in each neighbour row);
// the row table (start and last rank of every neighbour row) comes from
// scalar loads when the wavefront's cells lie in at most two rows (the usual
// case), else from vector loads.  The same slots as k_nbr are read and
// written, so the two forms give identical results.
// ---------------------------------------------------------------------------
template<class L, int P, int AD, bool DF>
__global__ void __launch_bounds__(256)
k_nbr_c(const InstDesc *__restrict__ insts, const InstState *__restrict__ st,
        const LocalWork *__restrict__ work, int nwork, Partial *__restrict__ part, Defer df)
{
    __shared__ InstState shS;
    int cur = -1;
    using T = typename Store<L, P>::T;
    constexpr int Q = L::Q;
    constexpr int NK = Q - 1;
    const ItemRange R = my_items(nwork);
    for(int it = R.begin; it < R.end; it += R.step) {
        const LocalWork w = work[it];
        int flux, vslot;
        if constexpr(DF) {
            const InstDesc &D0 = insts[w.inst];
            const StepFlags sf = step_flags<DF, true>(D0, st, w.inst, it, D0.np0, df, shS, cur);
            if(sf.phase >= PH_DONE) continue;
            flux = sf.flux;
            vslot = sf.vslot;
        } else {
            if(st[w.inst].phase >= PH_DONE) continue;
            flux = st[w.inst].flux_next;
            vslot = st[w.inst].vel_next;
        }
        const InstDesc &D = insts[w.inst];
        gptr<T> f = as_global<T>(D.f);
        const int c = (int)threadIdx.x;
        const bool act = c < w.ncell;
        const uint32_t i = (uint32_t)(w.first + c);

        double u[3] = {0.0, 0.0, 0.0};
        if(act) {
            // the cell's record, as 8-byte vector loads
            constexpr int NR = RowSlots<L>::N;
            constexpr int NW = 2 * NR;
            typedef uint32_t u2 __attribute__((ext_vector_type(2)));
            constexpr int NV = (int)(sizeof(CRec<NR>) / 8);
            union { u2 v[NV]; CRec<NR> c; } r;
            gptr<const u2> src = as_global<const u2>(const_cast<void *>(D.ctab)) + (int64_t)i * NV;
#pragma unroll
            for(int q = 0; q < NV; q++) r.v[q] = src[q];
            const uint32_t row = r.c.row;
            // the rows of this wavefront's active cells (monotonic in the cell order)
            const uint64_t lanes = __ballot(1);
            const int last = 63 - __builtin_clzll(lanes);
            const uint32_t row0 = uni(row);
            const uint32_t row1 = __builtin_amdgcn_readlane(row, last);
            const bool two = __ballot(row != row0 && row != row1) == 0ull;
            uint32_t tb[NW];
            if(two) {
                cptr<uint32_t> b0 = as_const<uint32_t>(D.rowbase) + (int64_t)row0 * NW;
                cptr<uint32_t> b1 = as_const<uint32_t>(D.rowbase) + (int64_t)row1 * NW;
                const bool in0 = row == row0;
#pragma unroll
                for(int w = 0; w < NW; w++) {
                    const uint32_t a0 = b0[w], a1 = b1[w];
                    tb[w] = in0 ? a0 : a1;
                }
            } else {
                gptr<const uint32_t> b = as_global<const uint32_t>((void *)const_cast<uint32_t *>(D.rowbase)) +
                                         (int64_t)row * NW;
#pragma unroll
                for(int w = 0; w < NW; w++) tb[w] = b[w];
            }
            uint32_t rk[NR], start[NR], lastr[NR];
#pragma unroll
            for(int m = 0; m < NR; m++) { start[m] = tb[m]; lastr[m] = tb[NR + m]; }
            rk[0] = i - start[0];
#pragma unroll
            for(int m = 1; m < NR; m++) rk[m] = r.c.rank[m - 1];
            using O = typename OffOf<AD>::T;
            O off[Q];
            compact_offsets<L, O>(i, r.c.flags, rk, start, lastr, off);
            T v[Q];
#pragma unroll
            for(int k = 0; k < Q; k++) v[k] = *pop_ptr<AD>(f, off[k]);
            node<L, P>(v, D, u);
#pragma unroll
            for(int k = 0; k < Q; k++) *pop_ptr<AD>(f, off[L::opp(k)]) = v[k];
            if(vslot) write_vel<L>(D, vslot, (int64_t)i, u);
        }
        if(flux) {
            double ch[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
            int bad = 0;
            if(act) fix_add<L::DIM>(u, ch, bad);
            block_flux<L::DIM>(ch, bad, &part[it]);
            if(it + R.step < R.end) __syncthreads();   // LDS reused by the next item
        }
    }
}

// sum of 128-bit values across a 256-thread workgroup (result in thread 0)
__device__ __forceinline__ void block_sum_i128(I128 (&s)[3], int &bad)
{
    __shared__ I128 acc[3][256];
    __shared__ int accbad[256];
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
    for(int k = 0; k < 3; k++) s[k] = acc[k][0];
    bad = accbad[0];
}

// ---------------------------------------------------------------------------
// After every step: the exact flux of every instance that produced one, then
// its convergence state machine, in one launch.  Each workgroup takes one
// reduction item: up to REDUCE_SPAN step-kernel partials of one instance,
// folded into 128-bit sums (on flux steps).  Every workgroup then counts
// itself in on a per-instance counter; the last one of an instance (so, after
// every workgroup of that instance has read the state) adds the instance's
// 128-bit sums, rounds q once, advances the state and resets the counter.
// An instance with a single reduction item skips the count.
// ---------------------------------------------------------------------------
__device__ void monitor_update(const InstCfg &c, InstState *st, const double q[3], int flux, int adv)
{
    InstState S = *st;
    advance_state_n(c, S, q, flux, adv);
    plan_next(c, S);
    *st = S;
}

__global__ void __launch_bounds__(256)
k_finish(const RedWork *__restrict__ work, const InstDesc *__restrict__ insts,
         const InstCfg *__restrict__ cfgs, InstState *__restrict__ st,
         const Partial *__restrict__ part, RPartial *__restrict__ red,
         unsigned int *__restrict__ counter, int local, int adv)
{
    const RedWork w = work[blockIdx.x];
    __shared__ int sh_phase, sh_flux, sh_last;
    if(threadIdx.x == 0) {
        sh_phase = st[w.inst].phase;
        sh_flux = st[w.inst].flux_next;
    }
    __syncthreads();
    if(sh_phase >= PH_DONE) return;                // the same for every item of the instance
    const int flux = sh_flux;
    const InstDesc &D = insts[w.inst];
    const int r0 = local ? D.rl0 : D.rn0, r1 = local ? D.rl1 : D.rn1;
    I128 s[3] = {i128_zero(), i128_zero(), i128_zero()};
    int bad = 0;
    if(flux) {
        for(int p = w.p0 + (int)threadIdx.x; p < w.p1; p += blockDim.x) {
            const Partial pp = part[p];
            for(int k = 0; k < 3; k++)
                s[k] = i128_add(s[k], i128_from_chunks(pp.c[k][0], pp.c[k][1], pp.c[k][2]));
            bad |= pp.bad;
        }
        block_sum_i128(s, bad);
    }
    if(threadIdx.x == 0) {
        if(flux) {
            RPartial r;
            for(int k = 0; k < 3; k++) r.s[k] = s[k];
            r.bad = bad;
            r.pad[0] = r.pad[1] = r.pad[2] = 0;
            red[blockIdx.x] = r;
        }
        if(r1 - r0 == 1) {
            // the instance's only item: no count (on small lattices the fence
            // and atomic were 0.3-0.5 us of a 7.9-11 us step)
            sh_last = 1;
        } else {
            __threadfence();                       // the sums are visible before the count
            const unsigned int n = atomicAdd(&counter[w.inst], 1u);
            sh_last = n == (unsigned int)(r1 - r0 - 1);
        }
    }
    __syncthreads();
    if(!sh_last) return;
    __threadfence();                               // every item's sums are visible now
    double q[3] = {0.0, 0.0, 0.0};
    if(flux) {
        s[0] = s[1] = s[2] = i128_zero();
        bad = 0;
        for(int p = r0 + (int)threadIdx.x; p < r1; p += blockDim.x) {
            const RPartial pp = red[p];
            for(int k = 0; k < 3; k++) s[k] = i128_add(s[k], pp.s[k]);
            bad |= pp.bad;
        }
        block_sum_i128(s, bad);
    }
    if(threadIdx.x == 0) {
        const InstCfg &c = cfgs[w.inst];
        if(flux) {
            for(int k = 0; k < 3; k++) q[k] = i128_to_double_scaled(s[k]) / c.denom[k];
            if(bad) q[0] = q[1] = q[2] = __builtin_nan("");
        }
        counter[w.inst] = 0u;                      // ready for the next step
        monitor_update(c, &st[w.inst], q, flux, adv);
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

static int device_cus()
{
    static int cus[64] = {0};
    int dev = 0;
    (void)hipGetDevice(&dev);
    if(dev < 0 || dev >= 64) dev = 0;
    if(!cus[dev]) {
        int v = 0;
        if(hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) v = 256;
        cus[dev] = v;
    }
    return cus[dev];
}

// workgroups launched for nwork items: blocks_per_cu resident per CU (0: one
// workgroup per item), at least 8 (one per XCD), never more than the items
static int grid_for(int nwork, int blocks_per_cu)
{
    if(blocks_per_cu <= 0) return nwork;
    const int64_t g = (int64_t)device_cus() * blocks_per_cu;
    return (int)std::max<int64_t>(std::min<int64_t>(nwork, g), std::min(nwork, 8));
}

// Dynamic LDS that holds a kernel to wgs resident workgroups per CU (0: no cap).
// The kernel never touches it: its only effect is that no more workgroups fit in
// a CU's LDS.  The footprint (the kernel's static LDS plus this) is a multiple of
// 5120 bytes, so allocation granules of 1024 or 1280 bytes do not round it up,
// and it lies in (lds/(wgs+1), lds/wgs] for the CU's lds bytes.  A cap that no
// such footprint reaches (wgs 1 or 7 and up with 160 KiB, or above what one
// workgroup may allocate) returns 0 and is not applied.
static size_t lds_for_wgs(const void *kern, int wgs)
{
    if(wgs <= 0) return 0;
    static int lds_cu[64] = {0}, lds_wg[64] = {0};
    int dev = 0;
    (void)hipGetDevice(&dev);
    if(dev < 0 || dev >= 64) return 0;
    if(!lds_cu[dev]) {
        int cu = 0, wg = 0;
        if(hipDeviceGetAttribute(&cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess ||
           hipDeviceGetAttribute(&wg, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
           cu <= 0 || wg <= 0)
            return 0;
        lds_wg[dev] = wg;
        lds_cu[dev] = cu;
    }
    hipFuncAttributes a;
    if(hipFuncGetAttributes(&a, kern) != hipSuccess) return 0;
    const size_t G = 5120, cu = (size_t)lds_cu[dev];
    const size_t F = std::min(cu / (size_t)wgs, (size_t)lds_wg[dev]) / G * G;
    if(F <= cu / (size_t)(wgs + 1) || F <= a.sharedSizeBytes) return 0;
    return F - a.sharedSizeBytes;
}

template<class L, int P, int CPT, bool NT>
static void launch_local(const Tuning &tu, bool wide, const InstDesc *insts, const InstState *st,
                         const LocalWork *lw, int nwork, Partial *part, const Defer &df, hipStream_t s)
{
    const dim3 g(grid_for(nwork, tu.local_bpc)), b(256);
    // the deferred finish runs on small lattices only (32-bit indices)
    if(df.cfg) {
        const void *k = (const void *)&k_local<L, P, CPT, NT, false, true>;
        hipLaunchKernelGGL((k_local<L, P, CPT, NT, false, true>), g, b, lds_for_wgs(k, tu.local_wgs), s,
                           insts, st, lw, nwork, part, df);
    } else if(wide) {
        const void *k = (const void *)&k_local<L, P, CPT, NT, true, false>;
        hipLaunchKernelGGL((k_local<L, P, CPT, NT, true, false>), g, b, lds_for_wgs(k, tu.local_wgs), s,
                           insts, st, lw, nwork, part, df);
    } else {
        const void *k = (const void *)&k_local<L, P, CPT, NT, false, false>;
        hipLaunchKernelGGL((k_local<L, P, CPT, NT, false, false>), g, b, lds_for_wgs(k, tu.local_wgs), s,
                           insts, st, lw, nwork, part, df);
    }
}

template<class L, int P, int AD, bool DF>
static void launch_nbr(const Tuning &tu, const InstDesc *insts, const InstState *st, const void *work,
                       int nwork, Partial *part, const Defer &df, hipStream_t s)
{
    // the neighbour kernel's scattered partial-line traffic needs the L2:
    // it never uses non-temporal accesses (measured 1.6-1.8x the bytes)
    if(tu.nbr_mode == 1) {
        const dim3 g(grid_for(nwork, tu.nbr_bpc)), b(256);
        const size_t lds = lds_for_wgs((const void *)&k_nbr_c<L, P, AD, DF>, tu.nbr_wgs);
        hipLaunchKernelGGL((k_nbr_c<L, P, AD, DF>), g, b, lds, s, insts, st, (const LocalWork *)work, nwork, part, df);
    } else {
        const dim3 g(grid_for(nwork, tu.nbr_bpc)), b(64 * tu.nbr_waves);
        const size_t lds = lds_for_wgs((const void *)&k_nbr<L, P, AD, DF>, tu.nbr_wgs);
        hipLaunchKernelGGL((k_nbr<L, P, AD, DF>), g, b, lds, s, insts, st, (const NbrWork *)work, nwork, part, df);
    }
}

template<class L, int P>
static void launch_step(bool local, const Tuning &tu, int ad, const InstDesc *insts, const InstState *st,
                        const void *work, int nwork, Partial *part, const Defer &df, hipStream_t s)
{
    if(nwork <= 0) return;
    if(local) {
        const LocalWork *lw = (const LocalWork *)work;
        const bool wide = ad == AD_E64;
        if(tu.nontemporal) {
            if(tu.local_cpt == 4)      launch_local<L, P, 4, true>(tu, wide, insts, st, lw, nwork, part, df, s);
            else if(tu.local_cpt == 2) launch_local<L, P, 2, true>(tu, wide, insts, st, lw, nwork, part, df, s);
            else                       launch_local<L, P, 1, true>(tu, wide, insts, st, lw, nwork, part, df, s);
        } else {
            if(tu.local_cpt == 4)      launch_local<L, P, 4, false>(tu, wide, insts, st, lw, nwork, part, df, s);
            else if(tu.local_cpt == 2) launch_local<L, P, 2, false>(tu, wide, insts, st, lw, nwork, part, df, s);
            else                       launch_local<L, P, 1, false>(tu, wide, insts, st, lw, nwork, part, df, s);
        }
    } else {
        if(df.cfg)            launch_nbr<L, P, AD_B32, true>(tu, insts, st, work, nwork, part, df, s);
        else if(ad == AD_B32) launch_nbr<L, P, AD_B32, false>(tu, insts, st, work, nwork, part, df, s);
        else if(ad == AD_E32) launch_nbr<L, P, AD_E32, false>(tu, insts, st, work, nwork, part, df, s);
        else                  launch_nbr<L, P, AD_E64, false>(tu, insts, st, work, nwork, part, df, s);
    }
}

// plain device copies, the bandwidth reference for the step kernels.
// k_copy: a persistent grid, 4 loads in flight per thread, 16 KiB per
// workgroup and iteration.  k_copy_flat: one workgroup per 8 KiB, two float4
// per thread, both loads issued before the stores; NT: loads and stores that
// do not allocate in the caches (the form of the local step's streams).
// k_copy_one: one workgroup per 4 KiB, one non-temporal float4 per thread
// (6.6 TB/s on 2 GiB, the fastest of these).
__global__ void __launch_bounds__(256)
k_copy(const float4 *__restrict__ a, float4 *__restrict__ b, int64_t n)
{
    const int64_t step = (int64_t)gridDim.x * 1024;
    for(int64_t base = (int64_t)blockIdx.x * 1024; base < n; base += step) {
        float4 v[4];
#pragma unroll
        for(int j = 0; j < 4; j++) {
            const int64_t i = base + j * 256 + threadIdx.x;
            if(i < n) v[j] = a[i];
        }
#pragma unroll
        for(int j = 0; j < 4; j++) {
            const int64_t i = base + j * 256 + threadIdx.x;
            if(i < n) b[i] = v[j];
        }
    }
}

template<bool NT>
__global__ void __launch_bounds__(256)
k_copy_flat(const float4 *__restrict__ a, float4 *__restrict__ b, int64_t n)
{
    using V4 = VecT<float, 4>::type;
    gptr<V4> src = (gptr<V4>)(void *)a;
    gptr<V4> dst = (gptr<V4>)(void *)b;
    const int64_t i0 = (int64_t)blockIdx.x * 512 + threadIdx.x, i1 = i0 + 256;
    V4 v0, v1;
    if(i0 < n) v0 = NT ? __builtin_nontemporal_load(src + i0) : src[i0];
    if(i1 < n) v1 = NT ? __builtin_nontemporal_load(src + i1) : src[i1];
    if(i0 < n) { if constexpr(NT) __builtin_nontemporal_store(v0, dst + i0); else dst[i0] = v0; }
    if(i1 < n) { if constexpr(NT) __builtin_nontemporal_store(v1, dst + i1); else dst[i1] = v1; }
}

__global__ void __launch_bounds__(256)
k_copy_one(const float4 *__restrict__ a, float4 *__restrict__ b, int64_t n)
{
    using V4 = VecT<float, 4>::type;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if(i < n) __builtin_nontemporal_store(__builtin_nontemporal_load((gptr<V4>)(void *)a + i),
                                          (gptr<V4>)(void *)b + i);
}

void kern_copy(int kind, const void *a, void *b, int64_t n16, hipStream_t s)
{
    if(kind == 3) {
        hipLaunchKernelGGL(k_copy_one, dim3((unsigned)((n16 + 255) / 256)), dim3(256), 0, s,
                           (const float4 *)a, (float4 *)b, n16);
        return;
    }
    if(kind == 0) {
        const int64_t blocks = std::min<int64_t>((n16 + 1023) / 1024, 256 * 32);
        hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, s, (const float4 *)a, (float4 *)b, n16);
        return;
    }
    const dim3 g((unsigned)((n16 + 511) / 512)), t(256);
    if(kind == 1) hipLaunchKernelGGL((k_copy_flat<false>), g, t, 0, s, (const float4 *)a, (float4 *)b, n16);
    else          hipLaunchKernelGGL((k_copy_flat<true>), g, t, 0, s, (const float4 *)a, (float4 *)b, n16);
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

void kern_step(int lattice, int prec, bool local, const Tuning &tu, int ad, const InstDesc *insts,
               const InstState *st, const void *work, int nwork, Partial *part, hipStream_t s,
               const Defer *defer)
{
    Defer df = {};
    if(defer) df = *defer;
    if(lattice == 3) {
        if(prec == PREC_F64) launch_step<D3Q19, PREC_F64>(local, tu, ad, insts, st, work, nwork, part, df, s);
        else                 launch_step<D3Q19, PREC_F32>(local, tu, ad, insts, st, work, nwork, part, df, s);
    } else {
        if(prec == PREC_F64)      launch_step<D2Q9, PREC_F64>(local, tu, ad, insts, st, work, nwork, part, df, s);
        else if(prec == PREC_F32) launch_step<D2Q9, PREC_F32>(local, tu, ad, insts, st, work, nwork, part, df, s);
        else                      launch_step<D2Q9, PREC_F32_REF2D>(local, tu, ad, insts, st, work, nwork, part, df, s);
    }
}

void kern_finish(int nwork, const RedWork *work, const InstDesc *insts, const InstCfg *cfgs,
                 InstState *st, const Partial *part, RPartial *red, unsigned int *counter,
                 bool local, hipStream_t s, int adv)
{
    if(nwork <= 0) return;
    hipLaunchKernelGGL(k_finish, dim3(nwork), dim3(256), 0, s, work, insts, cfgs, st, part, red,
                       counter, local ? 1 : 0, adv);
}

} // namespace fp
