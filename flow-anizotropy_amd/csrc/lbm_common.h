// lbm_common.h -- lattice tables, exact flux arithmetic and the device-side
// descriptors shared by the D3Q19 / D2Q9 kernels and the host engine.
//
// Lattices (population order is the reference's and fixes every array index):
//   D3Q19: LMB3d/lbm3d-perm.cpp:94-110  (rest, 6 faces, 12 edges)
//   D2Q9 : LMB2d/lbm.cpp:20-23
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define FP_HD __host__ __device__ __forceinline__

namespace fp {

// ---------------------------------------------------------------------------
// lattice tables
// ---------------------------------------------------------------------------
struct D3Q19 {
    static constexpr int Q = 19;
    static constexpr int DIM = 3;
    static FP_HD int ex(int k) { constexpr int v[19] = {0, 1,-1, 0, 0, 0, 0, 1,-1, 1,-1, 1,-1, 1,-1, 0, 0, 0, 0}; return v[k]; }
    static FP_HD int ey(int k) { constexpr int v[19] = {0, 0, 0, 1,-1, 0, 0, 1,-1,-1, 1, 0, 0, 0, 0, 1,-1, 1,-1}; return v[k]; }
    static FP_HD int ez(int k) { constexpr int v[19] = {0, 0, 0, 0, 0, 1,-1, 0, 0, 0, 0, 1,-1,-1, 1, 1,-1,-1, 1}; return v[k]; }
    // opp[k]: the reference finds it by search (:105-110); the result is this table
    static FP_HD int opp(int k) { constexpr int v[19] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17}; return v[k]; }
    static FP_HD double w(int k) {
        return k == 0 ? 1.0/3.0 : (k < 7 ? 1.0/18.0 : 1.0/36.0);
    }
};

struct D2Q9 {
    static constexpr int Q = 9;
    static constexpr int DIM = 2;
    static FP_HD int ex(int k) { constexpr int v[9] = {0, 1, 0,-1, 0, 1,-1,-1, 1}; return v[k]; }
    static FP_HD int ey(int k) { constexpr int v[9] = {0, 0, 1, 0,-1, 1, 1,-1,-1}; return v[k]; }
    static FP_HD int ez(int) { return 0; }
    static FP_HD int opp(int k) { constexpr int v[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6}; return v[k]; }
    // the reference keeps w as float (lbm.cpp:23); every 2D variant uses these values
    static FP_HD float wf(int k) {
        return k == 0 ? (float)(4.0/9.0) : (k < 5 ? (float)(1.0/9.0) : (float)(1.0/36.0));
    }
};

// ---------------------------------------------------------------------------
// exact flux sums
//
// A node velocity u (|u| < 1) contributes trunc(u * 2^96) to a 128-bit
// integer sum.  Integer addition is associative, so the total does not depend
// on block shapes, launch order or batch composition; it is rounded once to
// double.  On the device u is split exactly into three integer-valued doubles
//   u*2^96 = c0*2^88 + c1*2^44 + c2,  |c0| < 2^8, |c1|,|c2| < 2^44,
// all of the sign of u, which a wavefront adds exactly in double (64 terms of
// < 2^44 stay below 2^53) before they are widened to 64-bit integers.
// ---------------------------------------------------------------------------
struct I128 {
    uint64_t lo;
    int64_t  hi;
};

FP_HD I128 i128_zero() { I128 r; r.lo = 0; r.hi = 0; return r; }

FP_HD I128 i128_from_i64(int64_t v) { I128 r; r.lo = (uint64_t)v; r.hi = v < 0 ? -1 : 0; return r; }

FP_HD I128 i128_add(I128 a, I128 b)
{
    I128 r;
    r.lo = a.lo + b.lo;
    r.hi = a.hi + b.hi + (r.lo < a.lo ? 1 : 0);
    return r;
}

// v * 2^sh for 0 <= sh < 64
FP_HD I128 i128_shl(int64_t v, int sh)
{
    I128 r = i128_from_i64(v);
    if(sh == 0) return r;
    I128 o;
    o.lo = r.lo << sh;
    o.hi = (int64_t)(((uint64_t)r.hi << sh) | (r.lo >> (64 - sh)));
    return o;
}

// the three chunk sums of a block (or of a whole lattice) as one integer
FP_HD I128 i128_from_chunks(int64_t c0, int64_t c1, int64_t c2)
{
    // c0 * 2^88 = (c0 * 2^24) * 2^64
    I128 a; a.lo = 0; a.hi = c0 * (int64_t)(1 << 24);
    return i128_add(i128_add(a, i128_shl(c1, 44)), i128_from_i64(c2));
}

// round-to-nearest-even conversion of a 128-bit integer to double, then *2^-96
FP_HD double i128_to_double_scaled(I128 s)
{
    const bool neg = s.hi < 0;
    uint64_t hi = (uint64_t)s.hi, lo = s.lo;
    if(neg) {                                    // magnitude
        lo = ~lo + 1;
        hi = ~hi + (lo == 0 ? 1 : 0);
    }
    if(hi == 0 && lo == 0) return 0.0;
    int lz = hi ? __builtin_clzll(hi) : 64 + __builtin_clzll(lo);
    int top = 127 - lz;                          // index of the leading one
    // take 53 significant bits m = floor(|s| / 2^(top-52)), rest for rounding
    uint64_t m;
    bool round_up = false;
    int e = top - 52;                            // |s| ~ m * 2^e
    if(e <= 0) {
        m = lo << (-e);                          // exact (top <= 52, hi == 0)
    } else {
        // shift right by e (1..75)
        uint64_t mm, rem_hi_bit, sticky;
        if(e < 64) {
            mm = (lo >> e) | (e ? (hi << (64 - e)) : 0);
            rem_hi_bit = (lo >> (e - 1)) & 1u;
            uint64_t below = (e - 1) ? (lo & ((1ull << (e - 1)) - 1)) : 0;
            sticky = below != 0;
        } else {
            int e2 = e - 64;
            mm = hi >> e2;
            if(e2 == 0) {
                rem_hi_bit = lo >> 63;
                sticky = (lo & ((1ull << 63) - 1)) != 0;
            } else {
                rem_hi_bit = (hi >> (e2 - 1)) & 1u;
                uint64_t below_hi = (e2 - 1) ? (hi & ((1ull << (e2 - 1)) - 1)) : 0;
                sticky = (below_hi | lo) != 0;
            }
        }
        m = mm;
        round_up = rem_hi_bit && (sticky || (m & 1u));
    }
    if(round_up) {
        m += 1;
        if(m == (1ull << 53)) { m >>= 1; e += 1; }
    }
    double d = (double)m;                        // exact: m < 2^53
    d = ldexp(d, e - 96);
    return neg ? -d : d;
}

// split u*2^96 into three integer-valued doubles; *bad when |u| >= 1 or NaN
FP_HD void fix_split(double u, double &c0, double &c1, double &c2, int &bad)
{
    if(!(fabs(u) < 1.0)) { bad = 1; c0 = c1 = c2 = 0.0; return; }
    double a = u * 256.0;                        // 2^8
    c0 = trunc(a);
    double r = (a - c0) * 17592186044416.0;      // 2^44
    c1 = trunc(r);
    double r2 = (r - c1) * 17592186044416.0;
    c2 = trunc(r2);
}

// ---------------------------------------------------------------------------
// device descriptors
// ---------------------------------------------------------------------------
// Population storage.  Fluid cells are numbered in compact order; the cells
// are grouped in blocks of 128 and a block stores population 0 of its 128
// cells, then population 1, ..., then Q-1 (an array of structures of arrays).
// A step kernel's wavefront then streams one contiguous 19 x 512-byte region
// (FP32) instead of 19 arrays far apart: the local step runs at 5.9-6.0 TB/s
// instead of 5.6, as a copy over the two layouts does (6.3 against 5.7).
// The arrays hold a whole number of 1024-cell chunks, so blocks never split.
constexpr int POP_BLOCK_LOG2 = 7;
// the element index of population k of compact cell n (for Q * stride < 2^32)
template<int Q> FP_HD uint32_t pop_at(int k, uint32_t n)
{
    return (n >> POP_BLOCK_LOG2) * (uint32_t)(Q << POP_BLOCK_LOG2) + ((uint32_t)k << POP_BLOCK_LOG2) +
           (n & ((1u << POP_BLOCK_LOG2) - 1u));
}
// the same in 64 bits, for arrays of 2^32 elements and more (compact cell
// indices stay 32-bit: fewer than 2^32 fluid cells)
template<int Q> FP_HD uint64_t pop_at64(int k, uint64_t n)
{
    return (n >> POP_BLOCK_LOG2) * (uint64_t)(Q << POP_BLOCK_LOG2) + ((uint64_t)k << POP_BLOCK_LOG2) +
           (n & ((1u << POP_BLOCK_LOG2) - 1u));
}
// pop_at in the offset type O (uint32_t or uint64_t)
template<int Q, class O> FP_HD O pop_at_as(int k, uint32_t n)
{
    if constexpr(sizeof(O) == 8) return pop_at64<Q>(k, n);
    else return pop_at<Q>(k, n);
}

// ---------------------------------------------------------------------------
// one 64-cell run of a lattice row along x: fluid bits and the compact index
// of its first fluid cell.  Rows are ordered z-major then y, x fastest, the
// .raw order, so compact indices follow the reference's node order.
struct SegInfo {
    uint64_t mask;
    uint32_t start;
    uint32_t pad;
};

// Neighbour links of one 64-cell row segment, built once per structure.  For
// every direction k >= 1 (entry k - 1), the cell of lane l reads population
// opp(k) at the cell - e_k:
//   M[k-1]     bit l set when that cell is fluid (else the link bounces back),
//   base[k-1]  that cell's compact index is base + popcount(M & ((1 << l) - 1)),
// except for one lane at the periodic x-wrap (the first segment's lane 0 when
// e_k.x = +1, the last segment's lane of x = nx - 1 when e_k.x = -1), whose
// neighbour's index is kept per row: rowspec[row * (Q - 1) + k - 1].
template<int NK> struct LinkRec {
    uint64_t M[NK];
    uint32_t base[NK];
};

// Compact neighbour table (the neighbour step that packs 64 fluid cells per
// wavefront), one record per fluid cell.  A neighbour x - ex of row r' (the
// row at y - ey, z - ez) has compact index start(r') + its rank among the
// fluid cells of r'.  With P = the rank of x in r' (fluid cells of r' before
// x), that rank is P for ex = 0, P - 1 for x - 1 (when fluid) and
// P + fluid(x) for x + 1, and across the periodic x wrap the last rank of r'
// (x - 1 from x = 0) or 0 (x + 1 from x = nx - 1).  So a record keeps, per
// row slot (0 = the cell's own row, then the 8 rows at (ey, ez) != (0, 0) in
// 3D, 2 in 2D), three fluid bits (x - 1, x, x + 1) and the rank P, and per
// lattice row and slot a table holds start(r') and the last rank of r':
// 24 bytes per cell for D3Q19 instead of one 16-bit offset per direction.
template<class L> struct RowSlots {
    static constexpr int N = L::DIM == 3 ? 9 : 3;
    // the row slot direction k reads from
    static FP_HD int of(int k)
    {
        constexpr int py[9] = {0, 1, -1, 0, 0, 1, -1, 1, -1};
        constexpr int pz[9] = {0, 0, 0, 1, -1, 1, -1, -1, 1};
        for(int m = 0; m < 9; m++)
            if(py[m] == L::ey(k) && pz[m] == (L::DIM == 3 ? L::ez(k) : 0)) return m;
        return -1;
    }
    static FP_HD int ey(int m) { constexpr int py[9] = {0, 1, -1, 0, 0, 1, -1, 1, -1}; return py[m]; }
    static FP_HD int ez(int m) { constexpr int pz[9] = {0, 0, 0, 1, -1, 1, -1, -1, 1}; return pz[m]; }
};
constexpr uint32_t CFLAG_X0 = 1u << 30;      // the cell is at x = 0
constexpr uint32_t CFLAG_XLAST = 1u << 31;   // the cell is at x = nx - 1
template<int NR> struct alignas(8) CRec {
    uint32_t row;             // the cell's lattice row
    uint32_t flags;           // bit 3m + e: fluid at x - 1 + e in row slot m; CFLAG_X0, CFLAG_XLAST
    uint16_t rank[NR - 1];    // P of row slots 1..NR-1 (slot 0: the cell's own rank)
};
// per lattice row: start(r') of every slot, then the last rank of every slot
template<class L> FP_HD int rowtab_words() { return 2 * RowSlots<L>::N; }

// The slots a compact-form cell reads population k from (and writes opp(k)
// to): pop_at(opp k, neighbour) when the neighbour at cell - e_k is fluid,
// else (halfway bounce-back) pop_at(k, i).  P[m]: the rank of x in slot m's
// row (P[0]: the cell's own); start/last: the row table entries of its row.
template<class L, class O = uint32_t>
FP_HD void compact_offsets(uint32_t i, uint32_t flags, const uint32_t (&P)[RowSlots<L>::N],
                           const uint32_t (&start)[RowSlots<L>::N],
                           const uint32_t (&last)[RowSlots<L>::N], O (&off)[L::Q])
{
    off[0] = pop_at_as<L::Q, O>(0, i);
    const bool x0 = (flags & CFLAG_X0) != 0, xl = (flags & CFLAG_XLAST) != 0;
#pragma unroll
    for(int k = 1; k < L::Q; k++) {
        const int m = RowSlots<L>::of(k);
        const int e = 1 - L::ex(k);                      // x - ex = x - 1 + e
        const bool fluid = (flags >> (3 * m + e)) & 1u;
        uint32_t rank;
        if(L::ex(k) == 0) rank = P[m];
        else if(L::ex(k) > 0) rank = x0 ? last[m] : P[m] - 1u;
        else rank = xl ? 0u : P[m] + ((flags >> (3 * m + 1)) & 1u);
        off[k] = fluid ? pop_at_as<L::Q, O>(L::opp(k), start[m] + rank) : pop_at_as<L::Q, O>(k, i);
    }
}

// How the step kernels address populations.  AD_B32: 32-bit byte offsets from
// the instance's base (its arrays span less than 4 GiB); AD_E32: 32-bit
// element indices on a 64-bit pointer (fewer than 2^32 elements); AD_E64:
// 64-bit element indices (any size).  A batch uses the narrowest mode that
// covers all its instances.
enum Addr : int { AD_B32 = 0, AD_E32 = 1, AD_E64 = 2 };

enum Prec : int {
    PREC_F64 = 0,       // f in double, the reference's arithmetic
    PREC_F32 = 1,       // f - w in float (shifted), FP64 flux sums
    PREC_F32_REF2D = 2  // D2Q9 only: unshifted float, the 2D reference's arithmetic
};

struct InstDesc {
    const SegInfo *seg;     // [nrows * nseg]
    const void *links;      // LinkRec<Q-1>[nrows * nseg]
    const void *ctab;       // CRec[n_fluid] (compact neighbour step), or null
    const uint32_t *rowbase;   // [nrows * rowtab_words] (compact neighbour step)
    const uint32_t *rowspec;   // [nrows * (Q-1)]
    void    *f;             // populations, f[pop_at<Q>(k, i)]
    float   *vel[2];        // compact velocity fields (DIM floats per cell), or null
    int64_t  stride;        // padded cell count (a multiple of 1024): Q * stride elements
    int64_t  n_fluid;
    int      nx, ny, nz, nseg;
    double   F[3];          // body force
    float    Ff[3];         // the same as float (2D reference stores fx as float)
    int      lattice;       // 3 = D3Q19, 2 = D2Q9
    int      prec;
    // FP32 D3Q19 per-direction force terms 4.5 w_k (e_k.F) and 1.5 w_k (e_k.F)
    float    g45[19], h15[19];
    // partial-sum slots of this instance in the local / neighbour kernels,
    // and its second-level slots for each
    int32_t  lp0, lp1, np0, np1;
    int32_t  rl0, rl1, rn0, rn1;
};

// convergence state of one solve, advanced on the device after every step.
// Phase 1 / phase 2 follow LMB3d/lbm3d-perm.cpp:320-398 (kind 3) or
// LMB2d/main.cpp:99-235 (kind 2).
struct InstCfg {
    double  eps, qfloor, ksigma;
    int64_t lag, min_steps, max_steps, avg_steps, avg_sample;
    int     hold;
    int     kind;           // 3 or 2
    int     want_vel;       // write the velocity field on the final step
    int     pad;
    double  denom[3];       // q_c = sum_c / denom[c]  (NN in 3D; L*L0*L0 in 2D)
};

enum Phase : int { PH_RUN = 0, PH_AVG = 1, PH_DONE = 2, PH_FAILED = 3, PH_IDLE = 4 };

struct InstState {
    int64_t s;              // steps done
    int64_t tavg;           // steps done in the averaging phase
    int     phase;
    int     hold;
    int     have_prev;
    int     flux_next;      // the next step must produce the flux sums
    int     vel_next;       // the next step writes the velocity field: 0 no, 1 slot A, 2 slot B
    int     converged;
    int64_t fail_step;
    int64_t n;              // samples in the averaging phase
    double  qprev[3];
    double  sum[3], sumsq[3];
    // 2D block statistics (main.cpp:119-125)
    double  bs[2], bss[2], pm[2], pse[2];
    int64_t bn;
    double  q_last[3];      // flux of the last flux step
    // the last first-phase check, as the reference's progress line prints it
    // (LMB3d/lbm3d-perm.cpp:353-357, LMB2d/main.cpp:167-172): its step, the
    // flux (3D) or block means (2D), drift/tol (3D) or dx/tol, dy/tol (2D),
    // hold after it; chk_n counts the checks
    int64_t chk_s, chk_n;
    double  chk_q[3], chk_r[2];
    int     chk_hold;
    int     chk_pad;
    // results
    double  q[3], se[3];
    int64_t steps;
    int     status;         // 0 ok, 2 non-finite flux
    int     pad;
};

// ---------------------------------------------------------------------------
// the convergence state machine, shared by the device monitor and the host
// ---------------------------------------------------------------------------
FP_HD double dmax_ref(double a, double b) { return (a < b) ? b : a; }   // std::max

// which outputs the next step must produce
FP_HD void plan_next(const InstCfg &c, InstState &S)
{
    S.flux_next = 0;
    S.vel_next = 0;
    if(S.phase == PH_RUN) {
        const int64_t t = S.s + 1;
        const bool maybe_last = (t % c.lag == 0) || (t >= c.max_steps);
        if(c.kind == 3) {
            S.flux_next = (t % c.lag == 0) || (c.avg_steps <= 0 && t >= c.max_steps);
            if(c.want_vel && c.avg_steps <= 0 && maybe_last) S.vel_next = 2;
        } else {
            S.flux_next = (t % c.avg_sample == 0) || (c.avg_steps <= 0 && maybe_last);
            if(maybe_last) S.vel_next = 1;               // field snapshot candidate
        }
    } else if(S.phase == PH_AVG) {
        const int64_t ta = S.tavg + 1;
        const bool last = ta >= c.avg_steps;
        if(c.kind == 3) {
            S.flux_next = 1;
            if(c.want_vel && last) S.vel_next = 2;
        } else {
            S.flux_next = (ta % c.avg_sample == 0) || (last && c.avg_steps < c.avg_sample);
            if(last) S.vel_next = 2;
        }
    }
}

// solve(), end of the run: LMB3d/lbm3d-perm.cpp:380-392 / LMB2d/main.cpp:210-228
FP_HD void finish_state(const InstCfg &c, InstState &S)
{
    const int nd = c.kind == 3 ? 3 : 2;
    if(S.n == 0) {
        for(int k = 0; k < nd; k++) S.q[k] = S.q_last[k];
    } else {
        for(int k = 0; k < nd; k++) {
            S.q[k] = S.sum[k] / S.n;
            if(S.n > 1) {
                const double var = (S.sumsq[k] - S.n*S.q[k]*S.q[k]) / (S.n - 1);
                S.se[k] = sqrt(dmax_ref(var, 0.0) / S.n);
            }
        }
    }
    S.steps = S.s;
    S.phase = PH_DONE;
}

FP_HD void enter_avg(const InstCfg &c, InstState &S)
{
    S.phase = PH_AVG;
    S.tavg = 0;
    S.n = 0;
    for(int k = 0; k < 3; k++) { S.sum[k] = 0.0; S.sumsq[k] = 0.0; }
    if(c.avg_steps <= 0) finish_state(c, S);
}

FP_HD void fail_state(InstState &S)
{
    S.status = 2;
    S.phase = PH_FAILED;
    S.fail_step = S.s;
    S.steps = S.s;
}

FP_HD bool all_finite(const double *q, int nd)
{
    for(int k = 0; k < nd; k++) if(!__builtin_isfinite(q[k])) return false;
    return true;
}

// advance by the step just taken; q holds its flux when have_q
FP_HD void advance_state(const InstCfg &c, InstState &S, const double *q, int have_q)
{
    S.s += 1;
    const int64_t s = S.s;
    const int nd = c.kind == 3 ? 3 : 2;
    if(have_q) for(int k = 0; k < nd; k++) S.q_last[k] = q[k];

    if(c.kind == 3) {                                  // lbm3d-perm.cpp:332-390
        if(S.phase == PH_RUN) {
            bool end = false;
            if(s % c.lag == 0) {
                if(!all_finite(q, 3)) { fail_state(S); return; }
                const double norm = sqrt(q[0]*q[0] + q[1]*q[1] + q[2]*q[2]);
                double drift = 0.0;
                for(int k = 0; k < 3; k++) drift = dmax_ref(drift, fabs(q[k] - S.qprev[k]));
                if(S.have_prev && s >= c.min_steps && drift < c.eps * norm) S.hold++;
                else                                                       S.hold = 0;
                S.chk_s = s;
                S.chk_n += 1;
                for(int k = 0; k < 3; k++) S.chk_q[k] = q[k];
                S.chk_r[0] = norm > 0 ? drift / (c.eps * norm) : 0.0;
                S.chk_r[1] = 0.0;
                S.chk_hold = S.hold;
                for(int k = 0; k < 3; k++) S.qprev[k] = q[k];
                S.have_prev = 1;
                if(S.hold >= c.hold) end = true;
            }
            if(!end && s >= c.max_steps) end = true;
            if(end) {
                S.converged = (S.hold >= c.hold) ? 1 : 0;
                enter_avg(c, S);
            }
        } else if(S.phase == PH_AVG) {
            S.tavg += 1;
            for(int k = 0; k < 3; k++) { S.sum[k] += q[k]; S.sumsq[k] += q[k]*q[k]; }
            S.n += 1;
            if(S.tavg >= c.avg_steps) finish_state(c, S);
        }
    } else {                                           // main.cpp:127-226
        if(S.phase == PH_RUN) {
            if(s % c.avg_sample == 0) {
                if(!all_finite(q, 2)) { fail_state(S); return; }
                S.bs[0] += q[0]; S.bss[0] += q[0]*q[0];
                S.bs[1] += q[1]; S.bss[1] += q[1]*q[1];
                S.bn += 1;
            }
            bool end = false;
            if(!(s % c.lag || S.bn < 2)) {
                const int64_t bn = S.bn;
                const double mx = S.bs[0] / bn, my = S.bs[1] / bn;
                const double vx = (S.bss[0] - bn*mx*mx) / (bn - 1);
                const double vy = (S.bss[1] - bn*my*my) / (bn - 1);
                const double sex = sqrt(dmax_ref(vx, 0.0) / bn);
                const double sey = sqrt(dmax_ref(vy, 0.0) / bn);
                double dx = 0.0, dy = 0.0, tolx = 0.0, toly = 0.0;
                if(S.have_prev) {
                    const double qmag = sqrt(mx*mx + my*my);
                    const double flo  = c.qfloor * qmag;
                    dx = fabs(mx - S.pm[0]);
                    dy = fabs(my - S.pm[1]);
                    tolx = dmax_ref(c.eps * dmax_ref(fabs(mx), flo),
                                    c.ksigma * sqrt(sex*sex + S.pse[0]*S.pse[0]));
                    toly = dmax_ref(c.eps * dmax_ref(fabs(my), flo),
                                    c.ksigma * sqrt(sey*sey + S.pse[1]*S.pse[1]));
                    if(s >= c.min_steps && dx < tolx && dy < toly) S.hold++;
                    else                                           S.hold = 0;
                }
                S.chk_s = s;
                S.chk_n += 1;
                S.chk_q[0] = mx; S.chk_q[1] = my; S.chk_q[2] = 0.0;
                S.chk_r[0] = tolx > 0 ? dx / tolx : 0.0;
                S.chk_r[1] = toly > 0 ? dy / toly : 0.0;
                S.chk_hold = S.hold;
                S.pm[0] = mx; S.pm[1] = my; S.pse[0] = sex; S.pse[1] = sey; S.have_prev = 1;
                S.bs[0] = S.bs[1] = S.bss[0] = S.bss[1] = 0.0; S.bn = 0;
                if(S.hold >= c.hold) end = true;
            }
            if(!end && s >= c.max_steps) end = true;
            if(end) {
                S.converged = (S.hold >= c.hold) ? 1 : 0;
                enter_avg(c, S);
            }
        } else if(S.phase == PH_AVG) {
            S.tavg += 1;
            if(S.tavg % c.avg_sample == 0) {
                if(!all_finite(q, 2)) { fail_state(S); return; }
                S.sum[0] += q[0]; S.sumsq[0] += q[0]*q[0];
                S.sum[1] += q[1]; S.sumsq[1] += q[1]*q[1];
                S.n += 1;
            }
            if(S.tavg >= c.avg_steps) finish_state(c, S);
        }
    }
}

// Whether step t of a solve in its first phase (S.s becomes t) is quiet: the
// step sums no flux and writes no velocity (plan_next), and the update after it
// only counts it -- no block sample, no check, no end of the phase
// (advance_state).  A run of quiet steps needs no finish of its own: the next
// finish adds them (advance_state_n).
FP_HD bool quiet_run_step(const InstCfg &c, int64_t t)
{
    if(t >= c.max_steps || t % c.lag == 0) return false;
    if(c.kind == 2 && t % c.avg_sample == 0) return false;
    return true;
}

// Whether a check can raise hold at all (3D: drift < eps |q|; 2D: d < tol with
// tol from eps and ksigma): with a zero tolerance a first phase ends only at
// max_steps, or when the hold it needs is already reached.
FP_HD bool hold_can_grow(const InstCfg &c)
{
    return c.kind == 3 ? c.eps > 0.0 : (c.eps > 0.0 || c.ksigma > 0.0);
}

// Whether step ta of a 2D averaging phase (S.tavg becomes ta) is quiet: no
// flux sum, no velocity write, no sample and not the last step.  A 3D
// averaging phase sums the flux on every step and has no quiet step.
FP_HD bool quiet_avg_step(const InstCfg &c, int64_t ta)
{
    return c.kind == 2 && ta % c.avg_sample != 0 && ta < c.avg_steps;
}

// advance by adv steps: adv - 1 quiet steps (quiet_run_step, quiet_avg_step)
// and then the step just taken, whose flux q is given when have_q
FP_HD void advance_state_n(const InstCfg &c, InstState &S, const double *q, int have_q, int adv)
{
    if(adv > 1) {
        if(S.phase == PH_RUN) {
            S.s += adv - 1;
        } else if(S.phase == PH_AVG && c.kind == 2) {
            S.s += adv - 1;
            S.tavg += adv - 1;
        }
    }
    advance_state(c, S, q, have_q);
}

// state of a fresh solve
FP_HD void init_state(const InstCfg &c, InstState &S)
{
    InstState z = {};
    S = z;
    S.phase = PH_RUN;
    if(c.max_steps <= 0) {                  // phase 1 runs no step at all
        S.converged = (0 >= c.hold) ? 1 : 0;
        enter_avg(c, S);
    }
    plan_next(c, S);
}

// per-block exact partial sums: [component][chunk]
struct Partial {
    int64_t c[3][3];
    int32_t bad;
    int32_t pad;
};

// second level: up to REDUCE_SPAN step-kernel partials of one instance
// folded into 128-bit integers by one workgroup
constexpr int REDUCE_SPAN = 1024;
// the deferred finish (struct Defer) takes instances of up to this many
// partials per step kernel (every workgroup of such an instance sums them on
// the steps after a flux step)
constexpr int DEFER_MAX_PARTIALS = 4096;
// chunk graphs a batch keeps, one per finish plan (plan_finishes)
constexpr size_t MAX_PLAN_GRAPHS = 8;
// where the deferred finish applies, a chunk takes a finish plan only if it
// has at most one finish per this many steps
constexpr int PLAN_OVER_DEFER = 5;
struct RPartial {
    I128    s[3];
    int32_t bad;
    int32_t pad[3];
};
struct RedWork {
    int32_t inst;
    int32_t p0, p1;           // step-kernel partials [p0, p1)
    int32_t pad;
};

// The deferred finish.  When every instance of a batch has at most
// DEFER_MAX_PARTIALS partials per step kernel, a step kernel applies the previous
// step's finish itself instead of a finish launch after every step: each
// workgroup reduces its instance's partials of the previous step (on flux
// steps) and runs the convergence update, all workgroups of an instance
// computing the same state from the same inputs; the workgroup of the
// instance's first work item writes that state for the next step.  The
// states alternate between two arrays (a step reads one and writes the
// other), the partials between the local and the neighbour arrays; a chunk
// ends with one finish launch.  cfg == nullptr: every step is followed by a
// finish launch and the step kernel reads the current state.
struct Defer {
    const InstCfg *cfg;       // the instances' configs, or null (no deferral)
    InstState *sout;          // the state after the previous step, for the next step
    const Partial *prev;      // the previous step's partials (the other kernel kind's array)
    int32_t pending;          // 0: the state read is already current (first step of a chunk)
    int32_t slot0;            // partial slot of this launch's item 0
};

// launch shapes of the step kernels (chosen by the autotuner)
// The defaults are shapes the autotuner picks on MI355X for D3Q19 at 256^3 in
// both precisions (local 2 cells/thread, non-temporal, one workgroup per item;
// neighbour 4 stacked row segments, one workgroup per item): about 20% faster
// per step than 1 cell/thread with cached accesses and persistent grids.  With
// populations in blocks, one workgroup per item beats 16 resident workgroups per
// CU looping over items: local step 0.253 against 0.267 ms (FP32), 0.533
// against 0.572 (FP64).
struct Tuning {
    int local_cpt = 1;      // cells per thread in the local kernel: 1, 2 or 4
    int nbr_waves = 4;      // wavefronts per workgroup in the neighbour kernel: 1..8, one
                            // row segment each
    int nbr_segs = 4;       // of those, x-adjacent segments of one row (1, 2, 4 or 8, at
                            // most nbr_waves, and a divisor of the row's segments); the
                            // rest of the workgroup's wavefronts stack rows along z
                            // (along y in 2D)
    int nontemporal = 1;    // local kernel: population loads/stores bypass cache allocation
    int local_bpc = 0;      // resident workgroups per CU launched (each loops over items);
                            // 0: one workgroup per work item
    int nbr_bpc = 0;        // 0: one workgroup per work item
    int nbr_mode = 2;       // neighbour step: 0 = one wavefront per row segment (link
                            // records), 1 = 64 fluid cells per wavefront (compact table),
                            // 2 = per instance by porosity (COMPACT_BELOW_POROSITY)
    int local_wgs = 3;      // resident workgroups per CU the local kernel is held to (LDS
                            // it does not use, sized so that no more fit) in D3Q19 chunks
                            // that run a finish plan; 0: as many as its registers allow
    int nbr_wgs = 5;        // the same for the neighbour kernels
};

// Below this porosity the compact neighbour step is the faster form: it keeps
// every lane busy, while the row-segment step idles the lanes of solid cells.
// Compact time over row-segment time on MI355X (D3Q19, 192^3 ellipsoids, rows
// of 3 segments, one per workgroup, FP32 / FP64): 0.68 / 0.71 at porosity 0.30,
// 0.89 / 0.94 at 0.52, 0.94 / 0.97 at 0.62, 1.02 / 1.04 at 0.72, 1.05 / 1.06
// at 0.92.
constexpr double COMPACT_BELOW_POROSITY = 0.68;
// D3Q19 rows whose row-segment workgroups take 4 or more of their segments
// share the lines at the segment boundaries in one L1, and the crossover moves
// down (256^3 ellipsoids, compact over row-segment time, FP32 / FP64: 0.88 /
// 0.84 at porosity 0.40, 1.00 / 0.94 at 0.50, 1.05 / 0.98 at 0.58, 1.09 / 1.02
// at 0.66, 1.09 / 1.03 at 0.75)
constexpr double COMPACT_BELOW_POROSITY_WIDE_F32 = 0.50;
constexpr double COMPACT_BELOW_POROSITY_WIDE_F64 = 0.62;

// whether a structure of this porosity takes the compact form when the form
// is chosen per structure (nbr_mode 2); segs: the row segments per
// row-segment workgroup it would get
inline bool prefer_compact(double porosity, int segs, int lattice, int prec)
{
    double below = COMPACT_BELOW_POROSITY;
    if(lattice == 3 && segs >= 4)
        below = prec == PREC_F64 ? COMPACT_BELOW_POROSITY_WIDE_F64 : COMPACT_BELOW_POROSITY_WIDE_F32;
    return porosity < below;
}

// From this many bytes of populations a 3D instance deals its neighbour-step
// items to the 8 XCDs in y bands through all z instead of z ranges (the slabs
// either side of a boundary then share an L2).  Below it the z ranges stay, so
// that an XCD's neighbour step and local step cover the same cells.
constexpr double BAND_MIN_BYTES = 4e8;

// work of one workgroup of the local (in-place collide) step: a run of compact cells
struct LocalWork {
    int32_t inst;
    int32_t ncell;
    int64_t first;
};

// One neighbour-step work item: S = 2^sshift x-adjacent 64-cell segments
// (seg .. seg + S - 1) of each of nunit rows stacked along the slowest axis.
// Wave w takes segment seg + (w mod S) of the row at plane p0 + w / S.  3D:
// rows (y, z = p0 + w / S); 2D: rows y = p0 + w / S.  Segments past the row's
// end leave their wave idle.  seg, nunit and sshift share one 32-bit word
// (bits 0-15, 16-23, 24-31): every field is a whole dword, so the kernel reads
// the item with one scalar load (16-bit fields came as vector loads, which
// cost the FP32 neighbour step at 256^3 0.420 against 0.345 ms).
struct NbrWork {
    int32_t  inst;
    uint32_t shape;
    int32_t  y;
    int32_t  p0;
    FP_HD int seg() const { return (int)(shape & 0xffffu); }
    FP_HD int nunit() const { return (int)((shape >> 16) & 0xffu); }
    FP_HD int sshift() const { return (int)(shape >> 24); }
    FP_HD void set_shape(int seg_, int nunit_, int sshift_)
    {
        shape = (uint32_t)seg_ | ((uint32_t)nunit_ << 16) | ((uint32_t)sshift_ << 24);
    }
};

} // namespace fp
