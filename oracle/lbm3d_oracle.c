/*
 * oracle/lbm3d_oracle.c -- CPU restatement of the D3Q19 permeability solver of
 * LMB3d/lbm3d-perm.cpp, used ONLY as the checker for the HIP solver.
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only to check or time against it.
 * Nothing in the product path calls it.
 *
 * What it follows, function by function (reference = LMB3d/lbm3d-perm.cpp):
 *   lattice tables, opp search ......... :94-110
 *   constants TAU, CS2, MU, FX_BASE .... :112-115
 *   Config defaults .................... :118-130
 *   read_structure checks .............. :174-193   (or3_check_mask)
 *   pull neighbour tables .............. :158-172
 *   reset_distributions ................ :201-212
 *   collide_node (BGK + Guo) ........... :214-245   (same operation order)
 *   collide_only / stream_collide ...... :247-295   (AoS, two buffers, pull)
 *   volumeavg .......................... :297-303   (see "flux sum" below)
 *   solve() phase 1 + phase 2 .......... :319-398
 *   jacobi3 ............................ :400-443
 *   K assembly, reciprocity, FA ........ :522-553
 *   --validate tensor test ............. :555-572
 *   export_velocity payload ............ :305-317
 *
 * Built strictly: -O2 -ffp-contract=off, no -ffast-math, so every double
 * operation is the IEEE operation written here, in the order written.  The
 * reference binary is built with -ffast-math (LMB3d/make_perm.sh:15), which
 * lets the compiler reassociate and contract; the two therefore agree to
 * round-off (checked against reference-produced fixtures), not bitwise.
 *
 * Flux sum.  volumeavg() (:298-303) adds U over all NN nodes sequentially
 * (vectorised and reassociated under -ffast-math).  Here the sum is exact: each
 * velocity is truncated toward zero to a multiple of 2^-96, the integers are
 * added in 128-bit arithmetic (order-independent), and the total is rounded once
 * to double.  This differs from the reference's float sum by at most a few ulp
 * of q, and it makes the result independent of summation order, which the GPU
 * solver reproduces bit-for-bit.  |u| >= 1 (beyond any stable lattice: c_s = 0.577)
 * is reported as a non-finite flux, like NaN/Inf.
 *
 * Parity pins (tests/): the exact sum against Python's Fraction arithmetic;
 * the slit closed form K = h(h^2+1)/(12 nz); blocked direction = 0; reciprocity,
 * tensor test and axis-permutation covariance; and reference-binary fixtures in
 * tests/golden/ (K to 1e-10 with a trace floor, identical step counts).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define Q3 19

static const int EX[Q3] = {0, 1,-1, 0, 0, 0, 0, 1,-1, 1,-1, 1,-1, 1,-1, 0, 0, 0, 0};
static const int EY[Q3] = {0, 0, 0, 1,-1, 0, 0, 1,-1,-1, 1, 0, 0, 0, 0, 1,-1, 1,-1};
static const int EZ[Q3] = {0, 0, 0, 0, 0, 1,-1, 0, 0, 0, 0, 1,-1,-1, 1, 1,-1,-1, 1};
static const double WQ[Q3] = {
    1.0/3.0,
    1.0/18.0, 1.0/18.0, 1.0/18.0, 1.0/18.0, 1.0/18.0, 1.0/18.0,
    1.0/36.0, 1.0/36.0, 1.0/36.0, 1.0/36.0, 1.0/36.0, 1.0/36.0,
    1.0/36.0, 1.0/36.0, 1.0/36.0, 1.0/36.0, 1.0/36.0, 1.0/36.0};

static const double TAU = 1.0;
static const double CS2 = 1.0/3.0;
#define MU3 (CS2 * (TAU - 0.5))
static const double FX_BASE = 2.5e-07;

/* opp[k] by search, as lbm3d-perm.cpp:105-110 does it */
static void build_opp(int opp[Q3])
{
    for(int k = 0; k < Q3; k++)
        for(int m = 0; m < Q3; m++)
            if(EX[m] == -EX[k] && EY[m] == -EY[k] && EZ[m] == -EZ[k]) { opp[k] = m; break; }
}

/* ------------------------------------------------------------------------- */
/* exact flux sum                                                             */
/* ------------------------------------------------------------------------- */
typedef __int128 i128;

/* trunc(u * 2^96) as a 128-bit integer; *bad set when |u| >= 1 or not finite */
static inline i128 fix96(double u, int *bad)
{
    if(!(fabs(u) < 1.0)) { *bad = 1; return 0; }
    return (i128)(u * 0x1p96);          /* scaling by 2^96 is exact; cast truncates */
}

/* exact sum -> double: one correctly rounded conversion, then exact 2^-96 */
static inline double unfix96(i128 s)
{
    return ldexp((double)s, -96);
}

/* exported for the tests that pin the exact sum against Fraction arithmetic */
int or_fix96_sum(const double *u, long n, int64_t *hi, uint64_t *lo, double *as_double)
{
    int bad = 0;
    i128 s = 0;
    for(long i = 0; i < n; i++) s += fix96(u[i], &bad);
    *hi = (int64_t)(s >> 64);
    *lo = (uint64_t)s;
    *as_double = unfix96(s);
    return bad;
}

/* ------------------------------------------------------------------------- */
/* configuration and results                                                 */
/* ------------------------------------------------------------------------- */
typedef struct {
    double ff, eps;
    long   lag;
    int    hold;
    long   min_steps, max_steps, avg_steps;
} or3_cfg;

typedef struct {
    double q[3], se[3];
    long   steps;
    int    converged;
    int    status;        /* 0 ok, 2 non-finite flux during phase 1 */
    long   fail_step;
} or3_solve;

void or3_default_cfg(or3_cfg *c)              /* lbm3d-perm.cpp:118-130 */
{
    c->ff = 1.0; c->eps = 1e-5; c->lag = 2000; c->hold = 2;
    c->min_steps = 4000; c->max_steps = 200000; c->avg_steps = 2000;
}

double or3_mu(void)      { return MU3; }
double or3_fx_base(void) { return FX_BASE; }

/* read_structure checks (:182-192): 0 ok, 1 bad dims, 2 non-binary, 3 no pore */
int or3_check_mask(const uint8_t *solid, int nx, int ny, int nz, long *n_pore)
{
    if(nx <= 0 || ny <= 0 || nz <= 0) return 1;
    long nn = (long)nx * ny * nz, np = 0;
    for(long i = 0; i < nn; i++) { if(solid[i] > 1) return 2; if(!solid[i]) np++; }
    *n_pore = np;
    return np == 0 ? 3 : 0;
}

/* ------------------------------------------------------------------------- */
/* lattice state for one solve                                               */
/* ------------------------------------------------------------------------- */
typedef struct {
    int nx, ny, nz;
    long nn;
    const uint8_t *solid;
    double *f0, *f1;             /* AoS f[i*Q+k], two buffers, as :150 */
    double *U, *V, *W, *RHO;
    int *pmx, *pmy, *pmz;
    int opp[Q3];
    double Fx, Fy, Fz;
    int primed;
} lat3;

static inline long idx3(const lat3 *L, int x, int y, int z)
{
    return ((long)z * L->ny + y) * L->nx + x;
}

/* collide_node, lbm3d-perm.cpp:218-245 -- same expressions, same order */
static inline void collide_node(lat3 *L, long i, const double *g, double *out)
{
    double r = 0, ux = 0, uy = 0, uz = 0;
    for(int k = 0; k < Q3; k++)
    {
        r += g[k];
        ux += g[k]*EX[k]; uy += g[k]*EY[k]; uz += g[k]*EZ[k];
    }
    const double Fx = L->Fx, Fy = L->Fy, Fz = L->Fz;
    L->RHO[i] = r;
    ux = L->U[i] = (ux + 0.5*Fx)/r;
    uy = L->V[i] = (uy + 0.5*Fy)/r;
    uz = L->W[i] = (uz + 0.5*Fz)/r;

    const double u2 = ux*ux + uy*uy + uz*uz;
    const double pref = 1.0 - 0.5/TAU;
    for(int k = 0; k < Q3; k++)
    {
        const double eu = EX[k]*ux + EY[k]*uy + EZ[k]*uz;
        const double feq = WQ[k]*r*(1.0 + 3.0*eu + 4.5*eu*eu - 1.5*u2);
        const double sk = WQ[k] * pref *
            ( 3.0*((EX[k]-ux)*Fx + (EY[k]-uy)*Fy + (EZ[k]-uz)*Fz)
            + 9.0*eu*(EX[k]*Fx + EY[k]*Fy + EZ[k]*Fz) );
        out[k] = g[k] + ((feq - g[k]) / TAU + sk);
    }
}

static void collide_only(lat3 *L)                 /* :249-253 */
{
    for(long i = 0; i < L->nn; i++)
        if(!L->solid[i]) collide_node(L, i, &L->f0[i*Q3], &L->f0[i*Q3]);
}

static void stream_collide(lat3 *L)               /* :270-289 */
{
    double g[Q3];
    const int nx = L->nx, ny = L->ny, nz = L->nz;
    for(int z = 0; z < nz; z++)
    for(int y = 0; y < ny; y++)
    for(int x = 0; x < nx; x++)
    {
        const long i = idx3(L, x, y, z);
        if(L->solid[i]) continue;
        for(int k = 0; k < Q3; k++)
        {
            const long p = idx3(L, L->pmx[(size_t)k*nx + x],
                                   L->pmy[(size_t)k*ny + y],
                                   L->pmz[(size_t)k*nz + z]);
            g[k] = L->solid[p] ? L->f0[i*Q3 + L->opp[k]] : L->f0[p*Q3 + k];
        }
        collide_node(L, i, g, &L->f1[i*Q3]);
    }
    double *t = L->f0; L->f0 = L->f1; L->f1 = t;
}

static void lbm_step(lat3 *L)                     /* :291-295 */
{
    if(!L->primed) { collide_only(L); L->primed = 1; }
    else           stream_collide(L);
}

/* volumeavg (:298-303) with the exact sum; returns nonzero on |u|>=2^8 / NaN */
static int volumeavg(const lat3 *L, double q[3])
{
    int bad = 0;
    i128 su = 0, sv = 0, sw = 0;
    for(long i = 0; i < L->nn; i++)
    {
        su += fix96(L->U[i], &bad);
        sv += fix96(L->V[i], &bad);
        sw += fix96(L->W[i], &bad);
    }
    q[0] = unfix96(su)/(double)L->nn;
    q[1] = unfix96(sv)/(double)L->nn;
    q[2] = unfix96(sw)/(double)L->nn;
    if(bad) { q[0] = q[1] = q[2] = NAN; }
    return bad;
}

static int lat3_init(lat3 *L, const uint8_t *solid, int nx, int ny, int nz)
{
    memset(L, 0, sizeof(*L));
    L->nx = nx; L->ny = ny; L->nz = nz;
    L->nn = (long)nx * ny * nz;
    L->solid = solid;
    build_opp(L->opp);
    L->f0 = (double*)malloc(sizeof(double) * L->nn * Q3);
    L->f1 = (double*)malloc(sizeof(double) * L->nn * Q3);
    L->U = (double*)calloc(L->nn, sizeof(double));
    L->V = (double*)calloc(L->nn, sizeof(double));
    L->W = (double*)calloc(L->nn, sizeof(double));
    L->RHO = (double*)malloc(sizeof(double) * L->nn);
    L->pmx = (int*)malloc(sizeof(int) * Q3 * nx);
    L->pmy = (int*)malloc(sizeof(int) * Q3 * ny);
    L->pmz = (int*)malloc(sizeof(int) * Q3 * nz);
    if(!L->f0 || !L->f1 || !L->U || !L->V || !L->W || !L->RHO || !L->pmx || !L->pmy || !L->pmz)
        return -1;
    for(int k = 0; k < Q3; k++)                   /* :163-172 */
    {
        for(int x = 0; x < nx; x++) L->pmx[(size_t)k*nx + x] = ((x - EX[k]) % nx + nx) % nx;
        for(int y = 0; y < ny; y++) L->pmy[(size_t)k*ny + y] = ((y - EY[k]) % ny + ny) % ny;
        for(int z = 0; z < nz; z++) L->pmz[(size_t)k*nz + z] = ((z - EZ[k]) % nz + nz) % nz;
    }
    return 0;
}

static void lat3_free(lat3 *L)
{
    free(L->f0); free(L->f1); free(L->U); free(L->V); free(L->W); free(L->RHO);
    free(L->pmx); free(L->pmy); free(L->pmz);
}

static void reset_distributions(lat3 *L)          /* :203-212 */
{
    for(long i = 0; i < L->nn; i++)
        for(int k = 0; k < Q3; k++) L->f0[i*Q3+k] = L->f1[i*Q3+k] = WQ[k];
    for(long i = 0; i < L->nn; i++) { L->U[i] = 0.0; L->V[i] = 0.0; L->W[i] = 0.0; L->RHO[i] = 1.0; }
    L->primed = 0;
}

static inline double dmax(double a, double b) { return (a < b) ? b : a; }   /* std::max */

/* The flux solve() reads after step s: volumeavg of the lattice, or, with no
 * lattice (L == NULL), entry s of a planted sequence (seq[(s-1)*3 + c]; NaN past
 * its end), which tests use to drive the convergence rules alone. */
static void flux3(const lat3 *L, const double *seq, long nseq, long s, double q[3])
{
    if(L) { volumeavg(L, q); return; }
    for(int c = 0; c < 3; c++) q[c] = (s >= 1 && s <= nseq) ? seq[(s - 1) * 3 + c] : NAN;
}

/* solve(), lbm3d-perm.cpp:320-398, for one forcing vector (fx, fy, fz).
 * vel, when not NULL, receives NN*3 float32 (U,V,W per node, x fastest):
 * the payload export_velocity (:305-317) writes after its 16-byte header.
 * With L == NULL nothing is stepped and the flux comes from seq (flux3). */
static void solve_dir(lat3 *L, const or3_cfg *cfg, double fx, double fy, double fz,
                      or3_solve *out, float *vel, const double *seq, long nseq)
{
    memset(out, 0, sizeof(*out));
    if(L) {
        reset_distributions(L);
        L->Fx = fx; L->Fy = fy; L->Fz = fz;
    }

    long s = 0;
    int hold = 0;
    int have_prev = 0;
    double qprev[3] = {0, 0, 0};

    while(s < cfg->max_steps)
    {
        if(L) lbm_step(L);
        s++;
        if(s % cfg->lag) continue;

        double q[3];
        flux3(L, seq, nseq, s, q);
        for(int c = 0; c < 3; c++)
            if(!isfinite(q[c])) { out->status = 2; out->fail_step = s; out->steps = s; return; }

        const double norm = sqrt(q[0]*q[0] + q[1]*q[1] + q[2]*q[2]);
        double drift = 0.0;
        for(int c = 0; c < 3; c++) drift = dmax(drift, fabs(q[c] - qprev[c]));

        if(have_prev && s >= cfg->min_steps && drift < cfg->eps * norm) hold++;
        else                                                            hold = 0;

        for(int c = 0; c < 3; c++) qprev[c] = q[c];
        have_prev = 1;

        if(hold >= cfg->hold) break;
    }

    out->converged = (hold >= cfg->hold) ? 1 : 0;

    double sum[3] = {0,0,0}, sumsq[3] = {0,0,0};
    long n = 0;
    for(long i = 0; i < cfg->avg_steps; i++)
    {
        if(L) lbm_step(L);
        s++;
        double q[3]; flux3(L, seq, nseq, s, q);
        for(int c = 0; c < 3; c++) { sum[c] += q[c]; sumsq[c] += q[c]*q[c]; }
        n++;
    }

    if(n == 0) flux3(L, seq, nseq, s, out->q);
    else
        for(int c = 0; c < 3; c++)
        {
            out->q[c] = sum[c] / n;
            if(n > 1)
            {
                const double var = (sumsq[c] - n*out->q[c]*out->q[c]) / (n - 1);
                out->se[c] = sqrt(dmax(var, 0.0) / n);
            }
        }
    out->steps = s;

    if(vel && L)
        for(long i = 0; i < L->nn; i++)
        {
            vel[3*i+0] = (float)L->U[i];
            vel[3*i+1] = (float)L->V[i];
            vel[3*i+2] = (float)L->W[i];
        }
}

/* solve()'s convergence rules alone, over a planted flux sequence seq of
 * nseq x 3 values (the flux after step s is seq[(s-1)*3 ..]); returns 0 ok,
 * 2 non-finite flux */
int or3_replay(const or3_cfg *cfg, const double *seq, long nseq, or3_solve *out)
{
    solve_dir(NULL, cfg, 0, 0, 0, out, NULL, seq, nseq);
    return out->status;
}

/* one forcing direction on a mask; returns 0 ok, 2 non-finite flux, -1 OOM */
int or3_solve_dir(const uint8_t *solid, int nx, int ny, int nz, const or3_cfg *cfg,
                  double fx, double fy, double fz, or3_solve *out, float *vel)
{
    lat3 L;
    if(lat3_init(&L, solid, nx, ny, nz)) { lat3_free(&L); return -1; }
    solve_dir(&L, cfg, fx, fy, fz, out, vel, NULL, 0);
    lat3_free(&L);
    return out->status;
}

/* Run exactly `steps` lattice steps from rest under force (fx,fy,fz) and
 * return the populations in the reference's AoS order f[i*Q+k] (the buffer the
 * reference calls f0, holding post-collision values) and U,V,W.  Used by the
 * tests that compare lattice states step by step. */
int or3_run_steps(const uint8_t *solid, int nx, int ny, int nz, double fx, double fy, double fz,
                  long steps, double *f_out, double *uvw_out)
{
    lat3 L;
    if(lat3_init(&L, solid, nx, ny, nz)) { lat3_free(&L); return -1; }
    reset_distributions(&L);
    L.Fx = fx; L.Fy = fy; L.Fz = fz;
    for(long s = 0; s < steps; s++) lbm_step(&L);
    if(f_out) memcpy(f_out, L.f0, sizeof(double) * L.nn * Q3);
    if(uvw_out)
        for(long i = 0; i < L.nn; i++)
        { uvw_out[3*i] = L.U[i]; uvw_out[3*i+1] = L.V[i]; uvw_out[3*i+2] = L.W[i]; }
    lat3_free(&L);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* jacobi3, lbm3d-perm.cpp:401-443                                            */
/* ------------------------------------------------------------------------- */
static void jacobi3(const double Ain[3][3], double val[3], double vec[3][3])
{
    double a[3][3];
    memcpy(a, Ain, sizeof(a));
    for(int i = 0; i < 3; i++) for(int j = 0; j < 3; j++) vec[i][j] = (i == j) ? 1.0 : 0.0;

    for(int sweep = 0; sweep < 50; sweep++)
    {
        double off = fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[1][2]);
        if(off < 1e-18) break;
        for(int p = 0; p < 2; p++)
        for(int q = p+1; q < 3; q++)
        {
            if(fabs(a[p][q]) < 1e-300) continue;
            const double theta = (a[q][q] - a[p][p]) / (2.0*a[p][q]);
            const double t = (theta >= 0 ? 1.0 : -1.0) /
                             (fabs(theta) + sqrt(theta*theta + 1.0));
            const double c = 1.0/sqrt(t*t + 1.0), s = t*c;
            for(int k = 0; k < 3; k++)
            {
                const double akp = a[k][p], akq = a[k][q];
                a[k][p] = c*akp - s*akq;  a[k][q] = s*akp + c*akq;
            }
            for(int k = 0; k < 3; k++)
            {
                const double apk = a[p][k], aqk = a[q][k];
                a[p][k] = c*apk - s*aqk;  a[q][k] = s*apk + c*aqk;
                const double vkp = vec[k][p], vkq = vec[k][q];
                vec[k][p] = c*vkp - s*vkq; vec[k][q] = s*vkp + c*vkq;
            }
        }
    }
    for(int i = 0; i < 3; i++) val[i] = a[i][i];

    for(int i = 0; i < 2; i++)
    for(int j = i+1; j < 3; j++)
        if(val[j] > val[i])
        {
            double t = val[i]; val[i] = val[j]; val[j] = t;
            for(int k = 0; k < 3; k++) { t = vec[k][i]; vec[k][i] = vec[k][j]; vec[k][j] = t; }
        }
}

void or3_jacobi3(const double *A9, double *val3, double *vec9)
{
    double A[3][3], vec[3][3];
    memcpy(A, A9, sizeof(A));
    jacobi3(A, val3, vec);
    memcpy(vec9, vec, sizeof(vec));
}

/* ------------------------------------------------------------------------- */
/* the whole invocation: three solves + assembly (lbm3d-perm.cpp:507-572)     */
/* ------------------------------------------------------------------------- */
typedef struct {
    int    nx, ny, nz;
    long   n_pore;
    double porosity, force, mu;
    or3_solve run[3];
    double K[3][3];              /* K[flux][force] */
    double k_xy, k_xz, k_yz;
    double recip_xy, recip_xz, recip_yz, recip_resid;
    double kval[3], k_mean, fa;
    double vec[3][3];            /* vec[component][eigen index] */
    int    validated;
    or3_solve val;
    double val_pred[3], val_err;
    int    status;               /* 0 all converged, 3 some did not, 2 hard error */
} or3_tensor;

int or3_solve_tensor(const uint8_t *solid, int nx, int ny, int nz, const or3_cfg *cfg,
                     int validate, or3_tensor *T)
{
    memset(T, 0, sizeof(*T));
    long np = 0;
    int chk = or3_check_mask(solid, nx, ny, nz, &np);
    if(chk) { T->status = 2; return 2; }
    T->nx = nx; T->ny = ny; T->nz = nz; T->n_pore = np;

    lat3 L;
    if(lat3_init(&L, solid, nx, ny, nz)) { lat3_free(&L); T->status = 2; return 2; }

    const double porosity = (double)np / (double)L.nn;
    const double F = FX_BASE * cfg->ff;
    T->porosity = porosity; T->force = F; T->mu = MU3;

    solve_dir(&L, cfg, F, 0, 0, &T->run[0], NULL, NULL, 0);
    if(T->run[0].status) { lat3_free(&L); T->status = 2; return 2; }
    solve_dir(&L, cfg, 0, F, 0, &T->run[1], NULL, NULL, 0);
    if(T->run[1].status) { lat3_free(&L); T->status = 2; return 2; }
    solve_dir(&L, cfg, 0, 0, F, &T->run[2], NULL, NULL, 0);
    if(T->run[2].status) { lat3_free(&L); T->status = 2; return 2; }

    for(int c = 0; c < 3; c++)
    {
        T->K[c][0] = MU3 * T->run[0].q[c] / F;
        T->K[c][1] = MU3 * T->run[1].q[c] / F;
        T->K[c][2] = MU3 * T->run[2].q[c] / F;
    }
    double (*K)[3] = T->K;
    const double trace = K[0][0] + K[1][1] + K[2][2];
    T->k_xy = 0.5*(K[0][1] + K[1][0]);
    T->k_xz = 0.5*(K[0][2] + K[2][0]);
    T->k_yz = 0.5*(K[1][2] + K[2][1]);
    T->recip_xy = fabs(K[0][1] - K[1][0]) / fabs(trace);
    T->recip_xz = fabs(K[0][2] - K[2][0]) / fabs(trace);
    T->recip_yz = fabs(K[1][2] - K[2][1]) / fabs(trace);
    T->recip_resid = dmax(T->recip_xy, dmax(T->recip_xz, T->recip_yz));

    double Ks[3][3] = {{K[0][0], T->k_xy, T->k_xz},
                       {T->k_xy, K[1][1], T->k_yz},
                       {T->k_xz, T->k_yz, K[2][2]}};
    double *val = T->kval;
    jacobi3(Ks, val, T->vec);
    T->k_mean = trace/3.0;
    const double dv = (val[0]-val[1])*(val[0]-val[1])
                    + (val[1]-val[2])*(val[1]-val[2])
                    + (val[2]-val[0])*(val[2]-val[0]);
    const double nrm = val[0]*val[0] + val[1]*val[1] + val[2]*val[2];
    T->fa = nrm > 0 ? sqrt(0.5*dv/nrm) : NAN;

    if(validate)
    {
        const double c = 1.0/sqrt(3.0);
        solve_dir(&L, cfg, F*c, F*c, F*c, &T->val, NULL, NULL, 0);
        if(T->val.status) { lat3_free(&L); T->status = 2; return 2; }
        double err = 0, qn = 0;
        for(int r = 0; r < 3; r++)
        {
            T->val_pred[r] = (Ks[r][0]*F*c + Ks[r][1]*F*c + Ks[r][2]*F*c) / MU3;
            err += (T->val.q[r]-T->val_pred[r])*(T->val.q[r]-T->val_pred[r]);
            qn  += T->val.q[r]*T->val.q[r];
        }
        T->val_err = sqrt(err)/(qn > 0 ? sqrt(qn) : 1);
        T->validated = 1;
    }
    lat3_free(&L);
    T->status = (T->run[0].converged && T->run[1].converged && T->run[2].converged) ? 0 : 3;
    return T->status;
}
