/*
 * oracle/lbm2d_oracle.c -- CPU restatement of the D2Q9 permeability solver of
 * LMB2d/lbm.cpp + LMB2d/main.cpp, used ONLY as the checker for the HIP solver.
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library.  Nothing in the product calls it.
 *
 * What it follows (lbm.cpp unless noted):
 *   ex, ey, inv, w ...................... :20-23
 *   tau = 1, mu = 0.33*(tau-0.5) (float)  :44-45
 *   LX, LY, L0 .......................... lbm.h:24-27 (here run-time; L0 passed in)
 *   volumeavg ........................... :141-152 (exact sum, see below)
 *   snapshot_field / field_residual ..... :158-182
 *   getporosity_full .................... :188-195
 *   resetdistributions / parity ......... :200-211, :665-679
 *   macro / collision / transport ....... :516-585 (float, same operation order)
 *   solve() ............................. main.cpp:99-235
 *   Config defaults ..................... main.cpp:64-79
 *   assembly, angles, --validate ........ main.cpp:314-380
 *
 * The reference fixes the lattice at 256x256 at compile time and scales L0
 * with it (lbm.h:24-27: LX = 256*R, L0 = 4*R).  Here LX, LY and L0 are
 * arguments; at 256x256 with L0 = 4 this is the reference's arithmetic.
 *
 * Built strictly (-O2 -ffp-contract=off): the reference 2D build uses plain
 * -O3 without -march (LMB2d/make_perm.sh:6), i.e. SSE float with no FMA, so
 * its float operations are the IEEE operations written here.
 *
 * Flux sum: volumeavg (:141-152) adds float U into a double accumulator
 * sequentially.  Here each value is truncated to a multiple of 2^-96 and summed
 * exactly in 128-bit integers, then rounded once (the same rule as the 3D
 * oracle), which agrees with the sequential double sum to ~1e-14 relative and
 * is independent of summation order.
 *
 * The reference 2D solver cannot be built here (it needs Magick++), so this
 * restatement is pinned by mathematics instead: the slit closed form for this
 * forcing convention, the blocked-direction value, reciprocity, rot90
 * covariance, and a loop-by-loop Python restatement on tiny lattices (tests/).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Compiled twice: REAL = float (the reference's arithmetic, symbols or2_*)
 * and REAL = double (the same scheme in double precision, symbols or2d_*,
 * the checker of the GPU's FP64 D2Q9 variant).  In double the weights and
 * the force are the double values; mu stays the reference's float. */
#ifndef REAL
#define REAL float
#define PFX(n) or2_##n
#define W2V(v) ((float)(v))
#define FORCE(F) ((float)(F))
#else
#define PFX(n) or2d_##n
#define W2V(v) (v)
#define FORCE(F) (F)
#endif

static const int   EX2[9]  = {0,  1,0,-1, 0,  1,-1,-1, 1};
static const int   EY2[9]  = {0,  0,1, 0,-1,  1, 1,-1,-1};
static const int   INV2[9] = {0, 3,4, 1, 2,  7, 8, 5, 6};
static const REAL  W2[9]   = {W2V(4.0/9.0), W2V(1.0/9.0), W2V(1.0/9.0), W2V(1.0/9.0), W2V(1.0/9.0),
                              W2V(1.0/36.0), W2V(1.0/36.0), W2V(1.0/36.0), W2V(1.0/36.0)};
static const double FX_BASE2 = 2.5e-07;

typedef __int128 i128;

static inline i128 fix96(double u, int *bad)
{
    if(!(fabs(u) < 1.0)) { *bad = 1; return 0; }
    return (i128)(u * 0x1p96);
}
static inline double unfix96(i128 s) { return ldexp((double)s, -96); }

#ifndef OR2_TYPES
#define OR2_TYPES
typedef struct {
    double ff, eps;
    long   lag;
    int    hold;
    long   min_steps, max_steps, avg_steps, avg_sample;
    double qfloor, ksigma;
    double L0;
} or2_cfg;

typedef struct {
    double qx, qy, se_qx, se_qy;
    long   steps;
    int    converged;
    double fieldres;
    int    status;            /* 0 ok, 2 non-finite flux */
    long   fail_step;
} or2_solve;
#endif

void PFX(default_cfg)(or2_cfg *c)          /* main.cpp:64-79 */
{
    c->ff = 1.0; c->eps = 1e-3; c->lag = 5000; c->hold = 3;
    c->min_steps = 10000; c->max_steps = 500000; c->avg_steps = 20000;
    c->avg_sample = 50; c->qfloor = 1e-2; c->ksigma = 3.0; c->L0 = 4.0;
}

float PFX(mu)(void)
{
    float tau = 1;
    float mu = 0.33*(tau-0.5);           /* lbm.cpp:44-45: double product, stored as float */
    return mu;
}

typedef struct {
    int LX, LY;
    double L0;
    int *F;                              /* F[x*LY + y], 1 = solid */
    REAL *df;                            /* df[c][x][y][k] */
    REAL *U, *V, *R, *USNAP, *VSNAP;     /* [x*LY + y] */
    REAL fx, fy, tau;
    int lbm_c;
} lat2;

#define DF(L,c,i,j,k) (L)->df[((((size_t)(c)*(L)->LX + (i))*(L)->LY + (j))*9) + (k)]
#define AT(L,i,j) ((size_t)(i)*(L)->LY + (j))

static void macro2(lat2 *L, int c)                    /* :516-536 */
{
    for(int i = 0; i < L->LX; i++)
    for(int j = 0; j < L->LY; j++)
    if(L->F[AT(L,i,j)] == 0)
    {
        REAL rho = 0, ux = 0, uy = 0;
        for(int k = 0; k < 9; k++)
        {
            rho = rho + DF(L,c,i,j,k);
            ux  = ux  + DF(L,c,i,j,k) * EX2[k];
            uy  = uy  + DF(L,c,i,j,k) * EY2[k];
        }
        ux /= rho;
        uy /= rho;
        L->U[AT(L,i,j)] = ux;
        L->V[AT(L,i,j)] = uy;
        L->R[AT(L,i,j)] = rho;
    }
}

static void collision2(lat2 *L, int c)                /* :540-565 */
{
    const REAL tau = L->tau, fx = L->fx, fy = L->fy;
    for(int i = 0; i < L->LX; i++)
    for(int j = 0; j < L->LY; j++)
    if(L->F[AT(L,i,j)] == 0)
    {
        REAL ux = L->U[AT(L,i,j)];
        REAL uy = L->V[AT(L,i,j)];
        REAL rho = L->R[AT(L,i,j)];
        ux = ux + fx * tau / rho;
        uy = uy + fy * tau / rho;
        for(int k = 0; k < 9; k++)
        {
            REAL feq = W2[k] * rho * ((REAL)1.0 - ((REAL)3.0/(REAL)2.0) * (ux*ux + uy*uy)
                        + (REAL)3.0 * (EX2[k] * ux + EY2[k]*uy)
                        + ((REAL)9.0/(REAL)2.0) * (EX2[k] * ux + EY2[k]*uy) * (EX2[k] * ux + EY2[k]*uy));
            DF(L,c,i,j,k) = DF(L,c,i,j,k) - (1/tau) * (DF(L,c,i,j,k) - feq);
        }
    }
}

static void transport2(lat2 *L, int c)                /* :568-585 */
{
    const int LX = L->LX, LY = L->LY;
    for(int i = 0; i < LX; i++)
    for(int j = 0; j < LY; j++)
    if(L->F[AT(L,i,j)] == 0)
    {
        for(int k = 0; k < 9; k++)
        {
            int ip = (i + EX2[k] + LX) % LX;
            int jp = (j + EY2[k] + LY) % LY;
            if(L->F[AT(L,ip,jp)] == 1)
                DF(L,1-c,i,j,INV2[k]) = DF(L,c,i,j,k);
            else
                DF(L,1-c,ip,jp,k) = DF(L,c,i,j,k);
        }
    }
}

static void lbm2(lat2 *L)                             /* :672-679 */
{
    L->lbm_c = 1 - L->lbm_c;
    int c = L->lbm_c;
    macro2(L, c);
    collision2(L, c);
    transport2(L, c);
}

static int volumeavg2(const lat2 *L, double *qx, double *qy)   /* :141-152 */
{
    int bad = 0;
    i128 su = 0, sv = 0;
    for(int i = 0; i < L->LX; i++)
    for(int j = 0; j < L->LY; j++)
    {
        su += fix96((double)L->U[AT(L,i,j)], &bad);
        sv += fix96((double)L->V[AT(L,i,j)], &bad);
    }
    *qx = unfix96(su) / ((double)L->LX * L->L0 * L->L0);
    *qy = unfix96(sv) / ((double)L->LY * L->L0 * L->L0);
    if(bad) { *qx = *qy = NAN; }
    return bad;
}

static void snapshot_field2(lat2 *L)                  /* :158-167 */
{
    size_t n = (size_t)L->LX * L->LY;
    memcpy(L->USNAP, L->U, n * sizeof(REAL));
    memcpy(L->VSNAP, L->V, n * sizeof(REAL));
}

static float field_residual2(const lat2 *L)           /* :169-182 */
{
    double num = 0.0, den = 0.0;
    for(int i = 0; i < L->LX; i++)
    for(int j = 0; j < L->LY; j++)
    {
        double du = (float)L->U[AT(L,i,j)] - (float)L->USNAP[AT(L,i,j)];
        double dv = (float)L->V[AT(L,i,j)] - (float)L->VSNAP[AT(L,i,j)];
        num += du*du + dv*dv;
        den += (double)(float)L->U[AT(L,i,j)]*(float)L->U[AT(L,i,j)]
             + (double)(float)L->V[AT(L,i,j)]*(float)L->V[AT(L,i,j)];
    }
    return den > 0.0 ? (float)sqrt(num/den) : 10.0f;
}

static void reset2(lat2 *L)                           /* :200-211 */
{
    for(int i = 0; i < L->LX; i++)
    for(int j = 0; j < L->LY; j++)
    {
        for(int k = 0; k < 9; k++)
            DF(L,0,i,j,k) = DF(L,1,i,j,k) = W2[k];
        L->U[AT(L,i,j)] = L->V[AT(L,i,j)] = 0;
        L->R[AT(L,i,j)] = 1;
    }
    L->lbm_c = 1;
}

/* solid is [y][x] (row = solver y, column = solver x), 1 = solid */
static int lat2_init(lat2 *L, const uint8_t *solid, int LX, int LY, double L0)
{
    memset(L, 0, sizeof(*L));
    L->LX = LX; L->LY = LY; L->L0 = L0; L->tau = 1;
    size_t n = (size_t)LX * LY;
    L->F = (int*)malloc(n * sizeof(int));
    L->df = (REAL*)malloc(2 * n * 9 * sizeof(REAL));
    L->U = (REAL*)malloc(n * sizeof(REAL));
    L->V = (REAL*)malloc(n * sizeof(REAL));
    L->R = (REAL*)malloc(n * sizeof(REAL));
    L->USNAP = (REAL*)calloc(n, sizeof(REAL));
    L->VSNAP = (REAL*)calloc(n, sizeof(REAL));
    if(!L->F || !L->df || !L->U || !L->V || !L->R || !L->USNAP || !L->VSNAP) return -1;
    for(int y = 0; y < LY; y++)
        for(int x = 0; x < LX; x++)
            L->F[AT(L,x,y)] = solid[(size_t)y*LX + x] ? 1 : 0;
    return 0;
}

static void lat2_free(lat2 *L)
{
    free(L->F); free(L->df); free(L->U); free(L->V); free(L->R); free(L->USNAP); free(L->VSNAP);
}

static inline double dmax2(double a, double b) { return (a < b) ? b : a; }

/* The flux solve() reads after step s: volumeavg of the lattice, or, with no
 * lattice (L == NULL), entry s of a planted sequence (seq[(s-1)*2 + c]; NaN past
 * its end), which tests use to drive the convergence rules alone. */
static void flux2(const lat2 *L, const double *seq, long nseq, long s, double *qx, double *qy)
{
    if(L) { volumeavg2(L, qx, qy); return; }
    *qx = (s >= 1 && s <= nseq) ? seq[(s - 1) * 2] : NAN;
    *qy = (s >= 1 && s <= nseq) ? seq[(s - 1) * 2 + 1] : NAN;
}

/* main.cpp:99-235.  With L == NULL nothing is stepped and the flux comes from
 * seq (flux2); fieldres is then 0. */
static void solve2(lat2 *L, const or2_cfg *cfg, double Fx, double Fy, or2_solve *out, float *vel,
                   const double *seq, long nseq)
{
    memset(out, 0, sizeof(*out));
    if(L) {
        reset2(L);
        L->fx = FORCE(Fx);
        L->fy = FORCE(Fy);
    }

    long s = 0;
    int  hold = 0;
    int  have_prev = 0;
    double pm_x = 0, pm_y = 0, pse_x = 0, pse_y = 0;
    double bsx = 0, bsy = 0, bsxx = 0, bsyy = 0;
    long   bn = 0;

    while(s < cfg->max_steps)
    {
        if(L) lbm2(L);
        s++;

        if(s % cfg->avg_sample == 0)
        {
            double qx, qy;
            flux2(L, seq, nseq, s, &qx, &qy);
            if(!isfinite(qx) || !isfinite(qy)) { out->status = 2; out->fail_step = s; out->steps = s; return; }
            bsx += qx; bsxx += qx*qx;
            bsy += qy; bsyy += qy*qy;
            bn++;
        }

        if(s % cfg->lag || bn < 2) continue;

        const double mx = bsx / bn, my = bsy / bn;
        const double vx = (bsxx - bn*mx*mx) / (bn - 1);
        const double vy = (bsyy - bn*my*my) / (bn - 1);
        const double sex = sqrt(dmax2(vx, 0.0) / bn);
        const double sey = sqrt(dmax2(vy, 0.0) / bn);

        double dx = 0, dy = 0, tolx = 0, toly = 0;
        if(have_prev)
        {
            const double qmag = sqrt(mx*mx + my*my);
            const double flo  = cfg->qfloor * qmag;
            dx = fabs(mx - pm_x);
            dy = fabs(my - pm_y);
            tolx = dmax2(cfg->eps * dmax2(fabs(mx), flo),
                         cfg->ksigma * sqrt(sex*sex + pse_x*pse_x));
            toly = dmax2(cfg->eps * dmax2(fabs(my), flo),
                         cfg->ksigma * sqrt(sey*sey + pse_y*pse_y));
            if(s >= cfg->min_steps && dx < tolx && dy < toly) hold++;
            else                                              hold = 0;
        }

        pm_x = mx; pm_y = my; pse_x = sex; pse_y = sey; have_prev = 1;
        bsx = bsy = bsxx = bsyy = 0.0; bn = 0;

        if(hold >= cfg->hold) break;
    }

    out->converged = (hold >= cfg->hold) ? 1 : 0;

    if(L) snapshot_field2(L);

    double sx = 0, sy = 0, sxx = 0, syy = 0;
    long   n = 0;
    for(long t = 1; t <= cfg->avg_steps; t++)
    {
        if(L) lbm2(L);
        s++;
        if(t % cfg->avg_sample) continue;
        double qx, qy;
        flux2(L, seq, nseq, s, &qx, &qy);
        if(!isfinite(qx) || !isfinite(qy)) { out->status = 2; out->fail_step = s; out->steps = s; return; }
        sx += qx; sxx += qx*qx;
        sy += qy; syy += qy*qy;
        n++;
    }

    if(n == 0)
    {
        flux2(L, seq, nseq, s, &out->qx, &out->qy);
        n = 1;
    }
    else
    {
        out->qx = sx / n;
        out->qy = sy / n;
        if(n > 1)
        {
            double vx = (sxx - n*out->qx*out->qx) / (n - 1);
            double vy = (syy - n*out->qy*out->qy) / (n - 1);
            out->se_qx = sqrt(dmax2(vx, 0.0) / n);
            out->se_qy = sqrt(dmax2(vy, 0.0) / n);
        }
    }

    out->steps    = s;
    out->fieldres = L ? field_residual2(L) : 0.0;

    if(vel && L)                             /* [y][x] rows of (U, V) */
        for(int y = 0; y < L->LY; y++)
            for(int x = 0; x < L->LX; x++)
            {
                vel[2*((size_t)y*L->LX + x)]     = L->U[AT(L,x,y)];
                vel[2*((size_t)y*L->LX + x) + 1] = L->V[AT(L,x,y)];
            }
}

int PFX(solve_dir)(const uint8_t *solid, int LX, int LY, const or2_cfg *cfg,
                  double Fx, double Fy, or2_solve *out, float *vel)
{
    lat2 L;
    if(lat2_init(&L, solid, LX, LY, cfg->L0)) { lat2_free(&L); return -1; }
    solve2(&L, cfg, Fx, Fy, out, vel, NULL, 0);
    lat2_free(&L);
    return out->status;
}

/* solve()'s convergence rules alone, over a planted flux sequence seq of
 * nseq x 2 values (the flux after step s is seq[(s-1)*2 ..]); returns 0 ok,
 * 2 non-finite flux */
int PFX(replay)(const or2_cfg *cfg, const double *seq, long nseq, or2_solve *out)
{
    solve2(NULL, cfg, 0.0, 0.0, out, NULL, seq, nseq);
    return out->status;
}

/* the field residual (:169-182) of two [y][x] (U, V) float fields */
double PFX(field_residual)(int LX, int LY, const float *cur, const float *snap)
{
    lat2 L;
    memset(&L, 0, sizeof(L));
    L.LX = LX; L.LY = LY;
    size_t n = (size_t)LX * LY;
    L.U = (REAL*)malloc(n * sizeof(REAL)); L.V = (REAL*)malloc(n * sizeof(REAL));
    L.USNAP = (REAL*)malloc(n * sizeof(REAL)); L.VSNAP = (REAL*)malloc(n * sizeof(REAL));
    if(!L.U || !L.V || !L.USNAP || !L.VSNAP) { lat2_free(&L); return NAN; }
    for(int y = 0; y < LY; y++)
        for(int x = 0; x < LX; x++) {
            size_t d = 2 * ((size_t)y * LX + x);
            L.U[AT(&L,x,y)] = cur[d];  L.V[AT(&L,x,y)] = cur[d + 1];
            L.USNAP[AT(&L,x,y)] = snap[d]; L.VSNAP[AT(&L,x,y)] = snap[d + 1];
        }
    double r = field_residual2(&L);
    lat2_free(&L);
    return r;
}

/* Run `steps` steps from rest and return df of the buffer the next step reads
 * (post-stream state, df[1-c] after the last step) in [x][y][k] order, and
 * U,V as [y][x] pairs.  For step-by-step state comparison. */
int PFX(run_steps)(const uint8_t *solid, int LX, int LY, double Fx, double Fy, long steps,
                  REAL *df_out, float *uv_out)
{
    lat2 L;
    if(lat2_init(&L, solid, LX, LY, 4.0)) { lat2_free(&L); return -1; }
    reset2(&L);
    L.fx = FORCE(Fx); L.fy = FORCE(Fy);
    for(long s = 0; s < steps; s++) lbm2(&L);
    size_t n = (size_t)LX * LY;
    int next = 1 - L.lbm_c;               /* buffer the next lbm() will read */
    if(df_out) memcpy(df_out, &L.df[(size_t)next * n * 9], n * 9 * sizeof(REAL));
    if(uv_out)
        for(int y = 0; y < LY; y++)
            for(int x = 0; x < LX; x++)
            {
                uv_out[2*((size_t)y*LX + x)]     = L.U[AT(&L,x,y)];
                uv_out[2*((size_t)y*LX + x) + 1] = L.V[AT(&L,x,y)];
            }
    lat2_free(&L);
    return 0;
}

/* main.cpp:314-380 */
#ifndef OR2_TENSOR
#define OR2_TENSOR
typedef struct {
    int    LX, LY;
    double porosity, force, mu;
    or2_solve run[2];
    double k_xx, k_yx, k_xy, k_yy, k_off, recip;
    double k1, k2, k_mean, aniso, theta, alpha_x, alpha_y;
    int    validated;
    or2_solve val;
    double val_pred[2], val_err;
    int    status;
} or2_tensor;
#endif

int PFX(solve_tensor)(const uint8_t *solid, int LX, int LY, const or2_cfg *cfg,
                     double validate_deg, or2_tensor *T)
{
    memset(T, 0, sizeof(*T));
    lat2 L;
    if(lat2_init(&L, solid, LX, LY, cfg->L0)) { lat2_free(&L); T->status = 2; return 2; }
    T->LX = LX; T->LY = LY;

    const double F = FX_BASE2 * cfg->ff;
    int nfluid = 0;                                    /* getporosity_full :188-195 */
    for(int i = 0; i < LX; i++) for(int j = 0; j < LY; j++) if(L.F[AT(&L,i,j)] == 0) nfluid++;
    float por = nfluid / (float)(LX*LY);
    T->porosity = por;
    T->force = F;
    const float mu = PFX(mu)();
    T->mu = mu;

    solve2(&L, cfg, F, 0.0, &T->run[0], NULL, NULL, 0);
    if(T->run[0].status) { lat2_free(&L); T->status = 2; return 2; }
    solve2(&L, cfg, 0.0, F, &T->run[1], NULL, NULL, 0);
    if(T->run[1].status) { lat2_free(&L); T->status = 2; return 2; }

    const or2_solve *rx = &T->run[0], *ry = &T->run[1];
    const double k_xx = mu * rx->qx / F;
    const double k_yx = mu * rx->qy / F;
    const double k_xy = mu * ry->qx / F;
    const double k_yy = mu * ry->qy / F;
    const double k_off = 0.5 * (k_xy + k_yx);
    const double trace = k_xx + k_yy;
    const double recip = trace != 0.0 ? fabs(k_xy - k_yx) / fabs(trace) : NAN;
    const double km = 0.5 * (k_xx + k_yy);
    const double d1 = 0.5 * (k_xx - k_yy);
    const double d2 = k_off;
    const double r  = sqrt(d1*d1 + d2*d2);
    T->k_xx = k_xx; T->k_yx = k_yx; T->k_xy = k_xy; T->k_yy = k_yy;
    T->k_off = k_off; T->recip = recip;
    T->k1 = km + r; T->k2 = km - r; T->k_mean = km;
    T->aniso = km != 0.0 ? r / km : NAN;
    T->theta = 0.5 * atan2(d2, d1) * 180.0 / M_PI;
    T->alpha_x = atan2(rx->qy, rx->qx) * 180.0 / M_PI;
    T->alpha_y = atan2(ry->qy, ry->qx) * 180.0 / M_PI - 90.0;

    if(validate_deg >= 0.0)
    {
        const double th = validate_deg * M_PI / 180.0;
        const double Fx = F * cos(th), Fy = F * sin(th);
        solve2(&L, cfg, Fx, Fy, &T->val, NULL, NULL, 0);
        if(T->val.status) { lat2_free(&L); T->status = 2; return 2; }
        const double pred_qx = (k_xx * Fx + k_off * Fy) / mu;
        const double pred_qy = (k_off * Fx + k_yy * Fy) / mu;
        const double qn = sqrt(T->val.qx*T->val.qx + T->val.qy*T->val.qy);
        T->val_pred[0] = pred_qx; T->val_pred[1] = pred_qy;
        T->val_err = sqrt((T->val.qx-pred_qx)*(T->val.qx-pred_qx)
                        + (T->val.qy-pred_qy)*(T->val.qy-pred_qy)) / (qn > 0 ? qn : 1);
        T->validated = 1;
    }
    lat2_free(&L);
    T->status = (rx->converged && ry->converged) ? 0 : 3;
    return T->status;
}
