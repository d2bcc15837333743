// collide.h -- per-node moments + collision for every (lattice, precision).
//
// The whole library is compiled with -ffp-contract=off, so each expression
// below is exactly the IEEE operation sequence written.  The FP64 D3Q19 and
// the float D2Q9 "reference" variants therefore reproduce the reference
// arithmetic bit for bit (the CPU oracle is built the same way); the shifted
// FP32 variants are free to use explicit fmaf.
#pragma once
#include "lbm_common.h"

namespace fp {

// ---------------------------------------------------------------------------
// D3Q19, FP64: collide_node of LMB3d/lbm3d-perm.cpp:218-245.
//   rho = sum g;  u = (sum g e + F/2) / rho                         (Guo)
//   feq = w rho (1 + 3 e.u + 4.5 (e.u)^2 - 1.5 u^2)
//   S   = w (1 - 1/2tau) [3 (e - u).F + 9 (e.u)(e.F)]
//   out = g + ((feq - g)/tau + S),  tau = 1
// Terms with a zero lattice component are dropped from the sums: adding
// g*0 = +-0 to a running sum leaves it unchanged, so the result is the same.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void collide3_f64(double *g, const double F[3],
                                             double &uxo, double &uyo, double &uzo)
{
    using L = D3Q19;
    double r = 0, ux = 0, uy = 0, uz = 0;
#pragma unroll
    for(int k = 0; k < 19; k++) {
        r += g[k];
        if(L::ex(k) == 1) ux += g[k]; else if(L::ex(k) == -1) ux -= g[k];
        if(L::ey(k) == 1) uy += g[k]; else if(L::ey(k) == -1) uy -= g[k];
        if(L::ez(k) == 1) uz += g[k]; else if(L::ez(k) == -1) uz -= g[k];
    }
    const double Fx = F[0], Fy = F[1], Fz = F[2];
    ux = (ux + 0.5*Fx)/r;
    uy = (uy + 0.5*Fy)/r;
    uz = (uz + 0.5*Fz)/r;
    uxo = ux; uyo = uy; uzo = uz;

    const double u2 = ux*ux + uy*uy + uz*uz;
    const double pref = 0.5;                          // 1 - 0.5/tau at tau = 1
#pragma unroll
    for(int k = 0; k < 19; k++) {
        // e.u accumulated over the non-zero components, in x, y, z order
        double eu = 0.0;
        bool any = false;
        if(L::ex(k) != 0) { eu = L::ex(k) == 1 ? ux : -ux; any = true; }
        if(L::ey(k) != 0) { double t = L::ey(k) == 1 ? uy : -uy; eu = any ? eu + t : t; any = true; }
        if(L::ez(k) != 0) { double t = L::ez(k) == 1 ? uz : -uz; eu = any ? eu + t : t; any = true; }
        const double wq = L::w(k);
        const double feq = wq*r*(1.0 + 3.0*eu + 4.5*eu*eu - 1.5*u2);
        const double ef = (L::ex(k)*Fx + L::ey(k)*Fy) + L::ez(k)*Fz;
        const double sk = wq * pref *
            ( 3.0*(((L::ex(k)-ux)*Fx + (L::ey(k)-uy)*Fy) + (L::ez(k)-uz)*Fz)
            + 9.0*eu*ef );
        g[k] = g[k] + ((feq - g[k]) + sk);
    }
}

// ---------------------------------------------------------------------------
// D3Q19, FP32 shifted populations h = f - w, tau = 1.  Same model as above:
//   rho - 1 = sum h,  j = sum h e  (sum w e = 0),  u = (j + F/2)/rho
//   h'_k = feq_k - w_k + S_k
//        = w_k [ (rho-1) + rho(4.5 eu^2 - 1.5 u^2) - 1.5 u.F + 4.5 eu eF_k ]   (even in k)
//        + w_k [ 3 rho eu + 1.5 eF_k ]                                      (odd in k)
// so each opposite pair (k, opp k) shares the even part and flips the odd one.
// ---------------------------------------------------------------------------
// g45[k] = 4.5 w_k (e_k.F), h15[k] = 1.5 w_k (e_k.F) come precomputed per
// instance (InstDesc), so they sit in scalar registers.
__device__ __forceinline__ void collide3_f32(float *h, const float F[3], const float *g45,
                                             const float *h15, float &uxo, float &uyo, float &uzo)
{
    using L = D3Q19;
    float drho = ((h[0] + (h[1] + h[2])) + ((h[3] + h[4]) + (h[5] + h[6])))
               + (((h[7] + h[8]) + (h[9] + h[10])) + ((h[11] + h[12]) + (h[13] + h[14])))
               + ((h[15] + h[16]) + (h[17] + h[18]));
    float jx = ((h[1] - h[2]) + (h[7] - h[8])) + ((h[9] - h[10]) + (h[11] - h[12])) + (h[13] - h[14]);
    float jy = ((h[3] - h[4]) + (h[7] - h[8])) + ((h[10] - h[9]) + (h[15] - h[16])) + (h[17] - h[18]);
    float jz = ((h[5] - h[6]) + (h[11] - h[12])) + ((h[14] - h[13]) + (h[15] - h[16])) + (h[18] - h[17]);
    const float rho = 1.0f + drho;
    const float inv = 1.0f / rho;
    const float ux = fmaf(0.5f, F[0], jx) * inv;
    const float uy = fmaf(0.5f, F[1], jy) * inv;
    const float uz = fmaf(0.5f, F[2], jz) * inv;
    uxo = ux; uyo = uy; uzo = uz;
    const float u2 = fmaf(ux, ux, fmaf(uy, uy, uz*uz));
    const float uF = fmaf(ux, F[0], fmaf(uy, F[1], uz*F[2]));
    const float base = fmaf(-1.5f, fmaf(rho, u2, uF), drho);   // drho - 1.5 rho u2 - 1.5 uF
    const float w0 = 1.0f/3.0f, w1 = 1.0f/18.0f, w2 = 1.0f/36.0f;
    h[0] = w0 * base;
    const float C1 = w1 * base, C2 = w2 * base;
    const float A1 = 4.5f * w1 * rho, A2 = 4.5f * w2 * rho;
    const float B1 = 3.0f * w1 * rho, B2 = 3.0f * w2 * rho;
#pragma unroll
    for(int k = 1; k < 19; k += 2) {
        const int kb = k + 1;                        // opp(k) for this ordering
        float eu = 0.0f;
        if(L::ex(k)) eu = L::ex(k) > 0 ? ux : -ux;
        if(L::ey(k)) eu = L::ey(k) > 0 ? eu + uy : eu - uy;
        if(L::ez(k)) eu = L::ez(k) > 0 ? eu + uz : eu - uz;
        const bool face = k < 7;
        const float C = face ? C1 : C2, A = face ? A1 : A2, B = face ? B1 : B2;
        const float E = fmaf(A, eu*eu, fmaf(g45[k], eu, C));
        const float O = fmaf(B, eu, h15[k]);
        h[k]  = E + O;
        h[kb] = E - O;
    }
}

// ---------------------------------------------------------------------------
// D2Q9, the reference's own scheme (LMB2d/lbm.cpp:516-565), in T = float
// (the reference's arithmetic, unshifted) or T = double:
//   rho = sum f;  u = sum f e / rho                (raw moment, reported)
//   u_eq = u + F tau / rho                         (equilibrium-velocity shift)
//   f <- f - (1/tau)(f - feq(rho, u_eq)),  tau = 1
// ---------------------------------------------------------------------------
template<class T>
__device__ __forceinline__ void collide2_ref(T *f, T fx, T fy, T &uxo, T &uyo)
{
    using L = D2Q9;
    T rho = 0, ux = 0, uy = 0;
#pragma unroll
    for(int k = 0; k < 9; k++) {
        rho = rho + f[k];
        if(L::ex(k) == 1) ux = ux + f[k]; else if(L::ex(k) == -1) ux = ux - f[k];
        if(L::ey(k) == 1) uy = uy + f[k]; else if(L::ey(k) == -1) uy = uy - f[k];
    }
    ux /= rho;
    uy /= rho;
    uxo = ux; uyo = uy;
    const T tau = 1;
    ux = ux + fx * tau / rho;
    uy = uy + fy * tau / rho;
#pragma unroll
    for(int k = 0; k < 9; k++) {
        T eu;
        if(L::ex(k) != 0 && L::ey(k) != 0) {
            const T a = L::ex(k) == 1 ? ux : -ux;
            const T b = L::ey(k) == 1 ? uy : -uy;
            eu = a + b;
        } else if(L::ex(k) != 0) eu = L::ex(k) == 1 ? ux : -ux;
        else if(L::ey(k) != 0)   eu = L::ey(k) == 1 ? uy : -uy;
        else                     eu = 0;
        const T w = (T)((sizeof(T) == 4) ? (double)L::wf(k)
                                        : (k == 0 ? 4.0/9.0 : (k < 5 ? 1.0/9.0 : 1.0/36.0)));
        const T feq = w * rho * (T(1) - (T(3)/T(2)) * (ux*ux + uy*uy) + T(3) * eu
                                 + (T(9)/T(2)) * eu * eu);
        f[k] = f[k] - (T(1)/tau) * (f[k] - feq);
    }
}

// ---------------------------------------------------------------------------
// D2Q9 shifted FP32 (h = f - w), the same scheme at tau = 1: the output is
// feq(rho, u_eq) - w, evaluated pairwise like collide3_f32.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void collide2_f32(float *h, float fx, float fy, float &uxo, float &uyo)
{
    const float drho = (h[0] + ((h[1] + h[3]) + (h[2] + h[4]))) + ((h[5] + h[7]) + (h[6] + h[8]));
    const float jx = ((h[1] - h[3]) + (h[5] - h[7])) + (h[8] - h[6]);
    const float jy = ((h[2] - h[4]) + (h[5] - h[7])) + (h[6] - h[8]);
    const float rho = 1.0f + drho;
    const float inv = 1.0f / rho;
    const float ux = jx * inv, uy = jy * inv;        // raw moment (reported)
    uxo = ux; uyo = uy;
    const float vx = fmaf(fx, inv, ux), vy = fmaf(fy, inv, uy);   // u_eq
    const float u2 = fmaf(vx, vx, vy*vy);
    const float base = fmaf(-1.5f * rho, u2, drho);               // drho - 1.5 rho u2
    const float w0 = 4.0f/9.0f, w1 = 1.0f/9.0f, w2 = 1.0f/36.0f;
    h[0] = w0 * base;
    const float C1 = w1 * base, C2 = w2 * base;
    const float A1 = 4.5f * w1 * rho, A2 = 4.5f * w2 * rho;
    const float B1 = 3.0f * w1 * rho, B2 = 3.0f * w2 * rho;
    // pairs (1,3) (2,4) (5,7) (6,8)
    float e;
    e = vx;       { float E = fmaf(A1, e*e, C1), O = B1*e; h[1] = E + O; h[3] = E - O; }
    e = vy;       { float E = fmaf(A1, e*e, C1), O = B1*e; h[2] = E + O; h[4] = E - O; }
    e = vx + vy;  { float E = fmaf(A2, e*e, C2), O = B2*e; h[5] = E + O; h[7] = E - O; }
    e = vy - vx;  { float E = fmaf(A2, e*e, C2), O = B2*e; h[6] = E + O; h[8] = E - O; }
}

} // namespace fp
