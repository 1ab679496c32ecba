// dispersion.hip -- per-column fundamental-mode Rayleigh/Love dispersion on gfx950.
//
// Replaces, for every grid column, convert_to_layer (src/likelihood_surf.F90
// :523-629), the surfdisp96 branch of surfmodes (surfmodes/surfmodes.f90
// :79-95) and surfdisp96 itself (surfmodes/surfdisp96.f:52-382).
//
// Three launch shapes.  k_dispersion_sm (throughput): one lane per column,
// columns taken from a queue, surfdisp96's control flow as a per-lane state
// machine with one period-equation call site, so the 64 lanes of a wave share
// every evaluation whatever stage of the root search each is in.
// k_dispersion_lane: one lane per column running surfdisp96 as written (lanes
// wait for the wave's longest walk at every branch).  k_dispersion: one
// wave64 per column with the layer model in LDS (latency).  The expensive part of
// surfdisp96 is getsol's bracketing walk (:759-815): one period-equation
// evaluation per dc step, ~10^3 steps for the first period.  Here the 64 lanes
// evaluate 64 consecutive steps at once; every lane replays the walk's own
// sequence of FP64 additions (c2 = c1 + dc, the clow floor and direction
// rules) so its trial velocity is bit-identical to the reference's, and a
// wave ballot finds the first sign change or exit condition.  The Neville /
// bisection refinement (:903-1020) is sequential and runs in all lanes in
// lockstep.  Arithmetic types follow the Fortran (REAL*4 where the reference
// is implicitly typed real, FP64 in the period equations); FP contraction is
// off.  The period equations take exp, sin and cos from libm_glibc.h (the host
// C library's numbers, which the reference's build uses) and sqrt and division
// from the hardware's correctly rounded operations.
#include "mct_internal.h"
#include "libm_glibc.h"
#include <float.h>

#pragma clang fp contract(off)

#define DSP_NMAX 100           // NMAX layers per column (src/likelihood_surf.F90:30)
#define DSP_MAXP 60            // NP periods (surfmodes/surfdisp96.f:59)

// vp and rho of the layer that starts at node k0 (0-based) of a column whose
// Vs there is vsk: the Voronoi sites' own values (the surface half of the
// joint likelihood, datatype 3: src/likelihood.f90:83-88 reads model%vp and
// model%rho as kdtree_to_grid set them) or vs2vp_3d / vp2rho_3d of Vs
// (datatype 2: src/likelihood.f90:75-76, src/utils.f90:102-134).
__device__ __forceinline__ void dsp_vprho(const mct_dsp_args &A, long cbase, int k0, double vsk, double &vp,
                                          double &rho)
{
    if (A.vp) {
        vp = A.vp[cbase + k0];
        rho = A.rho[cbase + k0];
        return;
    }
    vp = vsk * (double)1.730f;   // POISSON, a default-real literal
    rho = (double)1.74f * pow(vp, 0.25);
}

// convert_to_layer (src/likelihood_surf.F90:523-629) for the column at cb into
// REAL*4 layer arrays of element stride ST, as surfmodes passes them on
// (real(thick,4) ..., surfmodes/surfmodes.f90:81-83).  With A.water > EPS a
// water layer comes first (vp = waterVel 1.5, vs = 0, rho = waterDensity 1,
// :32-33,546-551); a layer starts where vs changes by more than EPS and takes
// the values of its first node; the half-space (thickness 0) takes the bottom
// node's (either branch of :575-602).  Thicknesses are scaled, except that
// with A.water > 0 the first layer's is A.water itself (:616-617: the water
// layer's, or for 0 < water <= EPS the first solid layer's).  Returns the
// layer count (at most nz + 1).
template <int ST>
__device__ __forceinline__ int dsp_convert(const mct_dsp_args &A, long cb, const double *vs, float *sd, float *sa,
                                           float *sb, float *sr)
{
    const double EPS = (double)1.0e-10f;       // src/likelihood_surf.F90:37
    int nl = 0, last_k = 1;
    if (A.water > EPS) {
        sa[0] = 1.5f; sb[0] = 0.0f; sr[0] = 1.0f; sd[0] = (float)A.water;
        nl = 1;
    }
    double last_vs = vs[0];
    for (int k = 2; k <= A.nz; k++) {
        const double vk = vs[k - 1];
        if (fabs(vk - last_vs) > EPS) {
            double vpk, rhok;
            dsp_vprho(A, cb, last_k - 1, last_vs, vpk, rhok);
            sa[ST * nl] = (float)vpk;
            sb[ST * nl] = (float)last_vs;
            sr[ST * nl] = (float)rhok;
            sd[ST * nl] = (float)(((double)(k - last_k) * A.dz) / A.scaling);
            nl++;
            last_vs = vk;
            last_k = k;
        }
    }
    {
        double vpk, rhok;
        dsp_vprho(A, cb, A.nz - 1, vs[A.nz - 1], vpk, rhok);
        sa[ST * nl] = (float)vpk;
        sb[ST * nl] = (float)vs[A.nz - 1];
        sr[ST * nl] = (float)rhok;
        sd[ST * nl] = 0.0f;
    }
    if (A.water > 0.0) sd[0] = (float)A.water;
    return nl + 1;
}

// A column's layer model; ST is the element stride of the layer arrays (1 in
// LDS, 64 when lanes interleave their models).
template <int ST> struct dsp_model_t {
    static constexpr int st = ST;
    int mmax, llw;
    const float *d, *a, *b, *rho;
};

__device__ __forceinline__ double dsgn(double x) { return copysign(1.0, x); }

__device__ static float d_gtsolh(float a, float b)
{
    float c = 0.95f * b;
    for (int i = 0; i < 5; i++) {
        float gamma = b / a;
        float kappa = c / b;
        float k2 = kappa * kappa;
        float gk2 = (gamma * kappa) * (gamma * kappa);
        float fac1 = sqrtf(1.0f - gk2);
        float fac2 = sqrtf(1.0f - k2);
        float fr = (2.0f - k2) * (2.0f - k2) - 4.0f * fac1 * fac2;
        float frp = -(4.0f * (2.0f - k2) * kappa) + 4.0f * fac2 * gamma * gamma * kappa / fac1
                    + 4.0f * fac1 * kappa / fac2;
        frp = frp / b;
        c = c - fr / frp;
    }
    return c;
}

// SH period equation, dltar1 (surfdisp96.f:1056-1115)
template <class MD>
__device__ static double d_dltar1(double wvno, double omega, const MD &M)
{
    double beta1 = (double)M.b[M.st * (M.mmax - 1)];
    double rho1 = (double)M.rho[M.st * (M.mmax - 1)];
    double xkb = omega / beta1;
    double wvnop = wvno + xkb;
    double wvnom = fabs(wvno - xkb);
    double rb = sqrt(wvnop * wvnom);
    double e1 = rho1 * rb;
    double e2 = 1.0 / (beta1 * beta1);
    for (int m = M.mmax - 1; m >= M.llw; m--) {
        beta1 = (double)M.b[M.st * (m - 1)];
        rho1 = (double)M.rho[M.st * (m - 1)];
        double xmu = rho1 * beta1 * beta1;
        xkb = omega / beta1;
        wvnop = wvno + xkb;
        wvnom = fabs(wvno - xkb);
        rb = sqrt(wvnop * wvnom);
        double q = (double)M.d[M.st * (m - 1)] * rb;
        double sinq, y, z, cosq;
        if (wvno < xkb) {
            mct_libm::libm_sincos(q, &sinq, &cosq);
            y = sinq / rb;
            z = -(rb * sinq);
        } else if (wvno == xkb) {
            cosq = 1.0;
            y = (double)M.d[M.st * (m - 1)];
            z = 0.0;
        } else {
            double fac = 0.0;
            if (q < 16) fac = mct_libm::libm_exp(-2.0 * q);
            cosq = (1.0 + fac) * 0.5;
            sinq = (1.0 - fac) * 0.5;
            y = sinq / rb;
            z = rb * sinq;
        }
        double e10 = e1 * cosq + e2 * xmu * z;
        double e20 = e1 * y / xmu + e2 * cosq;
        double xnor = fabs(e10), ynor = fabs(e20);
        if (ynor > xnor) xnor = ynor;
        if (xnor < 1.0e-40) xnor = 1.0;
        e1 = e10 / xnor;
        e2 = e20 / xnor;
    }
    return e1;
}

// P-SV period equation, dltar4 with var and dnka inlined (surfdisp96.f:1119-1414)
template <class MD>
__device__ static double d_dltar4(double wvno, double omga, const MD &M)
{
    double omega = omga;
    if (omega < 1.0e-4) omega = 1.0e-4;
    const double one = 1.0, two = 2.0;
    double wvno2 = wvno * wvno;
    int mm = M.mmax - 1;
    double xka = omega / (double)M.a[M.st * (mm)];
    double xkb = omega / (double)M.b[M.st * (mm)];
    double wvnop = wvno + xka, wvnom = fabs(wvno - xka);
    double ra = sqrt(wvnop * wvnom);
    wvnop = wvno + xkb;
    wvnom = fabs(wvno - xkb);
    double rb = sqrt(wvnop * wvnom);
    double t = (double)M.b[M.st * (mm)] / omega;
    double gammk = 2.0 * t * t;
    double gam = gammk * wvno2;
    double gamm1 = gam - 1.0;
    double rho1 = (double)M.rho[M.st * (mm)];
    double e0 = rho1 * rho1 * (gamm1 * gamm1 - gam * gammk * ra * rb);
    double e1 = -(rho1 * ra);
    double e2 = rho1 * (gamm1 - gammk * ra * rb);
    double e3 = rho1 * rb;
    double e4 = wvno2 - ra * rb;
    for (int m = M.mmax - 1; m >= M.llw; m--) {
        int i0 = m - 1;
        xka = omega / (double)M.a[M.st * (i0)];
        xkb = omega / (double)M.b[M.st * (i0)];
        t = (double)M.b[M.st * (i0)] / omega;
        gammk = 2.0 * t * t;
        gam = gammk * wvno2;
        wvnop = wvno + xka;
        wvnom = fabs(wvno - xka);
        ra = sqrt(wvnop * wvnom);
        wvnop = wvno + xkb;
        wvnom = fabs(wvno - xkb);
        rb = sqrt(wvnop * wvnom);
        double dpth = (double)M.d[M.st * (i0)];
        rho1 = (double)M.rho[M.st * (i0)];
        double p = ra * dpth, q = rb * dpth;
        // ---- var
        double pex = 0.0, sex = 0.0, sinp, w = 0, x = 0, cosp = 0, sinq, y = 0, z = 0, cosq = 0;
        if (wvno < xka) {
            mct_libm::libm_sincos(p, &sinp, &cosp);
            w = sinp / ra;
            x = -(ra * sinp);
        } else if (wvno == xka) {
            cosp = 1.0;
            w = dpth;
            x = 0.0;
        } else if (wvno > xka) {
            pex = p;
            double fac = 0.0;
            if (p < 16) fac = mct_libm::libm_exp(-2.0 * p);
            cosp = (1.0 + fac) * 0.5;
            sinp = (1.0 - fac) * 0.5;
            w = sinp / ra;
            x = ra * sinp;
        }
        if (wvno < xkb) {
            mct_libm::libm_sincos(q, &sinq, &cosq);
            y = sinq / rb;
            z = -(rb * sinq);
        } else if (wvno == xkb) {
            cosq = 1.0;
            y = dpth;
            z = 0.0;
        } else if (wvno > xkb) {
            sex = q;
            double fac = 0.0;
            if (q < 16) fac = mct_libm::libm_exp(-2.0 * q);
            cosq = (1.0 + fac) * 0.5;
            sinq = (1.0 - fac) * 0.5;
            y = sinq / rb;
            z = rb * sinq;
        }
        double exa = pex + sex;
        double a0 = 0.0;
        if (exa < 60.0) a0 = mct_libm::libm_exp(-exa);
        double cpcq = cosp * cosq, cpy = cosp * y, cpz = cosp * z, cqw = cosq * w, cqx = cosq * x;
        double xy = x * y, xz = x * z, wy = w * y, wz = w * z;
        // ---- dnka
        double g1 = gam - one;
        double twgm1 = gam + g1;
        double gmgmk = gam * gammk;
        double gmgm1 = gam * g1;
        double gm1sq = g1 * g1;
        double rho2 = rho1 * rho1;
        double a0pq = a0 - cpcq;
        double c11 = cpcq - two * gmgm1 * a0pq - gmgmk * xz - wvno2 * gm1sq * wy;
        double c12 = (wvno2 * cpy - cqx) / rho1;
        double c13 = -((twgm1 * a0pq + gammk * xz + wvno2 * g1 * wy) / rho1);
        double c14 = (cpz - wvno2 * cqw) / rho1;
        double c15 = -((two * wvno2 * a0pq + xz + wvno2 * wvno2 * wy) / rho2);
        double c21 = (gmgmk * cpz - gm1sq * cqw) * rho1;
        double c22 = cpcq;
        double c23 = gammk * cpz - g1 * cqw;
        double c24 = -wz;
        double c25 = c14;
        double c41 = (gm1sq * cpy - gmgmk * cqx) * rho1;
        double c42 = -xy;
        double c43 = g1 * cpy - gammk * cqx;
        double c44 = c22;
        double c45 = c12;
        double c51 = -((two * gmgmk * gm1sq * a0pq + gmgmk * gmgmk * xz + gm1sq * gm1sq * wy) * rho2);
        double c52 = c41;
        double c53 = -((gammk * g1 * twgm1 * a0pq + gam * gammk * gammk * xz + g1 * gm1sq * wy) * rho1);
        double c54 = c21;
        double c55 = c11;
        double tt = -(two * wvno2);
        double c31 = tt * c53;
        double c32 = tt * c43;
        double c33 = a0 + two * (cpcq - c11);
        double c34 = tt * c23;
        double c35 = tt * c13;
        // ee(i) = sum_j e(j) ca(j,i), summed in j order from 0
        double ee0 = 0.0, ee1 = 0.0, ee2 = 0.0, ee3 = 0.0, ee4 = 0.0;
        ee0 = ee0 + e0 * c11; ee0 = ee0 + e1 * c21; ee0 = ee0 + e2 * c31; ee0 = ee0 + e3 * c41; ee0 = ee0 + e4 * c51;
        ee1 = ee1 + e0 * c12; ee1 = ee1 + e1 * c22; ee1 = ee1 + e2 * c32; ee1 = ee1 + e3 * c42; ee1 = ee1 + e4 * c52;
        ee2 = ee2 + e0 * c13; ee2 = ee2 + e1 * c23; ee2 = ee2 + e2 * c33; ee2 = ee2 + e3 * c43; ee2 = ee2 + e4 * c53;
        ee3 = ee3 + e0 * c14; ee3 = ee3 + e1 * c24; ee3 = ee3 + e2 * c34; ee3 = ee3 + e3 * c44; ee3 = ee3 + e4 * c54;
        ee4 = ee4 + e0 * c15; ee4 = ee4 + e1 * c25; ee4 = ee4 + e2 * c35; ee4 = ee4 + e3 * c45; ee4 = ee4 + e4 * c55;
        // ---- normc
        double t1 = 0.0;
        if (fabs(ee0) > t1) t1 = fabs(ee0);
        if (fabs(ee1) > t1) t1 = fabs(ee1);
        if (fabs(ee2) > t1) t1 = fabs(ee2);
        if (fabs(ee3) > t1) t1 = fabs(ee3);
        if (fabs(ee4) > t1) t1 = fabs(ee4);
        if (t1 < 1.0e-40) t1 = 1.0;
        e0 = ee0 / t1; e1 = ee1 / t1; e2 = ee2 / t1; e3 = ee3 / t1; e4 = ee4 / t1;
    }
    if (M.llw != 1) {   // water layer on top
        xka = omega / (double)M.a[M.st * (0)];
        wvnop = wvno + xka;
        wvnom = fabs(wvno - xka);
        ra = sqrt(wvnop * wvnom);
        double dpth = (double)M.d[M.st * (0)];
        rho1 = (double)M.rho[M.st * (0)];
        double p = ra * dpth;
        double w = 0, cosp = 0;
        if (wvno < xka) { double sp; mct_libm::libm_sincos(p, &sp, &cosp); w = sp / ra; }
        else if (wvno == xka) { cosp = 1.0; w = dpth; }
        else {
            double fac = 0.0;
            if (p < 16) fac = mct_libm::libm_exp(-2.0 * p);
            cosp = (1.0 + fac) * 0.5;
            w = ((1.0 - fac) * 0.5) / ra;
        }
        double w0 = -(rho1 * w);
        return cosp * e0 + w0 * e1;
    }
    return e0;
}

template <class MD>
__device__ __forceinline__ double d_dltar(double wvno, double omega, int ifunc, const MD &M)
{
    return ifunc == 1 ? d_dltar1(wvno, omega, M) : d_dltar4(wvno, omega, M);
}

// nevill (surfdisp96.f:903-1020), run identically by every lane of the wave
template <class MD>
__device__ static double d_nevill(double t, double c1, double c2, double del1, double del2, int ifunc,
                                  const MD &M, double twopi, int &ndl)
{
    double x[12], y[12];
    double omega = twopi / t;
    double c3 = 0.5 * (c1 + c2);
    double del3 = d_dltar(omega / c3, omega, ifunc, M);
    ndl++;
    int nev = 1, nctrl = 1, m = 1;
    for (;;) {
        nctrl = nctrl + 1;
        if (nctrl >= 100) break;
        if (c3 < fmin(c1, c2) || c3 > fmax(c1, c2)) {
            nev = 0;
            c3 = 0.5 * (c1 + c2);
            del3 = d_dltar(omega / c3, omega, ifunc, M);
            ndl++;
        }
        double s13 = del1 - del3;
        double s32 = del3 - del2;
        if (dsgn(del3) * dsgn(del1) < 0.0) { c2 = c3; del2 = del3; }
        else { c1 = c3; del1 = del3; }
        if (fabs(c1 - c2) <= 1.0e-6 * c1) break;
        if (dsgn(s13) != dsgn(s32)) nev = 0;
        double ss1 = fabs(del1);
        double s1 = (double)0.01f * ss1;
        double ss2 = fabs(del2);
        double s2 = (double)0.01f * ss2;
        bool halve = s1 > ss2 || s2 > ss1 || nev == 0;
        if (!halve) {
            if (nev == 2) {
                x[m + 1] = c3;
                y[m + 1] = del3;
            } else {
                x[1] = c1; y[1] = del1;
                x[2] = c2; y[2] = del2;
                m = 1;
            }
            bool bad = false;
            for (int kk = 1; kk <= m; kk++) {
                int j = m - kk + 1;
                double denom = y[m + 1] - y[j];
                if (fabs(denom) < 1.0e-10 * fabs(y[m + 1])) { bad = true; break; }
                x[j] = (-(y[j] * x[j + 1]) + y[m + 1] * x[j]) / denom;
            }
            if (!bad) {
                c3 = x[1];
                del3 = d_dltar(omega / c3, omega, ifunc, M);
                ndl++;
                nev = 2;
                m = m + 1;
                if (m > 10) m = 10;
                continue;
            }
        }
        c3 = 0.5 * (c1 + c2);
        del3 = d_dltar(omega / c3, omega, ifunc, M);
        ndl++;
        nev = 1;
        m = 1;
    }
    return c3;
}

// getsol (surfdisp96.f:734-827) with the bracketing walk spread over the wave.
template <class MD>
__device__ static int d_getsol(double t1, double &c1io, double clow, double dc, double cm, float betmx,
                               int ifunc, int ifirst, double &del1st, const MD &M, int &ndl)
{
    const int lane = threadIdx.x & 63;
    const double twopi = 2.0 * 3.141592653589793;
    double c1 = c1io;
    double omega = twopi / t1;
    double del1 = d_dltar(omega / c1, omega, ifunc, M);
    ndl++;
    if (ifirst == 1) del1st = del1;
    double plmn = dsgn(del1st) * dsgn(del1);
    int idir = (ifirst == 1) ? 1 : (plmn >= 0.0 ? 1 : -1);
    const double cmax = (double)betmx + dc;
    for (int pass = 0; pass < 4096; pass++) {
        // replay the walk up to this lane's step: c2 sequence and c1 before it
        double cc = c1, cprev = c1, cand = c1;
        int dd = idir;
        for (int s = 0; s <= lane; s++) {
            double c2;
            for (;;) {
                c2 = dd > 0 ? cc + dc : cc - dc;
                if (c2 <= clow) { dd = 1; cc = clow; continue; }
                break;
            }
            cprev = cc;
            cand = c2;
            cc = c2;
        }
        double omg = twopi / t1;
        double del2 = d_dltar(omg / cand, omg, ifunc, M);
        double delprev = __shfl_up(del2, 1, 64);
        if (lane == 0) delprev = del1;
        bool change = dsgn(delprev) != dsgn(del2);
        bool fail = !change && (cand < cm || cand >= cmax);
        unsigned long long ev = __ballot(change || fail);
        if (ev) {
            int e = __ffsll(ev) - 1;
            ndl += e + 1;             // the sequential walk stops at step e

            bool ch = __shfl(change ? 1 : 0, e, 64) != 0;
            double ce = __shfl(cand, e, 64);
            if (!ch) { c1io = ce; return -1; }
            double cb = __shfl(cprev, e, 64);
            double d2 = __shfl(del2, e, 64);
            double d1 = __shfl(delprev, e, 64);
            double cn = d_nevill(t1, cb, ce, d1, d2, ifunc, M, twopi, ndl);
            c1io = cn;
            if (cn > (double)betmx) return -1;
            return 1;
        }
        ndl += 64;
        // no event in these 64 steps: continue from the last one
        c1 = __shfl(cand, 63, 64);
        del1 = __shfl(del2, 63, 64);
        idir = __shfl(dd, 63, 64);
    }
    c1io = c1;
    return -1;
}


// One wave per column: check_model, convert_to_layer, surfdisp96.
__global__ __launch_bounds__(64) void k_dispersion(mct_dsp_args A)
{
    __shared__ float sd[DSP_NMAX + 1], sa[DSP_NMAX + 1], sb[DSP_NMAX + 1], sr[DSP_NMAX + 1];
    __shared__ int s_nl;
    const int col = blockIdx.x, c = blockIdx.y, lane = threadIdx.x;
    const long cb = ((long)c * A.ncol + col) * A.nz;
    const double *vs = A.vs + cb;
    // check_model (src/likelihood_surf.F90:631-646): a deeper vs below the top one
    bool below = false;
    for (int k = 1 + lane; k < A.nz; k += 64) below |= vs[k] < vs[0];
    if (__ballot(below) != 0ull && lane == 0) A.bad_model[c] = 1;
    if (lane == 0) s_nl = dsp_convert<1>(A, cb, vs, sd, sa, sb, sr);
    __syncthreads();
    dsp_model_t<1> M;
    M.mmax = s_nl;
    M.d = sd; M.a = sa; M.b = sb; M.rho = sr;
    M.llw = (sb[0] <= 0.0f) ? 2 : 1;
    // surfdisp96 (flat earth, fundamental mode)
    double *cp = A.phase + ((long)c * A.ncol + col) * A.np;
    double *cg = A.group + ((long)c * A.ncol + col) * A.np;
    int ierr = 0, ndl = 0;
    const float sone = 1.500f, h = 0.005f;
    const float ddc = (float)A.dphase;
    const double one = 1.0e-2;
    int jmn = 1, jsol = 1;
    float betmx = -1.e20f, betmn = 1.e20f;
    for (int i = 1; i <= M.mmax; i++) {
        float bi = sb[i - 1], ai = sa[i - 1];
        if (bi > 0.01f && bi < betmn) { betmn = bi; jmn = i; jsol = 1; }
        else if (bi <= 0.01f && ai < betmn) { betmn = ai; jmn = i; jsol = 0; }
        if (bi > betmx) betmx = bi;
    }
    const int ifunc = A.iwave;
    const double onea = (double)sone;
    float cc1 = jsol == 0 ? betmn : d_gtsolh(sa[jmn - 1], sb[jmn - 1]);
    cc1 = 0.95f * cc1;
    cc1 = 0.90f * cc1;
    const double cc = (double)cc1;
    const double dc = fabs((double)ddc);
    const double cm = cc;
    double cvec[DSP_MAXP], cbv[DSP_MAXP], del1st = 0.0;
    for (int i = 0; i < A.np; i++) { cvec[i] = 0.0; cbv[i] = 0.0; }
    int k;
    for (k = 1; k <= A.np; k++) {
        double t1 = A.periods[k - 1];
        float t1a, t1b = 0.0f;
        if (A.igr > 0) {
            t1a = (float)(t1 / (double)(1.0f + h));
            t1b = (float)(t1 / (double)(1.0f - h));
            t1 = (double)t1a;
        } else {
            t1a = (float)t1;
        }
        int ifirst;
        double c1, clow;
        if (k == 1) {
            c1 = cc; clow = cc; ifirst = 1;
        } else {
            ifirst = 0;
            c1 = cc;
            for (int previd = k - 1; previd >= 1; previd--)
                if (cvec[previd - 1] > 0) { c1 = cvec[previd - 1] - onea * dc; break; }
            clow = cm;
        }
        int iret = d_getsol(t1, c1, clow, dc, cm, betmx, ifunc, ifirst, del1st, M, ndl);
        if (iret == -1) break;
        cvec[k - 1] = c1;
        if (A.igr > 0) {
            t1 = (double)t1b;
            ifirst = 0;
            clow = cbv[k - 1] + one * dc;
            c1 = c1 - onea * dc;
            iret = d_getsol(t1, c1, clow, dc, cm, betmx, ifunc, ifirst, del1st, M, ndl);
            if (iret == -1) { c1 = cvec[k - 1]; ierr = 1; }
            cbv[k - 1] = c1;
        } else {
            c1 = 0.0;
        }
        float cc0 = (float)cvec[k - 1];
        float cc1k = (float)c1;
        if (lane == 0) {
            if (A.igr == 0) {
                cp[k - 1] = (double)cc0;
                cg[k - 1] = 100.0;
            } else {
                float gvel = (1.0f / t1a - 1.0f / t1b) / (1.0f / (t1a * cc0) - 1.0f / (t1b * cc1k));
                cg[k - 1] = (double)gvel;
                cp[k - 1] = (double)cc0;
                if (gvel < 0 || cvec[k - 1] == 0) ierr = 1;
            }
        }
    }
    if (k <= A.np) {   // no root at period k: the rest keep cp=100, cg=0 (labels 1700/1750)
        ierr = 1;
        if (lane == 0)
            for (int i = k; i <= A.np; i++) { cp[i - 1] = 100.0; cg[i - 1] = 0.0; }
    }
    if (lane == 0) {
        A.ierr[(long)c * A.ncol + col] = ierr;
        if (A.ndltar) A.ndltar[(long)c * A.ncol + col] = ndl;
        if (A.nlayer) A.nlayer[(long)c * A.ncol + col] = M.mmax;
    }
}

// getsol (surfdisp96.f:734-827) as the reference runs it: one lane walks the
// bracket step by step.
template <class MD>
__device__ static int d_getsol_seq(double t1, double &c1io, double clow, double dc, double cm, float betmx,
                                   int ifunc, int ifirst, double &del1st, const MD &M, int &ndl)
{
    const double twopi = 2.0 * 3.141592653589793;
    double c1 = c1io;
    double omega = twopi / t1;
    double del1 = d_dltar(omega / c1, omega, ifunc, M);
    ndl++;
    if (ifirst == 1) del1st = del1;
    double plmn = dsgn(del1st) * dsgn(del1);
    int idir = (ifirst == 1) ? 1 : (plmn >= 0.0 ? 1 : -1);
    const double cmax = (double)betmx + dc;
    for (int it = 0; it < (1 << 24); it++) {
        double c2;
        for (;;) {
            c2 = idir > 0 ? c1 + dc : c1 - dc;
            if (c2 <= clow) { idir = 1; c1 = clow; continue; }
            break;
        }
        omega = twopi / t1;
        double del2 = d_dltar(omega / c2, omega, ifunc, M);
        ndl++;
        if (dsgn(del1) != dsgn(del2)) {
            double cn = d_nevill(t1, c1, c2, del1, del2, ifunc, M, twopi, ndl);
            c1io = cn;
            if (cn > (double)betmx) return -1;
            return 1;
        }
        c1 = c2;
        del1 = del2;
        if (c1 < cm || c1 >= cmax) { c1io = c1; return -1; }
    }
    c1io = c1;
    return -1;
}

// One lane per column: check_model, convert_to_layer and surfdisp96 for 64
// columns per wave.  The layer model of lane i is stored at lay + i with a
// stride of 64 floats (planes d, a, b, rho of nlmax x 64 floats per wave).
__global__ __launch_bounds__(64) void k_dispersion_lane(mct_dsp_args A, float *lay, int nlmax)
{
    const int lane = threadIdx.x;
    const long col = (long)blockIdx.x * 64 + lane;
    const int c = blockIdx.y;
    if (col >= A.ncol) return;
    const long cb = ((long)c * A.ncol + col) * A.nz;
    const double *vs = A.vs + cb;
    const long plane = (long)nlmax * 64;
    float *base = lay + ((long)c * gridDim.x + blockIdx.x) * 4 * plane + lane;
    float *sd = base, *sa = base + plane, *sb = base + 2 * plane, *sr = base + 3 * plane;
    // check_model (src/likelihood_surf.F90:631-646)
    bool below = false;
    const double top = vs[0];
    for (int k = 1; k < A.nz; k++) below |= vs[k] < top;
    if (below) A.bad_model[c] = 1;
    dsp_model_t<64> M;
    M.mmax = dsp_convert<64>(A, cb, vs, sd, sa, sb, sr);
    M.d = sd; M.a = sa; M.b = sb; M.rho = sr;
    M.llw = (sb[0] <= 0.0f) ? 2 : 1;
    double *cp = A.phase + ((long)c * A.ncol + col) * A.np;
    double *cg = A.group + ((long)c * A.ncol + col) * A.np;
    int ierr = 0, ndl = 0;
    const float sone = 1.500f, h = 0.005f;
    const float ddc = (float)A.dphase;
    const double one = 1.0e-2;
    int jmn = 1, jsol = 1;
    float betmx = -1.e20f, betmn = 1.e20f;
    for (int i = 1; i <= M.mmax; i++) {
        float bi = sb[64 * (i - 1)], ai = sa[64 * (i - 1)];
        if (bi > 0.01f && bi < betmn) { betmn = bi; jmn = i; jsol = 1; }
        else if (bi <= 0.01f && ai < betmn) { betmn = ai; jmn = i; jsol = 0; }
        if (bi > betmx) betmx = bi;
    }
    const int ifunc = A.iwave;
    const double onea = (double)sone;
    float cc1 = jsol == 0 ? betmn : d_gtsolh(sa[64 * (jmn - 1)], sb[64 * (jmn - 1)]);
    cc1 = 0.95f * cc1;
    cc1 = 0.90f * cc1;
    const double cc = (double)cc1;
    const double dc = fabs((double)ddc);
    const double cm = cc;
    // the fundamental mode only: the previous period's root seeds the next
    // (cvec of surfdisp96), the group pass starts from cb(k) = 0
    double cprev = 0.0, del1st = 0.0;
    int k;
    for (k = 1; k <= A.np; k++) {
        double t1 = A.periods[k - 1];
        float t1a, t1b = 0.0f;
        if (A.igr > 0) {
            t1a = (float)(t1 / (double)(1.0f + h));
            t1b = (float)(t1 / (double)(1.0f - h));
            t1 = (double)t1a;
        } else {
            t1a = (float)t1;
        }
        int ifirst;
        double c1, clow;
        if (k == 1) {
            c1 = cc; clow = cc; ifirst = 1;
        } else {
            ifirst = 0;
            c1 = cprev > 0 ? cprev - onea * dc : cc;
            clow = cm;
        }
        int iret = d_getsol_seq(t1, c1, clow, dc, cm, betmx, ifunc, ifirst, del1st, M, ndl);
        if (iret == -1) break;
        const double croot = c1;
        cprev = croot;
        if (A.igr > 0) {
            t1 = (double)t1b;
            ifirst = 0;
            clow = 0.0 + one * dc;
            c1 = c1 - onea * dc;
            iret = d_getsol_seq(t1, c1, clow, dc, cm, betmx, ifunc, ifirst, del1st, M, ndl);
            if (iret == -1) { c1 = croot; ierr = 1; }
        } else {
            c1 = 0.0;
        }
        float cc0 = (float)croot;
        float cc1k = (float)c1;
        if (A.igr == 0) {
            cp[k - 1] = (double)cc0;
            cg[k - 1] = 100.0;
        } else {
            float gvel = (1.0f / t1a - 1.0f / t1b) / (1.0f / (t1a * cc0) - 1.0f / (t1b * cc1k));
            cg[k - 1] = (double)gvel;
            cp[k - 1] = (double)cc0;
            if (gvel < 0 || croot == 0) ierr = 1;
        }
    }
    if (k <= A.np) {   // no root at period k: the rest keep cp=100, cg=0 (labels 1700/1750)
        ierr = 1;
        for (int i = k; i <= A.np; i++) { cp[i - 1] = 100.0; cg[i - 1] = 0.0; }
    }
    A.ierr[(long)c * A.ncol + col] = ierr;
    if (A.ndltar) A.ndltar[(long)c * A.ncol + col] = ndl;
    if (A.nlayer) A.nlayer[(long)c * A.ncol + col] = M.mmax;
}

// One lane per column, columns handed out from a queue, and surfdisp96's
// control flow as a per-lane state machine with a single period-equation
// call site: every iteration each lane advances its own column (walk step,
// Neville step, next period, next column) until it needs dltar, then the
// whole wave evaluates dltar together.  Lanes never wait for one another's
// walks, and a lane that finishes a column takes the next one.  The sequence
// of operations per column is surfdisp96's (the same as k_dispersion_lane).
enum {
    SM_NEXT, SM_PERIOD, SM_G0, SM_G1, SM_N0, SM_NLOOP, SM_N1, SM_NBODY, SM_N2, SM_N3,
    SM_NDONE, SM_GDONE, SM_FINISH, SM_DONE
};

// Column order for the queue: by layer count, most layers first, so the lanes
// of a wave hold columns whose period equations loop over as many layers
// (a counting sort; columns of one count keep no particular order, which
// changes nothing since every column is computed on its own).
#define DSP_SORT_IPT 4
// the global column (chain * ncol + column) of queue position idx: all columns
// in grid order, or the columns of each chain's box
__device__ __forceinline__ long dsp_job_column(const mct_dsp_args &A, long idx)
{
    if (!A.box_off) return idx;
    int lo = 0, hi = A.nchain - 1;           // the chain whose box holds idx
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (A.box_off[mid] <= idx) lo = mid; else hi = mid - 1;
    }
    const int32_t *b = A.box + 4 * lo;
    const long local = idx - A.box_off[lo];
    const int w = b[3] - b[2] + 1;
    const int ix = b[0] + (int)(local / w), iy = b[2] + (int)(local % w);
    return (long)lo * A.ncol + (long)(ix - 1) * A.ny + (iy - 1);
}

__device__ __forceinline__ int dsp_layer_bin_v(const double *vs, int nz)
{
    const double EPS = (double)1.0e-10f;       // src/likelihood_surf.F90:37
    int nl = 0;
    double last = vs[0];
    for (int k = 1; k < nz; k++) {
        const double v = vs[k];
        if (fabs(v - last) > EPS) { nl++; last = v; }
    }
    return nz - 1 - nl;       // mmax = nl + 1 in [1, nz]: bin 0 holds the most layers
}

__device__ __forceinline__ int dsp_layer_bin(const mct_dsp_args &A, long g)
{
    return dsp_layer_bin_v(A.vs + g * A.nz, A.nz);
}

// ---- columns with identical profiles ------------------------------------------
// A column's dispersion curves depend on nothing but its Vs profile (check_model,
// convert_to_layer and surfdisp96 read only vs(:), dz and the periods), so
// columns of one chain whose profiles are equal bit for bit get the same
// phase and group velocities, ierr and walk counts.  Each chain keeps an
// open-addressing table of profile hashes; the smallest queue position holding
// a hash is its representative, a column whose profile equals its
// representative's copies that column's results, and a hash collision (a
// profile that differs) is solved on its own.
__device__ __forceinline__ unsigned long long dsp_profile_hash(const double *v, int nz)
{
    unsigned long long h = 0x9e3779b97f4a7c15ull;
    for (int k = 0; k < nz; k++) {
        h ^= (unsigned long long)__double_as_longlong(v[k]);
        h *= 0xff51afd7ed558ccdull;
        h ^= h >> 31;
    }
    return h | 1ull;                         // 0 marks an empty slot
}

__device__ __forceinline__ bool dsp_same_profile(const double *u, const double *v, int nz)
{
    for (int k = 0; k < nz; k++)
        if (__double_as_longlong(u[k]) != __double_as_longlong(v[k])) return false;
    return true;
}

// the profile of queue position idx (global column g) enters its chain's table;
// its layer-count bin is kept in keys[idx] for the sort
__device__ __forceinline__ void dsp_hash_one(const mct_dsp_args &A, long idx, long g, const double *v)
{
    const unsigned mask = (1u << A.hlog) - 1u;
    const long t0 = (g / A.ncol) << A.hlog;
    const unsigned long long h = dsp_profile_hash(v, A.nz);
    unsigned s = (unsigned)(h >> 17) & mask;
    // at most ncol keys per chain in >= 2 * ncol slots: an empty slot is always found
    for (;;) {
        const unsigned long long prev = atomicCAS(&A.hkey[t0 + s], 0ull, h);
        if (prev == 0ull || prev == h) break;
        s = (s + 1u) & mask;
    }
    atomicMin(&A.hrep[t0 + s], (int32_t)idx);
    A.dup[idx] = (int32_t)s;
    A.keys[idx] = dsp_layer_bin_v(v, A.nz);
}

// box calls: one lane per queue position reading its column in place
__global__ __launch_bounds__(256) void k_dsp_hash_insert(mct_dsp_args A, long total)
{
    const long base = (long)blockIdx.x * 256 * DSP_SORT_IPT;
#pragma unroll
    for (int i = 0; i < DSP_SORT_IPT; i++) {
        const long idx = base + i * 256 + threadIdx.x;
        if (idx >= total) continue;
        const long g = dsp_job_column(A, idx);
        dsp_hash_one(A, idx, g, A.vs + g * A.nz);
    }
}

// whole-grid calls (queue position = global column): a wave's 64 columns are
// contiguous, so they are staged through LDS with coalesced loads
__global__ __launch_bounds__(64) void k_dsp_hash_insert_grid(mct_dsp_args A, long total)
{
    extern __shared__ double dsp_sv[];         // 64 * nz
    const long idx0 = (long)blockIdx.x * 64;
    const int n = (int)(total - idx0 < 64 ? total - idx0 : 64);
    const double *src = A.vs + idx0 * A.nz;
    for (int i = threadIdx.x; i < n * A.nz; i += 64) dsp_sv[i] = src[i];
    __syncthreads();
    if ((int)threadIdx.x >= n) return;
    const long idx = idx0 + threadIdx.x;
    dsp_hash_one(A, idx, idx, dsp_sv + threadIdx.x * A.nz);
}

// results of the solved representative copied to each duplicate column
__global__ __launch_bounds__(256) void k_dsp_dup_copy(mct_dsp_args A, long total)
{
    const long idx = (long)blockIdx.x * 256 + threadIdx.x;
    if (idx >= total) return;
    const int d = A.dup[idx];
    if (d < 0) return;
    const long g = dsp_job_column(A, idx), r = dsp_job_column(A, d);
    for (int p = 0; p < A.np; p++) {
        A.phase[g * A.np + p] = A.phase[r * A.np + p];
        A.group[g * A.np + p] = A.group[r * A.np + p];
    }
    A.ierr[g] = A.ierr[r];
    if (A.ndltar) A.ndltar[g] = A.ndltar[r];
    if (A.nlayer) A.nlayer[g] = A.nlayer[r];
}

__global__ __launch_bounds__(256) void k_dsp_sort_count(mct_dsp_args A, long total)
{
    __shared__ int h[128];
    const int nb = A.nz;
    for (int b = threadIdx.x; b < nb; b += 256) h[b] = 0;
    __syncthreads();
    const long base = (long)blockIdx.x * 256 * DSP_SORT_IPT;
#pragma unroll
    for (int i = 0; i < DSP_SORT_IPT; i++) {
        const long idx = base + i * 256 + threadIdx.x;
        if (idx < total) {
            const long g = dsp_job_column(A, idx);
            int b = -1;
            if (A.dup) {
                // the slot k_dsp_hash_insert left here becomes the copy source or -1
                const long t0 = (g / A.ncol) << A.hlog;
                const int rep = A.hrep[t0 + A.dup[idx]];
                int d = -1;
                if (rep != idx && dsp_same_profile(A.vs + g * A.nz, A.vs + dsp_job_column(A, rep) * A.nz, A.nz))
                    d = rep;
                A.dup[idx] = d;
                if (d < 0) b = A.keys[idx];         // the bin k_dsp_hash_insert kept
            } else {
                b = dsp_layer_bin(A, g);
            }
            A.keys[idx] = b;
            if (b >= 0) atomicAdd(&h[b], 1);
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nb; b += 256)
        if (h[b]) atomicAdd(&A.hist[b], h[b]);
}

// exclusive prefix sum of the bin counts: the bins' first positions
__global__ void k_dsp_sort_scan(int32_t *hist, int nb, unsigned long long *queue, int32_t *njobs)
{
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int b = 0; b < nb; b++) { const int c = hist[b]; hist[b] = acc; acc += c; }
        queue[0] = 0ull;
        if (njobs) njobs[0] = acc;
    }
}

__global__ __launch_bounds__(256) void k_dsp_sort_place(mct_dsp_args A, long total)
{
    __shared__ int h[128], base[128];
    const int nb = A.nz;
    for (int b = threadIdx.x; b < nb; b += 256) h[b] = 0;
    __syncthreads();
    const long b0 = (long)blockIdx.x * 256 * DSP_SORT_IPT;
    int bin[DSP_SORT_IPT], rank[DSP_SORT_IPT];
#pragma unroll
    for (int i = 0; i < DSP_SORT_IPT; i++) {
        const long idx = b0 + i * 256 + threadIdx.x;
        bin[i] = idx < total ? A.keys[idx] : -1;
        rank[i] = bin[i] >= 0 ? atomicAdd(&h[bin[i]], 1) : 0;
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nb; b += 256) base[b] = h[b] ? atomicAdd(&A.hist[b], h[b]) : 0;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < DSP_SORT_IPT; i++)
        if (bin[i] >= 0) A.jobs[base[bin[i]] + rank[i]] = (int32_t)dsp_job_column(A, b0 + i * 256 + threadIdx.x);
}

__global__ void k_dsp_sort_reset(int32_t *hist, int nb)
{
    for (int b = threadIdx.x; b < nb; b += blockDim.x) hist[b] = 0;
}

__global__ void k_dsp_queue_reset(unsigned long long *q)
{
    if (threadIdx.x == 0) q[0] = 0ull;
}

__global__ __launch_bounds__(64) void k_dispersion_sm(mct_dsp_args A, float *lay, int nlmax, long total)
{
    const int lane = threadIdx.x;
    const long ntot = A.njobs ? (long)A.njobs[0] : total;   // duplicates are not queued
    const long plane = (long)nlmax * 64;
    float *base = lay + (long)blockIdx.x * 4 * plane + lane;
    float *sd = base, *sa = base + plane, *sb = base + 2 * plane, *sr = base + 3 * plane;
    const double twopi = 2.0 * 3.141592653589793;
    const float sone = 1.500f, h = 0.005f;
    const float ddc = (float)A.dphase;
    const double one = 1.0e-2, onea = (double)sone;
    const double dc = fabs((double)ddc);
    const int ifunc = A.iwave;

    dsp_model_t<64> M;
    M.d = sd; M.a = sa; M.b = sb; M.rho = sr;
    M.mmax = 1; M.llw = 1;
    // column state
    long job = -1;
    int c = 0;
    long col = 0;
    double *cp = nullptr, *cg = nullptr;
    int ierr = 0, ndl = 0, k = 1, pass = 0, ifirst = 0;
    float betmx = 0.0f;
    double cc = 0.0, cm = 0.0, cmax = 0.0, cprev = 0.0, del1st = 0.0, croot = 0.0;
    float t1a = 0.0f, t1b = 0.0f;
    double t1 = 0.0, omega = 0.0;
    // getsol / nevill state
    double c1 = 0.0, c2 = 0.0, del1 = 0.0, del2 = 0.0, clow = 0.0;
    int idir = 1, gret = 0, walk = 0;
    double gc = 0.0;
    double nc1 = 0.0, nc2 = 0.0, nd1 = 0.0, nd2 = 0.0, c3 = 0.0, del3 = 0.0;
    int nev = 0, nctrl = 0, m = 1;
    // Neville tables in LDS, lane-interleaved
    __shared__ double xs[12 * 64], ys[12 * 64];
#define X(j) xs[(j) * 64 + lane]
#define Y(j) ys[(j) * 64 + lane]
    double res = 0.0, ceval = 0.0;
    int pc = SM_NEXT;

    for (;;) {
        bool need = false;
        while (!need && pc != SM_DONE) {
            switch (pc) {
            case SM_NEXT: {
                job = (long)atomicAdd((unsigned long long *)A.queue, 1ull);
                if (job >= ntot) { pc = SM_DONE; break; }
                if (A.jobs) job = A.jobs[job];
                c = (int)(job / A.ncol);
                col = job - (long)c * A.ncol;
                const long cb = ((long)c * A.ncol + col) * A.nz;
                const double *vs = A.vs + cb;
                // check_model (src/likelihood_surf.F90:631-646)
                bool below = false;
                const double top = vs[0];
                for (int kk = 1; kk < A.nz; kk++) below |= vs[kk] < top;
                if (below) A.bad_model[c] = 1;
                M.mmax = dsp_convert<64>(A, cb, vs, sd, sa, sb, sr);
                M.llw = (sb[0] <= 0.0f) ? 2 : 1;
                cp = A.phase + ((long)c * A.ncol + col) * A.np;
                cg = A.group + ((long)c * A.ncol + col) * A.np;
                ierr = 0; ndl = 0;
                int jmn = 1, jsol = 1;
                betmx = -1.e20f;
                float betmn = 1.e20f;
                for (int ii = 1; ii <= M.mmax; ii++) {
                    float bi = sb[64 * (ii - 1)], ai = sa[64 * (ii - 1)];
                    if (bi > 0.01f && bi < betmn) { betmn = bi; jmn = ii; jsol = 1; }
                    else if (bi <= 0.01f && ai < betmn) { betmn = ai; jmn = ii; jsol = 0; }
                    if (bi > betmx) betmx = bi;
                }
                float cc1 = jsol == 0 ? betmn : d_gtsolh(sa[64 * (jmn - 1)], sb[64 * (jmn - 1)]);
                cc1 = 0.95f * cc1;
                cc1 = 0.90f * cc1;
                cc = (double)cc1;
                cm = cc;
                cmax = (double)betmx + dc;
                cprev = 0.0; del1st = 0.0;
                k = 1; pass = 0;
                pc = SM_PERIOD;
                break;
            }
            case SM_PERIOD: {
                t1 = A.periods[k - 1];
                if (A.igr > 0) {
                    t1a = (float)(t1 / (double)(1.0f + h));
                    t1b = (float)(t1 / (double)(1.0f - h));
                    t1 = (double)t1a;
                } else {
                    t1a = (float)t1;
                }
                if (k == 1) { c1 = cc; clow = cc; ifirst = 1; }
                else { ifirst = 0; c1 = cprev > 0 ? cprev - onea * dc : cc; clow = cm; }
                pass = 0;
                omega = twopi / t1;          // getsol: wvno = omega / c1
                ceval = c1; need = true; pc = SM_G0;
                break;
            }
            case SM_G0: {
                del1 = res; ndl++;
                if (ifirst == 1) del1st = del1;
                double plmn = dsgn(del1st) * dsgn(del1);
                idir = (ifirst == 1) ? 1 : (plmn >= 0.0 ? 1 : -1);
                for (;;) {
                    c2 = idir > 0 ? c1 + dc : c1 - dc;
                    if (c2 <= clow) { idir = 1; c1 = clow; continue; }
                    break;
                }
                walk = 0;
                omega = twopi / t1;
                ceval = c2; need = true; pc = SM_G1;
                break;
            }
            case SM_G1: {
                del2 = res; ndl++;
                if (dsgn(del1) != dsgn(del2)) {
                    nc1 = c1; nc2 = c2; nd1 = del1; nd2 = del2;
                    omega = twopi / t1;      // nevill (surfdisp96.f:903-1020)
                    c3 = 0.5 * (nc1 + nc2);
                    ceval = c3; need = true; pc = SM_N0;
                    break;
                }
                c1 = c2;
                del1 = del2;
                if (c1 < cm || c1 >= cmax || ++walk >= (1 << 24)) { gret = -1; gc = c1; pc = SM_GDONE; break; }
                for (;;) {
                    c2 = idir > 0 ? c1 + dc : c1 - dc;
                    if (c2 <= clow) { idir = 1; c1 = clow; continue; }
                    break;
                }
                omega = twopi / t1;
                ceval = c2; need = true; pc = SM_G1;
                break;
            }
            case SM_N0:
                del3 = res; ndl++;
                nev = 1; nctrl = 1; m = 1;
                pc = SM_NLOOP;
                break;
            case SM_NLOOP:
                nctrl = nctrl + 1;
                if (nctrl >= 100) { pc = SM_NDONE; break; }
                if (c3 < fmin(nc1, nc2) || c3 > fmax(nc1, nc2)) {
                    nev = 0;
                    c3 = 0.5 * (nc1 + nc2);
                    ceval = c3; need = true; pc = SM_N1;
                    break;
                }
                pc = SM_NBODY;
                break;
            case SM_N1:
                del3 = res; ndl++;
                pc = SM_NBODY;
                break;
            case SM_NBODY: {
                double s13 = nd1 - del3;
                double s32 = del3 - nd2;
                if (dsgn(del3) * dsgn(nd1) < 0.0) { nc2 = c3; nd2 = del3; }
                else { nc1 = c3; nd1 = del3; }
                if (fabs(nc1 - nc2) <= 1.0e-6 * nc1) { pc = SM_NDONE; break; }
                if (dsgn(s13) != dsgn(s32)) nev = 0;
                double ss1 = fabs(nd1);
                double s1 = (double)0.01f * ss1;
                double ss2 = fabs(nd2);
                double s2 = (double)0.01f * ss2;
                bool halve = s1 > ss2 || s2 > ss1 || nev == 0;
                if (!halve) {
                    if (nev == 2) {
                        X(m + 1) = c3;
                        Y(m + 1) = del3;
                    } else {
                        X(1) = nc1; Y(1) = nd1;
                        X(2) = nc2; Y(2) = nd2;
                        m = 1;
                    }
                    bool bad = false;
                    for (int kk = 1; kk <= m; kk++) {
                        int j = m - kk + 1;
                        double denom = Y(m + 1) - Y(j);
                        if (fabs(denom) < 1.0e-10 * fabs(Y(m + 1))) { bad = true; break; }
                        X(j) = (-(Y(j) * X(j + 1)) + Y(m + 1) * X(j)) / denom;
                    }
                    if (!bad) {
                        c3 = X(1);
                        ceval = c3; need = true; pc = SM_N2;
                        break;
                    }
                }
                c3 = 0.5 * (nc1 + nc2);
                ceval = c3; need = true; pc = SM_N3;
                break;
            }
            case SM_N2:
                del3 = res; ndl++;
                nev = 2;
                m = m + 1;
                if (m > 10) m = 10;
                pc = SM_NLOOP;
                break;
            case SM_N3:
                del3 = res; ndl++;
                nev = 1;
                m = 1;
                pc = SM_NLOOP;
                break;
            case SM_NDONE:
                gc = c3;
                gret = (c3 > (double)betmx) ? -1 : 1;
                pc = SM_GDONE;
                break;
            case SM_GDONE:
                if (pass == 0) {
                    if (gret == -1) {
                        // no root at period k: the rest keep cp=100, cg=0 (labels 1700/1750)
                        ierr = 1;
                        for (int ii = k; ii <= A.np; ii++) { cp[ii - 1] = 100.0; cg[ii - 1] = 0.0; }
                        pc = SM_FINISH;
                        break;
                    }
                    croot = gc;
                    cprev = croot;
                    if (A.igr > 0) {
                        pass = 1;
                        t1 = (double)t1b;
                        ifirst = 0;
                        clow = 0.0 + one * dc;
                        c1 = croot - onea * dc;
                        omega = twopi / t1;
                        ceval = c1; need = true; pc = SM_G0;
                        break;
                    }
                    cp[k - 1] = (double)(float)croot;
                    cg[k - 1] = 100.0;
                } else {
                    double cgr = gret == -1 ? croot : gc;
                    if (gret == -1) ierr = 1;
                    float cc0 = (float)croot, cc1k = (float)cgr;
                    float gvel = (1.0f / t1a - 1.0f / t1b) / (1.0f / (t1a * cc0) - 1.0f / (t1b * cc1k));
                    cg[k - 1] = (double)gvel;
                    cp[k - 1] = (double)cc0;
                    if (gvel < 0 || croot == 0) ierr = 1;
                    pass = 0;
                }
                k++;
                pc = k > A.np ? SM_FINISH : SM_PERIOD;
                break;
            case SM_FINISH:
                A.ierr[(long)c * A.ncol + col] = ierr;
                if (A.ndltar) A.ndltar[(long)c * A.ncol + col] = ndl;
                if (A.nlayer) A.nlayer[(long)c * A.ncol + col] = M.mmax;
                pc = SM_NEXT;
                break;
            default:
                pc = SM_DONE;
                break;
            }
        }
        if (__ballot(need) == 0ull) break;     // every lane has run out of columns
        if (need) res = d_dltar(omega / ceval, omega, ifunc, M);
    }
#undef X
#undef Y
}

int mct_launch_dispersion(const mct_dsp_args &a, hipStream_t s)
{
    if (a.np < 1 || a.np > DSP_MAXP || a.nz < 1 || a.nz > DSP_NMAX) return MCT_ERR_ARG;
    {
        int rc0 = mct_fill_i32(a.bad_model, a.nchain, 0, s);
        if (rc0) return rc0;
    }
    if (a.lay && a.queue) {
        // persistent waves, columns from a queue (ordered by layer count when
        // the order's buffers are given)
        mct_dsp_args b = a;
        const long total = a.box_off ? a.box_total : (long)a.ncol * a.nchain;
        if (a.box_off && !(b.jobs && b.keys && b.hist)) return MCT_ERR_ARG;
        if (total < 1) return MCT_OK;
        if (b.jobs && b.keys && b.hist && b.nz <= 128 && (long)a.ncol * a.nchain < 0x7fffffffL) {
            const unsigned nblk = (unsigned)((total + 256 * DSP_SORT_IPT - 1) / (256 * DSP_SORT_IPT));
            const bool dedup = b.hkey && b.hrep && b.dup && b.njobs && b.hlog > 0 && b.hlog < 31 &&
                               (2L * b.ncol) <= (1L << b.hlog);
            if (dedup) {
                const long slots = (long)b.nchain << b.hlog;
                int rc = mct_fill_i32((int32_t *)b.hkey, 2 * slots, 0, s);
                if (!rc) rc = mct_fill_i32(b.hrep, slots, 0x7fffffff, s);
                if (rc) return rc;
                if (b.box_off)
                    hipLaunchKernelGGL(k_dsp_hash_insert, dim3(nblk), dim3(256), 0, s, b, total);
                else
                    hipLaunchKernelGGL(k_dsp_hash_insert_grid, dim3((unsigned)((total + 63) / 64)), dim3(64),
                                       (size_t)64 * b.nz * sizeof(double), s, b, total);
            } else {
                b.dup = nullptr; b.njobs = nullptr;
            }
            hipLaunchKernelGGL(k_dsp_sort_reset, dim3(1), dim3(128), 0, s, b.hist, b.nz);
            hipLaunchKernelGGL(k_dsp_sort_count, dim3(nblk), dim3(256), 0, s, b, total);
            hipLaunchKernelGGL(k_dsp_sort_scan, dim3(1), dim3(64), 0, s, b.hist, b.nz, b.queue, b.njobs);
            hipLaunchKernelGGL(k_dsp_sort_place, dim3(nblk), dim3(256), 0, s, b, total);
            MCT_CHECK(hipGetLastError());
        } else {
            b.jobs = nullptr; b.dup = nullptr; b.njobs = nullptr;
            hipLaunchKernelGGL(k_dsp_queue_reset, dim3(1), dim3(64), 0, s, b.queue);
        }
        long nw = (total + 63) / 64;
        if (nw > b.lay_waves) nw = b.lay_waves;
        hipLaunchKernelGGL(k_dispersion_sm, dim3((unsigned)nw), dim3(64), 0, s, b, b.lay, b.nz + 1, total);
        if (b.dup) {
            MCT_CHECK(hipGetLastError());
            hipLaunchKernelGGL(k_dsp_dup_copy, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, b, total);
        }
    } else if (a.lay) {
        const unsigned nb = (unsigned)((a.ncol + 63) / 64);
        hipLaunchKernelGGL(k_dispersion_lane, dim3(nb, a.nchain), dim3(64), 0, s, a, a.lay, a.nz + 1);
    } else {
        hipLaunchKernelGGL(k_dispersion, dim3(a.ncol, a.nchain), dim3(64), 0, s, a);
    }
    MCT_CHECK(hipGetLastError());
    return MCT_OK;
}

// a column has at most nz + 1 layers (nz - 1 interfaces, the half-space and
// a water layer)
long mct_dispersion_lay_floats(int ncol, int nz, int nchain)
{
    return (long)nchain * ((ncol + 63) / 64) * 4L * (nz + 1) * 64;
}

// ---- C ABI: dispersion of explicit layer models (surfmodes' surfdisp96 call) ----
struct mct_layer_args {
    int nmod, nl_max, np;
    const int32_t *nlayer;        // (nmod)
    const float *layers;          // (nmod, 4, nl_max): thick, vp, vs, rho (REAL*4)
    const double *periods;        // (np)
    int iwave, igr;
    double dphase;
    double *phase, *group;        // (nmod, np)
    int32_t *ierr;                // (nmod)
};

__global__ __launch_bounds__(64) void k_dispersion_layers(mct_layer_args A)
{
    __shared__ float sd[DSP_NMAX * 2], sa[DSP_NMAX * 2], sb[DSP_NMAX * 2], sr[DSP_NMAX * 2];
    const int mod = blockIdx.x, lane = threadIdx.x;
    const int nl = A.nlayer[mod];
    const float *L = A.layers + (long)mod * 4 * A.nl_max;
    for (int i = lane; i < nl; i += 64) {
        sd[i] = L[i]; sa[i] = L[A.nl_max + i]; sb[i] = L[2 * A.nl_max + i]; sr[i] = L[3 * A.nl_max + i];
    }
    __syncthreads();
    dsp_model_t<1> M;
    M.mmax = nl;
    M.d = sd; M.a = sa; M.b = sb; M.rho = sr;
    M.llw = (sb[0] <= 0.0f) ? 2 : 1;
    double *cp = A.phase + (long)mod * A.np, *cg = A.group + (long)mod * A.np;
    int ierr = 0, ndl = 0;
    const float sone = 1.500f, h = 0.005f, ddc = (float)A.dphase;
    const double one = 1.0e-2, onea = (double)sone;
    int jmn = 1, jsol = 1;
    float betmx = -1.e20f, betmn = 1.e20f;
    for (int i = 1; i <= nl; i++) {
        float bi = sb[i - 1], ai = sa[i - 1];
        if (bi > 0.01f && bi < betmn) { betmn = bi; jmn = i; jsol = 1; }
        else if (bi <= 0.01f && ai < betmn) { betmn = ai; jmn = i; jsol = 0; }
        if (bi > betmx) betmx = bi;
    }
    float cc1 = jsol == 0 ? betmn : d_gtsolh(sa[jmn - 1], sb[jmn - 1]);
    cc1 = 0.95f * cc1;
    cc1 = 0.90f * cc1;
    const double cc = (double)cc1, dc = fabs((double)ddc), cm = cc;
    double cvec[DSP_MAXP], cbv[DSP_MAXP], del1st = 0.0;
    for (int i = 0; i < A.np; i++) { cvec[i] = 0.0; cbv[i] = 0.0; }
    if (lane == 0) for (int i = 0; i < A.np; i++) { cp[i] = 100.0; cg[i] = 100.0; }
    int k;
    for (k = 1; k <= A.np; k++) {
        double t1 = A.periods[k - 1];
        float t1a, t1b = 0.0f;
        if (A.igr > 0) {
            t1a = (float)(t1 / (double)(1.0f + h));
            t1b = (float)(t1 / (double)(1.0f - h));
            t1 = (double)t1a;
        } else {
            t1a = (float)t1;
        }
        int ifirst;
        double c1, clow;
        if (k == 1) { c1 = cc; clow = cc; ifirst = 1; }
        else {
            ifirst = 0;
            c1 = cc;
            for (int previd = k - 1; previd >= 1; previd--)
                if (cvec[previd - 1] > 0) { c1 = cvec[previd - 1] - onea * dc; break; }
            clow = cm;
        }
        int iret = d_getsol(t1, c1, clow, dc, cm, betmx, A.iwave, ifirst, del1st, M, ndl);
        if (iret == -1) break;
        cvec[k - 1] = c1;
        if (A.igr > 0) {
            t1 = (double)t1b;
            clow = cbv[k - 1] + one * dc;
            c1 = c1 - onea * dc;
            iret = d_getsol(t1, c1, clow, dc, cm, betmx, A.iwave, 0, del1st, M, ndl);
            if (iret == -1) { c1 = cvec[k - 1]; ierr = 1; }
            cbv[k - 1] = c1;
        } else {
            c1 = 0.0;
        }
        float cc0 = (float)cvec[k - 1], cc1k = (float)c1;
        if (lane == 0) {
            if (A.igr == 0) cp[k - 1] = (double)cc0;
            else {
                float gvel = (1.0f / t1a - 1.0f / t1b) / (1.0f / (t1a * cc0) - 1.0f / (t1b * cc1k));
                cg[k - 1] = (double)gvel;
                cp[k - 1] = (double)cc0;
                if (gvel < 0 || cvec[k - 1] == 0) ierr = 1;
            }
        }
    }
    if (k <= A.np) {
        ierr = 1;
        if (lane == 0) for (int i = k; i <= A.np; i++) cg[i - 1] = 0.0;
    }
    if (lane == 0) A.ierr[mod] = ierr;
}

extern "C" int mct_surfdisp96_batch(int nmod, int nl_max, const int32_t *nlayer, const float *layers,
                                    int np, const double *periods, int iwave, int igr, double dphase,
                                    double *phase, double *group, int32_t *ierr)
{
    if (nmod < 1 || nl_max < 1 || nl_max > 2 * DSP_NMAX || np < 1 || np > DSP_MAXP || !nlayer ||
        !layers || !periods || !phase || !group || !ierr || (iwave != 1 && iwave != 2))
        return MCT_ERR_ARG;
    for (int i = 0; i < nmod; i++)
        if (nlayer[i] < 1 || nlayer[i] > nl_max) return MCT_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MCT_ERR_HIP;
    void *dn = nullptr, *dl = nullptr, *dp = nullptr, *dph = nullptr, *dg = nullptr, *de = nullptr;
    int rc = MCT_OK;
    size_t sl = sizeof(float) * 4 * (size_t)nl_max * nmod;
    if (hipMalloc(&dn, sizeof(int32_t) * nmod) != hipSuccess || hipMalloc(&dl, sl) != hipSuccess ||
        hipMalloc(&dp, sizeof(double) * np) != hipSuccess ||
        hipMalloc(&dph, sizeof(double) * np * (size_t)nmod) != hipSuccess ||
        hipMalloc(&dg, sizeof(double) * np * (size_t)nmod) != hipSuccess ||
        hipMalloc(&de, sizeof(int32_t) * nmod) != hipSuccess)
        rc = MCT_ERR_NOMEM;
    if (!rc && (hipMemcpy(dn, nlayer, sizeof(int32_t) * nmod, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(dl, layers, sl, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(dp, periods, sizeof(double) * np, hipMemcpyHostToDevice) != hipSuccess))
        rc = MCT_ERR_HIP;
    if (!rc) {
        mct_layer_args a;
        a.nmod = nmod; a.nl_max = nl_max; a.np = np; a.nlayer = (const int32_t *)dn;
        a.layers = (const float *)dl; a.periods = (const double *)dp; a.iwave = iwave; a.igr = igr;
        a.dphase = dphase; a.phase = (double *)dph; a.group = (double *)dg; a.ierr = (int32_t *)de;
        hipLaunchKernelGGL(k_dispersion_layers, dim3(nmod), dim3(64), 0, 0, a);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = MCT_ERR_HIP;
    }
    if (!rc && (hipMemcpy(phase, dph, sizeof(double) * np * (size_t)nmod, hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(group, dg, sizeof(double) * np * (size_t)nmod, hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(ierr, de, sizeof(int32_t) * nmod, hipMemcpyDeviceToHost) != hipSuccess))
        rc = MCT_ERR_HIP;
    void *bufs[] = {dn, dl, dp, dph, dg, de};
    for (void *b : bufs) if (b) (void)hipFree(b);
    return rc;
}

// ---- C ABI: the likelihood's dispersion stage on explicit Vs columns -------------
extern "C" int mct_dispersion_columns(int ncol, int nz, double dz, double scaling, double water,
                                      const double *vs, int np, const double *periods, int iwave, int igr,
                                      double dphase, int mode, double *phase, double *group, int32_t *ierr,
                                      int32_t *bad_model)
{
    if (ncol < 1 || nz < 1 || nz > DSP_NMAX || np < 1 || np > DSP_MAXP || !vs || !periods || !phase ||
        !group || !ierr || !bad_model || (iwave != 1 && iwave != 2) || mode < 1 || mode > 3 ||
        !(dz > 0) || !(scaling > 0) || !(water >= 0.0))
        return MCT_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MCT_ERR_HIP;
    int hlog = 1;
    while ((1L << hlog) < 2L * ncol) hlog++;
    const long lay = mct_dispersion_lay_floats(ncol, nz, 1);
    const size_t sizes[] = {sizeof(double) * (size_t)ncol * nz, sizeof(double) * np,
                            sizeof(double) * (size_t)ncol * np, sizeof(double) * (size_t)ncol * np,
                            sizeof(int32_t) * (size_t)ncol, sizeof(int32_t) * 4, sizeof(float) * (size_t)lay,
                            256, sizeof(int32_t) * (size_t)ncol, sizeof(int32_t) * (size_t)ncol,
                            sizeof(int32_t) * 128, 8 * ((size_t)1 << hlog), 4 * ((size_t)1 << hlog),
                            sizeof(int32_t) * (size_t)ncol, 4 * 4};
    const int nb = (int)(sizeof sizes / sizeof sizes[0]);
    void *b[16] = {nullptr};
    int rc = MCT_OK;
    for (int i = 0; i < nb && !rc; i++)
        if (hipMalloc(&b[i], sizes[i]) != hipSuccess) rc = MCT_ERR_NOMEM;
    if (!rc && (hipMemcpy(b[0], vs, sizes[0], hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(b[1], periods, sizes[1], hipMemcpyHostToDevice) != hipSuccess ||
                hipMemset(b[7], 0, 256) != hipSuccess))
        rc = MCT_ERR_HIP;
    if (!rc) {
        mct_dsp_args a;
        memset(&a, 0, sizeof a);
        a.ncol = ncol; a.nz = nz; a.nchain = 1; a.vs = (const double *)b[0];
        a.dz = dz; a.scaling = scaling; a.water = water; a.periods = (const double *)b[1]; a.np = np;
        a.iwave = iwave; a.igr = igr; a.dphase = dphase;
        a.phase = (double *)b[2]; a.group = (double *)b[3]; a.ierr = (int32_t *)b[4];
        a.bad_model = (int32_t *)b[5];
        if (mode == 1 || mode == 3) a.lay = (float *)b[6];
        if (mode == 3) {
            a.queue = (unsigned long long *)b[7];
            a.lay_waves = (ncol + 63) / 64;
            a.jobs = (int32_t *)b[8]; a.keys = (int32_t *)b[9]; a.hist = (int32_t *)b[10];
            a.hlog = hlog; a.hkey = (unsigned long long *)b[11]; a.hrep = (int32_t *)b[12];
            a.dup = (int32_t *)b[13]; a.njobs = (int32_t *)b[14];
        }
        rc = mct_launch_dispersion(a, 0);
        if (!rc && hipDeviceSynchronize() != hipSuccess) rc = MCT_ERR_HIP;
    }
    if (!rc && (hipMemcpy(phase, b[2], sizes[2], hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(group, b[3], sizes[3], hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(ierr, b[4], sizes[4], hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(bad_model, b[5], sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess))
        rc = MCT_ERR_HIP;
    for (int i = 0; i < nb; i++)
        if (b[i]) (void)hipFree(b[i]);
    return rc;
}
