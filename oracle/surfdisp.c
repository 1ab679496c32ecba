/*
 * surfdisp.c -- oracle for fundamental-mode surface-wave dispersion.
 * TEST INFRASTRUCTURE ONLY (see mct_oracle.h).
 *
 * Restates surfdisp96 (surfmodes/surfdisp96.f:52-382) for a flat earth
 * (iflsph=0) and the fundamental mode (mode=1), with its helpers gtsolh
 * (:711-732), getsol (:734-827), nevill (:903-1020), half (:1022-1032),
 * dltar (:1036-1052), dltar1 (:1056-1115), dltar4 (:1119-1217), var
 * (:1220-1337), normc (:1341-1366) and dnka (:1370-1414).
 *
 * The Fortran source relies on implicit typing: names starting with a-h and
 * o-z are REAL*4 in surfdisp96 and gtsolh (betmx, betmn, cc1, cc0, t1a, t1b,
 * gvel, ...), while getsol, nevill, dltar* and var are double precision.
 * The float variables below are exactly those; every literal keeps the kind it
 * has in the Fortran (e.g. 0.01 in nevill is a default-real constant).
 * Compiled with -ffp-contract=off so no FMA contraction changes a rounding.
 */
#include "mct_oracle.h"
#include <math.h>
#include <string.h>

typedef struct {
    int mmax, llw;
    const float *d, *a, *b, *rho;
} sd_model;

static double sgn(double x) { return copysign(1.0, x); }   /* dsign(1.0d0, x) */

/* gtsolh: 5 Newton steps on the half-space Rayleigh equation, REAL*4 */
static float gtsolh(float a, float b)
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

/* SH period equation (Haskell-Thomson from the half-space up) */
static double dltar1(double wvno, double omega, const sd_model *M)
{
    double beta1 = (double)M->b[M->mmax - 1];
    double rho1 = (double)M->rho[M->mmax - 1];
    double xkb = omega / beta1;
    double wvnop = wvno + xkb;
    double wvnom = fabs(wvno - xkb);
    double rb = sqrt(wvnop * wvnom);
    double e1 = rho1 * rb;
    double e2 = 1.0 / (beta1 * beta1);
    for (int m = M->mmax - 1; m >= M->llw; m--) {      /* Fortran m = mmax-1 .. llw */
        beta1 = (double)M->b[m - 1];
        rho1 = (double)M->rho[m - 1];
        double xmu = rho1 * beta1 * beta1;
        xkb = omega / beta1;
        wvnop = wvno + xkb;
        wvnom = fabs(wvno - xkb);
        rb = sqrt(wvnop * wvnom);
        double q = (double)M->d[m - 1] * rb;
        double sinq, y, z, cosq;
        if (wvno < xkb) {
            sinq = sin(q);
            y = sinq / rb;
            z = -(rb * sinq);
            cosq = cos(q);
        } else if (wvno == xkb) {
            cosq = 1.0;
            y = (double)M->d[m - 1];
            z = 0.0;
        } else {
            double fac = 0.0;
            if (q < 16) fac = exp(-2.0 * q);
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

typedef struct { double w, cosp, a0, cpcq, cpy, cpz, cqw, cqx, xy, xz, wy, wz; } sd_var;

/* var: cosP, cosQ, sinP, sinQ and the products for the compound matrix */
static void var(double p, double q, double ra, double rb, double wvno, double xka, double xkb,
                double dpth, sd_var *o)
{
    double pex = 0.0, sex = 0.0, sinp, w = 0, x = 0, cosp = 0, sinq, y = 0, z = 0, cosq = 0;
    if (wvno < xka) {
        sinp = sin(p);
        w = sinp / ra;
        x = -(ra * sinp);
        cosp = cos(p);
    } else if (wvno == xka) {
        cosp = 1.0;
        w = dpth;
        x = 0.0;
    } else if (wvno > xka) {
        pex = p;
        double fac = 0.0;
        if (p < 16) fac = exp(-2.0 * p);
        cosp = (1.0 + fac) * 0.5;
        sinp = (1.0 - fac) * 0.5;
        w = sinp / ra;
        x = ra * sinp;
    }
    if (wvno < xkb) {
        sinq = sin(q);
        y = sinq / rb;
        z = -(rb * sinq);
        cosq = cos(q);
    } else if (wvno == xkb) {
        cosq = 1.0;
        y = dpth;
        z = 0.0;
    } else if (wvno > xkb) {
        sex = q;
        double fac = 0.0;
        if (q < 16) fac = exp(-2.0 * q);
        cosq = (1.0 + fac) * 0.5;
        sinq = (1.0 - fac) * 0.5;
        y = sinq / rb;
        z = rb * sinq;
    }
    double exa = pex + sex;
    o->a0 = 0.0;
    if (exa < 60.0) o->a0 = exp(-exa);
    o->cpcq = cosp * cosq;
    o->cpy = cosp * y;
    o->cpz = cosp * z;
    o->cqw = cosq * w;
    o->cqx = cosq * x;
    o->xy = x * y;
    o->xz = x * z;
    o->wy = w * y;
    o->wz = w * z;
    o->w = w;
    o->cosp = cosp;
}

/* dnka: Dunkin's compound matrix; ca[i][j] is the Fortran ca(i+1,j+1) */
static void dnka(double ca[5][5], double wvno2, double gam, double gammk, double rho, const sd_var *v)
{
    const double one = 1.0, two = 2.0;
    double gamm1 = gam - one;
    double twgm1 = gam + gamm1;
    double gmgmk = gam * gammk;
    double gmgm1 = gam * gamm1;
    double gm1sq = gamm1 * gamm1;
    double rho2 = rho * rho;
    double a0pq = v->a0 - v->cpcq;
    ca[0][0] = v->cpcq - two * gmgm1 * a0pq - gmgmk * v->xz - wvno2 * gm1sq * v->wy;
    ca[0][1] = (wvno2 * v->cpy - v->cqx) / rho;
    ca[0][2] = -((twgm1 * a0pq + gammk * v->xz + wvno2 * gamm1 * v->wy) / rho);
    ca[0][3] = (v->cpz - wvno2 * v->cqw) / rho;
    ca[0][4] = -((two * wvno2 * a0pq + v->xz + wvno2 * wvno2 * v->wy) / rho2);
    ca[1][0] = (gmgmk * v->cpz - gm1sq * v->cqw) * rho;
    ca[1][1] = v->cpcq;
    ca[1][2] = gammk * v->cpz - gamm1 * v->cqw;
    ca[1][3] = -v->wz;
    ca[1][4] = ca[0][3];
    ca[3][0] = (gm1sq * v->cpy - gmgmk * v->cqx) * rho;
    ca[3][1] = -v->xy;
    ca[3][2] = gamm1 * v->cpy - gammk * v->cqx;
    ca[3][3] = ca[1][1];
    ca[3][4] = ca[0][1];
    ca[4][0] = -((two * gmgmk * gm1sq * a0pq + gmgmk * gmgmk * v->xz + gm1sq * gm1sq * v->wy) * rho2);
    ca[4][1] = ca[3][0];
    ca[4][2] = -((gammk * gamm1 * twgm1 * a0pq + gam * gammk * gammk * v->xz + gamm1 * gm1sq * v->wy) * rho);
    ca[4][3] = ca[1][0];
    ca[4][4] = ca[0][0];
    double t = -(two * wvno2);
    ca[2][0] = t * ca[4][2];
    ca[2][1] = t * ca[3][2];
    ca[2][2] = v->a0 + two * (v->cpcq - ca[0][0]);
    ca[2][3] = t * ca[1][2];
    ca[2][4] = t * ca[0][2];
}

static void normc(double ee[5])
{
    double t1 = 0.0;
    for (int i = 0; i < 5; i++)
        if (fabs(ee[i]) > t1) t1 = fabs(ee[i]);
    if (t1 < 1.0e-40) t1 = 1.0;
    for (int i = 0; i < 5; i++) {
        double t2 = ee[i];
        t2 = t2 / t1;
        ee[i] = t2;
    }
}

/* P-SV period equation (Dunkin compound matrices from the half-space up) */
static double dltar4(double wvno, double omga, const sd_model *M)
{
    double omega = omga;
    if (omega < 1.0e-4) omega = 1.0e-4;
    double wvno2 = wvno * wvno;
    int mm = M->mmax - 1;
    double xka = omega / (double)M->a[mm];
    double xkb = omega / (double)M->b[mm];
    double wvnop = wvno + xka, wvnom = fabs(wvno - xka);
    double ra = sqrt(wvnop * wvnom);
    wvnop = wvno + xkb;
    wvnom = fabs(wvno - xkb);
    double rb = sqrt(wvnop * wvnom);
    double t = (double)M->b[mm] / omega;
    double gammk = 2.0 * t * t;
    double gam = gammk * wvno2;
    double gamm1 = gam - 1.0;
    double rho1 = (double)M->rho[mm];
    double e[5], ee[5], ca[5][5];
    e[0] = rho1 * rho1 * (gamm1 * gamm1 - gam * gammk * ra * rb);
    e[1] = -(rho1 * ra);
    e[2] = rho1 * (gamm1 - gammk * ra * rb);
    e[3] = rho1 * rb;
    e[4] = wvno2 - ra * rb;
    for (int m = M->mmax - 1; m >= M->llw; m--) {
        int i0 = m - 1;
        xka = omega / (double)M->a[i0];
        xkb = omega / (double)M->b[i0];
        t = (double)M->b[i0] / omega;
        gammk = 2.0 * t * t;
        gam = gammk * wvno2;
        wvnop = wvno + xka;
        wvnom = fabs(wvno - xka);
        ra = sqrt(wvnop * wvnom);
        wvnop = wvno + xkb;
        wvnom = fabs(wvno - xkb);
        rb = sqrt(wvnop * wvnom);
        double dpth = (double)M->d[i0];
        rho1 = (double)M->rho[i0];
        double p = ra * dpth, q = rb * dpth;
        sd_var v;
        var(p, q, ra, rb, wvno, xka, xkb, dpth, &v);
        dnka(ca, wvno2, gam, gammk, rho1, &v);
        for (int i = 0; i < 5; i++) {
            double cr = 0.0;
            for (int j = 0; j < 5; j++) cr = cr + e[j] * ca[j][i];
            ee[i] = cr;
        }
        normc(ee);
        for (int i = 0; i < 5; i++) e[i] = ee[i];
    }
    if (M->llw != 1) {   /* water layer on top */
        xka = omega / (double)M->a[0];
        wvnop = wvno + xka;
        wvnom = fabs(wvno - xka);
        ra = sqrt(wvnop * wvnom);
        double dpth = (double)M->d[0];
        rho1 = (double)M->rho[0];
        double p = ra * dpth;
        const double znul = 1.0e-05;
        sd_var v;
        var(p, znul, ra, znul, wvno, xka, znul, dpth, &v);
        double w0 = -(rho1 * v.w);
        return v.cosp * e[0] + w0 * e[1];
    }
    return e[0];
}

static double dltar(double wvno, double omega, int kk, const sd_model *M)
{
    return kk == 1 ? dltar1(wvno, omega, M) : dltar4(wvno, omega, M);
}

static void half(double c1, double c2, double *c3, double *del3, double omega, int ifunc,
                 const sd_model *M)
{
    *c3 = 0.5 * (c1 + c2);
    double wvno = omega / *c3;
    *del3 = dltar(wvno, omega, ifunc, M);
}

/* nevill: hybrid bisection / Neville iteration on a bracketed root */
static double nevill(double t, double c1, double c2, double del1, double del2, int ifunc,
                     const sd_model *M, double twopi, long *ndl)
{
    double x[21], y[21];
    double omega = twopi / t;
    double c3, del3;
    int m = 1;
    half(c1, c2, &c3, &del3, omega, ifunc, M); (*ndl)++;
    int nev = 1, nctrl = 1;
    for (;;) {
        nctrl = nctrl + 1;
        if (nctrl >= 100) break;
        if (c3 < fmin(c1, c2) || c3 > fmax(c1, c2)) {
            nev = 0;
            half(c1, c2, &c3, &del3, omega, ifunc, M); (*ndl)++;
        }
        double s13 = del1 - del3;
        double s32 = del3 - del2;
        if (sgn(del3) * sgn(del1) < 0.0) { c2 = c3; del2 = del3; }
        else { c1 = c3; del1 = del3; }
        if (fabs(c1 - c2) <= 1.0e-6 * c1) break;
        if (sgn(s13) != sgn(s32)) nev = 0;
        double ss1 = fabs(del1);
        double s1 = (double)0.01f * ss1;
        double ss2 = fabs(del2);
        double s2 = (double)0.01f * ss2;
        if (s1 > ss2 || s2 > ss1 || nev == 0) {
            half(c1, c2, &c3, &del3, omega, ifunc, M); (*ndl)++;
            nev = 1;
            m = 1;
        } else {
            if (nev == 2) {
                x[m + 1] = c3;
                y[m + 1] = del3;
            } else {
                x[1] = c1; y[1] = del1;
                x[2] = c2; y[2] = del2;
                m = 1;
            }
            int bad = 0;
            for (int kk = 1; kk <= m; kk++) {
                int j = m - kk + 1;
                double denom = y[m + 1] - y[j];
                if (fabs(denom) < 1.0e-10 * fabs(y[m + 1])) { bad = 1; break; }
                x[j] = (-(y[j] * x[j + 1]) + y[m + 1] * x[j]) / denom;
            }
            if (!bad) {
                c3 = x[1];
                double wvno = omega / c3;
                del3 = dltar(wvno, omega, ifunc, M); (*ndl)++;
                nev = 2;
                m = m + 1;
                if (m > 10) m = 10;
            } else {
                half(c1, c2, &c3, &del3, omega, ifunc, M); (*ndl)++;
                nev = 1;
                m = 1;
            }
        }
    }
    return c3;
}

/* getsol: bracket the root by stepping dc from c1, then refine with nevill.
 * del1st is the saved polarity of the first period's equation. */
static int getsol(double t1, double *c1p, double clow, double dc, double cm, float betmx,
                  int ifunc, int ifirst, double *del1st, const sd_model *M, long *ndl)
{
    const double twopi = 2.0 * 3.141592653589793;
    double c1 = *c1p, c2;
    double omega = twopi / t1;
    double wvno = omega / c1;
    double del1 = dltar(wvno, omega, ifunc, M), del2;
    (*ndl)++;
    if (ifirst == 1) *del1st = del1;
    double plmn = sgn(*del1st) * sgn(del1);
    int idir;
    if (ifirst == 1) idir = +1;
    else if (plmn >= 0.0) idir = +1;
    else idir = -1;
    for (;;) {
        if (idir > 0) c2 = c1 + dc;
        else c2 = c1 - dc;
        if (c2 <= clow) {
            idir = +1;
            c1 = clow;
            continue;
        }
        omega = twopi / t1;
        wvno = omega / c2;
        del2 = dltar(wvno, omega, ifunc, M);
        (*ndl)++;
        if (sgn(del1) != sgn(del2)) break;
        c1 = c2;
        del1 = del2;
        if (c1 < cm) { *c1p = c1; return -1; }
        if (c1 >= ((double)betmx + dc)) { *c1p = c1; return -1; }
    }
    double cn = nevill(t1, c1, c2, del1, del2, ifunc, M, twopi, ndl);
    c1 = cn;
    *c1p = c1;
    if (c1 > (double)betmx) return -1;
    return 1;
}

int orc_surfdisp96(const float *thkm, const float *vpm, const float *vsm, const float *rhom,
                   int nlayer, int iwave, int igr, int kmax, const double *t, double dphase,
                   double *cp, double *cg, long *n_dltar)
{
    int ierr = 0;
    long ndl = 0;
    for (int i = 0; i < kmax; i++) { cp[i] = 100.0; cg[i] = 100.0; }
    sd_model M;
    M.mmax = nlayer;
    M.d = thkm; M.a = vpm; M.b = vsm; M.rho = rhom;
    const float sone0 = 1.500f;
    const float ddc0 = (float)dphase;
    const float h0 = 0.005f;
    M.llw = 1;
    if (vsm[0] <= 0.0f) M.llw = 2;
    const double one = 1.0e-2;
    int jmn = 1, jsol = 1;
    float betmx = -1.e20f, betmn = 1.e20f;
    for (int i = 1; i <= nlayer; i++) {
        float bi = vsm[i - 1], ai = vpm[i - 1];
        if (bi > 0.01f && bi < betmn) { betmn = bi; jmn = i; jsol = 1; }
        else if (bi <= 0.01f && ai < betmn) { betmn = ai; jmn = i; jsol = 0; }
        if (bi > betmx) betmx = bi;
    }
    int ifunc = iwave;   /* idispl>0 only for Love (ifunc 1), idispr>0 only for Rayleigh (2) */
    {
        float ddc = ddc0, sone = sone0, h = h0;
        if (sone < 0.01f) sone = 2.0f;
        double onea = (double)sone;
        float cc1;
        if (jsol == 0) cc1 = betmn;
        else cc1 = gtsolh(vpm[jmn - 1], vsm[jmn - 1]);
        cc1 = 0.95f * cc1;
        cc1 = 0.90f * cc1;
        double cc = (double)cc1;
        double dc = (double)ddc;
        dc = fabs(dc);
        double c1 = cc, cm = cc, clow = cc;
        double c[64], cb[64], del1st = 0.0;
        for (int i = 0; i < kmax; i++) { cb[i] = 0.0; c[i] = 0.0; }
        int k;
        for (k = 1; k <= kmax; k++) {
            double t1 = t[k - 1];
            float t1a, t1b = 0.0f;
            if (igr > 0) {
                t1a = (float)(t1 / (double)(1.0f + h));
                t1b = (float)(t1 / (double)(1.0f - h));
                t1 = (double)t1a;
            } else {
                t1a = (float)t1;
            }
            int ifirst;
            if (k == 1) {
                c1 = cc; clow = cc; ifirst = 1;
            } else {
                ifirst = 0;
                c1 = cc;
                for (int previd = k - 1; previd >= 1; previd--)
                    if (c[previd - 1] > 0) { c1 = c[previd - 1] - onea * dc; break; }
                clow = cm;
            }
            int iret = getsol(t1, &c1, clow, dc, cm, betmx, ifunc, ifirst, &del1st, &M, &ndl);
            if (iret == -1) break;
            c[k - 1] = c1;
            if (igr > 0) {
                t1 = (double)t1b;
                ifirst = 0;
                clow = cb[k - 1] + one * dc;
                c1 = c1 - onea * dc;
                iret = getsol(t1, &c1, clow, dc, cm, betmx, ifunc, ifirst, &del1st, &M, &ndl);
                if (iret == -1) { c1 = c[k - 1]; ierr = 1; }
                cb[k - 1] = c1;
            } else {
                c1 = 0.0;
            }
            float cc0 = (float)c[k - 1];
            float cc1k = (float)c1;
            if (igr == 0) {
                cp[k - 1] = (double)cc0;
            } else {
                float gvel = (1.0f / t1a - 1.0f / t1b) / (1.0f / (t1a * cc0) - 1.0f / (t1b * cc1k));
                cg[k - 1] = (double)gvel;
                cp[k - 1] = (double)cc0;
                if (gvel < 0 || c[k - 1] == 0) ierr = 1;
            }
        }
        if (k <= kmax) {   /* no root found at period k: zero the rest (labels 1700/1750) */
            ierr = 1;
            for (int i = k; i <= kmax; i++) cg[i - 1] = 0.0;
        }
    }
    if (n_dltar) *n_dltar = ndl;
    return ierr;
}

/* surfmodes' surfdisp96 branch (surfmodes/surfmodes.f90:79-95): REAL*4 casts
 * of the layer model and periods t = 1/freqs.  raylov 1 Rayleigh, 0 Love. */
int orc_surfmodes(const double *thick, const double *vp, const double *vs, const double *rho,
                  int nlayer, const double *freqs, int np, int raylov, int phase_group,
                  double dc, double *phase, double *group, long *n_dltar)
{
    float d[200], a[200], b[200], r[200];
    double t[64];
    if (nlayer > 200 || np > 60) return -1;
    for (int i = 0; i < nlayer; i++) {
        d[i] = (float)thick[i]; a[i] = (float)vp[i]; b[i] = (float)vs[i]; r[i] = (float)rho[i];
    }
    for (int i = 0; i < np; i++) t[i] = 1 / freqs[i];
    return orc_surfdisp96(d, a, b, r, nlayer, raylov == 1 ? 2 : 1, phase_group, np, t, dc,
                          phase, group, n_dltar);
}
