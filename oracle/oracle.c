/*
 * oracle.c -- CPU restatement of the reference's real-time propagation step.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Plain C, FP64 throughout, one
 * loop per reference array statement, in the reference's order.  Each
 * function cites the reference routine it restates (paths relative to the
 * reference repository root).  The FFT is this file's own recursive
 * mixed-radix Cooley-Tukey (radix 2/3/4/5 and a generic odd radix), unrelated
 * to the product's GPU Stockham kernels, so that the two check each other.
 *
 * Pinned by: tests/golden (vectors written by the reference itself, compiled
 * from its sources with flang + MKL FFTW3 by oracle/Makefile.ref) and by the
 * closed-form / brute-force checks in tests/test_oracle.py.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define PI_L 3.141592653589793238462643383279502884L

/* ---------------------------------------------------------------- helpers */

static inline size_t ix3(int nx, int ny, int i, int j, int k) {
    return (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * (size_t)k);
}

/* ---------------------------------------------------------------- 1-D FFT */

typedef struct {
    int n;
    int nf;
    int f[40];
    double *tw; /* exp(-2 pi i m/n), m = 0..n-1, interleaved */
} fftplan;

static int factorize(int n, int *f) {
    int nf = 0;
    while (n % 4 == 0) { f[nf++] = 4; n /= 4; }
    while (n % 2 == 0) { f[nf++] = 2; n /= 2; }
    for (int r = 3; n > 1; r += 2)
        while (n % r == 0) { f[nf++] = r; n /= r; }
    return nf;
}

static fftplan *plan_make(int n) {
    fftplan *p = (fftplan *)calloc(1, sizeof(fftplan));
    p->n = n;
    p->nf = factorize(n, p->f);
    p->tw = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    for (int m = 0; m < n; ++m) {
        long double a = -2.0L * PI_L * (long double)m / (long double)n;
        p->tw[2 * m] = (double)cosl(a);
        p->tw[2 * m + 1] = (double)sinl(a);
    }
    return p;
}

static void plan_free(fftplan *p) {
    if (p) { free(p->tw); free(p); }
}

/* Decimation in time: out[0..n) = DFT_n of in[0], in[s], in[2s], ...
 * X[k + m t] = sum_q W_n^{q (k + m t)} Y_q[k], Y_q = DFT_m(x[q + r j]). */
static void fft_rec(const fftplan *p, int n, const double *in, int s, double *out,
                    int sign, const int *f, int twstep) {
    if (n == 1) { out[0] = in[0]; out[1] = in[1]; return; }
    int r = f[0], m = n / r;
    for (int q = 0; q < r; ++q)
        fft_rec(p, m, in + 2 * (size_t)q * s, s * r, out + 2 * (size_t)q * m, sign, f + 1, twstep * r);
    double t[2 * 64], u[2 * 64];
    const int N = p->n;
    for (int k = 0; k < m; ++k) {
        for (int q = 0; q < r; ++q) {
            t[2 * q] = out[2 * ((size_t)k + (size_t)m * q)];
            t[2 * q + 1] = out[2 * ((size_t)k + (size_t)m * q) + 1];
        }
        for (int tt = 0; tt < r; ++tt) {
            double sr = 0.0, si = 0.0;
            for (int q = 0; q < r; ++q) {
                long e = ((long)q * (long)(k + m * tt)) % n; /* W_n^e */
                long ei = (e * (long)twstep) % N;
                double wr = p->tw[2 * ei], wi = -sign * p->tw[2 * ei + 1];
                sr += t[2 * q] * wr - t[2 * q + 1] * wi;
                si += t[2 * q] * wi + t[2 * q + 1] * wr;
            }
            u[2 * tt] = sr;
            u[2 * tt + 1] = si;
        }
        for (int tt = 0; tt < r; ++tt) {
            out[2 * ((size_t)k + (size_t)m * tt)] = u[2 * tt];
            out[2 * ((size_t)k + (size_t)m * tt) + 1] = u[2 * tt + 1];
        }
    }
}

/* sign -1: exp(-2 pi i ...) (FFTW forward), +1: backward; unnormalised */
static void fft1d(const fftplan *p, const double *in, double *out, int sign) {
    fft_rec(p, p->n, in, 1, out, sign, p->f, 1);
}

/* 3-D r2c / c2r with FFTW conventions (fft.f90:50-64): real (nx,ny,nz) <->
 * complex (nx/2+1,ny,nz), x fastest.  c2r: complex backward along z and y,
 * then along x the canonical one-dimensional c2r (Hermitian extension of the
 * half line; imaginary parts of k=0 and, for even nx, k=nx/2 ignored). */
static void r2c3d(int nx, int ny, int nz, const double *in, double *out,
                  fftplan *px, fftplan *py, fftplan *pz) {
    int nxh = nx / 2 + 1;
    int nmax = nx > ny ? nx : ny;
    if (nz > nmax) nmax = nz;
    double *a = (double *)malloc(sizeof(double) * 2 * nmax);
    double *b = (double *)malloc(sizeof(double) * 2 * nmax);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j) {
            for (int i = 0; i < nx; ++i) { a[2 * i] = in[ix3(nx, ny, i, j, k)]; a[2 * i + 1] = 0.0; }
            fft1d(px, a, b, -1);
            for (int i = 0; i < nxh; ++i) {
                size_t o = ix3(nxh, ny, i, j, k);
                out[2 * o] = b[2 * i]; out[2 * o + 1] = b[2 * i + 1];
            }
        }
    for (int k = 0; k < nz; ++k)
        for (int i = 0; i < nxh; ++i) {
            for (int j = 0; j < ny; ++j) { size_t o = ix3(nxh, ny, i, j, k); a[2 * j] = out[2 * o]; a[2 * j + 1] = out[2 * o + 1]; }
            fft1d(py, a, b, -1);
            for (int j = 0; j < ny; ++j) { size_t o = ix3(nxh, ny, i, j, k); out[2 * o] = b[2 * j]; out[2 * o + 1] = b[2 * j + 1]; }
        }
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nxh; ++i) {
            for (int k = 0; k < nz; ++k) { size_t o = ix3(nxh, ny, i, j, k); a[2 * k] = out[2 * o]; a[2 * k + 1] = out[2 * o + 1]; }
            fft1d(pz, a, b, -1);
            for (int k = 0; k < nz; ++k) { size_t o = ix3(nxh, ny, i, j, k); out[2 * o] = b[2 * k]; out[2 * o + 1] = b[2 * k + 1]; }
        }
    free(a); free(b);
}

static void c2r3d(int nx, int ny, int nz, const double *in, double *out, double scale,
                  fftplan *px, fftplan *py, fftplan *pz, double *work /* 2*nxh*ny*nz */) {
    int nxh = nx / 2 + 1;
    int nmax = nx > ny ? nx : ny;
    if (nz > nmax) nmax = nz;
    double *a = (double *)malloc(sizeof(double) * 2 * nmax);
    double *b = (double *)malloc(sizeof(double) * 2 * nmax);
    memcpy(work, in, sizeof(double) * 2 * (size_t)nxh * ny * nz);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nxh; ++i) {
            for (int k = 0; k < nz; ++k) { size_t o = ix3(nxh, ny, i, j, k); a[2 * k] = work[2 * o]; a[2 * k + 1] = work[2 * o + 1]; }
            fft1d(pz, a, b, +1);
            for (int k = 0; k < nz; ++k) { size_t o = ix3(nxh, ny, i, j, k); work[2 * o] = b[2 * k]; work[2 * o + 1] = b[2 * k + 1]; }
        }
    for (int k = 0; k < nz; ++k)
        for (int i = 0; i < nxh; ++i) {
            for (int j = 0; j < ny; ++j) { size_t o = ix3(nxh, ny, i, j, k); a[2 * j] = work[2 * o]; a[2 * j + 1] = work[2 * o + 1]; }
            fft1d(py, a, b, +1);
            for (int j = 0; j < ny; ++j) { size_t o = ix3(nxh, ny, i, j, k); work[2 * o] = b[2 * j]; work[2 * o + 1] = b[2 * j + 1]; }
        }
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j) {
            for (int i = 0; i < nxh; ++i) {
                size_t o = ix3(nxh, ny, i, j, k);
                a[2 * i] = work[2 * o]; a[2 * i + 1] = work[2 * o + 1];
            }
            a[1] = 0.0;
            if (nx % 2 == 0) a[2 * (nx / 2) + 1] = 0.0;
            for (int i = nxh; i < nx; ++i) { a[2 * i] = a[2 * (nx - i)]; a[2 * i + 1] = -a[2 * (nx - i) + 1]; }
            fft1d(px, a, b, +1);
            for (int i = 0; i < nx; ++i) out[ix3(nx, ny, i, j, k)] = b[2 * i] * scale;
        }
    free(a); free(b);
}

/* ------------------------------------------------- finite-difference weights */

/* Centred npd-point weights of the id-th derivative (id = 1, 2) at unit
 * spacing, c[j] for offset j-m, m = npd/2.  These are the centred rows of the
 * reference's table (newder.f:99-128, mates.f:172-14043), whose values are
 * the Fornberg closed forms: d_k = (-1)^{k+1} (m!)^2/(k (m-k)!(m+k)!) for the
 * first derivative, c_k = 2 d_k / k and c_0 = -2 sum_k 1/k^2 for the second. */
static void fd_centred(int npd, int id, double *c) {
    int m = npd / 2;
    for (int j = 0; j < npd; ++j) c[j] = 0.0;
    long double s2 = 0.0L;
    for (int k = 1; k <= m; ++k) {
        long double ratio = 1.0L; /* (m!)^2 / ((m-k)! (m+k)!) */
        for (int j = 1; j <= k; ++j) ratio *= (long double)(m - k + j) / (long double)(m + j);
        long double sg = (k % 2 == 1) ? 1.0L : -1.0L;
        if (id == 1) {
            long double d = sg * ratio / (long double)k;
            c[m + k] = (double)d;
            c[m - k] = (double)(-d);
        } else {
            long double d = 2.0L * sg * ratio / ((long double)k * (long double)k);
            c[m + k] = (double)d;
            c[m - k] = (double)d;
            s2 += 1.0L / ((long double)k * (long double)k);
        }
    }
    if (id == 2) c[m] = (double)(-2.0L * s2);
}

/* Deriv3D_p / Deriv3D_pc (init_deriv_parallel.f90:373-648, 933-1209): the
 * reference multiplies the weights by Sto = 1/h**id, and Redef with icon=13
 * (redef.f:151-170 then :133-150) turns the boundary rows into the centred
 * stencil with zeros outside [1,n]; icon=8 wraps periodically, icon=18
 * wraps with a sign flip. */
static int deriv3d(int nx, int ny, int nz, int npd, int icon, int id, int iv, double h,
                   int cplx, const double *f, double *df) {
    if (npd < 3 || npd > 25 || npd % 2 == 0 || id < 1 || id > 2 || iv < 1 || iv > 3) return HETD_ERR_ARG;
    if (icon != 13 && icon != 8 && icon != 18) return HETD_ERR_ARG;
    int nn[3] = {nx, ny, nz};
    int n = nn[iv - 1];
    size_t ntot = (size_t)nx * ny * nz;
    int nc = cplx ? 2 : 1;
    if (n < npd) { memset(df, 0, sizeof(double) * nc * ntot); return 0; }
    double c[32], cd[32];
    fd_centred(npd, id, c);
    double sto = 1.0 / (id == 1 ? h : h * h);
    for (int j = 0; j < npd; ++j) cd[j] = c[j] * sto;
    int m = npd / 2;
    size_t stride = iv == 1 ? 1 : (iv == 2 ? (size_t)nx : (size_t)nx * ny);
    for (int k = 0; k < nz; ++k)
        for (int jj = 0; jj < ny; ++jj)
            for (int i = 0; i < nx; ++i) {
                int pos = iv == 1 ? i : (iv == 2 ? jj : k);
                size_t base = ix3(nx, ny, i, jj, k) - (size_t)pos * stride;
                for (int comp = 0; comp < nc; ++comp) {
                    double s = 0.0;
                    for (int j = 0; j < npd; ++j) {
                        int q = pos + j - m;
                        double sign = 1.0;
                        if (q < 0 || q >= n) {
                            if (icon == 13) continue;
                            if (icon == 18) sign = -1.0;
                            q = (q + n) % n;
                        }
                        s += cd[j] * sign * f[nc * (base + (size_t)q * stride) + comp];
                    }
                    df[nc * ix3(nx, ny, i, jj, k) + comp] = s;
                }
            }
    return 0;
}

/* ------------------------------------------------------- special functions */

/* pi/2 - Si(x) for x > 0 (the reference's sinint, fforma.f90:143-196, which
 * calls the NAG-style S13ADF, s13adf.f90).  Restated from the mathematics:
 * power series of Si for x <= 4; for x > 4 the continued fraction
 * E1(z) = e^{-z} / (z + 1 - 1^2/(z + 3 - 2^2/(z + 5 - ...))) at z = ix
 * evaluated by the modified Lentz method, with pi/2 - Si(x) = -Im E1(ix). */
static long double sinint_l(long double x) {
    if (x <= 4.0L) {
        long double term = x, sum = x, x2 = x * x;
        for (int k = 1; k < 60; ++k) {
            term *= -x2 / ((long double)(2 * k) * (long double)(2 * k + 1));
            long double add = term / (long double)(2 * k + 1);
            sum += add;
            if (fabsl(add) < 1e-22L * fabsl(sum)) break;
        }
        return PI_L / 2.0L - sum;
    }
    /* complex arithmetic on (re, im) pairs */
    long double br = 1.0L, bi = x;            /* b = 1 + i x */
    long double cr = 1e300L, ci = 0.0L;       /* c = 1/tiny */
    long double den = br * br + bi * bi;
    long double dr = br / den, di = -bi / den; /* d = 1/b */
    long double hr = dr, hi = di;
    for (int i = 2; i < 10000; ++i) {
        long double a = -(long double)(i - 1) * (long double)(i - 1);
        br += 2.0L;
        /* d = 1/(a d + b) */
        long double tr = a * dr + br, ti = a * di + bi;
        den = tr * tr + ti * ti;
        dr = tr / den; di = -ti / den;
        /* c = b + a/c */
        den = cr * cr + ci * ci;
        long double qr = a * cr / den, qi = -a * ci / den;
        cr = br + qr; ci = bi + qi;
        long double er = cr * dr - ci * di, ei = cr * di + ci * dr;
        long double nhr = hr * er - hi * ei, nhi = hr * ei + hi * er;
        hr = nhr; hi = nhi;
        if (fabsl(er - 1.0L) + fabsl(ei) < 1e-21L) break;
    }
    /* E1(ix) = (cos x - i sin x) h */
    long double cs = cosl(x), sn = sinl(x);
    long double e1i = cs * hi - sn * hr;
    return -e1i;
}

double orc_si(double x) {
    long double ax = fabsl((long double)x);
    long double s = PI_L / 2.0L - sinint_l(ax);
    return (double)(x < 0 ? -s : s);
}

/* Fourier transform of the Orsay-Trento-cored LJ potential (fforma.f90:40-62
 * with sinx11 :73-102, sinx5 :108-137, bforce :256-280); p in cycles/A. */
double orc_vlj_hat(double p, double eps, double sigma, double hc, double b_unused) {
    (void)b_unused;
    long double s6 = powl((long double)sigma, 6);
    if (fabs(p) < 1e-10) {
        long double pi4 = 4.0L * PI_L, eps4 = 4.0L * eps;
        long double bb2 = pi4 * eps4 * s6 * s6 / (9.0L * powl((long double)hc, 9));
        long double bb3 = -pi4 * eps4 * s6 / (3.0L * powl((long double)hc, 3));
        return (double)(bb2 + bb3);
    }
    long double a = 2.0L * PI_L * (long double)p, x = hc;
    long double y = a * x, y2 = y * y;
    long double si = sinint_l(y);
    long double cy = cosl(y), sy = sinl(y);
    /* integral_x^inf sin(a r)/r^11 dr */
    long double bd = 40320.0L - 720.0L * y2 + 24.0L * y2 * y2 - 2.0L * y2 * y2 * y2 + y2 * y2 * y2 * y2;
    long double cdd = 362880.0L - 5040.0L * y2 + 120.0L * y2 * y2 - 6.0L * y2 * y2 * y2 + y2 * y2 * y2 * y2;
    long double t1 = (a * bd * cy + cdd * sy / x) / (3628800.0L * powl(x, 9)) - powl(a, 10) * si / 3628800.0L;
    /* integral_x^inf sin(a r)/r^5 dr */
    long double bd5 = -2.0L + y2, cd5 = -6.0L + y2;
    long double t2 = -(a * bd5 * cy + cd5 * sy / x) / (24.0L * x * x * x) + powl(a, 4) * si / 24.0L;
    long double aux1 = 8.0L * eps;
    return (double)(aux1 * s6 * (s6 * t1 - t2) / (long double)p);
}

/* Tang-Toennies Xe-He pair potential (V_impur.f90:2728-2772) with the
 * reference's constants, K for r in A; and Select_Pot's cutoff rule
 * (V_impur.f90:757-770): umax for r < r_cutoff.  selec 0 = 'null'.
 * V = A e^{-b r} - sum_{n=3..9} f_2n(b r) C_2n / r^2n, f_2n the Tang-Toennies
 * damping 1 - e^{-x} sum_{k<=2n} x^k/k!, C_12.. from C_2n = (C_2n-2/C_2n-4)^3 C_2n-6. */
double orc_vimp(int selec, double r, double r_cutoff, double umax) {
    /* evaluated in double in the reference's order: B(i) = B(i-1)*b/i,
     * S(k) = sum_{i<=k} B(i) rr^i, V = A e^{-b rr} + sum_n ((S(2n)+1) e^{-b rr} - 1) C_2n / rr^2n */
    if (selec == 0) return 0.0;
    if (r < r_cutoff) return umax;
    const double AA = 74.9002, BB = 1.8175, rbohr = 0.5292, h2k = 27.21 * 11604.0;
    double B[19], C[8], S[19];
    B[1] = BB;
    for (int i = 2; i <= 18; ++i) B[i] = B[i - 1] * BB / (double)i;
    C[1] = 21.5674; C[2] = 430.148; C[3] = 10610.21;
    for (int i = 4; i <= 7; ++i) C[i] = pow(C[i - 1] / C[i - 2], 3) * C[i - 3];
    double rr = r / rbohr;
    double e = exp(-BB * rr);
    double st = rr;
    S[1] = B[1] * st;
    for (int i = 2; i <= 18; ++i) { st = st * rr; S[i] = S[i - 1] + B[i] * st; }
    double aux = AA * e;
    double s1 = 1.0 / (rr * rr);
    st = s1 * s1 * s1;
    for (int i = 1; i <= 7; ++i) {
        int k = 2 * i + 4;
        aux = aux + ((S[k] + 1.0) * e - 1.0) * C[i] * st;
        st = st * s1;
    }
    return aux * h2k;
}

/* ---------------------------------------------------------------- context */

struct orc_ctx {
    hetd_params p;
    int nx, ny, nz, nxh;
    size_t n, nh;
    double hx, hy, hz, dxyz, mAg;
    double *x, *y, *z;
    double *tcx, *tcy, *tcz; /* 1 + tanh((|x|-xmean)/xsurf) per axis */
    fftplan *px, *py, *pz;
    /* fields */
    double *psi, *den, *pot4, *dencg, *delj4, *ccor, *dxden, *dyden, *dzden, *uimp, *penalty;
    double *psiold[3], *hpsiold[2], *pc, *q;
    double *fden, *wk, *wk2, *wk3, *fvlj, *wcgk, *kalfs, *hwork;
    double *denalf, *falfs, *intx, *inty, *intz, *ualphas;
    double *s1c, *s2c, *s3c, *f4c, *s5c; /* complex scratch */
    double *s1, *s2, *s3, *s4, *s5, *s6; /* real scratch */
    double *potion; int npot; double delinter;
    /* impurity (module classicimp, modules.f90:282-320) */
    double rimp[3], vimp[3], aimp[3];
    double rimpold[3][3], vimpold[3][3], aimpold[2][3]; /* [slot][xyz] */
    double qr[3], qv[3], pcr[3], pcv[3];
    int ioldp[3], ioldh[2], ioldr[3], ioldv[3], iolda[2];
    int iter_since_set;
    char err[256];
};

static double *zalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

const char *orc_error(orc_ctx *c) { return c ? c->err : "null context"; }

orc_ctx *orc_create(const hetd_params *p) {
    orc_ctx *c = (orc_ctx *)calloc(1, sizeof(orc_ctx));
    c->p = *p;
    c->nx = p->nx; c->ny = p->ny; c->nz = p->nz; c->nxh = p->nx / 2 + 1;
    c->n = (size_t)c->nx * c->ny * c->nz;
    c->nh = (size_t)c->nxh * c->ny * c->nz;
    /* DFT4He3d.f90:274-288 */
    c->hx = 2.0 * fabs(p->xmax) / p->nx;
    c->hy = 2.0 * fabs(p->ymax) / p->ny;
    c->hz = 2.0 * fabs(p->zmax) / p->nz;
    c->dxyz = c->hx * c->hy * c->hz;
    c->mAg = p->mimpur * 0.02061484; /* mp_u (modules.f90:315) */
    c->x = zalloc(c->nx); c->y = zalloc(c->ny); c->z = zalloc(c->nz);
    for (int i = 0; i < c->nx; ++i) c->x[i] = -p->xmax + c->hx * i;
    for (int i = 0; i < c->ny; ++i) c->y[i] = -p->ymax + c->hy * i;
    for (int i = 0; i < c->nz; ++i) c->z[i] = -p->zmax + c->hz * i;
    c->tcx = zalloc(c->nx); c->tcy = zalloc(c->ny); c->tcz = zalloc(c->nz);
    for (int i = 0; i < c->nx; ++i) c->tcx[i] = tanh((fabs(c->x[i]) - p->txmean) / p->txsurf);
    for (int i = 0; i < c->ny; ++i) c->tcy[i] = tanh((fabs(c->y[i]) - p->tymean) / p->tysurf);
    for (int i = 0; i < c->nz; ++i) c->tcz[i] = tanh((fabs(c->z[i]) - p->tzmean) / p->tzsurf);
    c->px = plan_make(c->nx); c->py = plan_make(c->ny); c->pz = plan_make(c->nz);
    size_t n = c->n, nh = c->nh;
    c->psi = zalloc(2 * n); c->den = zalloc(n); c->pot4 = zalloc(n); c->dencg = zalloc(n);
    c->delj4 = zalloc(n); c->ccor = zalloc(n); c->dxden = zalloc(n); c->dyden = zalloc(n);
    c->dzden = zalloc(n); c->uimp = zalloc(n); c->penalty = zalloc(n);
    for (int i = 0; i < 3; ++i) c->psiold[i] = zalloc(2 * n);
    for (int i = 0; i < 2; ++i) c->hpsiold[i] = zalloc(2 * n);
    c->pc = zalloc(2 * n); c->q = zalloc(2 * n);
    c->fden = zalloc(2 * nh); c->wk = zalloc(2 * nh); c->wk2 = zalloc(2 * nh); c->wk3 = zalloc(2 * nh);
    c->hwork = zalloc(2 * nh);
    c->fvlj = zalloc(nh); c->wcgk = zalloc(nh); c->kalfs = zalloc(nh);
    c->denalf = zalloc(n); c->falfs = zalloc(n); c->intx = zalloc(n); c->inty = zalloc(n);
    c->intz = zalloc(n); c->ualphas = zalloc(n);
    c->s1c = zalloc(2 * n); c->s2c = zalloc(2 * n); c->s3c = zalloc(2 * n);
    c->f4c = zalloc(2 * n); c->s5c = zalloc(2 * n);
    c->s1 = zalloc(n); c->s2 = zalloc(n); c->s3 = zalloc(n);
    c->s4 = zalloc(n); c->s5 = zalloc(n); c->s6 = zalloc(n);
    /* p-grid and pmod (DFT4He3d.f90:320-361); form factors fforma
     * (fforma.f90:40-62), initcg (initcg.f90:32-49), kalfs (DFT4He3d.f90:592-596) */
    double hpx = 1.0 / (c->nx * c->hx), hpy = 1.0 / (c->ny * c->hy), hpz = 1.0 / (c->nz * c->hz);
    double pmx = 1.0 / (2.0 * c->hx), pmy = 1.0 / (2.0 * c->hy), pmz = 1.0 / (2.0 * c->hz);
    double hcg = p->lsolid ? p->h4 * 1.065 : p->h4;
    double twopi = 2.0 * (4.0 * atan(1.0)), pi = 4.0 * atan(1.0);
    for (int k = 0; k < c->nz; ++k) {
        double pz = (k + 1 <= c->nz / 2 + 1) ? hpz * k : -pmz + hpz * (k + 1 - (c->nz / 2 + 1));
        for (int j = 0; j < c->ny; ++j) {
            double py = (j + 1 <= c->ny / 2 + 1) ? hpy * j : -pmy + hpy * (j + 1 - (c->ny / 2 + 1));
            for (int i = 0; i < c->nxh; ++i) {
                double px = hpx * i;
                (void)pmx;
                double pm = sqrt(px * px + py * py + pz * pz);
                size_t o = ix3(c->nxh, c->ny, i, j, k);
                c->fvlj[o] = orc_vlj_hat(pm, p->eps4, p->sigma4, p->h4, 0.0);
                double w;
                if (pm > 0.0) {
                    double arg = twopi * pm * hcg;
                    double aux = sin(arg) - arg * cos(arg);
                    w = fabs(aux) < 1e-10 ? 1.0 : 3.0 * aux / (arg * arg * arg);
                } else {
                    w = 1.0;
                }
                c->wcgk[o] = w;
                c->kalfs[o] = exp(-(pi * p->l_alpha * pm) * (pi * p->l_alpha * pm));
            }
        }
    }
    c->ioldp[0] = 1; c->ioldp[1] = 2; c->ioldp[2] = 3;
    c->ioldh[0] = 1; c->ioldh[1] = 2;
    c->ioldr[0] = 1; c->ioldr[1] = 2; c->ioldr[2] = 3;
    c->ioldv[0] = 1; c->ioldv[1] = 2; c->ioldv[2] = 3;
    c->iolda[0] = 1; c->iolda[1] = 2;
    return c;
}

void orc_destroy(orc_ctx *c) {
    if (!c) return;
    free(c->x); free(c->y); free(c->z); free(c->tcx); free(c->tcy); free(c->tcz);
    plan_free(c->px); plan_free(c->py); plan_free(c->pz);
    double *arrs[] = {c->psi, c->den, c->pot4, c->dencg, c->delj4, c->ccor, c->dxden, c->dyden,
                      c->dzden, c->uimp, c->penalty, c->psiold[0], c->psiold[1], c->psiold[2],
                      c->hpsiold[0], c->hpsiold[1], c->pc, c->q, c->fden, c->wk, c->wk2, c->wk3,
                      c->hwork, c->fvlj, c->wcgk, c->kalfs, c->denalf, c->falfs, c->intx, c->inty,
                      c->intz, c->ualphas, c->s1c, c->s2c, c->s3c, c->f4c, c->s5c, c->s1, c->s2,
                      c->s3, c->s4, c->s5, c->s6, c->potion};
    for (size_t i = 0; i < sizeof(arrs) / sizeof(arrs[0]); ++i) free(arrs[i]);
    free(c);
}

int orc_set_potion(orc_ctx *c, const double *potion, int npot, double delinter) {
    free(c->potion);
    c->potion = zalloc(npot);
    memcpy(c->potion, potion, sizeof(double) * npot);
    c->npot = npot;
    c->delinter = delinter;
    return 0;
}

static void den_from_psi(orc_ctx *c) {
    for (size_t i = 0; i < c->n; ++i) {
        double a = hypot(c->psi[2 * i], c->psi[2 * i + 1]);
        c->den[i] = a * a;
    }
}

int orc_set_state(orc_ctx *c, const double *psi, const double r[3], const double v[3]) {
    memcpy(c->psi, psi, sizeof(double) * 2 * c->n);
    for (int d = 0; d < 3; ++d) { c->rimp[d] = r[d]; c->vimp[d] = v[d]; c->aimp[d] = 0.0; }
    den_from_psi(c);
    memset(c->q, 0, sizeof(double) * 2 * c->n);
    memset(c->pc, 0, sizeof(double) * 2 * c->n);
    for (int i = 0; i < 3; ++i) { memset(c->psiold[i], 0, sizeof(double) * 2 * c->n); }
    for (int i = 0; i < 2; ++i) { memset(c->hpsiold[i], 0, sizeof(double) * 2 * c->n); }
    memset(c->rimpold, 0, sizeof(c->rimpold)); memset(c->vimpold, 0, sizeof(c->vimpold));
    memset(c->aimpold, 0, sizeof(c->aimpold));
    for (int d = 0; d < 3; ++d) { c->qr[d] = c->qv[d] = c->pcr[d] = c->pcv[d] = 0.0; }
    c->iter_since_set = 0;
    return 0;
}

/* updatepoten (potenimp.f90:71-95) */
int orc_potenimp(orc_ctx *c, const double r[3]) {
    if (!c->potion) { snprintf(c->err, sizeof c->err, "potion not set"); return HETD_ERR_STATE; }
    double D = c->delinter;
    for (int k = 0; k < c->nz; ++k) {
        double zt = (c->z[k] - r[2]) * (c->z[k] - r[2]);
        for (int j = 0; j < c->ny; ++j) {
            double yt = (c->y[j] - r[1]) * (c->y[j] - r[1]) + zt;
            for (int i = 0; i < c->nx; ++i) {
                double rr = sqrt((c->x[i] - r[0]) * (c->x[i] - r[0]) + yt);
                long ir = (long)(rr / D);
                if (ir > c->npot - 2) ir = c->npot - 2;
                double w = fmod(rr, D) / D;
                c->uimp[ix3(c->nx, c->ny, i, j, k)] = c->potion[ir] * (1.0 - w) + c->potion[ir + 1] * w;
            }
        }
    }
    return 0;
}

static void deriv_field(orc_ctx *c, int id, int iv, int cplx, const double *f, double *df) {
    double h = iv == 1 ? c->hx : (iv == 2 ? c->hy : c->hz);
    deriv3d(c->nx, c->ny, c->nz, c->p.npd, c->p.icon, id, iv, h, cplx, f, df);
}

/* term_alfa (term_alfa.f90:28-113) */
static void term_alfa(orc_ctx *c) {
    size_t n = c->n, nh = c->nh;
    double renor = 1.0 / (double)(c->nx * c->ny * c->nz);
    for (size_t i = 0; i < n; ++i) c->falfs[i] = 1.0 - c->denalf[i] / c->p.den0s;
    for (size_t i = 0; i < n; ++i) {
        c->s1[i] = c->falfs[i] * c->dxden[i];
        c->s2[i] = c->falfs[i] * c->dyden[i];
        c->s3[i] = c->falfs[i] * c->dzden[i];
    }
    r2c3d(c->nx, c->ny, c->nz, c->s1, c->wk, c->px, c->py, c->pz);
    r2c3d(c->nx, c->ny, c->nz, c->s2, c->wk2, c->px, c->py, c->pz);
    r2c3d(c->nx, c->ny, c->nz, c->s3, c->wk3, c->px, c->py, c->pz);
    for (size_t i = 0; i < nh; ++i) {
        double k = c->kalfs[i];
        c->wk[2 * i] *= k; c->wk[2 * i + 1] *= k;
        c->wk2[2 * i] *= k; c->wk2[2 * i + 1] *= k;
        c->wk3[2 * i] *= k; c->wk3[2 * i + 1] *= k;
    }
    c2r3d(c->nx, c->ny, c->nz, c->wk, c->intx, renor, c->px, c->py, c->pz, c->hwork);
    c2r3d(c->nx, c->ny, c->nz, c->wk2, c->inty, renor, c->px, c->py, c->pz, c->hwork);
    c2r3d(c->nx, c->ny, c->nz, c->wk3, c->intz, renor, c->px, c->py, c->pz, c->hwork);
    for (size_t i = 0; i < n; ++i)
        c->s1[i] = (c->dxden[i] * c->intx[i]) + (c->dyden[i] * c->inty[i]) + (c->dzden[i] * c->intz[i]);
    r2c3d(c->nx, c->ny, c->nz, c->s1, c->wk, c->px, c->py, c->pz);
    for (size_t i = 0; i < nh; ++i) { c->wk[2 * i] *= c->kalfs[i]; c->wk[2 * i + 1] *= c->kalfs[i]; }
    c2r3d(c->nx, c->ny, c->nz, c->wk, c->ualphas, renor, c->px, c->py, c->pz, c->hwork);
    for (size_t i = 0; i < n; ++i) {
        c->ualphas[i] = c->ualphas[i] / c->p.den0s;
        c->s1[i] = c->falfs[i] * c->intx[i];
        c->s2[i] = c->falfs[i] * c->inty[i];
        c->s3[i] = c->falfs[i] * c->intz[i];
    }
    deriv_field(c, 1, 1, 0, c->s1, c->s4);
    deriv_field(c, 1, 2, 0, c->s2, c->s5);
    deriv_field(c, 1, 3, 0, c->s3, c->s6);
    for (size_t i = 0; i < n; ++i)
        c->ualphas[i] = (c->ualphas[i] + c->s4[i] + c->s5[i] + c->s6[i]) * c->p.h2o2m4 * c->p.alphas;
}

/* poten (poten.f90:5-118) with derden (derden.f90:26-28) */
int orc_poten(orc_ctx *c) {
    size_t n = c->n, nh = c->nh;
    const hetd_params *p = &c->p;
    double renor = 1.0 / (double)(c->nx * c->ny * c->nz);
    r2c3d(c->nx, c->ny, c->nz, c->den, c->fden, c->px, c->py, c->pz);
    for (size_t i = 0; i < nh; ++i) { c->wk[2 * i] = c->fden[2 * i] * c->wcgk[i]; c->wk[2 * i + 1] = c->fden[2 * i + 1] * c->wcgk[i]; }
    c2r3d(c->nx, c->ny, c->nz, c->wk, c->dencg, renor, c->px, c->py, c->pz, c->hwork);
    deriv_field(c, 1, 1, 0, c->den, c->dxden);
    deriv_field(c, 1, 2, 0, c->den, c->dyden);
    deriv_field(c, 1, 3, 0, c->den, c->dzden);
    for (size_t i = 0; i < nh; ++i) { c->wk[2 * i] = c->fden[2 * i] * c->fvlj[i]; c->wk[2 * i + 1] = c->fden[2 * i + 1] * c->fvlj[i]; }
    c2r3d(c->nx, c->ny, c->nz, c->wk, c->delj4, renor, c->px, c->py, c->pz, c->hwork);
    double a0 = p->cp4 / 2.0, a1 = p->cpp4 / 3.0;
    for (size_t i = 0; i < n; ++i) c->s1[i] = c->den[i] * c->dencg[i] * (p->cp4 + p->cpp4 * c->dencg[i]);
    r2c3d(c->nx, c->ny, c->nz, c->s1, c->wk, c->px, c->py, c->pz);
    for (size_t i = 0; i < nh; ++i) { c->wk[2 * i] *= c->wcgk[i]; c->wk[2 * i + 1] *= c->wcgk[i]; }
    c2r3d(c->nx, c->ny, c->nz, c->wk, c->ccor, renor, c->px, c->py, c->pz, c->hwork);
    if (p->lsolid) {
        for (size_t i = 0; i < n; ++i) {
            double t = tanh(p->beta_solid * (c->den[i] - p->den_m));
            c->penalty[i] = p->c_solid * (1.0 + t + p->beta_solid * c->den[i] * (1.0 - t * t));
        }
    }
    for (size_t i = 0; i < n; ++i)
        c->pot4[i] = c->delj4[i] + c->dencg[i] * c->dencg[i] * (a0 + a1 * c->dencg[i]) + c->ccor[i];
    if (!p->lsolid && p->core == HETD_CORE_OTC) {
        for (size_t i = 0; i < nh; ++i) { c->wk[2 * i] = c->fden[2 * i] * c->kalfs[i]; c->wk[2 * i + 1] = c->fden[2 * i + 1] * c->kalfs[i]; }
        c2r3d(c->nx, c->ny, c->nz, c->wk, c->denalf, renor, c->px, c->py, c->pz, c->hwork);
        term_alfa(c);
        for (size_t i = 0; i < n; ++i) c->pot4[i] = c->pot4[i] + c->ualphas[i];
    }
    if (p->lsolid)
        for (size_t i = 0; i < n; ++i) c->pot4[i] = c->pot4[i] + c->penalty[i];
    return 0;
}

/* forceimp (potenimp.f90:24-55) */
int orc_forceimp(orc_ctx *c, const double r[3], double F[3]) {
    (void)r;
    double s[3] = {0, 0, 0};
    for (size_t i = 0; i < c->n; ++i) {
        s[0] += c->dxden[i] * c->uimp[i];
        s[1] += c->dyden[i] * c->uimp[i];
        s[2] += c->dzden[i] * c->uimp[i];
    }
    for (int d = 0; d < 3; ++d) F[d] = -s[d] * c->dxyz;
    return 0;
}

static void potential_update(orc_ctx *c, const double r[3]) {
    double F[3];
    orc_potenimp(c, r);
    orc_poten(c);
    orc_forceimp(c, r, F);
    for (int d = 0; d < 3; ++d) c->aimp[d] = F[d] / c->mAg;
}

/* RHS (steprk.f90:74, steppc.f90:55,123):
 * f = timec*((d2x+d2y+d2z)*h2o2m4 - pot4*psi) - i*uimp*psi, into c->f4c */
static void rhs(orc_ctx *c) {
    deriv_field(c, 2, 1, 1, c->psi, c->s1c);
    deriv_field(c, 2, 2, 1, c->psi, c->s2c);
    deriv_field(c, 2, 3, 1, c->psi, c->s3c);
    const double h2 = c->p.h2o2m4, L = c->p.Lambdah;
    for (int k = 0; k < c->nz; ++k)
        for (int j = 0; j < c->ny; ++j)
            for (int i = 0; i < c->nx; ++i) {
                size_t o = ix3(c->nx, c->ny, i, j, k);
                double tr = L * (1.0 + c->tcx[i] + 1.0 + c->tcy[j] + 1.0 + c->tcz[k]);
                double lr = c->s1c[2 * o] + c->s2c[2 * o] + c->s3c[2 * o];
                double li = c->s1c[2 * o + 1] + c->s2c[2 * o + 1] + c->s3c[2 * o + 1];
                double pr = c->psi[2 * o], pi = c->psi[2 * o + 1];
                double ar = lr * h2 - c->pot4[o] * pr;
                double ai = li * h2 - c->pot4[o] * pi;
                double u = c->uimp[o];
                c->f4c[2 * o] = (tr * ar - 1.0 * ai) + u * pi;
                c->f4c[2 * o + 1] = (tr * ai + 1.0 * ar) - u * pr;
            }
}

static int guard(orc_ctx *c) {
    for (size_t i = 0; i < c->n; ++i)
        if (!(c->den[i] > 0.0)) {
            snprintf(c->err, sizeof c->err, "den not > 0 at %zu: %g", i, c->den[i]);
            return HETD_ERR_GUARD;
        }
    return 0;
}

/* steprk (steprk.f90:1-153) */
int orc_steprk(orc_ctx *c, double dt) {
    const double s2 = sqrt(2.0);
    double arun[4] = {0.5, 1.0 - 1.0 / s2, 1.0 + 1.0 / s2, 1.0 / 6.0};
    double brun[4] = {2.0, 1.0, 1.0, 2.0};
    double crun[4] = {0.5, 1.0 - 1.0 / s2, 1.0 + 1.0 / s2, 0.5};
    size_t n = c->n;
    for (int j = 0; j < 4; ++j) {
        rhs(c);
        double a = arun[j], b = brun[j], cc = crun[j];
        if (j == 0) {
            double *t = c->hpsiold[1]; c->hpsiold[1] = c->hpsiold[0]; c->hpsiold[0] = t;
            double *t3 = c->psiold[2]; c->psiold[2] = c->psiold[1]; c->psiold[1] = c->psiold[0]; c->psiold[0] = t3;
        }
        for (size_t i = 0; i < 2 * n; ++i) {
            double f = c->f4c[i];
            double s1 = a * (f - b * c->q[i]);
            c->q[i] = c->q[i] + 3.0 * s1 - cc * f;
            if (j == 0) { c->hpsiold[0][i] = f; c->psiold[0][i] = c->psi[i]; }
            c->psi[i] = c->psi[i] + dt * s1;
        }
        den_from_psi(c);
        int g = guard(c);
        if (g) return g;
        double aux[3];
        for (int d = 0; d < 3; ++d) {
            aux[d] = a * (c->vimp[d] - b * c->qr[d]);
            c->qr[d] = c->qr[d] + 3.0 * aux[d] - cc * c->vimp[d];
        }
        if (j == 0)
            for (int d = 0; d < 3; ++d) {
                c->rimpold[2][d] = c->rimpold[1][d]; c->rimpold[1][d] = c->rimpold[0][d]; c->rimpold[0][d] = c->rimp[d];
            }
        for (int d = 0; d < 3; ++d) c->rimp[d] = c->rimp[d] + dt * aux[d];
        for (int d = 0; d < 3; ++d) {
            aux[d] = a * (c->aimp[d] - b * c->qv[d]);
            c->qv[d] = c->qv[d] + 3.0 * aux[d] - cc * c->aimp[d];
        }
        if (j == 0)
            for (int d = 0; d < 3; ++d) {
                c->aimpold[1][d] = c->aimpold[0][d]; c->aimpold[0][d] = c->aimp[d];
                c->vimpold[2][d] = c->vimpold[1][d]; c->vimpold[1][d] = c->vimpold[0][d]; c->vimpold[0][d] = c->vimp[d];
            }
        for (int d = 0; d < 3; ++d) c->vimp[d] = c->vimp[d] + dt * aux[d];
        if (j <= 2) potential_update(c, c->rimp);
    }
    for (int i = 0; i < 3; ++i) { c->ioldp[i] = i + 1; c->ioldr[i] = i + 1; c->ioldv[i] = i + 1; }
    for (int i = 0; i < 2; ++i) { c->ioldh[i] = i + 1; c->iolda[i] = i + 1; }
    return 0;
}

/* steppc (steppc.f90:1-178) */
int orc_steppc(orc_ctx *c, double dt, double *errHe, double *errimp) {
    const double c112 = 112.0 / 121.0, c9 = 9.0 / 121.0, c4o3 = 4.0 / 3.0;
    size_t n = c->n;
    rhs(c);
    double *po3 = c->psiold[c->ioldp[2] - 1];
    double *h1 = c->hpsiold[c->ioldh[0] - 1], *h2 = c->hpsiold[c->ioldh[1] - 1];
    for (size_t i = 0; i < 2 * n; ++i) {
        double f = c->f4c[i];
        double P = po3[i] + c4o3 * dt * (2.0 * f - h1[i] + 2.0 * h2[i]);
        po3[i] = c->psi[i];
        c->psi[i] = P - c112 * c->pc[i];
        c->pc[i] = P;
        h2[i] = f;
    }
    den_from_psi(c);
    { int t = c->ioldh[1]; c->ioldh[1] = c->ioldh[0]; c->ioldh[0] = t; }
    double auxr[3];
    /* the R predictor uses R(n-3) and the v history (steppc.f90:77-83),
     * the corrector rimpold(ioldr(2)) (steppc.f90:148-149) */
    int r2 = c->ioldr[1] - 1, r3 = c->ioldr[2] - 1;
    int v1 = c->ioldv[0] - 1, v2 = c->ioldv[1] - 1, v3 = c->ioldv[2] - 1;
    int a1 = c->iolda[0] - 1, a2 = c->iolda[1] - 1;
    for (int d = 0; d < 3; ++d)
        auxr[d] = c->rimpold[r3][d] + c4o3 * dt * (2.0 * c->vimp[d] - c->vimpold[v1][d] + 2.0 * c->vimpold[v2][d]);
    for (int d = 0; d < 3; ++d) { c->rimpold[r3][d] = auxr[d] - c112 * c->pcr[d]; c->pcr[d] = auxr[d]; }
    for (int d = 0; d < 3; ++d)
        auxr[d] = c->vimpold[v3][d] + c4o3 * dt * (2.0 * c->aimp[d] - c->aimpold[a1][d] + 2.0 * c->aimpold[a2][d]);
    for (int d = 0; d < 3; ++d) { c->vimpold[v3][d] = auxr[d] - c112 * c->pcv[d]; c->pcv[d] = auxr[d]; }
    for (int d = 0; d < 3; ++d) c->aimpold[a2][d] = c->aimp[d];
    { int t = c->iolda[1]; c->iolda[1] = c->iolda[0]; c->iolda[0] = t; }
    a1 = c->iolda[0] - 1; a2 = c->iolda[1] - 1;
    double rp[3] = {c->rimpold[r3][0], c->rimpold[r3][1], c->rimpold[r3][2]};
    potential_update(c, rp);
    rhs(c);
    double *po2 = c->psiold[c->ioldp[1] - 1];
    h1 = c->hpsiold[c->ioldh[0] - 1]; h2 = c->hpsiold[c->ioldh[1] - 1];
    for (size_t i = 0; i < 2 * n; ++i) {
        double C = 0.125 * (9.0 * po3[i] - po2[i] + 3.0 * dt * (c->f4c[i] + 2.0 * h1[i] - h2[i]));
        c->pc[i] = c->pc[i] - C;
        c->psi[i] = C + c9 * c->pc[i];
    }
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) s += hypot(c->pc[2 * i], c->pc[2 * i + 1]);
    *errHe = c9 * s / (double)n;
    den_from_psi(c);
    { int t = c->ioldp[2]; c->ioldp[2] = c->ioldp[1]; c->ioldp[1] = c->ioldp[0]; c->ioldp[0] = t; }
    for (int d = 0; d < 3; ++d)
        auxr[d] = 0.125 * (9.0 * c->rimp[d] - c->rimpold[r2][d] + 3.0 * dt * (c->vimpold[v3][d] + 2.0 * c->vimp[d] - c->vimpold[v1][d]));
    double e = 0.0;
    for (int d = 0; d < 3; ++d) {
        c->pcr[d] = c->pcr[d] - auxr[d];
        c->rimpold[r3][d] = c->rimp[d];
        c->rimp[d] = auxr[d] + c9 * c->pcr[d];
    }
    for (int d = 0; d < 3; ++d) e += fabs(c9 * c->pcr[d]);
    *errimp = e * 0.3333333333;
    { int t = c->ioldr[2]; c->ioldr[2] = c->ioldr[1]; c->ioldr[1] = c->ioldr[0]; c->ioldr[0] = t; }
    for (int d = 0; d < 3; ++d)
        auxr[d] = 0.125 * (9.0 * c->vimp[d] - c->vimpold[v2][d] + 3.0 * dt * (c->aimp[d] + 2.0 * c->aimpold[a1][d] - c->aimpold[a2][d]));
    for (int d = 0; d < 3; ++d) {
        c->pcv[d] = c->pcv[d] - auxr[d];
        c->vimpold[v3][d] = c->vimp[d];
        c->vimp[d] = auxr[d] + c9 * c->pcv[d];
    }
    { int t = c->ioldv[2]; c->ioldv[2] = c->ioldv[1]; c->ioldv[1] = c->ioldv[0]; c->ioldv[0] = t; }
    return 0;
}

int orc_set_aimp(orc_ctx *c, const double a[3]) {
    for (int d = 0; d < 3; ++d) c->aimp[d] = a[d];
    return 0;
}

int orc_init(orc_ctx *c) {
    potential_update(c, c->rimp);
    return 0;
}

int orc_evolve(orc_ctx *c, double dt, int nsteps, double *traj) {
    for (int s = 0; s < nsteps; ++s) {
        double eh = 0.0, ei = 0.0;
        int rc;
        c->iter_since_set++;
        if (c->iter_since_set <= 3) rc = orc_steprk(c, dt);
        else rc = orc_steppc(c, dt, &eh, &ei);
        if (rc) return rc;
        potential_update(c, c->rimp);
        if (traj) {
            double *t = traj + (size_t)s * HETD_TRAJ_STRIDE;
            t[0] = eh; t[1] = ei;
            for (int d = 0; d < 3; ++d) { t[2 + d] = c->rimp[d]; t[5 + d] = c->vimp[d]; t[8 + d] = c->aimp[d]; }
        }
    }
    return 0;
}

/* energy (energy.f90:12-144) */
int orc_energy(orc_ctx *c, hetd_energies *e) {
    const hetd_params *p = &c->p;
    size_t n = c->n;
    deriv_field(c, 1, 1, 1, c->psi, c->s1c);
    deriv_field(c, 1, 2, 1, c->psi, c->s2c);
    deriv_field(c, 1, 3, 1, c->psi, c->s3c);
    double aux1 = 0.5 * p->cp4, aux2 = (1.0 / 3.0) * p->cpp4;
    double ekin = 0.0, elj = 0.0, ecor = 0.0;
    for (size_t i = 0; i < n; ++i) {
        double a3 = c->den[i], a4 = c->dencg[i];
        double a5 = a3 * a4 * a4 * (aux1 + aux2 * a4);
        double g1 = hypot(c->s1c[2 * i], c->s1c[2 * i + 1]);
        double g2 = hypot(c->s2c[2 * i], c->s2c[2 * i + 1]);
        double g3 = hypot(c->s3c[2 * i], c->s3c[2 * i + 1]);
        ekin = ekin + g1 * g1 + g2 * g2 + g3 * g3;
        elj = elj + c->delj4[i] * a3;
        ecor = ecor + a5;
    }
    double ealphas = 0.0;
    if (!p->lsolid && p->core == HETD_CORE_OTC) {
        for (size_t i = 0; i < n; ++i)
            ealphas += c->falfs[i] * (c->dxden[i] * c->intx[i] + c->dyden[i] * c->inty[i] + c->dzden[i] * c->intz[i]);
        ealphas = -p->h2o2m4 * 0.5 * p->alphas * ealphas * c->dxyz;
    }
    double esolid = 0.0;
    if (p->lsolid) {
        for (size_t i = 0; i < n; ++i) esolid += c->den[i] * (1.0 + tanh(p->beta_solid * (c->den[i] - p->den_m)));
        esolid = p->c_solid * esolid * c->dxyz;
    }
    e->ekin4 = p->h2o2m4 * ekin * c->dxyz;
    e->elj4 = 0.5 * elj * c->dxyz;
    e->ecor4 = ecor * c->dxyz;
    e->esolid = esolid;
    e->ealphas = ealphas;
    e->etot4 = e->ekin4 + e->elj4 + e->ecor4 + e->esolid;
    if (!p->lsolid && p->core == HETD_CORE_OTC) e->etot4 += ealphas;
    e->ekinx = 0.5 * c->mAg * (c->vimp[0] * c->vimp[0] + c->vimp[1] * c->vimp[1] + c->vimp[2] * c->vimp[2]);
    double ehx = 0.0;
    for (size_t i = 0; i < n; ++i) ehx += c->uimp[i] * c->den[i];
    e->eHeX = ehx * c->dxyz;
    e->eimpu = e->ekinx + e->eHeX;
    e->etot = e->etot4 + e->eimpu;
    return 0;
}

int orc_norm(orc_ctx *c, double *norm) {
    double s = 0.0;
    for (size_t i = 0; i < c->n; ++i) s += c->den[i];
    *norm = s * c->dxyz;
    return 0;
}

int orc_get_field(orc_ctx *c, int which, double *out) {
    size_t n = c->n, nh = c->nh;
    switch (which) {
    case HETD_F_PSI: memcpy(out, c->psi, sizeof(double) * 2 * n); return 0;
    case HETD_F_DEN: memcpy(out, c->den, sizeof(double) * n); return 0;
    case HETD_F_POT4: memcpy(out, c->pot4, sizeof(double) * n); return 0;
    case HETD_F_DENCG: memcpy(out, c->dencg, sizeof(double) * n); return 0;
    case HETD_F_DELJ4: memcpy(out, c->delj4, sizeof(double) * n); return 0;
    case HETD_F_CCOR: memcpy(out, c->ccor, sizeof(double) * n); return 0;
    case HETD_F_DXDEN: memcpy(out, c->dxden, sizeof(double) * n); return 0;
    case HETD_F_DYDEN: memcpy(out, c->dyden, sizeof(double) * n); return 0;
    case HETD_F_DZDEN: memcpy(out, c->dzden, sizeof(double) * n); return 0;
    case HETD_F_UIMP: memcpy(out, c->uimp, sizeof(double) * n); return 0;
    case HETD_F_PENALTY: memcpy(out, c->penalty, sizeof(double) * n); return 0;
    case HETD_F_UALPHAS: memcpy(out, c->ualphas, sizeof(double) * n); return 0;
    case HETD_F_DENALF: memcpy(out, c->denalf, sizeof(double) * n); return 0;
    case HETD_F_FVLJ: memcpy(out, c->fvlj, sizeof(double) * nh); return 0;
    case HETD_F_WCGK: memcpy(out, c->wcgk, sizeof(double) * nh); return 0;
    case HETD_F_PC: memcpy(out, c->pc, sizeof(double) * 2 * n); return 0;
    case HETD_F_Q: memcpy(out, c->q, sizeof(double) * 2 * n); return 0;
    case HETD_F_FDEN: memcpy(out, c->fden, sizeof(double) * 2 * nh); return 0;
    default: return HETD_ERR_ARG;
    }
}

int orc_get_imp(orc_ctx *c, double r[3], double v[3], double a[3]) {
    for (int d = 0; d < 3; ++d) { r[d] = c->rimp[d]; v[d] = c->vimp[d]; a[d] = c->aimp[d]; }
    return 0;
}

int orc_fft_r2c(int nx, int ny, int nz, const double *in, double *out) {
    fftplan *px = plan_make(nx), *py = plan_make(ny), *pz = plan_make(nz);
    r2c3d(nx, ny, nz, in, out, px, py, pz);
    plan_free(px); plan_free(py); plan_free(pz);
    return 0;
}

int orc_fft_c2r(int nx, int ny, int nz, const double *in, double *out) {
    fftplan *px = plan_make(nx), *py = plan_make(ny), *pz = plan_make(nz);
    double *w = zalloc(2 * (size_t)(nx / 2 + 1) * ny * nz);
    c2r3d(nx, ny, nz, in, out, 1.0, px, py, pz, w);
    free(w);
    plan_free(px); plan_free(py); plan_free(pz);
    return 0;
}

/* The energy block's diagnostics (DFT4He3d.f90:868-947): r_cm (r_cm.f90:23-55),
 * <Lx,Ly,Lz> about it with complex sums as written, <x2,y2,z2> and the norm. */
int orc_moments(orc_ctx *c, double npart, double out[10]) {
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    double xcm = 0.0, ycm = 0.0, zcm = 0.0;
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                const double aux = c->den[ix + (size_t)nx * (iy + (size_t)ny * iz)];
                xcm = xcm + aux * c->x[ix];
                ycm = ycm + aux * c->y[iy];
                zcm = zcm + aux * c->z[iz];
            }
    xcm = xcm * c->dxyz / npart;
    ycm = ycm * c->dxyz / npart;
    zcm = zcm * c->dxyz / npart;
    deriv_field(c, 1, 1, 1, c->psi, c->s1c);
    deriv_field(c, 1, 2, 1, c->psi, c->s2c);
    deriv_field(c, 1, 3, 1, c->psi, c->s3c);
    /* caux = sum i conj(psi) W, W per component as DFT4He3d.f90:894-895,907-908,920-921 */
    double cre[3] = {0, 0, 0}, cim[3] = {0, 0, 0};
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                const size_t i = ix + (size_t)nx * (iy + (size_t)ny * iz);
                const double xr = c->x[ix] - xcm, yr = c->y[iy] - ycm, zr = c->z[iz] - zcm;
                const double pr = c->psi[2 * i], pim = -c->psi[2 * i + 1]; /* conj(psi) */
                const double d1r = c->s1c[2 * i], d1i = c->s1c[2 * i + 1];
                const double d2r = c->s2c[2 * i], d2i = c->s2c[2 * i + 1];
                const double d3r = c->s3c[2 * i], d3i = c->s3c[2 * i + 1];
                double wr[3], wi[3];
                wr[2] = yr * d1r - xr * d2r; wi[2] = yr * d1i - xr * d2i; /* z */
                wr[1] = xr * d3r - zr * d1r; wi[1] = xr * d3i - zr * d1i; /* y */
                wr[0] = zr * d2r - yr * d3r; wi[0] = zr * d2i - yr * d3i; /* x */
                for (int k = 0; k < 3; ++k) {
                    /* i * conj(psi) * W */
                    const double tr = pr * wr[k] - pim * wi[k], ti = pr * wi[k] + pim * wr[k];
                    cre[k] += -ti;
                    cim[k] += tr;
                }
            }
    double a1 = 0.0, a2 = 0.0, a3 = 0.0, s = 0.0;
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                const double d = c->den[ix + (size_t)nx * (iy + (size_t)ny * iz)];
                a1 = a1 + d * c->x[ix] * c->x[ix];
                a2 = a2 + d * c->y[iy] * c->y[iy];
                a3 = a3 + d * c->z[iz] * c->z[iz];
                s += d;
            }
    (void)cim;
    out[0] = xcm; out[1] = ycm; out[2] = zcm;
    out[3] = a1 * c->dxyz; out[4] = a2 * c->dxyz; out[5] = a3 * c->dxyz;
    out[6] = s * c->dxyz;
    for (int k = 0; k < 3; ++k) out[7 + k] = cre[k] * c->dxyz; /* real(caux)*dxyz */
    return 0;
}

int orc_deriv(int nx, int ny, int nz, int npd, int icon, int id, int iv, double h,
              int is_complex, const double *f, double *df) {
    return deriv3d(nx, ny, nz, npd, icon, id, iv, h, is_complex, f, df);
}
