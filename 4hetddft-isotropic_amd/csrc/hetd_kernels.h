// hetd_kernels.h -- kernel argument blocks and host-side launchers.
#pragma once

#include "hetd_dev.h"

namespace hetd {

// ---------------------------------------------------------------- FFT passes

enum XMode { XMODE_R2C = 0, XMODE_C2R = 1, XMODE_CG_S = 2, XMODE_POT = 3 };
enum ZMode { ZMODE_FWD = 0, ZMODE_INV = 1, ZMODE_CONV = 2 };

template <typename T> struct XPassArgs {
    FftAxis ax;
    long nlines;  // ny*nz
    int ldh;      // row stride of the half spectra (>= nx/2+1, padded for aligned rows)
    int ppb;      // line pairs per block (set by the launcher)
    int mode;
    int split;    // register passes: exchange real and imaginary parts in turn (half the LDS)
    T scale;      // applied after an inverse transform (renor)
    const T* rin;        // R2C: input; CG_S: den; POT: den
    const T* rin2;       // POT: dencg
    const T* rin3;       // POT: delj4 (null: already in the spectrum)
    const cplx<T>* cin;  // inverse input half spectrum
    T* rout;             // C2R output; CG_S: dencg; POT: pot4 (null: not stored)
    T* rout2;            // POT: correlation term sto1 (optional)
    cplx<T>* cout;       // forward output half spectrum
    T cp4, cpp4, a0, a1;
    int lsolid;
    T csol, beta, den_m;
    // fold: CG_S writes the local part of pot4, dencg^2 (a0 + a1 dencg)
    // [+ penalty], instead of dencg, and POT reads it back as rin2 (no den)
    int fold;
};

// Pointwise work of the x passes after an inverse transform, on element i of
// a real line (shared by the Stockham and register passes).
// CG_S: g = dencg; stores dencg (or the folded local part); returns
// s = den dencg (cp4 + cpp4 dencg) (poten.f90:34-38, 62-64).
template <typename T> __device__ __forceinline__ T cgs_point(const XPassArgs<T>& a, long i, T g) {
    const T d = a.rin[i];
    if (a.fold) {
        T u = g * g * (a.a0 + a.a1 * g);
        if (a.lsolid) u = u + solid_penalty<T>(d, a.csol, a.beta, a.den_m);
        a.rout[i] = u;
    } else {
        a.rout[i] = g;
    }
    return d * g * (a.cp4 + a.cpp4 * g);
}

// POT: cv = c2r of the spectrum; pot4 = delj4 + dencg^2 (a0 + a1 dencg) +
// sto1 [+ penalty] (poten.f90:80-94, 110-114), or cv + the folded local part
template <typename T> __device__ __forceinline__ void pot_point(const XPassArgs<T>& a, long i, T cv) {
    T u;
    if (a.fold) {
        u = a.rin2[i] + cv;
    } else {
        const T g = a.rin2[i];
        u = (a.rin3 ? a.rin3[i] : (T)0) + g * g * (a.a0 + a.a1 * g) + cv;
        if (a.lsolid) u = u + solid_penalty<T>(a.rin[i], a.csol, a.beta, a.den_m);
    }
    if (a.rout) a.rout[i] = u;
    if (a.rout2) a.rout2[i] = cv;
}

template <typename T> struct YPassArgs {
    FftAxis ax;
    int nxh, ldh, tx, inv;  // kx count, row stride
    int split;              // register passes: exchange real and imaginary parts in turn (half the LDS)
    cplx<T>* buf[4];
};

template <typename T> struct ZPassArgs {
    FftAxis ax;
    int nxh, ldh, ny, tx, mode, nk;  // kx count, row stride
    const cplx<T>* in[4];
    cplx<T>* out[4][3];
    const T* kern[3];
    cplx<T>* fwd_out;  // optional copy of the forward spectrum (CONV)
    cplx<T>* kout;     // optional kkern * forward spectrum, not inverted (CONV)
    const T* kkern;
    const cplx<T>* addin;  // optional z-spectrum added to kern[0] * forward before its inverse (CONV)
};

template <typename T> hipError_t launch_xpass(XPassArgs<T> a, hipStream_t st);
template <typename T> hipError_t launch_ypass(YPassArgs<T> a, int nz, int nbuf, hipStream_t st);
template <typename T> hipError_t launch_zpass(ZPassArgs<T> a, int nbuf, hipStream_t st);
// the LDS Stockham passes can transform lines of this axis (block shapes that
// fit the thread count and the LDS, in FP64 and FP32)
bool stockham_plan_ok(const FftAxis& ax);
// register-resident passes for axis lengths 16, 64, 256 (fft_reg.hip)
bool reg_fft_supported(int n);
template <typename T> hipError_t launch_xreg(XPassArgs<T> a, hipStream_t st);
template <typename T> hipError_t launch_yreg(YPassArgs<T> a, int nz, int nbuf, hipStream_t st);
template <typename T> hipError_t launch_zreg(ZPassArgs<T> a, int nbuf, hipStream_t st);
// register-resident x/y/z passes for N = P*Q (fft_reg2.hip); z: FWD, INV, CONV with nk = 1
bool reg2_supported(int n);
template <typename T> hipError_t launch_x2(XPassArgs<T> a, hipStream_t st);
template <typename T> hipError_t launch_y2(YPassArgs<T> a, int nz, int nbuf, hipStream_t st);
template <typename T> hipError_t launch_z2(ZPassArgs<T> a, int nbuf, hipStream_t st);

// --------------------------------------------------------------- grid info

// Stencil weights already divided by h**id (init_deriv_parallel.f90:415-431).
struct StencilCoef {
    int npd, m, icon;
    double c1[3][kMaxNpd];  // first derivative per axis (weights / h)
    double c2[3][kMaxNpd];  // second derivative per axis (weights / h^2)
    double w1[kMaxNpd], w2[kMaxNpd];  // the weights themselves (centred, symmetric / antisymmetric)
    double h[3];
};

struct Grid {
    int nx, ny, nz;
    long n;
    double dxyz;
    const double* x;  // coordinates x(i) = -xmax + hx*(i-1)
    const double* y;
    const double* z;
    const double* tcx;  // tanh((|x|-txmean)/txsurf) per axis (DFT4He3d.f90:689-699)
    const double* tcy;
    const double* tcz;
    double lambdah;
    const double* potion;  // impurity table (potenimpini.f90)
    int npot;
    double delinter;
};

// -------------------------------------------------------- update kernels

enum StageKind { STAGE_RK = 0, STAGE_PRED = 1, STAGE_CORR = 2 };

struct UpdateArgs {
    Grid g;
    StencilCoef sc;
    double h2o2m4;
    int kind;
    int tile;  // tile variant of the tiled update kernels (0: the default or HETD_TILE_UPDATE)
    int rk_j;  // RK stage 0..3
    double dt;
    double ra, rb, rc;  // Gill coefficients of the stage
    const double2* psi;  // stencil input
    const double* pot4;
    const ImpState* imp;
    double2* psi_out;
    double* den;
    double2* q;       // RK
    double2* hnew;    // RK j=0: f -> hpsiold(1); PRED: f -> hpsiold(ioldh(2))
    const double2* po3;  // PRED / CORR: psiold(ioldp(3))
    const double2* po2;  // CORR: psiold(ioldp(2))
    const double2* h1;   // hpsiold(ioldh(1))
    const double2* h2;   // hpsiold(ioldh(2))
    double2* pc;
    double* partial;  // CORR: per-block sum |pc|
    int* guard;       // RK: set if den not > 0
    int force;           // also sum grad(|psi|^2) u_imp (the force of the input state), icon 13/8
    // mixed precision when denf is set: FP32 mean field pot4f in, FP32 den out
    // (no FP64 den), and hnew/h1/h2 hold complex64 RHS history
    const float* pot4f;
    float* denf;
};

hipError_t launch_update(const UpdateArgs& a, hipStream_t st, int* nblocks_out);

// --------------------------------------------------------- derivatives

struct DerivArgs {
    int nx, ny, nz;
    StencilCoef sc;
    int id, iv, cplx;
    const double* f;
    double* df;
};
hipError_t launch_deriv(const DerivArgs& a, hipStream_t st);

// ---------------------------------------------------------- reductions

struct ForceArgs {
    Grid g;
    StencilCoef sc;
    const double* den;
    const ImpState* imp;
    double* partial;  // 3 per block
};
hipError_t launch_force_partial(const ForceArgs& a, hipStream_t st, int* nblocks);

enum { EN_KIN = 0, EN_LJ, EN_COR, EN_SOLID, EN_HEX, EN_ALPHA, EN_NORM, EN_NCOMP };
struct EnergyArgs {
    Grid g;
    StencilCoef sc;
    const double2* psi;
    const double* den;
    const double* dencg;
    const double* delj4;
    const ImpState* imp;
    double aux1, aux2, beta, den_m;
    int lsolid, otc;
    const double* falfs;
    const double* dxden;
    const double* dyden;
    const double* dzden;
    const double* intx;
    const double* inty;
    const double* intz;
    double* partial;  // EN_NCOMP per block
    const float* dencgf;  // mixed precision fields (else dencg / delj4)
    const float* delj4f;
    int mom;          // 0: energy terms; 1: density moments; 2: angular momentum about cm
    double cm[3];
};
hipError_t launch_energy_partial(const EnergyArgs& a, hipStream_t st, int* nblocks);

// The diagnostics of the driver's energy block at one point (psi, its
// gradient g1..g3, coordinates x, y, z), summed in place of the energy terms:
//   mom 1: den x, den y, den z (r_cm.f90:38-50), den x^2, den y^2, den z^2
//          (DFT4He3d.f90:935-947), den;
//   mom 2: Re(i conj(psi) W) with W = (z-zcm) dpsi/dy - (y-ycm) dpsi/dz, and
//          the two rotations of it: <Lx, Ly, Lz> (DFT4He3d.f90:883-925).
__device__ __forceinline__ void moment_terms(const EnergyArgs& a, double* v, double2 ps, double2 g1, double2 g2,
                                             double2 g3, double x, double y, double z) {
    if (a.mom == 1) {
        const double d = abs2(ps);
        v[0] += d * x; v[1] += d * y; v[2] += d * z;
        v[3] += d * (x * x); v[4] += d * (y * y); v[5] += d * (z * z);
        v[6] += d;
    } else {
        const double xr = x - a.cm[0], yr = y - a.cm[1], zr = z - a.cm[2];
        const double2 wx = make_double2(zr * g2.x - yr * g3.x, zr * g2.y - yr * g3.y);
        const double2 wy = make_double2(xr * g3.x - zr * g1.x, xr * g3.y - zr * g1.y);
        const double2 wz = make_double2(yr * g1.x - xr * g2.x, yr * g1.y - xr * g2.y);
        // Re(i conj(psi) W) = -(psi.x W.y - psi.y W.x)
        v[0] -= ps.x * wx.y - ps.y * wx.x;
        v[1] -= ps.x * wy.y - ps.y * wy.x;
        v[2] -= ps.x * wz.y - ps.y * wz.x;
    }
}

// LDS-tiled versions (npd 9 and 13); one partial per tile
bool tiled_supported(const StencilCoef& sc);
hipError_t launch_update_tiled(const UpdateArgs& a, hipStream_t st, int* nblocks);
hipError_t launch_force_tiled(const ForceArgs& a, hipStream_t st, int* nblocks);
hipError_t launch_energy_tiled(const EnergyArgs& a, hipStream_t st, int* nblocks);

// out[c] = sum_b partial[b*ncomp + c], fixed order (bitwise reproducible)
hipError_t launch_reduce_final(const double* partial, int nblocks, int ncomp, double* out, hipStream_t st);
// partial[b] = sum over a fixed slice of u * |psi|^2, nblocks blocks (<= 1024)
hipError_t launch_dot_uext(const double2* psi, const double* u, long n, int nblocks, double* partial, hipStream_t st);

// ------------------------------------------------------------ impurity ODE

struct ImpRkArgs { ImpState* imp; int j; double dt, a, b, c; int set_rpot; };
struct ImpPcArgs { ImpState* imp; double dt; int r1, r2, r3, v1, v2, v3, a1, a2; int set_rpot; };
hipError_t launch_imp_rk(const ImpRkArgs& a, hipStream_t st);
hipError_t launch_imp_pred(const ImpPcArgs& a, hipStream_t st);   // sets rpot = predicted R
hipError_t launch_imp_corr(const ImpPcArgs& a, hipStream_t st);   // set_rpot: rpot = corrected R
// F = -sum*dxyz, a = F/mass; from reduced sums in imp->F (overwritten)
// F = -sum(partial[b*stride+off+d]) dxyz, a = F/mass (also into rec_a[0..2] if set)
hipError_t launch_force_final(ImpState* imp, const double* partial, int nblocks, int stride, int off, double dxyz,
                              double mass, double* rec_a, hipStream_t st);
hipError_t launch_errhe_final(ImpState* imp, const double* partial, int nblocks, int stride, double npoints,
                              hipStream_t st);
hipError_t launch_traj_record(const ImpState* imp, double* rec, hipStream_t st);

// ---------------------------------------------------------- pointwise

hipError_t launch_den_from_psi(const double2* psi, double* den, long n, hipStream_t st);
hipError_t launch_den32_from_psi(const double2* psi, float* den, long n, hipStream_t st);
hipError_t launch_to_double(const float* a, double* out, long n, hipStream_t st);
hipError_t launch_to_float(const double* a, float* out, long n, hipStream_t st);
hipError_t launch_uimp_field(const Grid& g, const ImpState* imp, double* out, hipStream_t st);
hipError_t launch_penalty_field(const double* den, double csol, double beta, double den_m, double* out, long n,
                                hipStream_t st);
// OTC helpers (term_alfa.f90)
hipError_t launch_falfs(const double* denalf, double den0s, double* falfs, long n, hipStream_t st);
hipError_t launch_mul(const double* a, const double* b, double* out, long n, hipStream_t st);
hipError_t launch_dot3(const double* a1, const double* b1, const double* a2, const double* b2,
                       const double* a3, const double* b3, double* out, long n, hipStream_t st);
hipError_t launch_ualpha_final(double* ua, const double* s4, const double* s5, const double* s6,
                               double den0s, double h2o2m4, double alphas, double* pot4, float* pot4f, long n, hipStream_t st);
hipError_t launch_scale(double* a, double s, long n, hipStream_t st);

}  // namespace hetd
