// api.hip -- the C ABI of include/hetd.h: device context, call sequencing of
// poten / steprk / steppc / energy on one HIP stream, and host copies.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/hetd.h"
#include "hetd_kernels.h"
#include "impurity.hpp"
#include "setup.hpp"

using namespace hetd;

struct hetd_ctx {
    hetd_params p{};
    int nx = 0, ny = 0, nz = 0, nxh = 0;
    int ldh = 0;    // row stride of half spectra: nxh rounded up to 16 (128-byte rows)
    long nhp = 0;   // ldh*ny*nz, the allocated half-spectrum size
    long n = 0, nh = 0;
    double hx = 0, hy = 0, hz = 0, dxyz = 0, mAg = 0, renor = 0;
    bool otc = false;
    hipStream_t st = nullptr;
    std::string err;
    std::vector<std::pair<void*, size_t>> allocs;  // every device buffer and its size
    // grid tables
    double *d_x = nullptr, *d_y = nullptr, *d_z = nullptr, *d_tcx = nullptr, *d_tcy = nullptr, *d_tcz = nullptr;
    double2* d_tw[3] = {nullptr, nullptr, nullptr};
    float2* d_tw32[3] = {nullptr, nullptr, nullptr};
    FftAxis ax[3];
    StencilCoef sc{};
    Grid grid{};
    // state (psi family rotates by pointer)
    double2 *psi = nullptr, *psiold[3] = {}, *spare = nullptr, *hpsiold[2] = {}, *pc = nullptr, *q = nullptr;
    double *den = nullptr, *pot4 = nullptr, *dencg = nullptr, *delj4 = nullptr, *tmp = nullptr;
    double2 *A = nullptr, *B = nullptr, *Cs = nullptr;
    double *fvlj = nullptr, *wcgk = nullptr, *kalfs = nullptr;
    // alpha_s term (OTC)
    double *dxden = nullptr, *dyden = nullptr, *dzden = nullptr, *denalf = nullptr, *falfs = nullptr;
    double *intx = nullptr, *inty = nullptr, *intz = nullptr, *ualphas = nullptr;
    double *s1 = nullptr, *s2 = nullptr, *s3 = nullptr, *s4 = nullptr, *s5 = nullptr, *s6 = nullptr;
    // mixed precision (HETD_MIXED): FP32 density, mean field and FFT work
    bool mixed = false;
    // delj4 of the last poten: DELJ4_DEN = not formed (make it from den),
    // DELJ4_OK = formed, DELJ4_SPEC = its z spectrum waits in B
    int delj4_state = 0;
    // force of the current den not formed yet: the next update kernel sums it
    // (fused), before anything reads the acceleration
    bool force_pending = false;
    bool den64_ok = true;  // mixed: the FP64 den array matches psi (updates write only FP32 den)
    bool dencg_ok = true;  // dencg holds dencg (false: the local part of pot4, see poten_t)
    double* pend_rec = nullptr;  // trajectory row waiting for that acceleration
    float *den32 = nullptr, *pot4_32 = nullptr, *dencg32 = nullptr, *delj4_32 = nullptr, *tmp32 = nullptr;
    float2 *A32 = nullptr, *B32 = nullptr, *C32 = nullptr;
    float *fvlj32 = nullptr, *wcgk32 = nullptr, *kalfs32 = nullptr;
    double* potion = nullptr;
    long partial_len = 0;    // doubles in partial
    double* uext = nullptr;  // external potential of the energy (lsurf), or none
    bool uext_on = false;
    int npot = 0;
    double delinter = 0;
    ImpState* imp = nullptr;
    double* partial = nullptr;
    double* red = nullptr;
    int* guard = nullptr;
    double* traj = nullptr;
    long traj_cap = 0;
    double* dtmp[2] = {nullptr, nullptr};  // scratch of hetd_deriv / hetd_fft_* (allocated on first use)
    int ioldp[3] = {1, 2, 3}, ioldh[2] = {1, 2}, ioldr[3] = {1, 2, 3}, ioldv[3] = {1, 2, 3}, iolda[2] = {1, 2};
    long iter_since_set = 0;
    bool ready = false;  // hetd_init completed
    bool have_state = false;
    // predictor-corrector steps replayed from a HIP graph (see step_graph)
    bool use_graphs = true;
    bool xsplit = false;  // x and y register passes with split LDS exchanges (HETD_XSPLIT)
    int upd_tile = 0;     // update-kernel tile variant (0: default / HETD_TILE_UPDATE), set by the autotuner
    hipGraphExec_t gexec = nullptr;
    int gper = 0;        // steps per replay (the period of the buffer rotation)
    double gdt = 0.0;
    std::vector<long> gsig;  // host state the replay starts from (and ends in)
    // kernel-class timing (hetd_timing)
    int tclass = 0;
    std::vector<hipEvent_t> tev;
    size_t tused = 0;
    long tlaunch = 0;
    double tbytes = 0;
};

// ------------------------------------------------------------------ helpers

static int fail(hetd_ctx* c, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HCHK(c, expr)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) return fail((c), HETD_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

template <typename P> static hipError_t dalloc(hetd_ctx* c, P** p, size_t bytes) {
    void* v = nullptr;
    hipError_t e = hipMalloc(&v, bytes ? bytes : 16);
    if (e == hipSuccess) {
        c->allocs.push_back({v, bytes ? bytes : 16});
        *p = reinterpret_cast<P*>(v);
    }
    return e;
}

static void fill_grid(hetd_ctx* c) {
    Grid& g = c->grid;
    g.nx = c->nx; g.ny = c->ny; g.nz = c->nz; g.n = c->n; g.dxyz = c->dxyz;
    g.x = c->d_x; g.y = c->d_y; g.z = c->d_z;
    g.tcx = c->d_tcx; g.tcy = c->d_tcy; g.tcz = c->d_tcz;
    g.lambdah = c->p.Lambdah;
    g.potion = c->potion; g.npot = c->npot; g.delinter = c->delinter;
}


// ------------------------------------------------------- kernel-class timing

static void tbeg(hetd_ctx* c, int k) {
    if (c->tclass != k) return;
    while (c->tused + 2 > c->tev.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) { c->tclass = 0; return; }
        c->tev.push_back(e);
    }
    (void)hipEventRecord(c->tev[c->tused], c->st);
}

static void tend(hetd_ctx* c, int k, double bytes) {
    if (c->tclass != k) return;
    (void)hipEventRecord(c->tev[c->tused + 1], c->st);
    c->tused += 2;
    c->tlaunch += 1;
    c->tbytes += bytes;
}

// launches with their algorithmic HBM bytes (one read of each input, one
// write of each output; half spectra have nxh*ny*nz points)
template <typename T> static hipError_t run_x(hetd_ctx* c, XPassArgs<T> x) {
    x.ldh = c->ldh;
    const double n = (double)c->n, nh = (double)c->nh, s = (double)sizeof(T);
    double b = 0;
    switch (x.mode) {
    case XMODE_R2C: b = s * n + 2 * s * nh; break;
    case XMODE_C2R: b = 2 * s * nh + s * n; break;
    case XMODE_CG_S: b = 4 * s * nh + 2 * s * n; break;
    default:  // POT: spectrum, rin2 (dencg or the folded local part), [delj4], [den]; pot4, [sto1]
        b = 2 * s * nh + s * n + (x.rin3 ? s * n : 0) + (!x.fold && x.lsolid ? s * n : 0) +
            (x.rout ? s * n : 0) + (x.rout2 ? s * n : 0);
        break;
    }
    tbeg(c, HETD_K_XPASS);
    x.split = c->xsplit ? 1 : 0;
    hipError_t e = reg_fft_supported(x.ax.n) ? launch_xreg(x, c->st)
                   : reg2_supported(x.ax.n)  ? launch_x2(x, c->st)
                                             : launch_xpass(x, c->st);
    tend(c, HETD_K_XPASS, b);
    return e;
}

template <typename T> static hipError_t run_y(hetd_ctx* c, YPassArgs<T> y, int nbuf) {
    y.ldh = c->ldh;
    y.split = c->xsplit ? 1 : 0;
    tbeg(c, HETD_K_YPASS);
    hipError_t e = reg_fft_supported(y.ax.n) ? launch_yreg(y, c->nz, nbuf, c->st)
                   : reg2_supported(y.ax.n)  ? launch_y2(y, c->nz, nbuf, c->st)
                                             : launch_ypass(y, c->nz, nbuf, c->st);
    tend(c, HETD_K_YPASS, 4.0 * sizeof(T) * c->nh * nbuf);
    return e;
}

template <typename T> static hipError_t run_z(hetd_ctx* c, ZPassArgs<T> z, int nbuf) {
    z.ldh = c->ldh;
    const double s = (double)sizeof(T);
    double per = 4 * s;
    if (z.mode == ZMODE_CONV)
        per = 2 * s + 3 * s * z.nk + (z.fwd_out ? 2 * s : 0.0) + (z.kout ? 3 * s : 0.0) + (z.addin ? 2 * s : 0.0);
    tbeg(c, HETD_K_ZPASS);
    const bool reg2 = reg2_supported(z.ax.n) && (z.mode != ZMODE_CONV || z.nk == 1);
    hipError_t e = reg_fft_supported(z.ax.n) ? launch_zreg(z, nbuf, c->st)
                   : reg2                    ? launch_z2(z, nbuf, c->st)
                                             : launch_zpass(z, nbuf, c->st);
    tend(c, HETD_K_ZPASS, per * c->nh * nbuf);
    return e;
}

static bool force_fusable(const hetd_ctx* c) { return tiled_supported(c->sc) && c->sc.icon != 18; }

static hipError_t run_update(hetd_ctx* c, UpdateArgs& u, int* nb) {
    u.force = c->force_pending ? 1 : 0;
    double per = 144.0;  // PRED / CORR (FP64 den and pot4)
    if (u.kind == STAGE_RK) per = u.rk_j == 0 ? 96.0 : 80.0;
    if (u.denf) {
        c->den64_ok = false;  // only the FP32 den is written
        per -= 4.0 + (u.kind == STAGE_RK ? 0.0 : 4.0);  // FP32 den written, FP32 pot4 read
        // complex64 history: PRED reads 2 and writes 1, CORR reads 2, RK stage 0 writes 1
        per -= u.kind == STAGE_PRED ? 24.0 : u.kind == STAGE_CORR ? 16.0 : (u.rk_j == 0 ? 8.0 : 0.0);
    }
    tbeg(c, HETD_K_UPDATE);
    int nbl = 0;
    hipError_t e = tiled_supported(c->sc) ? launch_update_tiled(u, c->st, &nbl) : launch_update(u, c->st, &nbl);
    tend(c, HETD_K_UPDATE, per * c->n);
    if (nb) *nb = nbl;
    if (e != hipSuccess || !u.force) return e;
    // forceimp of the input state (potenimp.f90:24-55), a = F/mAg_u (DFT4He3d.f90:852)
    c->force_pending = false;
    e = launch_force_final(c->imp, c->partial, nbl, 4, 1, c->dxyz, c->mAg, c->pend_rec, c->st);
    c->pend_rec = nullptr;
    return e;
}

static hipError_t run_force(hetd_ctx* c, const ForceArgs& f, int* nb) {
    tbeg(c, HETD_K_FORCE);
    hipError_t e = tiled_supported(c->sc) ? launch_force_tiled(f, c->st, nb) : launch_force_partial(f, c->st, nb);
    tend(c, HETD_K_FORCE, 8.0 * c->n);
    return e;
}

// ---------------------------------------------------------------- FFT chains

// r2c of a real device field into the half spectrum `out` (x, y, z passes)
static int r2c_dev(hetd_ctx* c, const double* in, double2* out) {
    XPassArgs<double> x{};
    x.ax = c->ax[0]; x.nlines = (long)c->ny * c->nz; x.mode = XMODE_R2C; x.rin = in; x.cout = out;
    HCHK(c, run_x(c, x));
    YPassArgs<double> y{};
    y.ax = c->ax[1]; y.nxh = c->nxh; y.inv = 0; y.buf[0] = out;
    HCHK(c, run_y(c, y, 1));
    ZPassArgs<double> z{};
    z.ax = c->ax[2]; z.nxh = c->nxh; z.ny = c->ny; z.mode = ZMODE_FWD; z.in[0] = out; z.out[0][0] = out;
    HCHK(c, run_z(c, z, 1));
    return 0;
}

// c2r (times scale) of the half spectrum `in` (destroyed) into `out`
static int c2r_dev(hetd_ctx* c, double2* in, double* out, double scale) {
    ZPassArgs<double> z{};
    z.ax = c->ax[2]; z.nxh = c->nxh; z.ny = c->ny; z.mode = ZMODE_INV; z.in[0] = in; z.out[0][0] = in;
    HCHK(c, run_z(c, z, 1));
    YPassArgs<double> y{};
    y.ax = c->ax[1]; y.nxh = c->nxh; y.inv = 1; y.buf[0] = in;
    HCHK(c, run_y(c, y, 1));
    XPassArgs<double> x{};
    x.ax = c->ax[0]; x.nlines = (long)c->ny * c->nz; x.mode = XMODE_C2R; x.cin = in; x.rout = out; x.scale = scale;
    HCHK(c, run_x(c, x));
    return 0;
}

// For nb real fields: r2c, multiply by the real half-spectrum kernel K, c2r
// times renor (the fftfw / *K / fftbk pattern of term_alfa.f90:41-80).
static int convolve_dev(hetd_ctx* c, int nb, const double* const* in, double2* const* work, double* const* out,
                        const double* K) {
    for (int b = 0; b < nb; ++b) {
        XPassArgs<double> x{};
        x.ax = c->ax[0]; x.nlines = (long)c->ny * c->nz; x.mode = XMODE_R2C; x.rin = in[b]; x.cout = work[b];
        HCHK(c, run_x(c, x));
    }
    YPassArgs<double> y{};
    y.ax = c->ax[1]; y.nxh = c->nxh; y.inv = 0;
    for (int b = 0; b < nb; ++b) y.buf[b] = work[b];
    HCHK(c, run_y(c, y, nb));
    ZPassArgs<double> z{};
    z.ax = c->ax[2]; z.nxh = c->nxh; z.ny = c->ny; z.mode = ZMODE_CONV; z.nk = 1; z.kern[0] = K;
    for (int b = 0; b < nb; ++b) { z.in[b] = work[b]; z.out[b][0] = work[b]; }
    HCHK(c, run_z(c, z, nb));
    y.inv = 1;
    HCHK(c, run_y(c, y, nb));
    for (int b = 0; b < nb; ++b) {
        XPassArgs<double> x{};
        x.ax = c->ax[0]; x.nlines = (long)c->ny * c->nz; x.mode = XMODE_C2R; x.cin = work[b]; x.rout = out[b];
        x.scale = c->renor;
        HCHK(c, run_x(c, x));
    }
    return 0;
}

static int deriv_dev(hetd_ctx* c, int id, int iv, int cplx, const double* f, double* df) {
    DerivArgs d{};
    d.nx = c->nx; d.ny = c->ny; d.nz = c->nz; d.sc = c->sc; d.id = id; d.iv = iv; d.cplx = cplx; d.f = f; d.df = df;
    HCHK(c, launch_deriv(d, c->st));
    return 0;
}

// term_alfa (term_alfa.f90:28-113), then pot4 += ualphas (poten.f90:103-106)
static int term_alfa_dev(hetd_ctx* c) {
    const long n = c->n;
    HCHK(c, launch_falfs(c->denalf, c->p.den0s, c->falfs, n, c->st));
    int rc;
    if ((rc = deriv_dev(c, 1, 1, 0, c->den, c->dxden))) return rc;
    if ((rc = deriv_dev(c, 1, 2, 0, c->den, c->dyden))) return rc;
    if ((rc = deriv_dev(c, 1, 3, 0, c->den, c->dzden))) return rc;
    HCHK(c, launch_mul(c->falfs, c->dxden, c->s1, n, c->st));
    HCHK(c, launch_mul(c->falfs, c->dyden, c->s2, n, c->st));
    HCHK(c, launch_mul(c->falfs, c->dzden, c->s3, n, c->st));
    {
        const double* in[3] = {c->s1, c->s2, c->s3};
        double2* wk[3] = {c->A, c->B, c->Cs};
        double* out[3] = {c->intx, c->inty, c->intz};
        if ((rc = convolve_dev(c, 3, in, wk, out, c->kalfs))) return rc;
    }
    HCHK(c, launch_dot3(c->dxden, c->intx, c->dyden, c->inty, c->dzden, c->intz, c->s1, n, c->st));
    {
        const double* in[1] = {c->s1};
        double2* wk[1] = {c->A};
        double* out[1] = {c->ualphas};
        if ((rc = convolve_dev(c, 1, in, wk, out, c->kalfs))) return rc;
    }
    HCHK(c, launch_mul(c->falfs, c->intx, c->s1, n, c->st));
    HCHK(c, launch_mul(c->falfs, c->inty, c->s2, n, c->st));
    HCHK(c, launch_mul(c->falfs, c->intz, c->s3, n, c->st));
    if ((rc = deriv_dev(c, 1, 1, 0, c->s1, c->s4))) return rc;
    if ((rc = deriv_dev(c, 1, 2, 0, c->s2, c->s5))) return rc;
    if ((rc = deriv_dev(c, 1, 3, 0, c->s3, c->s6))) return rc;
    HCHK(c, launch_ualpha_final(c->ualphas, c->s4, c->s5, c->s6, c->p.den0s, c->p.h2o2m4, c->p.alphas, c->pot4,
                                c->mixed ? c->pot4_32 : nullptr, n, c->st));
    return 0;
}

enum { DELJ4_DEN = 0, DELJ4_OK = 1, DELJ4_SPEC = 2 };

// the density / mean-field / FFT buffers of one precision
template <typename T> struct PF {
    T *den, *pot4, *dencg, *delj4, *tmp;
    cplx<T> *A, *B, *Cs;
    const T *fvlj, *wcgk, *kalfs;
};
static PF<double> pf64(hetd_ctx* c) {
    return PF<double>{c->den, c->pot4, c->dencg, c->delj4, c->tmp, c->A, c->B, c->Cs, c->fvlj, c->wcgk, c->kalfs};
}
static PF<float> pf32(hetd_ctx* c) {
    return PF<float>{c->den32, c->pot4_32, c->dencg32, c->delj4_32, c->tmp32, c->A32, c->B32, c->C32,
                     c->fvlj32, c->wcgk32, c->kalfs32};
}

// poten() (poten.f90:5-118) up to pot4 without the alpha_s term.
//
// Stepping form (full = false): the two linear convolutions that enter pot4
// additively, delj4 = IFFT(V_LJ fden) and sto1 = IFFT(W FFT(s)), are summed
// in the spectrum and inverted once: den -> A = W fden (inverted to dencg),
// B = V_LJ fden (kept as a z spectrum); s -> A = W FFT(s) + B -> pot4.
// delj4 itself is then not formed; ensure_delj4 makes it when asked for.
//
// Full form (full = true): the reference's sequence of transforms, storing
// delj4; want_ccor stores sto1 in tmp, want_fden the spectrum of den in Cs,
// keep_pot4 = false leaves pot4 untouched (fields recomputed for a copy).
template <typename T>
static int poten_t(hetd_ctx* c, const PF<T>& f, bool full, bool want_ccor, bool want_fden, bool keep_pot4) {
    const long nl = (long)c->ny * c->nz;
    XPassArgs<T> x{};
    x.ax = c->ax[0]; x.nlines = nl; x.mode = XMODE_R2C; x.rin = f.den; x.cout = f.A;
    HCHK(c, run_x(c, x));
    YPassArgs<T> y{};
    y.ax = c->ax[1]; y.nxh = c->nxh; y.inv = 0; y.buf[0] = f.A;
    HCHK(c, run_y(c, y, 1));
    // fden*W -> A, fden*V_LJ -> B [, fden*K_alpha -> Cs] (poten.f90:34-50, 96-101)
    ZPassArgs<T> z{};
    z.ax = c->ax[2]; z.nxh = c->nxh; z.ny = c->ny; z.mode = ZMODE_CONV; z.in[0] = f.A;
    z.kern[0] = f.wcgk; z.out[0][0] = f.A;
    if (full) {
        z.nk = c->otc ? 3 : 2;
        z.out[0][1] = f.B; z.out[0][2] = f.Cs;
        z.kern[1] = f.fvlj; z.kern[2] = f.kalfs;
        if (want_fden && c->otc) return fail(c, HETD_ERR_STATE, "fden copy not available with OTC");
        z.fwd_out = want_fden ? f.Cs : nullptr;
    } else {
        z.nk = 1;
        z.kout = f.B; z.kkern = f.fvlj;
    }
    HCHK(c, run_z(c, z, 1));
    y.inv = 1; y.buf[0] = f.A; y.buf[1] = f.B; y.buf[2] = f.Cs;
    HCHK(c, run_y(c, y, full ? (c->otc ? 3 : 2) : 1));
    if (full) {
        // delj4 = c2r(fden*V)*renor
        XPassArgs<T> xl{};
        xl.ax = c->ax[0]; xl.nlines = nl; xl.mode = XMODE_C2R; xl.cin = f.B; xl.rout = f.delj4;
        xl.scale = (T)c->renor;
        HCHK(c, run_x(c, xl));
    }
    c->delj4_state = full ? DELJ4_OK : DELJ4_SPEC;
    // dencg = c2r(fden*W)*renor, s = den*dencg*(cp4+cpp4*dencg), A = x-r2c(s)
    // (stepping form: the local part of pot4 is formed here, where den is
    // read anyway, and stored in place of dencg)
    XPassArgs<T> xc{};
    xc.ax = c->ax[0]; xc.nlines = nl; xc.mode = XMODE_CG_S; xc.cin = f.A; xc.rin = f.den; xc.rout = f.dencg;
    xc.cout = f.A; xc.scale = (T)c->renor; xc.cp4 = (T)c->p.cp4; xc.cpp4 = (T)c->p.cpp4;
    xc.fold = full ? 0 : 1;
    xc.a0 = (T)(c->p.cp4 / 2.0); xc.a1 = (T)(c->p.cpp4 / 3.0);
    xc.lsolid = c->p.lsolid && !c->otc;
    xc.csol = (T)c->p.c_solid; xc.beta = (T)c->p.beta_solid; xc.den_m = (T)c->p.den_m;
    HCHK(c, run_x(c, xc));
    c->dencg_ok = full;
    y.inv = 0; y.buf[0] = f.A;
    HCHK(c, run_y(c, y, 1));
    ZPassArgs<T> z2{};
    z2.ax = c->ax[2]; z2.nxh = c->nxh; z2.ny = c->ny; z2.mode = ZMODE_CONV; z2.nk = 1; z2.in[0] = f.A;
    z2.out[0][0] = f.A; z2.kern[0] = f.wcgk;
    if (!full) z2.addin = f.B;
    HCHK(c, run_z(c, z2, 1));
    y.inv = 1;
    HCHK(c, run_y(c, y, 1));
    // pot4 = delj4 + dencg^2 (cp4/2 + cpp4/3 dencg) + sto1 [+ penalty] (poten.f90:80-94, 110-114)
    XPassArgs<T> xp{};
    xp.ax = c->ax[0]; xp.nlines = nl; xp.mode = XMODE_POT; xp.cin = f.A; xp.rin = f.den; xp.rin2 = f.dencg;
    xp.rin3 = full ? f.delj4 : nullptr; xp.rout = keep_pot4 ? f.pot4 : nullptr;
    xp.rout2 = want_ccor ? f.tmp : nullptr; xp.scale = (T)c->renor;
    xp.a0 = (T)(c->p.cp4 / 2.0); xp.a1 = (T)(c->p.cpp4 / 3.0);
    xp.lsolid = c->p.lsolid && !c->otc;
    xp.csol = (T)c->p.c_solid; xp.beta = (T)c->p.beta_solid; xp.den_m = (T)c->p.den_m;
    xp.fold = full ? 0 : 1;
    HCHK(c, run_x(c, xp));
    return 0;
}

// dencg = c2r(W r2c(den)) * renor (poten.f90:27-38) from the current den,
// for the energy and field copies after stepping-form potentials (which keep
// the local part of pot4 in the dencg buffer)
template <typename T> static int dencg_t(hetd_ctx* c, const PF<T>& f) {
    const long nl = (long)c->ny * c->nz;
    XPassArgs<T> x{};
    x.ax = c->ax[0]; x.nlines = nl; x.mode = XMODE_R2C; x.rin = f.den; x.cout = f.A;
    HCHK(c, run_x(c, x));
    YPassArgs<T> y{};
    y.ax = c->ax[1]; y.nxh = c->nxh; y.inv = 0; y.buf[0] = f.A;
    HCHK(c, run_y(c, y, 1));
    ZPassArgs<T> z{};
    z.ax = c->ax[2]; z.nxh = c->nxh; z.ny = c->ny; z.mode = ZMODE_CONV; z.nk = 1; z.in[0] = f.A;
    z.out[0][0] = f.A; z.kern[0] = f.wcgk;
    HCHK(c, run_z(c, z, 1));
    y.inv = 1;
    HCHK(c, run_y(c, y, 1));
    XPassArgs<T> xb{};
    xb.ax = c->ax[0]; xb.nlines = nl; xb.mode = XMODE_C2R; xb.cin = f.A; xb.rout = f.dencg; xb.scale = (T)c->renor;
    HCHK(c, run_x(c, xb));
    return 0;
}

// delj4 = c2r(V_LJ r2c(den)) * renor (poten.f90:38-45) for the energy and
// field copies after stepping-form potentials: from the spectrum V_LJ fden
// the last poten left in B, or (no poten since the state was set) from den
template <typename T> static int delj4_t(hetd_ctx* c, const PF<T>& f, bool from_b) {
    const long nl = (long)c->ny * c->nz;
    YPassArgs<T> y{};
    y.ax = c->ax[1]; y.nxh = c->nxh; y.buf[0] = f.B;
    ZPassArgs<T> z{};
    z.ax = c->ax[2]; z.nxh = c->nxh; z.ny = c->ny; z.in[0] = f.B; z.out[0][0] = f.B;
    if (from_b) {
        z.mode = ZMODE_INV;
    } else {
        XPassArgs<T> x{};
        x.ax = c->ax[0]; x.nlines = nl; x.mode = XMODE_R2C; x.rin = f.den; x.cout = f.B;
        HCHK(c, run_x(c, x));
        y.inv = 0;
        HCHK(c, run_y(c, y, 1));
        z.mode = ZMODE_CONV; z.nk = 1; z.kern[0] = f.fvlj;
    }
    HCHK(c, run_z(c, z, 1));
    y.inv = 1;
    HCHK(c, run_y(c, y, 1));
    XPassArgs<T> xl{};
    xl.ax = c->ax[0]; xl.nlines = nl; xl.mode = XMODE_C2R; xl.cin = f.B; xl.rout = f.delj4; xl.scale = (T)c->renor;
    HCHK(c, run_x(c, xl));
    return 0;
}

// (After a steppc called on its own, den has moved on from the middle poten;
// the reference's energy() at that point would use the middle poten's dencg,
// this one the current den's.  hetd_poten forms the full fields, and
// hetd_evolve ends with a poten of the current den.)
static int ensure_dencg(hetd_ctx* c) {
    if (c->dencg_ok) return 0;
    const int rc = c->mixed ? dencg_t<float>(c, pf32(c)) : dencg_t<double>(c, pf64(c));
    if (rc) return rc;
    c->dencg_ok = true;
    return 0;
}

static int ensure_delj4(hetd_ctx* c) {
    if (c->delj4_state == DELJ4_OK) return 0;
    const bool from_b = c->delj4_state == DELJ4_SPEC;
    const int rc = c->mixed ? delj4_t<float>(c, pf32(c), from_b) : delj4_t<double>(c, pf64(c), from_b);
    if (rc) return rc;
    c->delj4_state = DELJ4_OK;
    return 0;
}

static int ensure_den64(hetd_ctx* c);

static int poten_dev(hetd_ctx* c, bool want_ccor = false, bool want_fden = false, bool keep_pot4 = true,
                     bool full_form = false) {
    if (c->mixed) {
        // the alpha_s term (OTC) stays FP64: coarse-grained density
        // IFFT(K_alpha FFT(den)) and term_alfa from the FP64 den of psi,
        // added to the FP32 mean field
        int rc = poten_t<float>(c, pf32(c), want_ccor || want_fden || full_form, want_ccor, want_fden, keep_pot4);
        if (rc || !c->otc || !keep_pot4) return rc;
        if ((rc = ensure_den64(c))) return rc;
        const double* in[1] = {c->den};
        double2* wk[1] = {c->Cs};
        double* out[1] = {c->denalf};
        if ((rc = convolve_dev(c, 1, in, wk, out, c->kalfs))) return rc;
        return term_alfa_dev(c);
    }
    // OTC keeps the full form: its K_alpha spectrum rides on the first z pass
    const bool full = want_ccor || want_fden || c->otc || full_form;
    int rc = poten_t<double>(c, pf64(c), full, want_ccor, want_fden, keep_pot4);
    if (rc) return rc;
    if (c->otc && keep_pot4) {
        XPassArgs<double> xa{};
        xa.ax = c->ax[0]; xa.nlines = (long)c->ny * c->nz; xa.mode = XMODE_C2R; xa.cin = c->Cs; xa.rout = c->denalf;
        xa.scale = c->renor;
        HCHK(c, run_x(c, xa));
        rc = term_alfa_dev(c);
        if (rc) return rc;
    }
    return 0;
}

// FP64 den from psi where the mixed path has written only the FP32 copy
static int ensure_den64(hetd_ctx* c) {
    if (c->den64_ok) return 0;
    HCHK(c, launch_den_from_psi(c->psi, c->den, c->n, c->st));
    c->den64_ok = true;
    return 0;
}

// forceimp + aimp = F/mAg_u on the device (potenimp.f90:24-55, DFT4He3d.f90:852)
static int force_dev(hetd_ctx* c) {
    c->force_pending = false;
    if (int rc = ensure_den64(c)) return rc;
    ForceArgs f{};
    // FP64 den in both modes: the force is a sum with heavy cancellation
    f.g = c->grid; f.sc = c->sc; f.den = c->den; f.imp = c->imp; f.partial = c->partial;
    int nb = 0;
    HCHK(c, run_force(c, f, &nb));
    HCHK(c, launch_force_final(c->imp, c->partial, nb, 3, 0, c->dxyz, c->mAg, c->pend_rec, c->st));
    c->pend_rec = nullptr;
    return 0;
}

// potenimp (rpot already set) + poten + forceimp.  With defer, the force is
// left to the next update kernel, which reads the same state.
static int potential_update(hetd_ctx* c, bool defer) {
    int rc = poten_dev(c, false, false);
    if (rc) return rc;
    if (defer && force_fusable(c)) {
        c->force_pending = true;
        return 0;
    }
    return force_dev(c);
}

// an acceleration left to the next update is formed now (end of a call)
static int settle_force(hetd_ctx* c) { return c->force_pending ? force_dev(c) : 0; }

static UpdateArgs update_base(hetd_ctx* c, double dt) {
    UpdateArgs u{};
    u.g = c->grid; u.sc = c->sc; u.h2o2m4 = c->p.h2o2m4; u.dt = dt; u.psi = c->psi; u.pot4 = c->pot4;
    u.imp = c->imp; u.den = c->den; u.guard = c->guard; u.partial = c->partial;
    if (c->mixed) { u.pot4f = c->pot4_32; u.denf = c->den32; }
    u.tile = c->upd_tile;
    return u;
}

static int steprk_dev(hetd_ctx* c, double dt, bool last_sets_rpot) {
    const double s2 = std::sqrt(2.0);
    const double ar[4] = {0.5, 1.0 - 1.0 / s2, 1.0 + 1.0 / s2, 1.0 / 6.0};
    const double br[4] = {2.0, 1.0, 1.0, 2.0};
    const double cr[4] = {0.5, 1.0 - 1.0 / s2, 1.0 + 1.0 / s2, 0.5};
    for (int j = 0; j < 4; ++j) {
        UpdateArgs u = update_base(c, dt);
        u.kind = STAGE_RK; u.rk_j = j; u.ra = ar[j]; u.rb = br[j]; u.rc = cr[j];
        u.q = c->q; u.psi_out = c->spare;
        u.hnew = (j == 0) ? c->hpsiold[1] : nullptr;
        HCHK(c, run_update(c, u, nullptr));
        if (j == 0) {
            std::swap(c->hpsiold[0], c->hpsiold[1]);
            double2* t = c->psiold[2];
            c->psiold[2] = c->psiold[1];
            c->psiold[1] = c->psiold[0];
            c->psiold[0] = c->psi;
            c->psi = c->spare;
            c->spare = t;
        } else {
            std::swap(c->psi, c->spare);
        }
        ImpRkArgs ia{c->imp, j, dt, ar[j], br[j], cr[j], (j <= 2 || last_sets_rpot) ? 1 : 0};
        HCHK(c, launch_imp_rk(ia, c->st));
        if (j <= 2) {
            int rc = potential_update(c, true);
            if (rc) return rc;
        }
    }
    for (int i = 0; i < 3; ++i) c->ioldp[i] = c->ioldr[i] = c->ioldv[i] = i + 1;
    for (int i = 0; i < 2; ++i) c->ioldh[i] = c->iolda[i] = i + 1;
    return 0;
}

static int steppc_dev(hetd_ctx* c, double dt, bool sets_rpot) {
    const int p3 = c->ioldp[2] - 1, p2 = c->ioldp[1] - 1;
    {
        UpdateArgs u = update_base(c, dt);
        u.kind = STAGE_PRED;
        u.po3 = c->psiold[p3]; u.psi_out = c->psiold[p3];
        u.h1 = c->hpsiold[c->ioldh[0] - 1]; u.h2 = c->hpsiold[c->ioldh[1] - 1]; u.hnew = c->hpsiold[c->ioldh[1] - 1];
        u.pc = c->pc;
        HCHK(c, run_update(c, u, nullptr));
        std::swap(c->psi, c->psiold[p3]);
        std::swap(c->ioldh[0], c->ioldh[1]);
    }
    ImpPcArgs ia{c->imp, dt, c->ioldr[0] - 1, c->ioldr[1] - 1, c->ioldr[2] - 1, c->ioldv[0] - 1, c->ioldv[1] - 1,
                 c->ioldv[2] - 1, c->iolda[0] - 1, c->iolda[1] - 1, 0};
    HCHK(c, launch_imp_pred(ia, c->st));
    std::swap(c->iolda[0], c->iolda[1]);
    int rc = potential_update(c, true);
    if (rc) return rc;
    {
        UpdateArgs u = update_base(c, dt);
        u.kind = STAGE_CORR;
        u.po3 = c->psiold[p3]; u.po2 = c->psiold[p2];
        u.h1 = c->hpsiold[c->ioldh[0] - 1]; u.h2 = c->hpsiold[c->ioldh[1] - 1];
        u.pc = c->pc; u.psi_out = c->spare;
        int nb = 0;
        HCHK(c, run_update(c, u, &nb));
        HCHK(c, launch_errhe_final(c->imp, c->partial, nb, 4, (double)c->n, c->st));
        std::swap(c->psi, c->spare);
        const int t = c->ioldp[2];
        c->ioldp[2] = c->ioldp[1];
        c->ioldp[1] = c->ioldp[0];
        c->ioldp[0] = t;
    }
    ImpPcArgs ic{c->imp, dt, c->ioldr[0] - 1, c->ioldr[1] - 1, c->ioldr[2] - 1, c->ioldv[0] - 1, c->ioldv[1] - 1,
                 c->ioldv[2] - 1, c->iolda[0] - 1, c->iolda[1] - 1, sets_rpot ? 1 : 0};
    HCHK(c, launch_imp_corr(ic, c->st));
    int t = c->ioldr[2]; c->ioldr[2] = c->ioldr[1]; c->ioldr[1] = c->ioldr[0]; c->ioldr[0] = t;
    t = c->ioldv[2]; c->ioldv[2] = c->ioldv[1]; c->ioldv[1] = c->ioldv[0]; c->ioldv[0] = t;
    return 0;
}

static int check_guard(hetd_ctx* c) {
    int g = 0;
    HCHK(c, hipMemcpyAsync(&g, c->guard, sizeof(int), hipMemcpyDeviceToHost, c->st));
    HCHK(c, hipStreamSynchronize(c->st));
    if (g) return fail(c, HETD_ERR_GUARD, "den not > 0 after a Runge-Kutta stage (steprk.f90:89-92)");
    return 0;
}

// ------------------------------------------------------------ step graphs
//
// A steady-state iteration (steppc + the driver's potential update) is the
// same ~25 launches every step, on buffers that rotate with a short period
// (psi/psiold/spare pointers, the ring indices, the deferred-force flag).
// The launches of one full period are captured once into a HIP graph and
// replayed: one launch per period instead of ~25 per step, which is what
// small grids (and an ensemble of them) are bound by.  The graph holds no
// memset or host synchronisation; replays are bitwise identical to the
// launches they record.

// every host-side value a step reads or rotates
static std::vector<long> step_sig(const hetd_ctx* c) {
    std::vector<long> v;
    v.push_back((long)(intptr_t)c->psi);
    for (int i = 0; i < 3; ++i) v.push_back((long)(intptr_t)c->psiold[i]);
    v.push_back((long)(intptr_t)c->spare);
    for (int i = 0; i < 2; ++i) v.push_back((long)(intptr_t)c->hpsiold[i]);
    for (int i = 0; i < 3; ++i) v.push_back(c->ioldp[i]);
    for (int i = 0; i < 2; ++i) v.push_back(c->ioldh[i]);
    for (int i = 0; i < 3; ++i) v.push_back(c->ioldr[i]);
    for (int i = 0; i < 3; ++i) v.push_back(c->ioldv[i]);
    for (int i = 0; i < 2; ++i) v.push_back(c->iolda[i]);
    v.push_back(c->force_pending);
    v.push_back(c->delj4_state);
    v.push_back(c->den64_ok);
    v.push_back(c->dencg_ok);
    return v;
}

static void restore_sig(hetd_ctx* c, const std::vector<long>& v) {
    size_t k = 0;
    c->psi = (double2*)(intptr_t)v[k++];
    for (int i = 0; i < 3; ++i) c->psiold[i] = (double2*)(intptr_t)v[k++];
    c->spare = (double2*)(intptr_t)v[k++];
    for (int i = 0; i < 2; ++i) c->hpsiold[i] = (double2*)(intptr_t)v[k++];
    for (int i = 0; i < 3; ++i) c->ioldp[i] = (int)v[k++];
    for (int i = 0; i < 2; ++i) c->ioldh[i] = (int)v[k++];
    for (int i = 0; i < 3; ++i) c->ioldr[i] = (int)v[k++];
    for (int i = 0; i < 3; ++i) c->ioldv[i] = (int)v[k++];
    for (int i = 0; i < 2; ++i) c->iolda[i] = (int)v[k++];
    c->force_pending = v[k++] != 0;
    c->delj4_state = (int)v[k++];
    c->den64_ok = v[k++] != 0;
    c->dencg_ok = v[k++] != 0;
}

// one steady-state iteration of the driver loop (DFT4He3d.f90:844-852)
static int pc_iteration(hetd_ctx* c, double dt) {
    int rc = steppc_dev(c, dt, true);
    if (rc) return rc;
    return potential_update(c, true);
}

static void drop_graph(hetd_ctx* c) {
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
    c->gper = 0;
    c->gsig.clear();
}

// Capture one period of PC iterations from the current host state (nothing
// executes; the host state is left as it was).  Returns 0 with c->gexec set,
// or 0 with c->gexec null when no period of at most 12 steps was found.
static int build_step_graph(hetd_ctx* c, double dt) {
    drop_graph(c);
    const std::vector<long> s0 = step_sig(c);
    HCHK(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
    int per = 0, rc = 0;
    for (int k = 1; k <= 12 && !rc; ++k) {
        rc = pc_iteration(c, dt);
        if (!rc && step_sig(c) == s0) { per = k; break; }
    }
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->st, &g);
    restore_sig(c, s0);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    HCHK(c, e);
    if (per > 0) {
        hipGraphExec_t x = nullptr;
        const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HCHK(c, ei);
        c->gexec = x;
        c->gper = per;
        c->gdt = dt;
        c->gsig = s0;
    } else {
        (void)hipGraphDestroy(g);
    }
    return 0;
}

// Advance by one unit of work without waiting: a graph replay of gper PC
// iterations when that is allowed and at least gper steps remain, else one
// iteration issued kernel by kernel.  *done = steps advanced.  traj_row, when
// set, records the impurity after an eagerly issued step (graphs are used
// only without a trajectory).
static int evolve_unit(hetd_ctx* c, double dt, int remaining, double* traj_row, int* done) {
    *done = 1;
    c->iter_since_set++;
    if (c->iter_since_set <= 3) {
        // errHe/errimp are zero for RK steps (the driver prints them only after PC)
        HCHK(c, hipMemsetAsync(&c->imp->errHe, 0, 2 * sizeof(double), c->st));
        int rc = steprk_dev(c, dt, true);
        if (rc) return rc;
        rc = potential_update(c, true);
        if (rc) return rc;
    } else {
        const bool graphs = c->use_graphs && !traj_row && c->tclass == 0 && c->iter_since_set > 4;
        if (graphs && (!c->gexec || c->gdt != dt)) {
            int rc = build_step_graph(c, dt);
            if (rc) return rc;
        }
        if (graphs && c->gexec && remaining >= c->gper && step_sig(c) == c->gsig) {
            HCHK(c, hipGraphLaunch(c->gexec, c->st));
            c->iter_since_set += c->gper - 1;
            *done = c->gper;
            return 0;
        }
        int rc = pc_iteration(c, dt);
        if (rc) return rc;
    }
    if (traj_row) {
        HCHK(c, launch_traj_record(c->imp, traj_row, c->st));
        if (c->force_pending) c->pend_rec = traj_row + 8;  // aimp, filled when the force is formed
    }
    return 0;
}

// --------------------------------------------------------------------- ABI

extern "C" {

void hetd_default_params(hetd_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->nx = p->ny = p->nz = 64;
    p->npd = 13;
    p->icon = 13;
    p->core = HETD_CORE_OT;
    p->lsolid = 1;
    p->precision = HETD_FP64;
    p->xmax = p->ymax = p->zmax = 64.0;
    p->eps4 = 10.22; p->sigma4 = 2.556; p->h4 = 2.190323;
    p->cp4 = -2.41186e4; p->cpp4 = 1.85850e6; p->h2o2m4 = 6.05969638298;
    p->alphas = 54.31; p->l_alpha = 1.0; p->den0s = 0.04;
    p->c_solid = 3.1577504e4; p->beta_solid = 40.0; p->den_m = 0.37;
    p->mimpur = 131.29;
    p->Lambdah = 2.0;
    p->txmean = p->tymean = p->tzmean = 1000.0;
    p->txsurf = p->tysurf = p->tzsurf = 2.0;
}

const char* hetd_last_error(hetd_ctx* c) { return c ? c->err.c_str() : "null context"; }

void* hetd_stream(hetd_ctx* c) { return c ? (void*)c->st : nullptr; }

int hetd_vimp(const char* selec, double r, double r_cutoff, double umax, double* v) {
    if (!selec || !v) return HETD_ERR_ARG;
    return imp::select_pot(selec, r, r_cutoff, umax, v) ? HETD_OK : HETD_ERR_ARG;
}

int hetd_build_potion(const char* selec, double r_cutoff, double umax, int32_t nmax, double xmaxbox, double* potion,
                      int32_t* npot, double* delinter) {
    if (!selec || !npot || !delinter || nmax < 1) return HETD_ERR_ARG;
    const int np = 100 * nmax + 1;
    const double rmax = 4.0 * xmaxbox;
    const double d = rmax / (double)(np - 1);
    *npot = np;
    *delinter = d;
    if (!potion) {  // the query checks the name too
        double v;
        return hetd_vimp(selec, 0.0, r_cutoff, umax, &v);
    }
    for (int i = 0; i < np; ++i) {
        const double r = (double)i * d;
        if (hetd_vimp(selec, r, r_cutoff, umax, &potion[i])) return HETD_ERR_ARG;
    }
    return HETD_OK;
}

int hetd_init(const hetd_params* p, hetd_ctx** out) {
    if (!p || !out) return HETD_ERR_ARG;
    *out = nullptr;
    hetd_ctx* c = new hetd_ctx();
    c->p = *p;
    auto bad = [&](const char* m) { c->err = m; return HETD_ERR_ARG; };
    int rc = 0;
    if (p->nx < 1 || p->ny < 1 || p->nz < 1) rc = bad("grid sizes must be positive");
    else if (p->npd < 3 || p->npd > 25 || p->npd % 2 == 0) rc = bad("npd must be odd, 3..25 (newder.f:75-76)");
    else if (p->icon != 13 && p->icon != 8 && p->icon != 18) rc = bad("icon must be 13, 8 or 18");
    else if (p->precision != HETD_FP64 && p->precision != HETD_MIXED) rc = bad("precision must be HETD_FP64 or HETD_MIXED");
    else if (p->precision == HETD_MIXED && p->npd != 9 && p->npd != 13) rc = bad("HETD_MIXED needs npd 9 or 13");
    else if (p->nx > 1024 || p->ny > 1024 || p->nz > 1024) rc = bad("grid edge above 1024");
    if (rc) { *out = c; return rc; }
    c->nx = p->nx; c->ny = p->ny; c->nz = p->nz; c->nxh = p->nx / 2 + 1;
    c->n = (long)c->nx * c->ny * c->nz;
    c->nh = (long)c->nxh * c->ny * c->nz;
    c->ldh = (c->nxh + 15) / 16 * 16;
    c->nhp = (long)c->ldh * c->ny * c->nz;
    c->hx = 2.0 * std::fabs(p->xmax) / p->nx;
    c->hy = 2.0 * std::fabs(p->ymax) / p->ny;
    c->hz = 2.0 * std::fabs(p->zmax) / p->nz;
    c->dxyz = c->hx * c->hy * c->hz;
    c->mAg = p->mimpur * 0.02061484;  // mp_u (modules.f90:315)
    c->renor = 1.0 / (double)(c->nx * c->ny * c->nz);  // fft.f90:36
    c->otc = (!p->lsolid && p->core == HETD_CORE_OTC);
    c->mixed = p->precision == HETD_MIXED;
    *out = c;
    const int nn[3] = {c->nx, c->ny, c->nz};
    for (int a = 0; a < 3; ++a) {
        if (!setup::radix_plan(nn[a], c->ax[a].rad, &c->ax[a].nst, kMaxStages))
            return fail(c, HETD_ERR_ARG, "grid edge %d has a prime factor above 5", nn[a]);
        c->ax[a].n = nn[a];
        if (!stockham_plan_ok(c->ax[a]))
            return fail(c, HETD_ERR_ARG, "grid edge %d: no FFT block shape fits", nn[a]);
    }
    HCHK(c, hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
    if (const char* e = getenv("HETD_GRAPHS")) c->use_graphs = e[0] != '0';
    if (const char* e = getenv("HETD_XSPLIT")) c->xsplit = e[0] != '0';
    // grid tables (DFT4He3d.f90:309-317, 689-699)
    const double hh[3] = {c->hx, c->hy, c->hz}, mx[3] = {p->xmax, p->ymax, p->zmax};
    const double tmean[3] = {p->txmean, p->tymean, p->tzmean}, tsurf[3] = {p->txsurf, p->tysurf, p->tzsurf};
    double** dco[3] = {&c->d_x, &c->d_y, &c->d_z};
    double** dtc[3] = {&c->d_tcx, &c->d_tcy, &c->d_tcz};
    for (int a = 0; a < 3; ++a) {
        std::vector<double> co(nn[a]), tc(nn[a]);
        for (int i = 0; i < nn[a]; ++i) {
            co[i] = -mx[a] + hh[a] * i;
            tc[i] = std::tanh((std::fabs(co[i]) - tmean[a]) / tsurf[a]);
        }
        HCHK(c, dalloc(c, dco[a], sizeof(double) * nn[a]));
        HCHK(c, dalloc(c, dtc[a], sizeof(double) * nn[a]));
        HCHK(c, hipMemcpy(*dco[a], co.data(), sizeof(double) * nn[a], hipMemcpyHostToDevice));
        HCHK(c, hipMemcpy(*dtc[a], tc.data(), sizeof(double) * nn[a], hipMemcpyHostToDevice));
        std::vector<double> t64;
        std::vector<float> t32;
        setup::twiddles(nn[a], t64, t32);
        HCHK(c, dalloc(c, &c->d_tw[a], sizeof(double) * t64.size()));
        HCHK(c, dalloc(c, &c->d_tw32[a], sizeof(float) * t32.size()));
        HCHK(c, hipMemcpy(c->d_tw[a], t64.data(), sizeof(double) * t64.size(), hipMemcpyHostToDevice));
        HCHK(c, hipMemcpy(c->d_tw32[a], t32.data(), sizeof(float) * t32.size(), hipMemcpyHostToDevice));
        c->ax[a].tw64 = c->d_tw[a];
        c->ax[a].tw32 = c->d_tw32[a];
    }
    // stencil weights * 1/h**id (init_deriv_parallel.f90:415-431)
    c->sc.npd = p->npd; c->sc.m = p->npd / 2; c->sc.icon = p->icon;
    double w1[kMaxNpd], w2[kMaxNpd];
    setup::fd_weights(p->npd, 1, w1);
    setup::fd_weights(p->npd, 2, w2);
    for (int j = 0; j < p->npd; ++j) { c->sc.w1[j] = w1[j]; c->sc.w2[j] = w2[j]; }
    for (int a = 0; a < 3; ++a) {
        c->sc.h[a] = hh[a];
        const double s1 = 1.0 / hh[a], s2 = 1.0 / (hh[a] * hh[a]);
        for (int j = 0; j < p->npd; ++j) {
            c->sc.c1[a][j] = w1[j] * s1;
            c->sc.c2[a][j] = w2[j] * s2;
        }
    }
    // fields
    const size_t nR = sizeof(double) * c->n, nC = sizeof(double2) * c->n, nH = sizeof(double2) * c->nhp;
    HCHK(c, dalloc(c, &c->psi, nC));
    for (int i = 0; i < 3; ++i) HCHK(c, dalloc(c, &c->psiold[i], nC));
    HCHK(c, dalloc(c, &c->spare, nC));
    for (int i = 0; i < 2; ++i) HCHK(c, dalloc(c, &c->hpsiold[i], nC));
    HCHK(c, dalloc(c, &c->pc, nC));
    HCHK(c, dalloc(c, &c->q, nC));
    HCHK(c, dalloc(c, &c->den, nR));
    HCHK(c, dalloc(c, &c->pot4, nR));
    HCHK(c, dalloc(c, &c->dencg, nR));
    HCHK(c, dalloc(c, &c->delj4, nR));
    HCHK(c, dalloc(c, &c->tmp, nR));
    HCHK(c, dalloc(c, &c->A, nH));
    HCHK(c, dalloc(c, &c->B, nH));
    HCHK(c, dalloc(c, &c->Cs, nH));
    HCHK(c, dalloc(c, &c->fvlj, sizeof(double) * c->nhp));
    HCHK(c, dalloc(c, &c->wcgk, sizeof(double) * c->nhp));
    HCHK(c, dalloc(c, &c->kalfs, sizeof(double) * c->nhp));
    if (c->otc) {
        double** arr[] = {&c->dxden, &c->dyden, &c->dzden, &c->denalf, &c->falfs, &c->intx, &c->inty, &c->intz,
                          &c->ualphas, &c->s1, &c->s2, &c->s3, &c->s4, &c->s5, &c->s6};
        for (auto a : arr) HCHK(c, dalloc(c, a, nR));
    }
    if (c->mixed) {
        const size_t fR = sizeof(float) * c->n, fH = sizeof(float2) * c->nhp, fK = sizeof(float) * c->nhp;
        HCHK(c, dalloc(c, &c->den32, fR));
        HCHK(c, dalloc(c, &c->pot4_32, fR));
        HCHK(c, dalloc(c, &c->dencg32, fR));
        HCHK(c, dalloc(c, &c->delj4_32, fR));
        HCHK(c, dalloc(c, &c->tmp32, fR));
        HCHK(c, dalloc(c, &c->A32, fH));
        HCHK(c, dalloc(c, &c->B32, fH));
        HCHK(c, dalloc(c, &c->C32, fH));
        HCHK(c, dalloc(c, &c->fvlj32, fK));
        HCHK(c, dalloc(c, &c->wcgk32, fK));
        HCHK(c, dalloc(c, &c->kalfs32, fK));
    }
    HCHK(c, dalloc(c, &c->imp, sizeof(ImpState)));
    HCHK(c, dalloc(c, &c->guard, sizeof(int)));
    const long nblocks = (long)((c->nx + 63) / 64) * ((c->ny + 3) / 4) * c->nz;
    HCHK(c, dalloc(c, &c->partial, sizeof(double) * nblocks * EN_NCOMP));
    c->partial_len = nblocks * EN_NCOMP;
    HCHK(c, dalloc(c, &c->red, sizeof(double) * 16));
    HCHK(c, hipMemset(c->imp, 0, sizeof(ImpState)));
    HCHK(c, hipMemset(c->guard, 0, sizeof(int)));
    // form factors on the reference p-grid (fforma.f90, initcg.f90, DFT4He3d.f90:592-596)
    {
        std::vector<double> fv(c->nhp, 0.0), wc(c->nhp, 0.0), ka(c->nhp, 0.0);
        const double hcg = p->lsolid ? p->h4 * 1.065 : p->h4;
        const double pi = 4.0 * std::atan(1.0);
        const int nxh = c->nxh, ldh = c->ldh, ny = c->ny, nz = c->nz;
        const double hx = c->hx, hy = c->hy, hz = c->hz;
        // The three factors depend on |p| only, and a grid has few distinct
        // |p| (33 743 of 8.5 M half-spectrum points at 256^3): each is
        // evaluated once and looked up, which gives the same bits.
        auto pmod = [&](int i, int j, int k) {
            const double px = (1.0 / (c->nx * hx)) * i;  // hpx*(ipx-1), ipx <= nx/2+1
            const double py = setup::pgrid(j, ny, hy), pz = setup::pgrid(k, nz, hz);
            return std::sqrt(px * px + py * py + pz * pz);
        };
        std::vector<std::vector<double>> plane(nz);
        setup::parallel_for(nz, [&](int k) {
            std::vector<double>& u = plane[k];
            u.reserve((size_t)ny * nxh);
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nxh; ++i) u.push_back(pmod(i, j, k));
            std::sort(u.begin(), u.end());
            u.erase(std::unique(u.begin(), u.end()), u.end());
        });
        std::vector<double> pu;
        for (auto& u : plane) {
            pu.insert(pu.end(), u.begin(), u.end());
            std::vector<double>().swap(u);
        }
        std::sort(pu.begin(), pu.end());
        pu.erase(std::unique(pu.begin(), pu.end()), pu.end());
        const int nu = (int)pu.size();
        std::vector<double> ufv(nu), uwc(nu), uka(nu);
        setup::parallel_for(nu, [&](int q) {
            const double pm = pu[q];
            ufv[q] = setup::vlj_hat(pm, p->eps4, p->sigma4, p->h4);
            uwc[q] = setup::wcg(pm, hcg);
            uka[q] = std::exp(-(pi * p->l_alpha * pm) * (pi * p->l_alpha * pm));
        });
        setup::parallel_for(nz, [&](int k) {
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nxh; ++i) {
                    const double pm = pmod(i, j, k);
                    const int q = (int)(std::lower_bound(pu.begin(), pu.end(), pm) - pu.begin());
                    const size_t o = (size_t)i + (size_t)ldh * ((size_t)j + (size_t)ny * k);
                    fv[o] = ufv[q];
                    wc[o] = uwc[q];
                    ka[o] = uka[q];
                }
        });
        HCHK(c, hipMemcpy(c->fvlj, fv.data(), sizeof(double) * c->nhp, hipMemcpyHostToDevice));
        HCHK(c, hipMemcpy(c->wcgk, wc.data(), sizeof(double) * c->nhp, hipMemcpyHostToDevice));
        HCHK(c, hipMemcpy(c->kalfs, ka.data(), sizeof(double) * c->nhp, hipMemcpyHostToDevice));
        if (c->mixed) {
            HCHK(c, launch_to_float(c->fvlj, c->fvlj32, c->nhp, c->st));
            HCHK(c, launch_to_float(c->wcgk, c->wcgk32, c->nhp, c->st));
            HCHK(c, launch_to_float(c->kalfs, c->kalfs32, c->nhp, c->st));
            HCHK(c, hipStreamSynchronize(c->st));
        }
    }
    fill_grid(c);
    c->ready = true;
    return HETD_OK;
}

int hetd_finalize(hetd_ctx* c) {
    if (!c) return HETD_ERR_ARG;
    if (c->st) (void)hipStreamSynchronize(c->st);
    for (auto& a : c->allocs) (void)hipFree(a.first);
    if (c->traj) (void)hipFree(c->traj);
    if (c->dtmp[0]) (void)hipFree(c->dtmp[0]);
    if (c->dtmp[1]) (void)hipFree(c->dtmp[1]);
    for (hipEvent_t e : c->tev) (void)hipEventDestroy(e);
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
    return HETD_OK;
}

int hetd_set_potion(hetd_ctx* c, const double* potion, int32_t npot, double delinter) {
    if (!c || !potion || npot < 2 || !(delinter > 0)) return HETD_ERR_ARG;
    if (!c->ready) return HETD_ERR_STATE;  // hetd_init failed; hetd_last_error says why
    if (c->potion && c->npot != npot) return fail(c, HETD_ERR_ARG, "table size changed");
    if (!c->potion) HCHK(c, dalloc(c, &c->potion, sizeof(double) * npot));
    HCHK(c, hipMemcpyAsync(c->potion, potion, sizeof(double) * npot, hipMemcpyHostToDevice, c->st));
    HCHK(c, hipStreamSynchronize(c->st));
    c->npot = npot;
    c->delinter = delinter;
    fill_grid(c);
    drop_graph(c);  // the captured launches hold the old grid arguments
    return HETD_OK;
}

int hetd_vsurf(int32_t lsurf3d, double x, double y, double z, double* v) {
    if (!v) return HETD_ERR_ARG;
    *v = lsurf3d ? imp::morse_tio2_3d(x, y, z) : imp::morse_tio2_1d(z);
    return HETD_OK;
}

int hetd_set_uext(hetd_ctx* c, const double* uext) {
    if (!c) return HETD_ERR_ARG;
    if (!c->ready) return HETD_ERR_STATE;  // hetd_init failed; hetd_last_error says why
    if (!uext) {
        c->uext_on = false;
        return HETD_OK;
    }
    if (!c->uext) HCHK(c, dalloc(c, &c->uext, sizeof(double) * c->n));
    HCHK(c, hipMemcpyAsync(c->uext, uext, sizeof(double) * c->n, hipMemcpyHostToDevice, c->st));
    HCHK(c, hipStreamSynchronize(c->st));
    c->uext_on = true;
    return HETD_OK;
}

int hetd_set_surface(hetd_ctx* c, int32_t lsurf3d, double zsurf) {
    if (!c) return HETD_ERR_ARG;
    if (!c->ready) return HETD_ERR_STATE;  // hetd_init failed; hetd_last_error says why
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const hetd_params& p = c->p;
    std::vector<double> u((size_t)c->n);
    setup::parallel_for(nz, [&](int k) {
        const double z = -p.zmax + c->hz * k;
        for (int j = 0; j < ny; ++j) {
            const double y = -p.ymax + c->hy * j;
            for (int i = 0; i < nx; ++i) {
                const double x = -p.xmax + c->hx * i;
                const double v = lsurf3d ? imp::morse_tio2_3d(x, y, z - zsurf) : imp::morse_tio2_1d(z - zsurf);
                u[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k)] = std::min(3000.0, v);
            }
        }
    });
    return hetd_set_uext(c, u.data());
}

int hetd_set_state(hetd_ctx* c, const double* psi, const double r[3], const double v[3]) {
    if (!c || !psi || !r || !v) return HETD_ERR_ARG;
    if (!c->ready) return HETD_ERR_STATE;  // hetd_init failed; hetd_last_error says why
    if (!c->potion) return fail(c, HETD_ERR_STATE, "hetd_set_potion must precede hetd_set_state");
    drop_graph(c);
    const size_t nC = sizeof(double2) * c->n;
    HCHK(c, hipMemcpyAsync(c->psi, psi, nC, hipMemcpyHostToDevice, c->st));
    HCHK(c, launch_den_from_psi(c->psi, c->den, c->n, c->st));
    if (c->mixed) HCHK(c, launch_den32_from_psi(c->psi, c->den32, c->n, c->st));
    c->delj4_state = DELJ4_DEN;
    c->force_pending = false;
    c->pend_rec = nullptr;
    c->den64_ok = true;
    HCHK(c, hipMemsetAsync(c->q, 0, nC, c->st));
    HCHK(c, hipMemsetAsync(c->pc, 0, nC, c->st));
    for (int i = 0; i < 3; ++i) HCHK(c, hipMemsetAsync(c->psiold[i], 0, nC, c->st));
    for (int i = 0; i < 2; ++i) HCHK(c, hipMemsetAsync(c->hpsiold[i], 0, nC, c->st));
    ImpState s;
    std::memset(&s, 0, sizeof s);
    for (int d = 0; d < 3; ++d) { s.r[d] = r[d]; s.v[d] = v[d]; s.rpot[d] = r[d]; }
    HCHK(c, hipMemcpyAsync(c->imp, &s, sizeof s, hipMemcpyHostToDevice, c->st));
    HCHK(c, hipMemsetAsync(c->guard, 0, sizeof(int), c->st));
    HCHK(c, hipStreamSynchronize(c->st));
    for (int i = 0; i < 3; ++i) c->ioldp[i] = c->ioldr[i] = c->ioldv[i] = i + 1;
    for (int i = 0; i < 2; ++i) c->ioldh[i] = c->iolda[i] = i + 1;
    c->iter_since_set = 0;
    c->have_state = true;
    return HETD_OK;
}

// ----------------------------------------------------- integrator history
//
// Blob: header, the device ImpState, then psiold[0..2], hpsiold[0..1]
// (complex64 in mixed precision), pc and q, each nx*ny*nz values.  With the
// current psi of a mode-2 file it restarts a run bit for bit (the reference
// keeps no history and restarts with Runge-Kutta steps).

namespace {
struct HistHeader {
    char magic[8];
    int32_t version, nx, ny, nz, precision;
    int32_t ioldp[3], ioldh[2], ioldr[3], ioldv[3], iolda[2];
    int32_t pad;
    int64_t iter_since_set, bytes;
};
constexpr char kHistMagic[8] = {'H', 'E', 'T', 'D', 'H', 'I', 'S', 'T'};
}  // namespace

static size_t history_bytes(const hetd_ctx* c) {
    const size_t nC = sizeof(double2) * c->n, nHist = (c->mixed ? sizeof(float2) : sizeof(double2)) * c->n;
    return sizeof(HistHeader) + sizeof(ImpState) + 5 * nC + 2 * nHist;
}

int64_t hetd_history_size(hetd_ctx* c) { return c && c->ready ? (int64_t)history_bytes(c) : -1; }

int hetd_save_history(hetd_ctx* c, void* host) {
    if (!c || !host) return HETD_ERR_ARG;
    if (!c->ready) return HETD_ERR_STATE;  // hetd_init failed; hetd_last_error says why
    if (!c->have_state) return fail(c, HETD_ERR_STATE, "no state");
    if (int rc = settle_force(c)) return rc;
    HistHeader h{};
    std::memcpy(h.magic, kHistMagic, 8);
    h.version = 1; h.nx = c->nx; h.ny = c->ny; h.nz = c->nz; h.precision = c->mixed ? HETD_MIXED : HETD_FP64;
    for (int i = 0; i < 3; ++i) { h.ioldp[i] = c->ioldp[i]; h.ioldr[i] = c->ioldr[i]; h.ioldv[i] = c->ioldv[i]; }
    for (int i = 0; i < 2; ++i) { h.ioldh[i] = c->ioldh[i]; h.iolda[i] = c->iolda[i]; }
    h.iter_since_set = c->iter_since_set;
    h.bytes = (int64_t)history_bytes(c);
    char* p = static_cast<char*>(host);
    std::memcpy(p, &h, sizeof h);
    p += sizeof h;
    const size_t nC = sizeof(double2) * c->n, nHist = (c->mixed ? sizeof(float2) : sizeof(double2)) * c->n;
    HCHK(c, hipMemcpyAsync(p, c->imp, sizeof(ImpState), hipMemcpyDeviceToHost, c->st));
    p += sizeof(ImpState);
    for (int i = 0; i < 3; ++i, p += nC) HCHK(c, hipMemcpyAsync(p, c->psiold[i], nC, hipMemcpyDeviceToHost, c->st));
    for (int i = 0; i < 2; ++i, p += nHist)
        HCHK(c, hipMemcpyAsync(p, c->hpsiold[i], nHist, hipMemcpyDeviceToHost, c->st));
    HCHK(c, hipMemcpyAsync(p, c->pc, nC, hipMemcpyDeviceToHost, c->st));
    p += nC;
    HCHK(c, hipMemcpyAsync(p, c->q, nC, hipMemcpyDeviceToHost, c->st));
    HCHK(c, hipStreamSynchronize(c->st));
    return HETD_OK;
}

int hetd_load_history(hetd_ctx* c, const void* host, int64_t bytes) {
    if (!c || !host) return HETD_ERR_ARG;
    if (!c->ready) return HETD_ERR_STATE;  // hetd_init failed; hetd_last_error says why
    if (!c->have_state) return fail(c, HETD_ERR_STATE, "hetd_set_state must precede hetd_load_history");
    HistHeader h;
    if (bytes < (int64_t)sizeof h) return fail(c, HETD_ERR_ARG, "history blob too short");
    std::memcpy(&h, host, sizeof h);
    if (std::memcmp(h.magic, kHistMagic, 8) != 0 || h.version != 1)
        return fail(c, HETD_ERR_ARG, "not a history blob of this library");
    if (h.nx != c->nx || h.ny != c->ny || h.nz != c->nz || h.precision != (c->mixed ? HETD_MIXED : HETD_FP64) ||
        h.bytes != (int64_t)history_bytes(c) || bytes < h.bytes)
        return fail(c, HETD_ERR_ARG, "history blob is for another grid or precision");
    drop_graph(c);
    const char* p = static_cast<const char*>(host) + sizeof h;
    const size_t nC = sizeof(double2) * c->n, nHist = (c->mixed ? sizeof(float2) : sizeof(double2)) * c->n;
    HCHK(c, hipMemcpyAsync(c->imp, p, sizeof(ImpState), hipMemcpyHostToDevice, c->st));
    p += sizeof(ImpState);
    for (int i = 0; i < 3; ++i, p += nC) HCHK(c, hipMemcpyAsync(c->psiold[i], p, nC, hipMemcpyHostToDevice, c->st));
    for (int i = 0; i < 2; ++i, p += nHist)
        HCHK(c, hipMemcpyAsync(c->hpsiold[i], p, nHist, hipMemcpyHostToDevice, c->st));
    HCHK(c, hipMemcpyAsync(c->pc, p, nC, hipMemcpyHostToDevice, c->st));
    p += nC;
    HCHK(c, hipMemcpyAsync(c->q, p, nC, hipMemcpyHostToDevice, c->st));
    for (int i = 0; i < 3; ++i) { c->ioldp[i] = h.ioldp[i]; c->ioldr[i] = h.ioldr[i]; c->ioldv[i] = h.ioldv[i]; }
    for (int i = 0; i < 2; ++i) { c->ioldh[i] = h.ioldh[i]; c->iolda[i] = h.iolda[i]; }
    c->iter_since_set = h.iter_since_set;
    c->force_pending = false;  // the saved acceleration is that of the saved state
    // the mean field of the restored state, as the last step left it
    if (int rc = poten_dev(c)) return rc;
    HCHK(c, hipStreamSynchronize(c->st));
    return HETD_OK;
}

int hetd_potenimp(hetd_ctx* c, const double r[3]) {
    if (!c || !r) return HETD_ERR_ARG;
    if (!c->ready) return HETD_ERR_STATE;  // hetd_init failed; hetd_last_error says why
    HCHK(c, hipMemcpyAsync(c->imp->rpot, r, 3 * sizeof(double), hipMemcpyHostToDevice, c->st));
    HCHK(c, hipStreamSynchronize(c->st));
    return HETD_OK;
}

int hetd_poten(hetd_ctx* c) {
    if (!c) return HETD_ERR_ARG;
    if (!c->ready) return HETD_ERR_STATE;  // hetd_init failed; hetd_last_error says why
    if (!c->have_state) return fail(c, HETD_ERR_STATE, "no state");
    // called on its own: the full form, so that every field of poten is formed
    int rc = poten_dev(c, false, false, true, true);
    if (rc) return rc;
    HCHK(c, hipStreamSynchronize(c->st));
    return HETD_OK;
}

int hetd_forceimp(hetd_ctx* c, const double r[3], double F[3]) {
    (void)r;  // the reference's forceimp uses the u_imp of the last potenimp
    if (!c || !F) return HETD_ERR_ARG;
    if (!c->ready) return HETD_ERR_STATE;  // hetd_init failed; hetd_last_error says why
    if (!c->have_state) return fail(c, HETD_ERR_STATE, "no state");
    int rc = force_dev(c);
    if (rc) return rc;
    HCHK(c, hipMemcpyAsync(F, c->imp->F, 3 * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HCHK(c, hipStreamSynchronize(c->st));
    return HETD_OK;
}

int hetd_set_aimp(hetd_ctx* c, const double a[3]) {
    if (!c || !a) return HETD_ERR_ARG;
    if (!c->ready) return HETD_ERR_STATE;  // hetd_init failed; hetd_last_error says why
    HCHK(c, hipMemcpyAsync(c->imp->a, a, 3 * sizeof(double), hipMemcpyHostToDevice, c->st));
    HCHK(c, hipStreamSynchronize(c->st));
    return HETD_OK;
}

int hetd_steprk(hetd_ctx* c, double dt) {
    if (!c) return HETD_ERR_ARG;
    if (!c->ready) return HETD_ERR_STATE;  // hetd_init failed; hetd_last_error says why
    if (!c->have_state) return fail(c, HETD_ERR_STATE, "no state");
    int rc = steprk_dev(c, dt, false);
    if (rc) return rc;
    c->iter_since_set++;
    return check_guard(c);
}

int hetd_steppc(hetd_ctx* c, double dt, double* errHe, double* errimp) {
    if (!c || !errHe || !errimp) return HETD_ERR_ARG;
    if (!c->ready) return HETD_ERR_STATE;  // hetd_init failed; hetd_last_error says why
    if (!c->have_state) return fail(c, HETD_ERR_STATE, "no state");
    int rc = steppc_dev(c, dt, false);
    if (rc) return rc;
    c->iter_since_set++;
    double e[2];
    HCHK(c, hipMemcpyAsync(e, &c->imp->errHe, 2 * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HCHK(c, hipStreamSynchronize(c->st));
    *errHe = e[0];
    *errimp = e[1];
    return HETD_OK;
}

int hetd_evolve(hetd_ctx* c, double dt, int32_t nsteps, double* traj) {
    if (!c || nsteps < 0) return HETD_ERR_ARG;
    if (!c->ready) return HETD_ERR_STATE;  // hetd_init failed; hetd_last_error says why
    if (!c->have_state) return fail(c, HETD_ERR_STATE, "no state");
    if (traj && c->traj_cap < (long)nsteps * HETD_TRAJ_STRIDE) {
        if (c->traj) {
            HCHK(c, hipStreamSynchronize(c->st));
            HCHK(c, hipFree(c->traj));
            c->traj = nullptr;
        }
        HCHK(c, hipMalloc(&c->traj, sizeof(double) * (size_t)nsteps * HETD_TRAJ_STRIDE));
        c->traj_cap = (long)nsteps * HETD_TRAJ_STRIDE;
    }
    for (int s = 0; s < nsteps;) {
        int done = 0;
        const int rc = evolve_unit(c, dt, nsteps - s, traj ? c->traj + (long)s * HETD_TRAJ_STRIDE : nullptr, &done);
        if (rc) return rc;
        s += done;
    }
    if (int rc = settle_force(c)) return rc;
    if (traj && nsteps > 0)
        HCHK(c, hipMemcpyAsync(traj, c->traj, sizeof(double) * (size_t)nsteps * HETD_TRAJ_STRIDE,
                               hipMemcpyDeviceToHost, c->st));
    return check_guard(c);
}

int hetd_pool_evolve(hetd_ctx* const* ctxs, int32_t n, double dt, int32_t nsteps) {
    if (!ctxs || n < 0 || nsteps < 0) return HETD_ERR_ARG;
    for (int i = 0; i < n; ++i) {
        if (!ctxs[i]) return HETD_ERR_ARG;
        if (!ctxs[i]->ready) return HETD_ERR_STATE;
        if (!ctxs[i]->have_state) return fail(ctxs[i], HETD_ERR_STATE, "no state");
    }
    // units of work are issued round-robin, so every member's stream always
    // has queued work and the members run side by side on the GPU
    std::vector<int> left(n, nsteps);
    for (bool busy = true; busy;) {
        busy = false;
        for (int i = 0; i < n; ++i) {
            if (left[i] <= 0) continue;
            int done = 0;
            const int rc = evolve_unit(ctxs[i], dt, left[i], nullptr, &done);
            if (rc) return rc;
            left[i] -= done;
            busy = true;
        }
    }
    for (int i = 0; i < n; ++i)
        if (int rc = settle_force(ctxs[i])) return rc;
    for (int i = 0; i < n; ++i)
        if (int rc = check_guard(ctxs[i])) return rc;
    return HETD_OK;
}

// ------------------------------------------------------------ autotuner
//
// Chooses, by timing on the context's own grid and state, among kernel
// variants that compute the same bits, so that tuning never changes a result:
// the x/y register passes with a complex or a split LDS exchange, and the
// update kernels with one or two LDS planes, the latter also with
// non-temporal loads and stores of the operands used once.  Every device buffer is copied
// aside first and copied back after each trial, with the host-side step state.

int hetd_autotune(hetd_ctx* c, double dt, int32_t steps, int32_t* choice) {
    if (!c || steps < 1 || !(dt > 0.0)) return HETD_ERR_ARG;
    if (!c->ready) return HETD_ERR_STATE;  // hetd_init failed; hetd_last_error says why
    if (!c->have_state) return fail(c, HETD_ERR_STATE, "hetd_set_state must precede hetd_autotune");
    if (c->tclass) return fail(c, HETD_ERR_STATE, "hetd_autotune with kernel-class timing on");
    HCHK(c, hipStreamSynchronize(c->st));
    std::vector<void*> snap;
    hipEvent_t ev[2] = {nullptr, nullptr};
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(c->st);
        for (void* p : snap) (void)hipFree(p);
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    };
    const size_t nsnap = c->allocs.size();
    for (size_t i = 0; i < nsnap; ++i) {
        void* p = nullptr;
        if (hipMalloc(&p, c->allocs[i].second) != hipSuccess) {
            cleanup();
            return fail(c, HETD_ERR_HIP, "hetd_autotune: no device memory for the state copy");
        }
        snap.push_back(p);
        if (hipMemcpyAsync(p, c->allocs[i].first, c->allocs[i].second, hipMemcpyDeviceToDevice, c->st) != hipSuccess) {
            cleanup();
            return fail(c, HETD_ERR_HIP, "hetd_autotune: state copy failed");
        }
    }
    const std::vector<long> sig = step_sig(c);
    const long iters = c->iter_since_set;
    double* const prec = c->pend_rec;
    const bool xs0 = c->xsplit;
    const int ut0 = c->upd_tile;
    auto restore = [&]() -> int {
        for (size_t i = 0; i < nsnap; ++i)
            HCHK(c, hipMemcpyAsync(c->allocs[i].first, snap[i], c->allocs[i].second, hipMemcpyDeviceToDevice, c->st));
        restore_sig(c, sig);
        c->iter_since_set = iters;
        c->pend_rec = prec;
        HCHK(c, hipStreamSynchronize(c->st));
        return 0;
    };
    auto give_up = [&](int rc) {
        c->xsplit = xs0;
        c->upd_tile = ut0;
        drop_graph(c);
        (void)restore();
        cleanup();
        return rc;
    };
    for (hipEvent_t& e : ev)
        if (hipEventCreate(&e) != hipSuccess) return give_up(fail(c, HETD_ERR_HIP, "hetd_autotune: events"));
    float best = 0.0f;
    int bestv = 0;
    static const int kTile[3] = {0, 4, 5};  // default, two LDS planes, and non-temporal operand streams
    for (int v = 0; v < 6; ++v) {
        c->xsplit = (v & 1) != 0;
        c->upd_tile = kTile[v >> 1];
        drop_graph(c);
        // an untimed pass first: it builds this variant's step graph
        if (int rc = hetd_evolve(c, dt, steps, nullptr)) return give_up(rc);
        if (int rc = restore()) return give_up(rc);
        if (hipEventRecord(ev[0], c->st) != hipSuccess) return give_up(fail(c, HETD_ERR_HIP, "event"));
        if (int rc = hetd_evolve(c, dt, steps, nullptr)) return give_up(rc);
        if (hipEventRecord(ev[1], c->st) != hipSuccess) return give_up(fail(c, HETD_ERR_HIP, "event"));
        if (hipStreamSynchronize(c->st) != hipSuccess) return give_up(fail(c, HETD_ERR_HIP, "sync"));
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess) return give_up(fail(c, HETD_ERR_HIP, "timing"));
        if (int rc = restore()) return give_up(rc);
        if (v == 0 || ms < best) {
            best = ms;
            bestv = v;
        }
    }
    c->xsplit = (bestv & 1) != 0;
    c->upd_tile = kTile[bestv >> 1];
    drop_graph(c);
    cleanup();
    if (choice) *choice = bestv;
    return HETD_OK;
}

int hetd_energy(hetd_ctx* c, hetd_energies* e) {
    if (!c || !e) return HETD_ERR_ARG;
    if (!c->ready) return HETD_ERR_STATE;  // hetd_init failed; hetd_last_error says why
    if (!c->have_state) return fail(c, HETD_ERR_STATE, "no state");
    if (int rc = ensure_delj4(c)) return rc;
    if (int rc = ensure_dencg(c)) return rc;
    EnergyArgs a{};
    a.g = c->grid; a.sc = c->sc; a.psi = c->psi; a.den = c->den; a.dencg = c->dencg; a.delj4 = c->delj4;
    a.imp = c->imp; a.aux1 = 0.5 * c->p.cp4; a.aux2 = (1.0 / 3.0) * c->p.cpp4; a.beta = c->p.beta_solid;
    a.den_m = c->p.den_m; a.lsolid = c->p.lsolid; a.otc = c->otc ? 1 : 0;
    a.falfs = c->falfs; a.dxden = c->dxden; a.dyden = c->dyden; a.dzden = c->dzden;
    a.intx = c->intx; a.inty = c->inty; a.intz = c->intz; a.partial = c->partial;
    if (c->mixed) { a.dencgf = c->dencg32; a.delj4f = c->delj4_32; }
    int nb = 0;
    HCHK(c, tiled_supported(c->sc) ? launch_energy_tiled(a, c->st, &nb) : launch_energy_partial(a, c->st, &nb));
    HCHK(c, launch_reduce_final(c->partial, nb, EN_NCOMP, c->red, c->st));
    double s[EN_NCOMP];
    ImpState is;
    HCHK(c, hipMemcpyAsync(s, c->red, sizeof s, hipMemcpyDeviceToHost, c->st));
    HCHK(c, hipMemcpyAsync(&is, c->imp, sizeof is, hipMemcpyDeviceToHost, c->st));
    HCHK(c, hipStreamSynchronize(c->st));
    const hetd_params& p = c->p;
    // energy.f90:99-132
    e->ealphas = c->otc ? -p.h2o2m4 * 0.5 * p.alphas * s[EN_ALPHA] * c->dxyz : 0.0;
    e->esolid = p.lsolid ? p.c_solid * s[EN_SOLID] * c->dxyz : 0.0;
    e->ekin4 = p.h2o2m4 * s[EN_KIN] * c->dxyz;
    e->elj4 = 0.5 * s[EN_LJ] * c->dxyz;
    e->ecor4 = s[EN_COR] * c->dxyz;
    e->etot4 = e->ekin4 + e->elj4 + e->ecor4 + e->esolid;
    if (c->otc) e->etot4 = e->etot4 + e->ealphas;
    if (c->uext_on) {  // energy.f90:124
        const int cap = (int)std::min<long>(1024, c->partial_len);
        HCHK(c, launch_dot_uext(c->psi, c->uext, c->n, cap, c->partial, c->st));
        HCHK(c, launch_reduce_final(c->partial, cap, 1, c->red + 8, c->st));
        double su = 0.0;
        HCHK(c, hipMemcpyAsync(&su, c->red + 8, sizeof su, hipMemcpyDeviceToHost, c->st));
        HCHK(c, hipStreamSynchronize(c->st));
        e->etot4 = e->etot4 + su * c->dxyz;
    }
    e->ekinx = 0.5 * c->mAg * (is.v[0] * is.v[0] + is.v[1] * is.v[1] + is.v[2] * is.v[2]);
    e->eHeX = s[EN_HEX] * c->dxyz;
    e->eimpu = e->ekinx + e->eHeX;
    e->etot = e->etot4 + e->eimpu;
    return HETD_OK;
}

int hetd_moments(hetd_ctx* c, double npart, double out[10]) {
    if (!c || !out || !(npart != 0.0)) return HETD_ERR_ARG;
    if (!c->ready) return HETD_ERR_STATE;  // hetd_init failed; hetd_last_error says why
    if (!c->have_state) return fail(c, HETD_ERR_STATE, "no state");
    EnergyArgs a{};
    a.g = c->grid; a.sc = c->sc; a.psi = c->psi; a.den = c->den; a.dencg = c->dencg; a.delj4 = c->delj4;
    a.imp = c->imp; a.partial = c->partial;
    double s[EN_NCOMP];
    for (int pass = 1; pass <= 2; ++pass) {
        a.mom = pass;
        int nb = 0;
        HCHK(c, tiled_supported(c->sc) ? launch_energy_tiled(a, c->st, &nb) : launch_energy_partial(a, c->st, &nb));
        HCHK(c, launch_reduce_final(c->partial, nb, EN_NCOMP, c->red, c->st));
        HCHK(c, hipMemcpyAsync(s, c->red, sizeof s, hipMemcpyDeviceToHost, c->st));
        HCHK(c, hipStreamSynchronize(c->st));
        if (pass == 1) {
            // r_cm.f90:52-54 (divided by npart), DFT4He3d.f90:944-947, 869
            for (int d = 0; d < 3; ++d) {
                out[d] = s[d] * c->dxyz / npart;
                out[3 + d] = s[3 + d] * c->dxyz;
                a.cm[d] = out[d];
            }
            out[6] = s[6] * c->dxyz;
        } else {
            for (int d = 0; d < 3; ++d) out[7 + d] = s[d] * c->dxyz;  // DFT4He3d.f90:899,912,925
        }
    }
    return HETD_OK;
}

int hetd_norm(hetd_ctx* c, double* norm) {
    if (!c || !norm) return HETD_ERR_ARG;
    if (!c->ready) return HETD_ERR_STATE;  // hetd_init failed; hetd_last_error says why
    EnergyArgs a{};
    a.g = c->grid; a.sc = c->sc; a.psi = c->psi; a.den = c->den; a.dencg = c->dencg; a.delj4 = c->delj4;
    a.imp = c->imp; a.partial = c->partial;
    if (c->mixed) { a.dencgf = c->dencg32; a.delj4f = c->delj4_32; }
    int nb = 0;
    HCHK(c, tiled_supported(c->sc) ? launch_energy_tiled(a, c->st, &nb) : launch_energy_partial(a, c->st, &nb));
    HCHK(c, launch_reduce_final(c->partial, nb, EN_NCOMP, c->red, c->st));
    double s[EN_NCOMP];
    HCHK(c, hipMemcpyAsync(s, c->red, sizeof s, hipMemcpyDeviceToHost, c->st));
    HCHK(c, hipStreamSynchronize(c->st));
    *norm = s[EN_NORM] * c->dxyz;
    return HETD_OK;
}

int hetd_get_imp(hetd_ctx* c, double r[3], double v[3], double a[3]) {
    if (!c || !r || !v || !a) return HETD_ERR_ARG;
    if (!c->ready) return HETD_ERR_STATE;  // hetd_init failed; hetd_last_error says why
    ImpState is;
    HCHK(c, hipMemcpyAsync(&is, c->imp, sizeof is, hipMemcpyDeviceToHost, c->st));
    HCHK(c, hipStreamSynchronize(c->st));
    for (int d = 0; d < 3; ++d) { r[d] = is.r[d]; v[d] = is.v[d]; a[d] = is.a[d]; }
    return HETD_OK;
}

// FP32 field of the mixed path widened into the FP64 scratch tmp
static const double* widen(hetd_ctx* c, const float* f, long n, double* dst) {
    if (launch_to_double(f, dst, n, c->st) != hipSuccess) return nullptr;
    return dst;
}

int hetd_get_field(hetd_ctx* c, int32_t which, double* host) {
    if (!c || !host) return HETD_ERR_ARG;
    if (!c->ready) return HETD_ERR_STATE;  // hetd_init failed; hetd_last_error says why
    const size_t nR = sizeof(double) * c->n, nC = sizeof(double2) * c->n;
    const void* src = nullptr;
    size_t bytes = nR;
    size_t hrow = 0;  // element size of a half-spectrum field (rows padded to ldh)
    int rc = 0;
    const bool mx = c->mixed;
    if ((rc = ensure_den64(c))) return rc;
    switch (which) {
    case HETD_F_PSI: src = c->psi; bytes = nC; break;
    case HETD_F_DEN: src = c->den; break;
    case HETD_F_POT4: src = mx ? widen(c, c->pot4_32, c->n, c->tmp) : c->pot4; break;
    case HETD_F_DENCG:
        if ((rc = ensure_dencg(c))) return rc;
        src = mx ? widen(c, c->dencg32, c->n, c->tmp) : c->dencg;
        break;
    case HETD_F_DELJ4:
        if ((rc = ensure_delj4(c))) return rc;
        src = mx ? widen(c, c->delj4_32, c->n, c->tmp) : c->delj4;
        break;
    case HETD_F_PC: src = c->pc; bytes = nC; break;
    case HETD_F_Q: src = c->q; bytes = nC; break;
    case HETD_F_FVLJ: src = c->fvlj; hrow = sizeof(double); break;
    case HETD_F_WCGK: src = c->wcgk; hrow = sizeof(double); break;
    case HETD_F_CCOR:
    case HETD_F_FDEN: {
        // recompute the last poten in its full form (den is unchanged since)
        // keeping the intermediate; dencg and delj4 come out bit for bit as
        // the stepping form makes them, pot4 is left as it is (OTC, whose
        // stepping form is the full one, rewrites it bit for bit)
        if (!c->have_state) return fail(c, HETD_ERR_STATE, "no state");
        if (which == HETD_F_FDEN && c->otc) return fail(c, HETD_ERR_ARG, "FDEN not kept with OTC");
        if ((rc = poten_dev(c, which == HETD_F_CCOR, which == HETD_F_FDEN, c->otc))) return rc;
        if (which == HETD_F_CCOR) {
            src = mx ? widen(c, c->tmp32, c->n, c->tmp) : (const void*)c->tmp;
        } else {
            src = mx ? widen(c, reinterpret_cast<const float*>(c->C32), 2 * c->nhp, reinterpret_cast<double*>(c->A))
                     : (const void*)c->Cs;
            hrow = sizeof(double2);
        }
        break;
    }
    case HETD_F_DXDEN:
    case HETD_F_DYDEN:
    case HETD_F_DZDEN:
        if ((rc = deriv_dev(c, 1, which - HETD_F_DXDEN + 1, 0, c->den, c->tmp))) return rc;
        src = c->tmp;
        break;
    case HETD_F_UIMP:
        HCHK(c, launch_uimp_field(c->grid, c->imp, c->tmp, c->st));
        src = c->tmp;
        break;
    case HETD_F_PENALTY:
        if (!c->p.lsolid) return fail(c, HETD_ERR_ARG, "no penalty without lsolid");
        HCHK(c, launch_penalty_field(c->den, c->p.c_solid, c->p.beta_solid, c->p.den_m, c->tmp, c->n, c->st));
        src = c->tmp;
        break;
    case HETD_F_UALPHAS:
    case HETD_F_DENALF:
        if (!c->otc) return fail(c, HETD_ERR_ARG, "alpha_s fields exist only with core OTC");
        src = which == HETD_F_UALPHAS ? c->ualphas : c->denalf;
        break;
    default:
        return fail(c, HETD_ERR_ARG, "unknown field %d", which);
    }
    if (!src) return fail(c, HETD_ERR_HIP, "conversion launch failed");
    if (hrow)  // half spectrum: drop the row padding
        HCHK(c, hipMemcpy2DAsync(host, hrow * c->nxh, src, hrow * c->ldh, hrow * c->nxh, (size_t)c->ny * c->nz,
                                 hipMemcpyDeviceToHost, c->st));
    else
        HCHK(c, hipMemcpyAsync(host, src, bytes, hipMemcpyDeviceToHost, c->st));
    HCHK(c, hipStreamSynchronize(c->st));
    return HETD_OK;
}

int hetd_fft_r2c(hetd_ctx* c, const double* in, double* out) {
    if (!c || !in || !out) return HETD_ERR_ARG;
    if (!c->ready) return HETD_ERR_STATE;  // hetd_init failed; hetd_last_error says why
    HCHK(c, hipMemcpyAsync(c->tmp, in, sizeof(double) * c->n, hipMemcpyHostToDevice, c->st));
    int rc = r2c_dev(c, c->tmp, c->A);
    if (rc) return rc;
    HCHK(c, hipMemcpy2DAsync(out, sizeof(double2) * c->nxh, c->A, sizeof(double2) * c->ldh, sizeof(double2) * c->nxh,
                             (size_t)c->ny * c->nz, hipMemcpyDeviceToHost, c->st));
    HCHK(c, hipStreamSynchronize(c->st));
    return HETD_OK;
}

int hetd_fft_c2r(hetd_ctx* c, const double* in, double* out) {
    if (!c || !in || !out) return HETD_ERR_ARG;
    if (!c->ready) return HETD_ERR_STATE;  // hetd_init failed; hetd_last_error says why
    HCHK(c, hipMemcpy2DAsync(c->A, sizeof(double2) * c->ldh, in, sizeof(double2) * c->nxh, sizeof(double2) * c->nxh,
                             (size_t)c->ny * c->nz, hipMemcpyHostToDevice, c->st));
    int rc = c2r_dev(c, c->A, c->tmp, c->renor);
    if (rc) return rc;
    HCHK(c, hipMemcpyAsync(out, c->tmp, sizeof(double) * c->n, hipMemcpyDeviceToHost, c->st));
    HCHK(c, hipStreamSynchronize(c->st));
    return HETD_OK;
}

int hetd_deriv(hetd_ctx* c, int32_t id, int32_t iv, int32_t is_complex, const double* f, double* df) {
    if (!c || !f || !df || id < 1 || id > 2 || iv < 1 || iv > 3) return HETD_ERR_ARG;
    if (!c->ready) return HETD_ERR_STATE;  // hetd_init failed; hetd_last_error says why
    const size_t bytes = (is_complex ? sizeof(double2) : sizeof(double)) * c->n;
    for (int i = 0; i < 2; ++i)
        if (!c->dtmp[i]) HCHK(c, hipMalloc(&c->dtmp[i], sizeof(double2) * c->n));
    HCHK(c, hipMemcpyAsync(c->dtmp[0], f, bytes, hipMemcpyHostToDevice, c->st));
    int rc = deriv_dev(c, id, iv, is_complex ? 1 : 0, c->dtmp[0], c->dtmp[1]);
    if (rc) return rc;
    HCHK(c, hipMemcpyAsync(df, c->dtmp[1], bytes, hipMemcpyDeviceToHost, c->st));
    HCHK(c, hipStreamSynchronize(c->st));
    return HETD_OK;
}

int hetd_timing(hetd_ctx* c, int32_t kclass) {
    if (!c || kclass < HETD_K_OFF || kclass > HETD_K_FORCE) return HETD_ERR_ARG;
    if (!c->ready) return HETD_ERR_STATE;  // hetd_init failed; hetd_last_error says why
    HCHK(c, hipStreamSynchronize(c->st));
    c->tclass = kclass;
    c->tused = 0;
    c->tlaunch = 0;
    c->tbytes = 0;
    return HETD_OK;
}

int hetd_timing_read(hetd_ctx* c, double* total_ms, int64_t* launches, double* bytes) {
    if (!c || !total_ms || !launches || !bytes) return HETD_ERR_ARG;
    if (!c->ready) return HETD_ERR_STATE;  // hetd_init failed; hetd_last_error says why
    HCHK(c, hipStreamSynchronize(c->st));
    double t = 0;
    for (size_t i = 0; i + 1 < c->tused; i += 2) {
        float ms = 0;
        HCHK(c, hipEventElapsedTime(&ms, c->tev[i], c->tev[i + 1]));
        t += ms;
    }
    *total_ms = t;
    *launches = c->tlaunch;
    *bytes = c->tbytes;
    return HETD_OK;
}

}  // extern "C"
