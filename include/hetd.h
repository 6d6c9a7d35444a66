/*
 * hetd.h -- C ABI of the MI355X (gfx950) real-time propagation step of the
 * 4He TDDFT solver (Orsay-Trento functional + one classical impurity).
 *
 * The entry points replace the reference's Fortran routines for this path
 * one for one; each declaration names the routine it replaces (file:line in
 * the reference repository).  The reference keeps its state in module
 * globals; here it lives in an opaque context whose fields stay resident in
 * device memory (HBM) between calls.  Host pointers are used only for
 * explicit copies.  Arrays use the reference's layout: Fortran column-major
 * (x fastest), complex values interleaved (re, im) as complex(8).
 *
 * Every function returns 0 on success.  A nonzero status mirrors a STOP in
 * the reference; hetd_last_error() then describes it.  A context is not
 * re-entrant: one host thread drives it, and all its work is ordered on one
 * HIP stream.
 */
#ifndef HETD_H
#define HETD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define HETD_OK 0
#define HETD_ERR_ARG 1      /* invalid argument or configuration */
#define HETD_ERR_HIP 2      /* HIP runtime error (allocation, launch) */
#define HETD_ERR_GUARD 3    /* rho not > 0 somewhere (steprk.f90:89-92) */
#define HETD_ERR_STATE 4    /* call made in the wrong order */

/* arithmetic of the fields (the integrator state is FP64 in both) */
#define HETD_FP64 0
#define HETD_MIXED 1

/* core4: 'OT ' (core only) or 'OTC' (with the alpha_s term) (modules.f90:214) */
#define HETD_CORE_OT 0
#define HETD_CORE_OTC 1

/* field identifiers for hetd_get_field */
#define HETD_F_PSI 0      /* complex(8) psi */
#define HETD_F_DEN 1      /* den = |psi|^2 */
#define HETD_F_POT4 2     /* mean field U (poten.f90:90-114) */
#define HETD_F_DENCG 3    /* coarse-grained density (poten.f90:34-38) */
#define HETD_F_DELJ4 4    /* LJ convolution (poten.f90:46-50) */
#define HETD_F_CCOR 5     /* correlation convolution sto1 (poten.f90:62-74) */
#define HETD_F_DXDEN 6    /* d den/dx (derden.f90:26) */
#define HETD_F_DYDEN 7
#define HETD_F_DZDEN 8
#define HETD_F_UIMP 9     /* impurity field u_imp (potenimp.f90:71-95) */
#define HETD_F_PENALTY 10 /* solid penalty (poten.f90:80-83), in the cancellation-free
                             form the mean field uses (1 + tanh x as 2e^{2x}/(1+e^{2x})) */
#define HETD_F_UALPHAS 11 /* alpha_s field (term_alfa.f90) */
#define HETD_F_DENALF 12  /* Gaussian-smoothed density (poten.f90:96-101) */
#define HETD_F_FVLJ 13    /* half spectrum: LJ form factor (fforma.f90) */
#define HETD_F_WCGK 14    /* half spectrum: coarse-graining kernel (initcg.f90) */
#define HETD_F_PC 15      /* complex: predictor-corrector difference pc */
#define HETD_F_Q 16       /* complex: RK-Gill accumulator q */
#define HETD_F_FDEN 17    /* complex half spectrum: r2c of den (poten.f90:27) */

/* per-step trajectory record written by hetd_evolve */
#define HETD_TRAJ_STRIDE 11 /* errHe, errimp, rimp[3], vimp[3], aimp[3] */

typedef struct hetd_params {
    int32_t nx, ny, nz;   /* grid points (namelist nx, ny, nz) */
    int32_t npd;          /* stencil points, odd 3..25 (modules.f90:30) */
    int32_t icon;         /* 13: zero outside the box; 8: periodic; 18: antiperiodic (modules.f90:34) */
    int32_t core;         /* HETD_CORE_OT / HETD_CORE_OTC */
    int32_t lsolid;       /* solid penalty; forces core OT (DFT4He3d.f90:476) */
    int32_t precision;    /* HETD_FP64 or HETD_MIXED (npd 9 or 13) */
    double xmax, ymax, zmax;      /* box half-widths, h = 2|xmax|/nx (DFT4He3d.f90:274) */
    /* Orsay-Trento functional (modules.f90:130-150, 204-216) */
    double eps4, sigma4, h4;      /* LJ well, size, core radius (h4ot) */
    double cp4, cpp4, h2o2m4;     /* correlation constants, hbar^2/2m4 */
    double alphas, l_alpha, den0s;/* alpha_s term */
    double c_solid, beta_solid, den_m; /* solid penalty C, beta, den_m */
    double mimpur;                /* impurity mass (amu) */
    /* time constant timec (DFT4He3d.f90:689-699) */
    double Lambdah, txmean, txsurf, tymean, tysurf, tzmean, tzsurf;
} hetd_params;

typedef struct hetd_energies {   /* energy.f90:112-132 */
    double etot4, ekin4, elj4, ecor4, esolid, ealphas;
    double ekinx, eHeX, eimpu, etot;
} hetd_energies;

typedef struct hetd_ctx hetd_ctx;

/* Fill *p with the reference defaults (modules.f90, DFT4He3d.f90:52-145). */
void hetd_default_params(hetd_params *p);

/* Allocate the device context: replaces dimen (dimen.f90:29-127), fftini
 * (fft.f90:5-76), Init_deriv_p (init_deriv_parallel.f90:51-83), fforma
 * (fforma.f90:11-62), initcg (initcg.f90:5-51) and kalfs
 * (DFT4He3d.f90:592-596).  On failure *ctx still receives the context:
 * hetd_last_error says why, hetd_finalize releases it, and every other call
 * on it returns HETD_ERR_STATE. */
int hetd_init(const hetd_params *p, hetd_ctx **ctx);
/* Release everything (end of DFT4He3d.f90). */
int hetd_finalize(hetd_ctx *ctx);
const char *hetd_last_error(hetd_ctx *ctx);

/* Impurity pair potential V(r) = Select_Pot(selec, r, r_cutoff, umax)
 * (V_impur.f90:84-929); selec is any name Select_Pot accepts ("Xe_He",
 * "null", "RB", "Ba_plus_gs", ...).  Returns 0 and writes *v, or
 * HETD_ERR_ARG for an unknown selec (the reference STOPs,
 * V_impur.f90:925-927). */
int hetd_vimp(const char *selec, double r, double r_cutoff, double umax, double *v);
/* Interpolation table of potenimpini (potenimpini.f90:11-25):
 * npot = 100*max(nx,ny,nz)+1, delinter = 4*max(xmax,ymax,zmax)/(npot-1),
 * potion[i] = V(i*delinter).  potion must hold npot values (or be NULL to
 * query npot).  HETD_ERR_ARG for an unknown selec, also on the query. */
int hetd_build_potion(const char *selec, double r_cutoff, double umax,
                      int32_t nmax, double xmaxbox, double *potion,
                      int32_t *npot, double *delinter);
/* Upload the table (module interpol, modules.f90:363-369). */
int hetd_set_potion(hetd_ctx *ctx, const double *potion, int32_t npot, double delinter);

/* He-TiO2 surface potential (K) at a point z above the surface: the
 * laterally averaged Morse_HeTiO2_1D(z) (morse.f90:1-16) for lsurf3d = 0,
 * the corrugated Morse_HeTiO2_3D(x, y, z) (morse.f90:36-99) otherwise. */
int hetd_vsurf(int32_t lsurf3d, double x, double y, double z, double *v);
/* External potential uext (nx*ny*nz, Fortran order; NULL removes it).  As in
 * the reference it enters only the energy: etot4 += sum(uext*den)*dxyz
 * (energy.f90:124); the dynamics do not see it. */
int hetd_set_uext(hetd_ctx *ctx, const double *uext);
/* lsurf: uext = min(3000, hetd_vsurf(lsurf3d, x, y, z - zsurf)) on the grid
 * (DFT4He3d.f90:758-784), set with hetd_set_uext. */
int hetd_set_surface(hetd_ctx *ctx, int32_t lsurf3d, double zsurf);

/* Upload psi (complex(8), nx*ny*nz) and the impurity R, v (internal units,
 * v in A*K); sets den = |psi|^2 (readenc.f90:104-113, DFT4He3d.f90:637),
 * zeroes q, pc and the impurity histories (dimen.f90:117-118,
 * DFT4He3d.f90:796-802). */
int hetd_set_state(hetd_ctx *ctx, const double *psi, const double r[3], const double v[3]);

/* potenimp(rimp): u_imp for an impurity at r (potenimp.f90:1-20,71-95). */
int hetd_potenimp(hetd_ctx *ctx, const double r[3]);
/* poten(): mean field from den (poten.f90:5-118, with term_alfa.f90 for OTC
 * and derden.f90). */
int hetd_poten(hetd_ctx *ctx);
/* forceimp(rimp,F): F = -sum(grad den * u_imp) dV (potenimp.f90:24-55). */
int hetd_forceimp(hetd_ctx *ctx, const double r[3], double F[3]);
/* The driver's aimp = F/mAg_u (DFT4He3d.f90:644,852): set the device aimp. */
int hetd_set_aimp(hetd_ctx *ctx, const double a[3]);
/* steprk(deltat): one Runge-Kutta-Gill step of psi and (R, v)
 * (steprk.f90:1-153), including the potenimp/poten/forceimp of stages 1-3.
 * Returns HETD_ERR_GUARD if den is not > 0 somewhere (steprk.f90:89-92). */
int hetd_steprk(hetd_ctx *ctx, double deltat);
/* steppc(deltat,errHe,errimp): one Hamming predictor-modifier-corrector
 * step (steppc.f90:1-178), including the middle potenimp/poten/forceimp. */
int hetd_steppc(hetd_ctx *ctx, double deltat, double *errHe, double *errimp);
/* energy(): energy.f90:12-144 (uses the fields of the last poten). */
int hetd_energy(hetd_ctx *ctx, hetd_energies *e);
/* sum(den)*dxyz (DFT4He3d.f90:869). */
int hetd_norm(hetd_ctx *ctx, double *norm);
/* Autotuner (no counterpart in the reference; for callers that run many
 * steps on one grid).  Times, over `steps` steps from the current state on
 * this context's own grid, the kernel variants that compute the same bits:
 * x/y register FFT passes with a complex or a split LDS exchange, and update
 * kernels with one or two LDS planes (and, with two, with non-temporal
 * operand streams).  Keeps the fastest for this context and leaves the state
 * exactly as it was: every device buffer is copied aside during the call, so
 * the context's device memory is needed twice.  *choice (may be NULL): bit 0
 * = split exchanges; bits 1-2 = update kernels (0 default, 1 two planes,
 * 2 two planes and non-temporal streams). */
int hetd_autotune(hetd_ctx *ctx, double deltat, int32_t steps, int32_t *choice);
/* The diagnostics of the driver's energy block (DFT4He3d.f90:868-947), as
 * device reductions: out[0..2] = centre of mass sum(den*x)*dxyz/npart
 * (r_cm.f90:23-55), out[3..5] = sum(den*x^2)*dxyz (not about the centre),
 * out[6] = sum(den)*dxyz, out[7..9] = <Lx,Ly,Lz> =
 * Re(sum(i conj(psi) ((y-ycm) d/dz ... )))*dxyz about that centre of mass,
 * with the first-derivative stencil of derivnD. */
int hetd_moments(hetd_ctx *ctx, double npart, double out[10]);
/* nsteps iterations of the driver loop body (DFT4He3d.f90:840-852): steprk
 * for the first 3 iterations after hetd_set_state, steppc afterwards, then
 * potenimp/poten/forceimp and aimp = F/mAg_u at the new state -- entirely on
 * the device, no host synchronisation inside.  traj (host, may be NULL)
 * receives nsteps*HETD_TRAJ_STRIDE doubles. */
int hetd_evolve(hetd_ctx *ctx, double deltat, int32_t nsteps, double *traj);

/* Integrator history for a bit-identical restart (no reference counterpart:
 * the reference's checkpoint holds psi, R and v only, and a restart begins
 * again with Runge-Kutta steps, DFT4He3d.f90:840).  The blob holds the ring
 * of psiold/hpsiold, pc, q, the ring indices, the step count and the
 * impurity state with its histories; hetd_history_size gives its size in
 * bytes.  hetd_save_history forms a pending impurity force first.
 * hetd_load_history follows hetd_set_state with the saved psi, R, v and
 * recomputes the mean field; the run then continues as the saved one would
 * have after its last hetd_evolve. */
int64_t hetd_history_size(hetd_ctx *ctx);
int hetd_save_history(hetd_ctx *ctx, void *host);
int hetd_load_history(hetd_ctx *ctx, const void *host, int64_t bytes);

/* Ensemble of independent runs (no reference counterpart: the reference
 * runs one droplet per process).  Advances each of the n contexts by nsteps
 * iterations exactly as hetd_evolve(ctx[i], deltat, nsteps, NULL) would --
 * bitwise the same states -- with the members' launches interleaved over
 * their own streams so that they share the GPU, and steady-state
 * predictor-corrector iterations replayed from per-member HIP graphs.
 * Contexts may differ in grid, precision and parameters. */
int hetd_pool_evolve(hetd_ctx *const *ctx, int32_t n, double deltat, int32_t nsteps);

/* Copy a field to / from host memory (HETD_F_*; complex fields are
 * 2*nx*ny*nz doubles, half spectra (nx/2+1)*ny*nz doubles). */
int hetd_get_field(hetd_ctx *ctx, int32_t which, double *host);
int hetd_get_imp(hetd_ctx *ctx, double r[3], double v[3], double a[3]);

/* FFT-level entry points with the reference's conventions (fft.f90:81-153):
 * unnormalised r2c (exp(-2 pi i k.x/n)) to the (nx/2+1,ny,nz) half spectrum,
 * and c2r (exp(+...)) followed by *renor = 1/(nx*ny*nz). Host buffers. */
int hetd_fft_r2c(hetd_ctx *ctx, const double *in, double *out);
int hetd_fft_c2r(hetd_ctx *ctx, const double *in, double *out);
/* derivnD(id,nn,h,iv,f,df,icon): id-th derivative along axis iv (1..3)
 * with the context's npd/icon (init_deriv_parallel.f90:373-648 real,
 * 933-1209 complex).  Host buffers. */
int hetd_deriv(hetd_ctx *ctx, int32_t id, int32_t iv, int32_t is_complex,
               const double *f, double *df);

/* Device stream of the context (a hipStream_t), for callers that time or
 * order their own work against the library's. */
void *hetd_stream(hetd_ctx *ctx);

/* Kernel-class timing with HIP events on the context stream.  No counterpart
 * in the reference (which only times its whole loop, DFT4He3d.f90:829-831,
 * 1006-1015); it serves benchmarks and roofline accounting.  While kclass is
 * nonzero every launch of that class is bracketed by a pair of events and
 * its algorithmic HBM bytes (one read of each input, one write of each
 * output) are accumulated; hetd_timing_read returns the summed event time,
 * the launch count and the bytes since the last hetd_timing call. */
#define HETD_K_OFF 0
#define HETD_K_UPDATE 1 /* fused stencil + right-hand side + integrator update */
#define HETD_K_XPASS 2  /* x-axis FFT pass with its fused pointwise work */
#define HETD_K_YPASS 3  /* y-axis FFT pass */
#define HETD_K_ZPASS 4  /* z-axis FFT pass with the fused spectral multiply */
#define HETD_K_FORCE 5  /* density gradient + impurity-force reduction */
int hetd_timing(hetd_ctx *ctx, int32_t kclass);
int hetd_timing_read(hetd_ctx *ctx, double *total_ms, int64_t *launches, double *bytes);

#ifdef __cplusplus
}
#endif
#endif
