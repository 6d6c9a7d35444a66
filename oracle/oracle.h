/*
 * oracle.h -- CPU restatement of the reference's real-time propagation step.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker.
 * The product (lib/libhetd.so) never links, loads or calls it.
 *
 * Every routine restates the reference routine named in its comment
 * (file:line, relative to the reference repository root).  Arrays are
 * Fortran column-major (x fastest), complex values interleaved (re, im).
 */
#ifndef HETD_ORACLE_H
#define HETD_ORACLE_H

#include "../include/hetd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

orc_ctx *orc_create(const hetd_params *p);
void orc_destroy(orc_ctx *c);
const char *orc_error(orc_ctx *c);

/* impurity table P[i] = V((i-1)*delinter) (potenimpini.f90:11-25) */
int orc_set_potion(orc_ctx *c, const double *potion, int npot, double delinter);
int orc_set_state(orc_ctx *c, const double *psi, const double r[3], const double v[3]);
int orc_init(orc_ctx *c); /* first potenimp/poten/forceimp/energy (DFT4He3d.f90:637-645) */

int orc_potenimp(orc_ctx *c, const double r[3]);
int orc_poten(orc_ctx *c);
int orc_forceimp(orc_ctx *c, const double r[3], double F[3]);
int orc_set_aimp(orc_ctx *c, const double a[3]); /* the driver's aimp = F/mAg_u */
int orc_steprk(orc_ctx *c, double dt);
int orc_steppc(orc_ctx *c, double dt, double *errHe, double *errimp);
int orc_energy(orc_ctx *c, hetd_energies *e);
int orc_norm(orc_ctx *c, double *norm);
/* DFT4He3d.f90:868-947: cm (r_cm.f90), <x2,y2,z2>, norm, <Lx,Ly,Lz> (out[0..9]) */
int orc_moments(orc_ctx *c, double npart, double out[10]);
/* one iteration of the driver loop body (DFT4He3d.f90:840-852) */
int orc_evolve(orc_ctx *c, double dt, int nsteps, double *traj /* nsteps*HETD_TRAJ_STRIDE or NULL */);

int orc_get_field(orc_ctx *c, int which, double *out);
int orc_get_imp(orc_ctx *c, double r[3], double v[3], double a[3]);

/* free-standing pieces used by unit tests */
int orc_fft_r2c(int nx, int ny, int nz, const double *in, double *out);
int orc_fft_c2r(int nx, int ny, int nz, const double *in, double *out); /* unnormalised */
int orc_deriv(int nx, int ny, int nz, int npd, int icon, int id, int iv, double h,
              int is_complex, const double *f, double *df);
double orc_si(double x);
double orc_vlj_hat(double p, double eps, double sigma, double hc, double b_unused);
double orc_vimp(int selec, double r, double r_cutoff, double umax);

#ifdef __cplusplus
}
#endif
#endif
