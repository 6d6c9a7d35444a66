// imp.hip -- the classical impurity's equations of motion kept on the device
// (one-thread kernels, so a whole step needs no host round trip), the
// fixed-order final reductions, and small pointwise kernels.
#include "hetd_kernels.h"

namespace hetd {

// --------------------------------------------------------- final reductions

// Sums partial[b*stride + off + c] over b in a fixed order: thread t takes
// b = t, t+256, ... then a fixed shuffle tree.  Bitwise reproducible.
template <int NC>
__device__ void reduce_fixed(const double* __restrict__ partial, int nblocks, double (&out)[NC], int stride = NC,
                             int off = 0) {
    __shared__ double red[NC * 4];
    double v[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) v[c] = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
#pragma unroll
        for (int c = 0; c < NC; ++c) v[c] += partial[(long)b * stride + off + c];
    block_sum<NC>(v, red);
#pragma unroll
    for (int c = 0; c < NC; ++c) out[c] = v[c];
}

template <int NC>
__global__ void __launch_bounds__(256) k_reduce_final(const double* partial, int nblocks, double* out) {
    double v[NC];
    reduce_fixed<NC>(partial, nblocks, v);
    if (threadIdx.x == 0)
#pragma unroll
        for (int c = 0; c < NC; ++c) out[c] = v[c];
}

hipError_t launch_reduce_final(const double* partial, int nblocks, int ncomp, double* out, hipStream_t st) {
    switch (ncomp) {
    case 1: hipLaunchKernelGGL(k_reduce_final<1>, dim3(1), dim3(256), 0, st, partial, nblocks, out); break;
    case 3: hipLaunchKernelGGL(k_reduce_final<3>, dim3(1), dim3(256), 0, st, partial, nblocks, out); break;
    case EN_NCOMP: hipLaunchKernelGGL(k_reduce_final<EN_NCOMP>, dim3(1), dim3(256), 0, st, partial, nblocks, out); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// sum(uext * den) of energy.f90:124: each block sums a fixed strided slice
// and a fixed tree, so the result is the same run to run
__global__ void __launch_bounds__(256) k_dot_uext(const double2* psi, const double* u, long n, double* partial) {
    __shared__ double sh[256];
    double s = 0.0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) s += u[i] * abs2(psi[i]);
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

hipError_t launch_dot_uext(const double2* psi, const double* u, long n, int nblocks, double* partial, hipStream_t st) {
    if (nblocks < 1 || nblocks > 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_dot_uext, dim3(nblocks), dim3(256), 0, st, psi, u, n, partial);
    return hipGetLastError();
}

// F = -sum(grad den * u_imp) * dxyz (potenimp.f90:44-46); aimp = F/mAg_u
// (DFT4He3d.f90:852)
__global__ void __launch_bounds__(256) k_force_final(ImpState* imp, const double* partial, int nblocks, int stride,
                                                     int off, double dxyz, double mass, double* rec_a) {
    double v[3];
    reduce_fixed<3>(partial, nblocks, v, stride, off);
    if (threadIdx.x == 0) {
        for (int d = 0; d < 3; ++d) {
            const double F = -v[d] * dxyz;
            imp->F[d] = F;
            imp->a[d] = F / mass;
            if (rec_a) rec_a[d] = imp->a[d];
        }
    }
}

hipError_t launch_force_final(ImpState* imp, const double* partial, int nblocks, int stride, int off, double dxyz,
                              double mass, double* rec_a, hipStream_t st) {
    hipLaunchKernelGGL(k_force_final, dim3(1), dim3(256), 0, st, imp, partial, nblocks, stride, off, dxyz, mass,
                       rec_a);
    return hipGetLastError();
}

// errHe = c9*sum(abs(pc)) / nxyz (steppc.f90:131,134)
__global__ void __launch_bounds__(256) k_errhe_final(ImpState* imp, const double* partial, int nblocks, int stride,
                                                     double c9, double n) {
    double v[1];
    reduce_fixed<1>(partial, nblocks, v, stride, 0);
    if (threadIdx.x == 0) imp->errHe = (c9 * v[0]) / n;
}

hipError_t launch_errhe_final(ImpState* imp, const double* partial, int nblocks, int stride, double n,
                              hipStream_t st) {
    hipLaunchKernelGGL(k_errhe_final, dim3(1), dim3(256), 0, st, imp, partial, nblocks, stride, 9.0 / 121.0, n);
    return hipGetLastError();
}

// -------------------------------------------------------------- impurity ODE

// Gill stage for R (velocity v) and v (acceleration a) (steprk.f90:101-125)
__global__ void k_imp_rk(ImpRkArgs a) {
    ImpState* s = a.imp;
    double aux[3];
    for (int d = 0; d < 3; ++d) {
        aux[d] = a.a * (s->v[d] - a.b * s->qr[d]);
        s->qr[d] = (s->qr[d] + 3.0 * aux[d]) - a.c * s->v[d];
    }
    if (a.j == 0)
        for (int d = 0; d < 3; ++d) {
            s->rold[2][d] = s->rold[1][d];
            s->rold[1][d] = s->rold[0][d];
            s->rold[0][d] = s->r[d];
        }
    for (int d = 0; d < 3; ++d) s->r[d] = s->r[d] + a.dt * aux[d];
    for (int d = 0; d < 3; ++d) {
        aux[d] = a.a * (s->a[d] - a.b * s->qv[d]);
        s->qv[d] = (s->qv[d] + 3.0 * aux[d]) - a.c * s->a[d];
    }
    if (a.j == 0)
        for (int d = 0; d < 3; ++d) {
            s->aold[1][d] = s->aold[0][d];
            s->aold[0][d] = s->a[d];
            s->vold[2][d] = s->vold[1][d];
            s->vold[1][d] = s->vold[0][d];
            s->vold[0][d] = s->v[d];
        }
    for (int d = 0; d < 3; ++d) s->v[d] = s->v[d] + a.dt * aux[d];
    if (a.set_rpot)
        for (int d = 0; d < 3; ++d) s->rpot[d] = s->r[d];
}

// predictor + modifier for (R, v) (steppc.f90:77-95); the middle potential
// is evaluated at the modified R (steppc.f90:100-107).  Indices 0-based,
// taken before the iolda swap; a1/a2 here are iolda(1)-1, iolda(2)-1.
__global__ void k_imp_pred(ImpPcArgs a) {
    ImpState* s = a.imp;
    const double c112 = 112.0 / 121.0, c4 = (4.0 / 3.0) * a.dt;
    double aux[3];
    for (int d = 0; d < 3; ++d) aux[d] = s->rold[a.r3][d] + c4 * ((2.0 * s->v[d] - s->vold[a.v1][d]) + 2.0 * s->vold[a.v2][d]);
    for (int d = 0; d < 3; ++d) {
        s->rold[a.r3][d] = aux[d] - c112 * s->pcr[d];
        s->pcr[d] = aux[d];
    }
    for (int d = 0; d < 3; ++d) aux[d] = s->vold[a.v3][d] + c4 * ((2.0 * s->a[d] - s->aold[a.a1][d]) + 2.0 * s->aold[a.a2][d]);
    for (int d = 0; d < 3; ++d) {
        s->vold[a.v3][d] = aux[d] - c112 * s->pcv[d];
        s->pcv[d] = aux[d];
    }
    for (int d = 0; d < 3; ++d) s->aold[a.a2][d] = s->a[d];
    for (int d = 0; d < 3; ++d) s->rpot[d] = s->rold[a.r3][d];
}

// corrector for (R, v) (steppc.f90:148-174); a1/a2 are the swapped
// iolda(1)-1, iolda(2)-1.  set_rpot: rpot = R, the driver's potenimp(rimp).
__global__ void k_imp_corr(ImpPcArgs a) {
    ImpState* s = a.imp;
    const double c9 = 9.0 / 121.0, t3 = 3.0 * a.dt;
    double aux[3];
    for (int d = 0; d < 3; ++d)
        aux[d] = 0.125 * ((9.0 * s->r[d] - s->rold[a.r2][d]) + t3 * ((s->vold[a.v3][d] + 2.0 * s->v[d]) - s->vold[a.v1][d]));
    double e = 0.0;
    for (int d = 0; d < 3; ++d) {
        s->pcr[d] = s->pcr[d] - aux[d];
        s->rold[a.r3][d] = s->r[d];
        s->r[d] = aux[d] + c9 * s->pcr[d];
    }
    for (int d = 0; d < 3; ++d) e += fabs(c9 * s->pcr[d]);
    s->errimp = e * 0.3333333333;
    for (int d = 0; d < 3; ++d)
        aux[d] = 0.125 * ((9.0 * s->v[d] - s->vold[a.v2][d]) + t3 * ((s->a[d] + 2.0 * s->aold[a.a1][d]) - s->aold[a.a2][d]));
    for (int d = 0; d < 3; ++d) {
        s->pcv[d] = s->pcv[d] - aux[d];
        s->vold[a.v3][d] = s->v[d];
        s->v[d] = aux[d] + c9 * s->pcv[d];
    }
    if (a.set_rpot)
        for (int d = 0; d < 3; ++d) s->rpot[d] = s->r[d];
}

__global__ void k_traj_record(const ImpState* s, double* rec) {
    rec[0] = s->errHe;
    rec[1] = s->errimp;
    for (int d = 0; d < 3; ++d) {
        rec[2 + d] = s->r[d];
        rec[5 + d] = s->v[d];
        rec[8 + d] = s->a[d];
    }
}

hipError_t launch_imp_rk(const ImpRkArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_imp_rk, dim3(1), dim3(1), 0, st, a);
    return hipGetLastError();
}
hipError_t launch_imp_pred(const ImpPcArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_imp_pred, dim3(1), dim3(1), 0, st, a);
    return hipGetLastError();
}
hipError_t launch_imp_corr(const ImpPcArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_imp_corr, dim3(1), dim3(1), 0, st, a);
    return hipGetLastError();
}
hipError_t launch_traj_record(const ImpState* imp, double* rec, hipStream_t st) {
    hipLaunchKernelGGL(k_traj_record, dim3(1), dim3(1), 0, st, imp, rec);
    return hipGetLastError();
}

// ----------------------------------------------------------------- pointwise

static inline dim3 grid1(long n) { return dim3((unsigned)((n + 255) / 256)); }

__global__ void k_den_from_psi(const double2* psi, double* den, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        den[i] = abs2(psi[i]);
    }
}
hipError_t launch_den_from_psi(const double2* psi, double* den, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_den_from_psi, grid1(n), dim3(256), 0, st, psi, den, n);
    return hipGetLastError();
}

__global__ void k_den32_from_psi(const double2* psi, float* den, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        den[i] = (float)abs2(psi[i]);
    }
}
hipError_t launch_den32_from_psi(const double2* psi, float* den, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_den32_from_psi, grid1(n), dim3(256), 0, st, psi, den, n);
    return hipGetLastError();
}

__global__ void k_to_double(const float* a, double* out, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (double)a[i];
}
hipError_t launch_to_double(const float* a, double* out, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_to_double, grid1(n), dim3(256), 0, st, a, out, n);
    return hipGetLastError();
}

__global__ void k_to_float(const double* a, float* out, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)a[i];
}
hipError_t launch_to_float(const double* a, float* out, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_to_float, grid1(n), dim3(256), 0, st, a, out, n);
    return hipGetLastError();
}

__global__ void k_uimp_field(Grid g, const ImpState* imp, double* out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.n) return;
    const int ix = (int)(i % g.nx), iy = (int)((i / g.nx) % g.ny), iz = (int)(i / ((long)g.nx * g.ny));
    const double* R = imp->rpot;
    const double dz = g.z[iz] - R[2], dy = g.y[iy] - R[1];
    out[i] = uimp_at(g.potion, g.npot, g.delinter, g.x[ix] - R[0], dy * dy + dz * dz);
}
hipError_t launch_uimp_field(const Grid& g, const ImpState* imp, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_uimp_field, grid1(g.n), dim3(256), 0, st, g, imp, out);
    return hipGetLastError();
}

// the solid penalty field of poten.f90:80-94, with the device function the
// mean field adds (a field copy for callers; the step never stores it)
__global__ void k_penalty_field(const double* den, double csol, double beta, double den_m, double* out, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = solid_penalty<double>(den[i], csol, beta, den_m);
}
hipError_t launch_penalty_field(const double* den, double csol, double beta, double den_m, double* out, long n,
                                hipStream_t st) {
    hipLaunchKernelGGL(k_penalty_field, grid1(n), dim3(256), 0, st, den, csol, beta, den_m, out, n);
    return hipGetLastError();
}

__global__ void k_falfs(const double* denalf, double den0s, double* falfs, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) falfs[i] = 1.0 - denalf[i] / den0s;
}
hipError_t launch_falfs(const double* denalf, double den0s, double* falfs, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_falfs, grid1(n), dim3(256), 0, st, denalf, den0s, falfs, n);
    return hipGetLastError();
}

__global__ void k_mul(const double* a, const double* b, double* out, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a[i] * b[i];
}
hipError_t launch_mul(const double* a, const double* b, double* out, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_mul, grid1(n), dim3(256), 0, st, a, b, out, n);
    return hipGetLastError();
}

__global__ void k_dot3(const double* a1, const double* b1, const double* a2, const double* b2, const double* a3,
                       const double* b3, double* out, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = ((a1[i] * b1[i]) + (a2[i] * b2[i])) + (a3[i] * b3[i]);
}
hipError_t launch_dot3(const double* a1, const double* b1, const double* a2, const double* b2, const double* a3,
                       const double* b3, double* out, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_dot3, grid1(n), dim3(256), 0, st, a1, b1, a2, b2, a3, b3, out, n);
    return hipGetLastError();
}

// ualphas = (ualphas/den0s + s4 + s5 + s6)*h2o2m4*alphas; pot4 += ualphas
// (term_alfa.f90:98-113, poten.f90:103-106)
// pot4 (FP64) or pot4f (the FP32 mean field of the mixed path) += u_alpha
__global__ void k_ualpha_final(double* ua, const double* s4, const double* s5, const double* s6, double den0s,
                               double h2, double al, double* pot4, float* pot4f, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const double u = (((((ua[i] / den0s) + s4[i]) + s5[i]) + s6[i]) * h2) * al;
        ua[i] = u;
        if (pot4f) pot4f[i] = (float)((double)pot4f[i] + u);
        else pot4[i] = pot4[i] + u;
    }
}
hipError_t launch_ualpha_final(double* ua, const double* s4, const double* s5, const double* s6, double den0s,
                               double h2o2m4, double alphas, double* pot4, float* pot4f, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_ualpha_final, grid1(n), dim3(256), 0, st, ua, s4, s5, s6, den0s, h2o2m4, alphas, pot4,
                       pot4f, n);
    return hipGetLastError();
}

__global__ void k_scale(double* a, double s, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] *= s;
}
hipError_t launch_scale(double* a, double s, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_scale, grid1(n), dim3(256), 0, st, a, s, n);
    return hipGetLastError();
}

}  // namespace hetd
