// stencil.hip -- finite-difference kernels for gfx950: the fused
// Laplacian + right-hand side + integrator update (steprk / steppc), the
// generic derivative, the density gradient fused with the impurity-force
// reduction, and the energy reduction.
//
// Stencils are the reference's centred npd-point rows scaled by 1/h**id
// (init_deriv_parallel.f90:373-648, 933-1209); icon = 13 treats points
// outside the box as zero (Redef, redef.f:133-170), icon = 8 wraps
// periodically and icon = 18 wraps with a sign flip.
#include "hetd_kernels.h"

namespace hetd {

constexpr int TBX = 64, TBY = 4;

template <typename V> struct vops;
template <> struct vops<double> {
    static __device__ __forceinline__ double zero() { return 0.0; }
    static __device__ __forceinline__ void fma_(double& s, double c, double v) { s = fma(c, v, s); }
};
template <> struct vops<double2> {
    static __device__ __forceinline__ double2 zero() { return make_double2(0.0, 0.0); }
    static __device__ __forceinline__ void fma_(double2& s, double c, double2 v) {
        s.x = fma(c, v.x, s.x);
        s.y = fma(c, v.y, s.y);
    }
};

// sum_j c[j] f(pos + j - m) along one axis (stride), with the boundary rule
template <typename V>
__device__ __forceinline__ V stencil_axis(const V* __restrict__ f, long base, int pos, int n, long stride,
                                          const double* c, int npd, int m, int icon) {
    V s = vops<V>::zero();
    if (n < npd) return s;  // the reference returns df = 0 (init_deriv_parallel.f90:391-395)
    if (pos - m >= 0 && pos + m < n) {
        const V* p = f + base + (long)(pos - m) * stride;
        for (int j = 0; j < npd; ++j) vops<V>::fma_(s, c[j], p[(long)j * stride]);
    } else {
        for (int j = 0; j < npd; ++j) {
            int q = pos + j - m;
            double sg = 1.0;
            if (q < 0 || q >= n) {
                if (icon == 13) continue;
                if (icon == 18) sg = -1.0;
                q = q < 0 ? q + n : q - n;
            }
            vops<V>::fma_(s, c[j] * sg, f[base + (long)q * stride]);
        }
    }
    return s;
}

__device__ __forceinline__ double sq_abs(double2 v) { return abs2(v); }

// ------------------------------------------------------- fused update

__global__ void __launch_bounds__(TBX* TBY) k_update(UpdateArgs a) {
    const int ix = blockIdx.x * TBX + threadIdx.x % TBX;
    const int iy = blockIdx.y * TBY + threadIdx.x / TBX;
    const int iz = blockIdx.z;
    const Grid& g = a.g;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    double pcabs = 0.0;
    if (ix < nx && iy < ny) {
        const long row = (long)nx * (iy + (long)ny * iz);
        const long o = row + ix;
        const StencilCoef& sc = a.sc;
        // Laplacian: d2x + d2y + d2z (steprk.f90:67-69, steppc.f90:45-47)
        double2 lx = stencil_axis<double2>(a.psi, row, ix, nx, 1, sc.c2[0], sc.npd, sc.m, sc.icon);
        double2 ly = stencil_axis<double2>(a.psi, o - (long)nx * iy, iy, ny, nx, sc.c2[1], sc.npd, sc.m, sc.icon);
        double2 lz = stencil_axis<double2>(a.psi, o - (long)nx * ny * iz, iz, nz, (long)nx * ny, sc.c2[2], sc.npd,
                                           sc.m, sc.icon);
        const double lr = (lx.x + ly.x) + lz.x, li = (lx.y + ly.y) + lz.y;
        const double2 ps = a.psi[o];
        const double pot = a.pot4[o];
        const double* R = a.imp->rpot;
        const double dz = g.z[iz] - R[2], dy = g.y[iy] - R[1];
        const double u = uimp_at(g.potion, g.npot, g.delinter, g.x[ix] - R[0], dy * dy + dz * dz);
        const double tr = g.lambdah * ((((((1.0 + g.tcx[ix]) + 1.0) + g.tcy[iy]) + 1.0) + g.tcz[iz]));
        const double ar = lr * a.h2o2m4 - pot * ps.x;
        const double ai = li * a.h2o2m4 - pot * ps.y;
        // f = timec*(lap*h2o2m4 - pot4*psi) - i*uimp*psi (steprk.f90:74)
        double2 f;
        f.x = (tr * ar - ai) + u * ps.y;
        f.y = (tr * ai + ar) - u * ps.x;
        double2 pn;
        if (a.kind == STAGE_RK) {
            // Gill stage (steprk.f90:75-85)
            const double2 q = a.q[o];
            double2 s1, qn;
            s1.x = a.ra * (f.x - a.rb * q.x);
            s1.y = a.ra * (f.y - a.rb * q.y);
            qn.x = (q.x + 3.0 * s1.x) - a.rc * f.x;
            qn.y = (q.y + 3.0 * s1.y) - a.rc * f.y;
            a.q[o] = qn;
            if (a.rk_j == 0) a.hnew[o] = f;
            pn.x = ps.x + a.dt * s1.x;
            pn.y = ps.y + a.dt * s1.y;
            const double d = sq_abs(pn);
            a.psi_out[o] = pn;
            a.den[o] = d;
            if (!(d > 0.0)) atomicOr(a.guard, 1);
        } else if (a.kind == STAGE_PRED) {
            // predictor + modifier (steppc.f90:55-65)
            const double c4 = (4.0 / 3.0) * a.dt;
            const double2 p3 = a.po3[o], h1 = a.h1[o], h2 = a.h2[o], pc = a.pc[o];
            double2 P;
            P.x = p3.x + c4 * ((2.0 * f.x - h1.x) + 2.0 * h2.x);
            P.y = p3.y + c4 * ((2.0 * f.y - h1.y) + 2.0 * h2.y);
            const double c112 = 112.0 / 121.0;
            pn.x = P.x - c112 * pc.x;
            pn.y = P.y - c112 * pc.y;
            a.psi_out[o] = pn;  // the buffer of psiold(ioldp(3))
            a.pc[o] = P;
            a.hnew[o] = f;      // hpsiold(ioldh(2))
            a.den[o] = sq_abs(pn);
        } else {
            // corrector (steppc.f90:123-134)
            const double2 p3 = a.po3[o], p2 = a.po2[o], h1 = a.h1[o], h2 = a.h2[o], pc = a.pc[o];
            const double t3 = 3.0 * a.dt;
            double2 Cc, pcn;
            Cc.x = 0.125 * ((9.0 * p3.x - p2.x) + t3 * ((f.x + 2.0 * h1.x) - h2.x));
            Cc.y = 0.125 * ((9.0 * p3.y - p2.y) + t3 * ((f.y + 2.0 * h1.y) - h2.y));
            pcn.x = pc.x - Cc.x;
            pcn.y = pc.y - Cc.y;
            const double c9 = 9.0 / 121.0;
            pn.x = Cc.x + c9 * pcn.x;
            pn.y = Cc.y + c9 * pcn.y;
            a.pc[o] = pcn;
            a.psi_out[o] = pn;
            a.den[o] = sq_abs(pn);
            pcabs = cabs_(pcn);
        }
    }
    if (a.kind == STAGE_CORR) {
        __shared__ double red[TBX * TBY / 64];
        double v[1] = {pcabs};
        block_sum<1>(v, red);
        if (threadIdx.x == 0) {
            const long b = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
            a.partial[4 * b] = v[0];  // same layout as the tiled kernel
        }
    }
}

hipError_t launch_update(const UpdateArgs& a, hipStream_t st, int* nblocks_out) {
    dim3 grid((a.g.nx + TBX - 1) / TBX, (a.g.ny + TBY - 1) / TBY, a.g.nz);
    if (nblocks_out) *nblocks_out = (int)(grid.x * grid.y * grid.z);
    hipLaunchKernelGGL(k_update, grid, dim3(TBX * TBY), 0, st, a);
    return hipGetLastError();
}

// ---------------------------------------------------------- derivative

__global__ void __launch_bounds__(TBX* TBY) k_deriv(DerivArgs a) {
    const int ix = blockIdx.x * TBX + threadIdx.x % TBX;
    const int iy = blockIdx.y * TBY + threadIdx.x / TBX;
    const int iz = blockIdx.z;
    if (ix >= a.nx || iy >= a.ny) return;
    const long row = (long)a.nx * (iy + (long)a.ny * iz), o = row + ix;
    const int ax = a.iv - 1;
    const double* c = a.id == 1 ? a.sc.c1[ax] : a.sc.c2[ax];
    const int n = ax == 0 ? a.nx : (ax == 1 ? a.ny : a.nz);
    const int pos = ax == 0 ? ix : (ax == 1 ? iy : iz);
    const long stride = ax == 0 ? 1 : (ax == 1 ? (long)a.nx : (long)a.nx * a.ny);
    const long base = o - (long)pos * stride;
    if (a.cplx) {
        const double2* f = reinterpret_cast<const double2*>(a.f);
        reinterpret_cast<double2*>(a.df)[o] = stencil_axis<double2>(f, base, pos, n, stride, c, a.sc.npd, a.sc.m, a.sc.icon);
    } else {
        a.df[o] = stencil_axis<double>(a.f, base, pos, n, stride, c, a.sc.npd, a.sc.m, a.sc.icon);
    }
}

hipError_t launch_deriv(const DerivArgs& a, hipStream_t st) {
    dim3 grid((a.nx + TBX - 1) / TBX, (a.ny + TBY - 1) / TBY, a.nz);
    hipLaunchKernelGGL(k_deriv, grid, dim3(TBX * TBY), 0, st, a);
    return hipGetLastError();
}

// --------------------------------------- gradient + impurity force

// partial[b] = sum over the block of d(den)/dx_d * u_imp (potenimp.f90:44-46,
// with derden.f90:26-28 evaluated on the fly)
__global__ void __launch_bounds__(TBX* TBY) k_force(ForceArgs a) {
    const int ix = blockIdx.x * TBX + threadIdx.x % TBX;
    const int iy = blockIdx.y * TBY + threadIdx.x / TBX;
    const int iz = blockIdx.z;
    const Grid& g = a.g;
    double v[3] = {0.0, 0.0, 0.0};
    if (ix < g.nx && iy < g.ny) {
        const long row = (long)g.nx * (iy + (long)g.ny * iz), o = row + ix;
        const StencilCoef& sc = a.sc;
        const double gx = stencil_axis<double>(a.den, row, ix, g.nx, 1, sc.c1[0], sc.npd, sc.m, sc.icon);
        const double gy = stencil_axis<double>(a.den, o - (long)g.nx * iy, iy, g.ny, g.nx, sc.c1[1], sc.npd, sc.m, sc.icon);
        const double gz = stencil_axis<double>(a.den, o - (long)g.nx * g.ny * iz, iz, g.nz, (long)g.nx * g.ny, sc.c1[2],
                                               sc.npd, sc.m, sc.icon);
        const double* R = a.imp->rpot;
        const double dz = g.z[iz] - R[2], dy = g.y[iy] - R[1];
        const double u = uimp_at(g.potion, g.npot, g.delinter, g.x[ix] - R[0], dy * dy + dz * dz);
        v[0] = gx * u;
        v[1] = gy * u;
        v[2] = gz * u;
    }
    __shared__ double red[3 * TBX * TBY / 64];
    block_sum<3>(v, red);
    if (threadIdx.x == 0) {
        const long b = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
        a.partial[3 * b + 0] = v[0];
        a.partial[3 * b + 1] = v[1];
        a.partial[3 * b + 2] = v[2];
    }
}

hipError_t launch_force_partial(const ForceArgs& a, hipStream_t st, int* nblocks) {
    dim3 grid((a.g.nx + TBX - 1) / TBX, (a.g.ny + TBY - 1) / TBY, a.g.nz);
    *nblocks = (int)(grid.x * grid.y * grid.z);
    hipLaunchKernelGGL(k_force, grid, dim3(TBX * TBY), 0, st, a);
    return hipGetLastError();
}

// ---------------------------------------------------------- energy

// energy.f90:51-110 and the norm of DFT4He3d.f90:869, one pass
__global__ void __launch_bounds__(TBX* TBY) k_energy(EnergyArgs a) {
    const int ix = blockIdx.x * TBX + threadIdx.x % TBX;
    const int iy = blockIdx.y * TBY + threadIdx.x / TBX;
    const int iz = blockIdx.z;
    const Grid& g = a.g;
    double v[EN_NCOMP];
#pragma unroll
    for (int c = 0; c < EN_NCOMP; ++c) v[c] = 0.0;
    if (ix < g.nx && iy < g.ny) {
        const long row = (long)g.nx * (iy + (long)g.ny * iz), o = row + ix;
        const StencilCoef& sc = a.sc;
        const double2 g1 = stencil_axis<double2>(a.psi, row, ix, g.nx, 1, sc.c1[0], sc.npd, sc.m, sc.icon);
        const double2 g2 = stencil_axis<double2>(a.psi, o - (long)g.nx * iy, iy, g.ny, g.nx, sc.c1[1], sc.npd, sc.m, sc.icon);
        const double2 g3 = stencil_axis<double2>(a.psi, o - (long)g.nx * g.ny * iz, iz, g.nz, (long)g.nx * g.ny,
                                                 sc.c1[2], sc.npd, sc.m, sc.icon);
        if (a.mom) {
            moment_terms(a, v, a.psi[o], g1, g2, g3, g.x[ix], g.y[iy], g.z[iz]);
        } else {
            v[EN_KIN] = (sq_abs(g1) + sq_abs(g2)) + sq_abs(g3);
            const double d = a.den[o], cg = a.dencg[o];
            v[EN_LJ] = a.delj4[o] * d;
            v[EN_COR] = d * cg * cg * (a.aux1 + a.aux2 * cg);
            if (a.lsolid) v[EN_SOLID] = d * (1.0 + tanh(a.beta * (d - a.den_m)));
            const double* R = a.imp->rpot;
            const double dz = g.z[iz] - R[2], dy = g.y[iy] - R[1];
            v[EN_HEX] = uimp_at(g.potion, g.npot, g.delinter, g.x[ix] - R[0], dy * dy + dz * dz) * d;
            if (a.otc)
                v[EN_ALPHA] = a.falfs[o] * ((a.dxden[o] * a.intx[o] + a.dyden[o] * a.inty[o]) + a.dzden[o] * a.intz[o]);
            v[EN_NORM] = d;
        }
    }
    __shared__ double red[EN_NCOMP * TBX * TBY / 64];
    block_sum<EN_NCOMP>(v, red);
    if (threadIdx.x == 0) {
        const long b = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
#pragma unroll
        for (int c = 0; c < EN_NCOMP; ++c) a.partial[EN_NCOMP * b + c] = v[c];
    }
}

hipError_t launch_energy_partial(const EnergyArgs& a, hipStream_t st, int* nblocks) {
    dim3 grid((a.g.nx + TBX - 1) / TBX, (a.g.ny + TBY - 1) / TBY, a.g.nz);
    *nblocks = (int)(grid.x * grid.y * grid.z);
    hipLaunchKernelGGL(k_energy, grid, dim3(TBX * TBY), 0, st, a);
    return hipGetLastError();
}

}  // namespace hetd
