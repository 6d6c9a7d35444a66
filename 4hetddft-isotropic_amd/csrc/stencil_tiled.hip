// stencil_tiled.hip -- LDS-tiled 2.5-D stencil engine for gfx950.
//
// A workgroup owns a 64 (x) x TTY (y) column of the grid and marches it
// through a chunk of z planes.  Per plane the tile plus an M-point halo in x
// and y sits in LDS (x and y stencils read LDS); each thread keeps the z
// column of its TPY points in a register queue of 2M+1 values (the z stencil
// reads registers), so every grid value is fetched from HBM about once.  Tiles are
// ordered so that consecutive tiles, which share halos, run on the same XCD
// (blocks are dealt to the 8 XCDs round-robin).
//
// Three consumers of the stencils share the engine:
//   UPDATE: Laplacian of psi (npd-point second derivative, steprk.f90:67-69,
//           steppc.f90:45-47) + right-hand side + RK-Gill stage / predictor /
//           corrector update;
//   FORCE:  gradient of den (derden.f90:26-28) times u_imp, summed
//           (potenimp.f90:44-46);
//   ENERGY: gradient of psi and the energy densities (energy.f90:51-110).
#include <type_traits>

#include "hetd_kernels.h"

namespace hetd {

constexpr int TTX = 64;

template <typename V> struct vz;
template <> struct vz<double> {
    static __device__ __forceinline__ double zero() { return 0.0; }
    static __device__ __forceinline__ double add(double a, double b) { return a + b; }
    static __device__ __forceinline__ double sub(double a, double b) { return a - b; }
    // v, -v or +0 for mode 1, -1, 0, by bit operations on v (so that a load
    // feeding it is never moved under a branch)
    static __device__ __forceinline__ double mode(double v, int m) {
        unsigned long long b = (unsigned long long)__double_as_longlong(v);
        b &= 0ull - (unsigned long long)(m != 0);
        b ^= (unsigned long long)(m < 0) << 63;
        return __longlong_as_double((long long)b);
    }
    static __device__ __forceinline__ double scale(double v, double s) { return v * s; }
    static __device__ __forceinline__ void fma_(double& s, double c, double v) { s = fma(c, v, s); }
};
template <> struct vz<double2> {
    static __device__ __forceinline__ double2 zero() { return make_double2(0.0, 0.0); }
    static __device__ __forceinline__ double2 add(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
    static __device__ __forceinline__ double2 sub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
    static __device__ __forceinline__ double2 mode(double2 v, int m) {
        return make_double2(vz<double>::mode(v.x, m), vz<double>::mode(v.y, m));
    }
    static __device__ __forceinline__ double2 scale(double2 v, double s) { return make_double2(v.x * s, v.y * s); }
    static __device__ __forceinline__ void fma_(double2& s, double c, double2 v) {
        s.x = fma(c, v.x, s.x);
        s.y = fma(c, v.y, s.y);
    }
};

enum TKind { TK_UPDATE = 0, TK_FORCE = 1, TK_ENERGY = 2 };

template <int KIND> struct TArgs;
// UPDATE partials: [sum |pc| (corrector), force x, y, z (when fused)]
template <> struct TArgs<TK_UPDATE> { using V = double2; using A = UpdateArgs; static constexpr int NC = 4; };
template <> struct TArgs<TK_FORCE> { using V = double; using A = ForceArgs; static constexpr int NC = 3; };
template <> struct TArgs<TK_ENERGY> { using V = double2; using A = EnergyArgs; static constexpr int NC = EN_NCOMP; };

// Stencils in folded form: the centred weights are symmetric (second
// derivative) or antisymmetric (first), so
//   d2 f = s2 * (w2h[0] f_0 + sum_k w2h[k] (f_k + f_-k)),
//   d1 f = s1 * sum_k w1h[k] (f_k - f_-k),
// with one set of weights for all axes and s1 = 1/h, s2 = 1/h^2 per axis.
// (The reference applies weights/h^id term by term; the two differ only in
// rounding.)  Few distinct constants keep the kernel's scalar registers free.
struct TiledCommon {
    Grid g;
    int icon;
    int zc;          // z planes per workgroup
    int ntx, nty, ntz;
    double w1h[kMaxNpd / 2 + 1], w2h[kMaxNpd / 2 + 1];
    double s1[3], s2[3];
};

__device__ __forceinline__ double dens(double2 v) { return abs2(v); }

// RHS history hpsiold: complex128, or complex64 in mixed precision
__device__ __forceinline__ double2 widen2(float2 v) { return make_double2((double)v.x, (double)v.y); }
__device__ __forceinline__ double2 widen2(double2 v) { return v; }
template <bool MIX, bool NT = false> __device__ __forceinline__ void put_hist(double2* h, long o, double2 v) {
    if constexpr (MIX) {
        const float2 f = make_float2((float)v.x, (float)v.y);
        if constexpr (NT) st_nt(reinterpret_cast<float2*>(h) + o, f);
        else reinterpret_cast<float2*>(h)[o] = f;
    } else {
        if constexpr (NT) st_nt(h + o, v);
        else h[o] = v;
    }
}
// a store of a once-used stream (NT: non-temporal)
template <bool NT> __device__ __forceinline__ void put2(double2* p, long o, double2 v) {
    if constexpr (NT) st_nt(p + o, v);
    else p[o] = v;
}
template <bool NT, typename T> __device__ __forceinline__ T get2(const T* p, long o) {
    if constexpr (NT) return ld_nt(p + o);
    else return p[o];
}
template <bool MIX> __device__ __forceinline__ void put_den(const UpdateArgs& a, long o, double d) {
    if constexpr (MIX) a.denf[o] = (float)d;
    else a.den[o] = d;
}

// TPY points per thread along y, TNT threads: tile 64 x TTY with
// TTY = TPY*TNT/64; PF: software-pipelined halo / queue loads.
// FRC (UPDATE only): the density |psi|^2 of the input state is kept beside
// psi (LDS plane and z queue) and grad(den) * u_imp summed, which is
// forceimp (potenimp.f90:24-55) of the state the update starts from.
// MIX (UPDATE only): FP32 mean field in, FP32 den and complex64 RHS history.
// SK (UPDATE only): the stage, STAGE_RK / STAGE_PRED / STAGE_CORR.
// OPT (UPDATE variants): OPT_PP ping-pong LDS planes, OPT_NT non-temporal
// loads and stores of the once-used update operands.
enum { OPT_PP = 1, OPT_NT = 2 };
template <int KIND, int M, int TPY, int TNT, bool PF, bool FRC, bool MIX, int SK, int OPT>
__global__ void __launch_bounds__(TNT) k_tiled(TiledCommon tc, typename TArgs<KIND>::A args) {
    constexpr bool PP = (OPT & OPT_PP) != 0, NT = (OPT & OPT_NT) != 0;
    using V = typename TArgs<KIND>::V;
    constexpr int NC = TArgs<KIND>::NC;
    constexpr int RS = TNT / TTX;        // thread rows
    constexpr int TTY = TPY * RS;
    constexpr int LX = TTX + 2 * M, LY = TTY + 2 * M, W = 2 * M + 1;
    // PP: two LDS planes used in turn, so a plane is written while the
    // previous one may still be read, and one barrier per plane suffices
    __shared__ V lds_pl[PP ? 2 : 1][LX * LY];
    __shared__ double ldsd_pl[PP ? 2 : 1][FRC ? LX * LY : 1];
    __shared__ double red[NC * TNT / 64];

    const Grid& g = tc.g;
    const int nx = g.nx, ny = g.ny, nz = g.nz, icon = tc.icon;
    // XCD-aware tile order: the grid is padded to 8*per blocks and the blocks
    // dealt to one XCD (b % 8 equal) take a contiguous range of per tiles
    const int nb = tc.ntx * tc.nty * tc.ntz;
    const int b = blockIdx.x;
    const int per = (nb + 7) / 8;
    const int t = (b % 8) * per + b / 8;
    if (t >= nb) return;  // padding block: the whole workgroup leaves
    const int tile_x = t % tc.ntx, tile_y = (t / tc.ntx) % tc.nty, tile_z = t / (tc.ntx * tc.nty);
    const int ix0 = tile_x * TTX, iy0 = tile_y * TTY, iz0 = tile_z * tc.zc;
    const int izend = min(iz0 + tc.zc, nz);
    const int tx = threadIdx.x % TTX, ty = threadIdx.x / TTX;
    const int ix = ix0 + tx;
    const V* __restrict__ f;
    if constexpr (KIND == TK_FORCE) f = args.den;
    else f = args.psi;
    const bool zok = nz >= W;  // one periodic wrap reaches every z neighbour
    double acc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = 0.0;

    // Every boundary decision is taken once here, not per plane: a cell of
    // the LDS plane is (offset within a z plane, mode) with mode 0 = zero,
    // 1 = value, -1 = negated value (icon 18 across one face); the z queue of
    // a point needs a rule only for planes beyond the box (uniform per block).
    const long pl = (long)nx * ny;
    const bool ragged = ix0 + TTX > nx || iy0 + TTY > ny;
    int iy[TPY];
    bool valid[TPY];
    long c0[TPY];     // in-plane offset of the point (valid) or of its cell
    int cm[TPY];      // mode of that cell when the point is not valid
    V q[TPY][W];
    double dq[FRC ? TPY : 1][W];  // |psi|^2 of the z queue
    auto cell_rule = [&](int gx, int gy, long& off, int& mode) {
        // column (gx, gy) of an LDS cell: inside the box, or across one face
        // within one halo width on an axis at least as long as the stencil
        // (zero outside for icon 13, wrapped for 8 and 18, negated for 18);
        // anything else is zero, as no valid point reads it (corners) or the
        // reference's derivative along that axis is zero
        const bool ox = gx < 0 || gx >= nx, oy = gy < 0 || gy >= ny;
        mode = 1;
        if ((ox && oy) || (ox && (nx < W || gx < -M || gx >= nx + M)) ||
            (oy && (ny < W || gy < -M || gy >= ny + M)) || ((ox || oy) && icon == 13)) {
            mode = 0;
            off = 0;
            return;
        }
        if (ox) gx = gx < 0 ? gx + nx : gx - nx;
        if (oy) gy = gy < 0 ? gy + ny : gy - ny;
        if ((ox || oy) && icon == 18) mode = -1;
        off = (long)gx + (long)nx * gy;
    };
    // Loads are unconditional (a zero-mode cell reads a valid address and
    // discards it): a load under a branch makes the compiler wait for it at
    // the end of the branch, which would serialise the prefetches.
    auto fetch = [&](long off, int mode) -> V { return vz<V>::mode(f[off], mode); };
    // the point's own column at plane zz (any zz): zero, or one z wrap.
    // zsrc gives the address and sets the mode; the mode is applied where the
    // value is used, so that a prefetch is not waited for where it is issued.
    auto zsrc = [&](int p, int zz, int& m) -> long {
        int zw = zz, zm = 1;
        if (zz < 0 || zz >= nz) {  // uniform over the workgroup
            if (icon == 13 || !zok) {
                zm = 0;
                zw = 0;
            } else {
                zw = zz < 0 ? zz + nz : zz - nz;
                if (icon == 18) zm = -1;
            }
        }
        m = valid[p] ? zm : 0;
        return c0[p] + pl * zw;
    };
    auto zcol = [&](int p, int zz) -> V {
        int m;
        const long o = zsrc(p, zz, m);
        return fetch(o, m);
    };
#pragma unroll
    for (int p = 0; p < TPY; ++p) {
        iy[p] = iy0 + ty + RS * p;
        valid[p] = ix < nx && iy[p] < ny;
        cm[p] = 1;
        if (valid[p]) c0[p] = (long)ix + (long)nx * iy[p];
        else cell_rule(ix, iy[p], c0[p], cm[p]);
#pragma unroll
        for (int k = 0; k < W - 1; ++k) {
            q[p][k] = zcol(p, iz0 - M + k);
            if constexpr (FRC) dq[p][k] = dens(q[p][k]);
        }
    }

    // halo cells of this thread (slot k: cell threadIdx.x + k*TNT of the
    // ring); y halo = M rows above and below the tile columns first, then the
    // x halo = M columns on each side of the tile rows (corners are never read)
    constexpr int NHY = 2 * M * TTX, NH = NHY + 2 * M * TTY, KH = (NH + TNT - 1) / TNT;
    int hl[KH];  // LDS index of slot k, -1 if unused
    long hoff[KH];
    int hm[KH];
#pragma unroll
    for (int k = 0; k < KH; ++k) {
        const int h = threadIdx.x + k * TNT;
        hl[k] = -1;
        hoff[k] = 0;
        hm[k] = 0;
        if (h < NH) {
            int lx, ly;
            if (h < NHY) {
                const int r = h / TTX, cc = h % TTX;
                lx = cc + M;
                ly = r < M ? r : TTY + r;
            } else {
                const int hh = h - NHY;
                const int r = hh / (2 * M), cc = hh % (2 * M);
                ly = r + M;
                lx = cc < M ? cc : TTX + cc;
            }
            hl[k] = ly * LX + lx;
            cell_rule(ix0 + lx - M, iy0 + ly - M, hoff[k], hm[k]);
        }
    }
    // software pipeline: the halo of plane iz and the queue fronts are
    // loaded one iteration ahead, while plane iz-1 is being computed
    // UPDATE operands of a point (pot4, and q / psiold / hpsiold / pc by stage),
    // loaded one plane ahead with the pipelined variants
    using HT = typename std::conditional<MIX, float2, double2>::type;  // RHS history
    using PT = typename std::conditional<MIX, float, double>::type;    // mean field
    double2 opA[TPY], opB[TPY], opD[TPY];
    HT opH1[TPY], opH2[TPY];
    PT opP[TPY];
    auto load_ops = [&](int p, long o) {
        if constexpr (KIND == TK_UPDATE) {
            const UpdateArgs& a = args;
            if constexpr (MIX) opP[p] = a.pot4f[o];
            else opP[p] = a.pot4[o];
            if constexpr (SK == STAGE_RK) {
                opA[p] = get2<NT>(a.q, o);
            } else {
                opA[p] = get2<NT>(a.po3, o);
                opD[p] = get2<NT>(a.pc, o);
                opH1[p] = get2<NT>(reinterpret_cast<const HT*>(a.h1), o);
                opH2[p] = get2<NT>(reinterpret_cast<const HT*>(a.h2), o);
                if constexpr (SK == STAGE_CORR) opB[p] = get2<NT>(a.po2, o);
            }
        }
    };
    constexpr bool OPF = PF && KIND == TK_UPDATE;
    // (unconditional: an invalid point's c0 is an in-box column, and a load
    // under a lane-dependent branch would be waited for at the branch end)
    if constexpr (OPF)
#pragma unroll
        for (int p = 0; p < TPY; ++p)
            if (iz0 < izend) load_ops(p, c0[p] + pl * iz0);

    V hraw[KH], qraw[TPY];  // prefetched raw values; modes hm[] and qm[]
    int qm[TPY];
#pragma unroll
    for (int k = 0; k < KH; ++k) hraw[k] = PF && iz0 < izend ? f[hoff[k] + pl * iz0] : vz<V>::zero();
#pragma unroll
    for (int p = 0; p < TPY; ++p) {
        qm[p] = 0;
        qraw[p] = vz<V>::zero();
        if (PF) qraw[p] = f[zsrc(p, iz0 + M, qm[p])];
    }

    for (int iz = iz0; iz < izend; ++iz) {
        // with PP, plane iz goes to buffer iz & 1: the barrier of plane iz - 1
        // follows every read of plane iz - 2, the last user of that buffer
        V* lds = lds_pl[PP ? (iz & 1) : 0];
        double* ldsd = ldsd_pl[PP ? (iz & 1) : 0];
        (void)ldsd;
#pragma unroll
        for (int p = 0; p < TPY; ++p) {
            const V qn = PF ? vz<V>::mode(qraw[p], qm[p]) : zcol(p, iz + M);
            q[p][W - 1] = qn;
            if constexpr (FRC) dq[p][W - 1] = dens(qn);
        }
        if constexpr (!PP) __syncthreads();
        // plane iz into LDS: interior from the queues, halo from the prefetch
#pragma unroll
        for (int p = 0; p < TPY; ++p) {
            const int li = (ty + RS * p + M) * LX + tx + M;
            if (ragged) {  // block-uniform: some of its points lie beyond the box
                const V v = valid[p] ? q[p][M] : fetch(c0[p] + pl * iz, cm[p]);
                lds[li] = v;
                if constexpr (FRC) ldsd[li] = valid[p] ? dq[p][M] : dens(v);
            } else {
                lds[li] = q[p][M];
                if constexpr (FRC) ldsd[li] = dq[p][M];
            }
        }
#pragma unroll
        for (int k = 0; k < KH; ++k)
            if (hl[k] >= 0) {
                const V v = PF ? vz<V>::mode(hraw[k], hm[k]) : fetch(hoff[k] + pl * iz, hm[k]);
                lds[hl[k]] = v;
                if constexpr (FRC) ldsd[hl[k]] = dens(v);
            }
        __syncthreads();
        if (PF && iz + 1 < izend) {
#pragma unroll
            for (int k = 0; k < KH; ++k) hraw[k] = f[hoff[k] + pl * (iz + 1)];
#pragma unroll
            for (int p = 0; p < TPY; ++p) qraw[p] = f[zsrc(p, iz + 1 + M, qm[p])];
        }
#pragma unroll
        for (int p = 0; p < TPY; ++p) {
            if (!valid[p]) continue;
            const int ly = ty + RS * p + M;
            V sx = vz<V>::zero(), sy = vz<V>::zero(), sz = vz<V>::zero();
            const int cx = ly * LX + tx + M;  // LDS index of the point
            if constexpr (KIND == TK_UPDATE) {
                // second derivative (Laplacian) of psi
                if (nx >= W) {
                    V t = vz<V>::scale(lds[cx], tc.w2h[0]);
#pragma unroll
                    for (int k = 1; k <= M; ++k) vz<V>::fma_(t, tc.w2h[k], vz<V>::add(lds[cx + k], lds[cx - k]));
                    sx = vz<V>::scale(t, tc.s2[0]);
                }
                if (ny >= W) {
                    V t = vz<V>::scale(lds[cx], tc.w2h[0]);
#pragma unroll
                    for (int k = 1; k <= M; ++k)
                        vz<V>::fma_(t, tc.w2h[k], vz<V>::add(lds[cx + k * LX], lds[cx - k * LX]));
                    sy = vz<V>::scale(t, tc.s2[1]);
                }
                if (nz >= W) {
                    V t = vz<V>::scale(q[p][M], tc.w2h[0]);
#pragma unroll
                    for (int k = 1; k <= M; ++k) vz<V>::fma_(t, tc.w2h[k], vz<V>::add(q[p][M + k], q[p][M - k]));
                    sz = vz<V>::scale(t, tc.s2[2]);
                }
            } else {
                // gradient of den (FORCE) or psi (ENERGY)
                if (nx >= W) {
#pragma unroll
                    for (int k = 1; k <= M; ++k) vz<V>::fma_(sx, tc.w1h[k], vz<V>::sub(lds[cx + k], lds[cx - k]));
                    sx = vz<V>::scale(sx, tc.s1[0]);
                }
                if (ny >= W) {
#pragma unroll
                    for (int k = 1; k <= M; ++k)
                        vz<V>::fma_(sy, tc.w1h[k], vz<V>::sub(lds[cx + k * LX], lds[cx - k * LX]));
                    sy = vz<V>::scale(sy, tc.s1[1]);
                }
                if (nz >= W) {
#pragma unroll
                    for (int k = 1; k <= M; ++k) vz<V>::fma_(sz, tc.w1h[k], vz<V>::sub(q[p][M + k], q[p][M - k]));
                    sz = vz<V>::scale(sz, tc.s1[2]);
                }
            }
            const long o = (long)ix + (long)nx * ((long)iy[p] + (long)ny * iz);
            const double* R = args.imp->rpot;
            const double dz = g.z[iz] - R[2], dy = g.y[iy[p]] - R[1];
            const double u = uimp_at(g.potion, g.npot, g.delinter, g.x[ix] - R[0], dy * dy + dz * dz);
            if constexpr (KIND == TK_UPDATE) {
                const UpdateArgs& a = args;
                if constexpr (!OPF) load_ops(p, o);
                if constexpr (FRC) {
                    // grad den of the input state, as the FORCE kind forms it
                    double gx = 0.0, gy = 0.0, gz = 0.0;
                    if (nx >= W) {
#pragma unroll
                        for (int k = 1; k <= M; ++k) gx = fma(tc.w1h[k], ldsd[cx + k] - ldsd[cx - k], gx);
                        gx = gx * tc.s1[0];
                    }
                    if (ny >= W) {
#pragma unroll
                        for (int k = 1; k <= M; ++k) gy = fma(tc.w1h[k], ldsd[cx + k * LX] - ldsd[cx - k * LX], gy);
                        gy = gy * tc.s1[1];
                    }
                    if (nz >= W) {
#pragma unroll
                        for (int k = 1; k <= M; ++k) gz = fma(tc.w1h[k], dq[p][M + k] - dq[p][M - k], gz);
                        gz = gz * tc.s1[2];
                    }
                    acc[1] += gx * u;
                    acc[2] += gy * u;
                    acc[3] += gz * u;
                }
                const double lr = (sx.x + sy.x) + sz.x, li = (sx.y + sy.y) + sz.y;
                const double2 ps = q[p][M];
                const double pot = (double)opP[p];
                const double tr = g.lambdah * ((((((1.0 + g.tcx[ix]) + 1.0) + g.tcy[iy[p]]) + 1.0) + g.tcz[iz]));
                const double ar = lr * a.h2o2m4 - pot * ps.x;
                const double ai = li * a.h2o2m4 - pot * ps.y;
                double2 fv;
                fv.x = (tr * ar - ai) + u * ps.y;
                fv.y = (tr * ai + ar) - u * ps.x;
                double2 pn;
                if constexpr (SK == STAGE_RK) {
                    const double2 qq = opA[p];
                    double2 s1, qn;
                    s1.x = a.ra * (fv.x - a.rb * qq.x);
                    s1.y = a.ra * (fv.y - a.rb * qq.y);
                    qn.x = (qq.x + 3.0 * s1.x) - a.rc * fv.x;
                    qn.y = (qq.y + 3.0 * s1.y) - a.rc * fv.y;
                    put2<NT>(a.q, o, qn);
                    if (a.rk_j == 0) put_hist<MIX, NT>(a.hnew, o, fv);
                    pn.x = ps.x + a.dt * s1.x;
                    pn.y = ps.y + a.dt * s1.y;
                    const double d = abs2(pn);
                    put2<NT>(a.psi_out, o, pn);
                    put_den<MIX>(a, o, d);
                    if (!(d > 0.0)) atomicOr(a.guard, 1);
                } else if constexpr (SK == STAGE_PRED) {
                    const double c4 = (4.0 / 3.0) * a.dt;
                    const double2 p3 = opA[p], h1 = widen2(opH1[p]), h2 = widen2(opH2[p]);
                    const double2 pc = opD[p];
                    double2 P;
                    P.x = p3.x + c4 * ((2.0 * fv.x - h1.x) + 2.0 * h2.x);
                    P.y = p3.y + c4 * ((2.0 * fv.y - h1.y) + 2.0 * h2.y);
                    const double c112 = 112.0 / 121.0;
                    pn.x = P.x - c112 * pc.x;
                    pn.y = P.y - c112 * pc.y;
                    put2<NT>(a.psi_out, o, pn);
                    put2<NT>(a.pc, o, P);
                    put_hist<MIX, NT>(a.hnew, o, fv);
                    const double d = abs2(pn);
                    put_den<MIX>(a, o, d);
                } else {
                    const double2 p3 = opA[p], p2 = opB[p], pc = opD[p];
                    const double2 h1 = widen2(opH1[p]), h2 = widen2(opH2[p]);
                    const double t3 = 3.0 * a.dt;
                    double2 Cc, pcn;
                    Cc.x = 0.125 * ((9.0 * p3.x - p2.x) + t3 * ((fv.x + 2.0 * h1.x) - h2.x));
                    Cc.y = 0.125 * ((9.0 * p3.y - p2.y) + t3 * ((fv.y + 2.0 * h1.y) - h2.y));
                    pcn.x = pc.x - Cc.x;
                    pcn.y = pc.y - Cc.y;
                    const double c9 = 9.0 / 121.0;
                    pn.x = Cc.x + c9 * pcn.x;
                    pn.y = Cc.y + c9 * pcn.y;
                    put2<NT>(a.pc, o, pcn);
                    put2<NT>(a.psi_out, o, pn);
                    const double d = abs2(pn);
                    put_den<MIX>(a, o, d);
                    acc[0] += cabs_(pcn);
                }
            } else if constexpr (KIND == TK_FORCE) {
                acc[0] += (double)sx * u;
                acc[1] += (double)sy * u;
                acc[2] += (double)sz * u;
            } else {
                const EnergyArgs& a = args;
                if (a.mom) {
                    moment_terms(a, acc, q[p][M], sx, sy, sz, g.x[ix], g.y[iy[p]], g.z[iz]);
                    continue;
                }
                acc[EN_KIN] += (abs2(sx) + abs2(sy)) + abs2(sz);
                // den = |psi|^2 of the FP64 state (equal to the den array in FP64)
                const double d = abs2(q[p][M]);
                const double cg = a.dencgf ? (double)a.dencgf[o] : a.dencg[o];
                acc[EN_LJ] += (a.delj4f ? (double)a.delj4f[o] : a.delj4[o]) * d;
                acc[EN_COR] += d * cg * cg * (a.aux1 + a.aux2 * cg);
                if (a.lsolid) acc[EN_SOLID] += d * (1.0 + tanh(a.beta * (d - a.den_m)));
                acc[EN_HEX] += u * d;
                if (a.otc)
                    acc[EN_ALPHA] += a.falfs[o] * ((a.dxden[o] * a.intx[o] + a.dyden[o] * a.inty[o]) +
                                                   a.dzden[o] * a.intz[o]);
                acc[EN_NORM] += d;
            }
        }
        // the next plane's update operands, once this plane's are consumed
        if constexpr (OPF)
            if (iz + 1 < izend)
#pragma unroll
                for (int p = 0; p < TPY; ++p) load_ops(p, c0[p] + pl * (iz + 1));
        // shift the z queues
#pragma unroll
        for (int p = 0; p < TPY; ++p)
#pragma unroll
            for (int k = 0; k < W - 1; ++k) {
                q[p][k] = q[p][k + 1];
                if constexpr (FRC) dq[p][k] = dq[p][k + 1];
            }
    }
    bool reduce = true;
    if constexpr (KIND == TK_UPDATE) reduce = FRC || SK == STAGE_CORR;
    if (reduce) {
        block_sum<NC>(acc, red);
        if (threadIdx.x == 0) {
#pragma unroll
            for (int c = 0; c < NC; ++c) args.partial[(long)t * NC + c] = acc[c];
        }
    }
}

// ------------------------------------------------------------- launchers

static TiledCommon tiled_common(const Grid& g, const StencilCoef& sc, int tty) {
    TiledCommon tc;
    tc.g = g;
    tc.icon = sc.icon;
    for (int k = 0; k <= sc.m; ++k) {
        tc.w1h[k] = sc.w1[sc.m + k];
        tc.w2h[k] = sc.w2[sc.m + k];
    }
    for (int a = 0; a < 3; ++a) {
        tc.s1[a] = 1.0 / sc.h[a];
        tc.s2[a] = 1.0 / (sc.h[a] * sc.h[a]);
    }
    tc.ntx = (g.nx + TTX - 1) / TTX;
    tc.nty = (g.ny + tty - 1) / tty;
    const long nbxy = (long)tc.ntx * tc.nty;
    // aim at >= ~1024 workgroups (256 CUs), z chunks of 4..64 planes
    long zc = (g.nz * nbxy + 1023) / 1024;
    if (zc < 4) zc = 4;
    if (zc > 64) zc = 64;
    if (zc > g.nz) zc = g.nz;
    const int nchunk = (int)((g.nz + zc - 1) / zc);
    tc.zc = (g.nz + nchunk - 1) / nchunk;
    tc.ntz = (g.nz + tc.zc - 1) / tc.zc;
    return tc;
}

bool tiled_supported(const StencilCoef& sc) { return sc.m == 4 || sc.m == 6; }

// tile variants (points per thread, threads, pipelined loads, ping-pong LDS
// planes).  Measured on MI355X at 256^3: one point per thread wins for every
// kind (the complex register queue of two points caps a SIMD at 1-2 waves);
// for UPDATE, 256 threads (64x4 tiles, several workgroups per CU, so one
// workgroup's loads overlap another's arithmetic) with the next plane's loads
// issued before the current plane is computed beat 512-thread 64x8 tiles by
// 0.15 ms/step.  HETD_TILE_FORCE / HETD_TILE_ENERGY = 0..4 select another
// variant for those kinds; the update kernels (one per stage) are built for
// variant 3 and, selected by HETD_TILE_UPDATE=4 or 5, variant 4 (variant 3
// with ping-pong planes: one barrier per plane instead of two) and variant 5
// (4 with non-temporal loads and stores of the once-used operands).
struct TileVariant { int tpy, tnt; bool pf; int opt; };
static constexpr int kNVariants = 6;
static constexpr TileVariant kVariants[kNVariants] = {{1, 512, false, 0}, {1, 256, false, 0},
                                                      {1, 512, true, 0},  {1, 256, true, 0},
                                                      {1, 256, true, OPT_PP}, {1, 256, true, OPT_PP | OPT_NT}};

static constexpr int kUpdateVariant = 3;  // the default update kernels

static int variant_for(int kind) {
    static const char* names[3] = {"HETD_TILE_UPDATE", "HETD_TILE_FORCE", "HETD_TILE_ENERGY"};
    const char* e = getenv(names[kind]);
    if (kind == TK_UPDATE) return e && (e[0] == '4' || e[0] == '5') ? e[0] - '0' : kUpdateVariant;
    if (e && e[0] >= '0' && e[0] < '0' + kNVariants) return e[0] - '0';
    return 0;
}

template <int KIND, int M, int V, bool FRC, bool MIX, int SK>
static void launch_k(const TiledCommon& tc, const typename TArgs<KIND>::A& a, int nlaunch, hipStream_t st) {
    constexpr TileVariant tv = kVariants[V];
    // the NT option concerns the update operands only
    constexpr int opt = KIND == TK_UPDATE ? tv.opt : (tv.opt & OPT_PP);
    hipLaunchKernelGGL((k_tiled<KIND, M, tv.tpy, tv.tnt, tv.pf, FRC, MIX, SK, opt>), dim3(nlaunch), dim3(tv.tnt), 0,
                       st, tc, a);
}

template <int M, int V, bool FRC, bool MIX>
static void launch_stage(const TiledCommon& tc, const UpdateArgs& a, int nlaunch, hipStream_t st) {
    switch (a.kind) {
    case STAGE_RK: launch_k<TK_UPDATE, M, V, FRC, MIX, STAGE_RK>(tc, a, nlaunch, st); break;
    case STAGE_PRED: launch_k<TK_UPDATE, M, V, FRC, MIX, STAGE_PRED>(tc, a, nlaunch, st); break;
    default: launch_k<TK_UPDATE, M, V, FRC, MIX, STAGE_CORR>(tc, a, nlaunch, st); break;
    }
}

template <int KIND, int M, int V>
static void launch_v(const TiledCommon& tc, const typename TArgs<KIND>::A& a, int nlaunch, hipStream_t st) {
    if constexpr (KIND == TK_UPDATE) {
        const bool mix = a.denf != nullptr;
        if (a.force) {
            if (mix) launch_stage<M, V, true, true>(tc, a, nlaunch, st);
            else launch_stage<M, V, true, false>(tc, a, nlaunch, st);
        } else {
            if (mix) launch_stage<M, V, false, true>(tc, a, nlaunch, st);
            else launch_stage<M, V, false, false>(tc, a, nlaunch, st);
        }
    } else {
        launch_k<KIND, M, V, false, false, 0>(tc, a, nlaunch, st);
    }
}

template <int KIND, int M>
static void launch_m(int v, const TiledCommon& tc, const typename TArgs<KIND>::A& a, int nlaunch, hipStream_t st) {
    if constexpr (KIND == TK_UPDATE) {
        if (v == 4) launch_v<KIND, M, 4>(tc, a, nlaunch, st);
        else if (v == 5) launch_v<KIND, M, 5>(tc, a, nlaunch, st);
        else launch_v<KIND, M, kUpdateVariant>(tc, a, nlaunch, st);
    } else {
        switch (v) {
        case 0: launch_v<KIND, M, 0>(tc, a, nlaunch, st); break;
        case 1: launch_v<KIND, M, 1>(tc, a, nlaunch, st); break;
        case 2: launch_v<KIND, M, 2>(tc, a, nlaunch, st); break;
        case 3: launch_v<KIND, M, 3>(tc, a, nlaunch, st); break;
        default: launch_v<KIND, M, 4>(tc, a, nlaunch, st); break;
        }
    }
}

template <int KIND>
static hipError_t launch_tiled(const typename TArgs<KIND>::A& a, hipStream_t st, int* nblocks) {
    int v = variant_for(KIND);
    if constexpr (KIND == TK_UPDATE)
        if (a.tile == 4 || a.tile == 5 || a.tile == kUpdateVariant) v = a.tile;
    const TileVariant tv = kVariants[v];
    TiledCommon tc = tiled_common(a.g, a.sc, tv.tpy * tv.tnt / TTX);
    const int nb = tc.ntx * tc.nty * tc.ntz;
    const int nlaunch = 8 * ((nb + 7) / 8);
    if (nblocks) *nblocks = nb;
    switch (a.sc.m) {
    case 4: launch_m<KIND, 4>(v, tc, a, nlaunch, st); break;
    case 6: launch_m<KIND, 6>(v, tc, a, nlaunch, st); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_update_tiled(const UpdateArgs& a, hipStream_t st, int* nblocks) {
    return launch_tiled<TK_UPDATE>(a, st, nblocks);
}
hipError_t launch_force_tiled(const ForceArgs& a, hipStream_t st, int* nblocks) {
    return launch_tiled<TK_FORCE>(a, st, nblocks);
}
hipError_t launch_energy_tiled(const EnergyArgs& a, hipStream_t st, int* nblocks) {
    return launch_tiled<TK_ENERGY>(a, st, nblocks);
}

}  // namespace hetd
