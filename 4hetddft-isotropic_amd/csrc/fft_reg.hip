// fft_reg.hip -- register-resident FFT passes for axis lengths N = R*R
// (R = 4, 8, 16: N = 16, 64, 256) on gfx950.
//
// Four-step split of one length-N line over R threads (x[c + R j], j < R):
//   column phase: thread c does a DFT_R over j in registers, multiplies by
//                 W_N^(c k1), and stores to LDS;
//   row phase:    thread r reads row r and does a second DFT_R in registers,
//                 giving X[r + R k2].
// One LDS exchange per transform; the inverse of a spectrum held by rows
// (X[r + R k]) back to columns (x[c + R j]) is the mirror image, so the
// z-convolution (forward, multiply by the form factor, inverse) keeps the
// spectrum in registers and exchanges through LDS twice.  Twiddles W_N^e
// come from a per-block LDS copy of the exp(-2 pi i e/N) table.
//
// The passes compute exactly what the Stockham passes of fft.hip compute
// (same modes, same fused pointwise work); the library chooses this path
// when the axis length is a square of 4, 8 or 16.
#include "hetd_kernels.h"

namespace hetd {

// exp(-2 pi i e/32) = (kc32[e], -ks32[e]), e = 0..31 (double-rounded)
__device__ __forceinline__ double cos32(int e) {
    constexpr double t[32] = {
        +1.000000000000000000000e+00, +9.807852804032304305792e-01, +9.238795325112867384831e-01,
        +8.314696123025452356714e-01, +7.071067811865475727373e-01, +5.555702330196021776487e-01,
        +3.826834323650897817792e-01, +1.950903220161282758394e-01, 0.0,
        -1.950903220161282758394e-01, -3.826834323650897817792e-01, -5.555702330196021776487e-01,
        -7.071067811865475727373e-01, -8.314696123025452356714e-01, -9.238795325112867384831e-01,
        -9.807852804032304305792e-01, -1.000000000000000000000e+00, -9.807852804032304305792e-01,
        -9.238795325112867384831e-01, -8.314696123025452356714e-01, -7.071067811865475727373e-01,
        -5.555702330196021776487e-01, -3.826834323650897817792e-01, -1.950903220161282758394e-01, 0.0,
        +1.950903220161282758394e-01, +3.826834323650897817792e-01, +5.555702330196021776487e-01,
        +7.071067811865475727373e-01, +8.314696123025452356714e-01, +9.238795325112867384831e-01,
        +9.807852804032304305792e-01};
    return t[e & 31];
}
__device__ __forceinline__ double sin32(int e) { return cos32(e - 8); }

// w = W_R^e (forward exp(-2 pi i e/R)), conjugate for the inverse
template <typename C, int R>
__device__ __forceinline__ C wconst(int e, bool inv) {
    using T = decltype(C::x);
    const int e32 = (e * (32 / R)) & 31;
    const T c = (T)cos32(e32), s = (T)sin32(e32);
    return cmk<C>(c, inv ? s : -s);
}

// in-register DFT of length R (natural order in and out)
template <typename C, int R>
__device__ __forceinline__ void rdft(C* v, bool inv) {
    if constexpr (R == 2) {
        dft2(v);
    } else if constexpr (R == 4) {
        dft4(v, inv);
    } else if constexpr (R == 8) {
        dft8(v, inv);
    } else {  // R = 16 = 4 x 4
        C a[4][4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
#pragma unroll
            for (int l = 0; l < 4; ++l) a[j][l] = v[j + 4 * l];
            dft4(a[j], inv);
        }
#pragma unroll
        for (int j = 1; j < 4; ++j)
#pragma unroll
            for (int k = 1; k < 4; ++k) a[j][k] = cmul(a[j][k], wconst<C, 16>(j * k, inv));
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            C b[4] = {a[0][k], a[1][k], a[2][k], a[3][k]};
            dft4(b, inv);
#pragma unroll
            for (int m = 0; m < 4; ++m) v[k + 4 * m] = b[m];
        }
    }
}

// Column phase of a forward/inverse transform whose input is held by
// columns (v[j] = x[c + R j]); result goes to the LDS matrix M[k1][c] at
// lds[k1*rs + c*cs].  tw: LDS table exp(-2 pi i e/N).
template <typename C, int R>
__device__ __forceinline__ void col_phase(C* v, int c, C* lds, int rs, int cs, const C* tw, bool inv) {
    constexpr int N = R * R;
    rdft<C, R>(v, inv);
#pragma unroll
    for (int k = 0; k < R; ++k) {
        C w = tw[(c * k) & (N - 1)];
        if (inv) w.y = -w.y;
        lds[k * rs + c * cs] = k == 0 ? v[k] : cmul(v[k], w);
    }
}

// Row phase: thread r reads M[r][n2] and returns X[r + R k2] in v[k2].
template <typename C, int R>
__device__ __forceinline__ void row_phase(C* v, int r, const C* lds, int rs, int cs, bool inv) {
#pragma unroll
    for (int n = 0; n < R; ++n) v[n] = lds[r * rs + n * cs];
    rdft<C, R>(v, inv);
}

// Transform of data held by rows (v[j] = y[r + R j]) into columns: thread r
// does DFT_R over j, twiddles by W_N^(m2 r), stores M[m2][r]; thread c then
// reads M[c][.] and returns Y[c + R m1] in v[m1].
template <typename C, int R>
__device__ __forceinline__ void rowdata_phase1(C* v, int r, C* lds, int rs, int cs, const C* tw, bool inv) {
    constexpr int N = R * R;
    rdft<C, R>(v, inv);
#pragma unroll
    for (int m = 0; m < R; ++m) {
        C w = tw[(m * r) & (N - 1)];
        if (inv) w.y = -w.y;
        lds[m * rs + r * cs] = m == 0 ? v[m] : cmul(v[m], w);
    }
}

// ------------------------------------------------------------------ x pass
// Block: TL line pairs x R threads (thread c of pair l: tid = l*R + c).
// LDS per pair: an R x (R+1) matrix (padded rows) = R*(R+1) values.

template <typename T, int R, int TL>
__global__ void __launch_bounds__(R* TL) k_xreg(XPassArgs<T> a) {
    using C = cplx<T>;
    constexpr int N = R * R, NH = N / 2 + 1, RS = R + 1, MS = R * RS;
    __shared__ C lds[TL * MS];
    __shared__ C tw[N];
    const int tid = threadIdx.x, c = tid % R, l = tid / R;
    for (int e = tid; e < N; e += R * TL) tw[e] = twiddle<T>(a.ax, e);
    const long nlines = a.nlines;
    const long pair = (long)blockIdx.x * TL + l;
    const long la = 2 * pair, lb = la + 1;
    const bool va = la < nlines, vb = lb < nlines;
    C* M = lds + l * MS;
    C v[R];
    // the line loads are issued before the barrier that publishes the
    // twiddle table, so the two memory latencies overlap
    if (a.mode == XMODE_R2C) {
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const int x = c + R * j;
            v[j] = cmk<C>(va ? a.rin[la * N + x] : (T)0, vb ? a.rin[lb * N + x] : (T)0);
        }
    } else {
        // c2r input: Z_k = A_k + i B_k with the Hermitian extension along x;
        // imaginary parts of k = 0 and k = N/2 are not used
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const int k = c + R * j;
            const int kk = k < NH ? k : N - k;
            C A = va ? a.cin[la * a.ldh + kk] : cmk<C>((T)0, (T)0);
            C B = vb ? a.cin[lb * a.ldh + kk] : cmk<C>((T)0, (T)0);
            if (kk == 0 || 2 * kk == N) { A.y = (T)0; B.y = (T)0; }
            if (k != kk) { A.y = -A.y; B.y = -B.y; }
            v[j] = cmk<C>(A.x - B.y, A.y + B.x);
        }
    }
    __syncthreads();
    const bool inv = a.mode != XMODE_R2C;
    // columns -> rows: v[k2] = Z[r + R k2] (r = c)
    col_phase<C, R>(v, c, M, RS, 1, tw, inv);
    __syncthreads();
    row_phase<C, R>(v, c, M, RS, 1, inv);
    __syncthreads();
    if (inv) {
        // real lines a (re) and b (im) at x = c + R k2, times renor
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const int x = c + R * k;
            const long ia = la * N + x, ib = lb * N + x;
            const T ra = v[k].x * a.scale, rb = v[k].y * a.scale;
            if (a.mode == XMODE_C2R) {
                if (va) a.rout[ia] = ra;
                if (vb) a.rout[ib] = rb;
            } else if (a.mode == XMODE_CG_S) {
                const T sa = va ? cgs_point<T>(a, ia, ra) : (T)0;
                const T sb = vb ? cgs_point<T>(a, ib, rb) : (T)0;
                v[k] = cmk<C>(sa, sb);
            } else {  // POT
                if (va) pot_point<T>(a, ia, ra);
                if (vb) pot_point<T>(a, ib, rb);
            }
        }
        if (a.mode != XMODE_CG_S) return;
        // forward transform of s held by rows (x = c + R k): into columns
        rowdata_phase1<C, R>(v, c, M, RS, 1, tw, false);
        __syncthreads();
        row_phase<C, R>(v, c, M, RS, 1, false);  // v[m] = S[c + R m]
        __syncthreads();
        // natural order into LDS for the (k, N-k) split
#pragma unroll
        for (int m = 0; m < R; ++m) M[c + R * m + ((c + R * m) / R)] = v[m];
    } else {
#pragma unroll
        for (int k = 0; k < R; ++k) M[c + R * k + ((c + R * k) / R)] = v[k];
    }
    __syncthreads();
    // split Z = A + iB: A_k = (Z_k + conj Z_{N-k})/2, B_k = (Z_k - conj Z_{N-k})/2i
    const T h = (T)0.5;
#pragma unroll
    for (int j = 0; j <= R / 2; ++j) {
        const int k = c + R * j;
        if (k < NH) {
            const int km = (N - k) & (N - 1);
            const C zk = M[k + k / R], zn = M[km + km / R];
            if (va) a.cout[la * a.ldh + k] = cmk<C>(h * (zk.x + zn.x), h * (zk.y - zn.y));
            if (vb) a.cout[lb * a.ldh + k] = cmk<C>(h * (zk.y + zn.y), -h * (zk.x - zn.x));
        }
    }
}

// x pass with the LDS exchanges done in two halves (real parts, then
// imaginary parts) through a buffer of scalars: half the LDS per line pair,
// so more workgroups fit on a CU, for twice the barriers.  Same arithmetic
// as k_xreg, so the same bits.

// w[m] goes to slot wi(m), r[n] comes from slot ri(n); r may alias w.  Ends
// with a barrier, so the buffer is free again on return.
template <typename C, int NW, int NR, typename WI, typename RI>
__device__ __forceinline__ void xchg_split(const C* w, C* r, decltype(C::x)* L, WI wi, RI ri) {
    using T = decltype(C::x);
    T t[NR];
#pragma unroll
    for (int m = 0; m < NW; ++m) L[wi(m)] = w[m].x;
    __syncthreads();
#pragma unroll
    for (int n = 0; n < NR; ++n) t[n] = L[ri(n)];
    __syncthreads();
#pragma unroll
    for (int m = 0; m < NW; ++m) L[wi(m)] = w[m].y;
    __syncthreads();
#pragma unroll
    for (int n = 0; n < NR; ++n) r[n] = cmk<C>(t[n], L[ri(n)]);
    __syncthreads();
}

// v[k] *= W_N^(e k) for k >= 1 (the twiddles of col_phase / rowdata_phase1)
template <typename C, int R>
__device__ __forceinline__ void twiddle_regs(C* v, int e, const C* tw, bool inv) {
    constexpr int N = R * R;
#pragma unroll
    for (int k = 1; k < R; ++k) {
        C w = tw[(e * k) & (N - 1)];
        if (inv) w.y = -w.y;
        v[k] = cmul(v[k], w);
    }
}

template <typename T, int R, int TL>
__global__ void __launch_bounds__(R* TL) k_xreg_s(XPassArgs<T> a) {
    using C = cplx<T>;
    constexpr int N = R * R, NH = N / 2 + 1, RS = R + 1, MS = R * RS, NJ = R / 2 + 1;
    __shared__ T lds[TL * MS];
    __shared__ C tw[N];
    const int tid = threadIdx.x, c = tid % R, l = tid / R;
    for (int e = tid; e < N; e += R * TL) tw[e] = twiddle<T>(a.ax, e);
    const long nlines = a.nlines;
    const long pair = (long)blockIdx.x * TL + l;
    const long la = 2 * pair, lb = la + 1;
    const bool va = la < nlines, vb = lb < nlines;
    T* M = lds + l * MS;
    C v[R];
    if (a.mode == XMODE_R2C) {
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const int x = c + R * j;
            v[j] = cmk<C>(va ? a.rin[la * N + x] : (T)0, vb ? a.rin[lb * N + x] : (T)0);
        }
    } else {
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const int k = c + R * j;
            const int kk = k < NH ? k : N - k;
            C A = va ? a.cin[la * a.ldh + kk] : cmk<C>((T)0, (T)0);
            C B = vb ? a.cin[lb * a.ldh + kk] : cmk<C>((T)0, (T)0);
            if (kk == 0 || 2 * kk == N) { A.y = (T)0; B.y = (T)0; }
            if (k != kk) { A.y = -A.y; B.y = -B.y; }
            v[j] = cmk<C>(A.x - B.y, A.y + B.x);
        }
    }
    __syncthreads();
    const bool inv = a.mode != XMODE_R2C;
    auto to_rows = [=](int k) { return k * RS + c; };
    auto from_row = [=](int n) { return c * RS + n; };
    rdft<C, R>(v, inv);
    twiddle_regs<C, R>(v, c, tw, inv);
    xchg_split<C, R, R>(v, v, M, to_rows, from_row);
    rdft<C, R>(v, inv);  // v[k2] = Z[c + R k2]
    if (inv) {
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const int x = c + R * k;
            const long ia = la * N + x, ib = lb * N + x;
            const T ra = v[k].x * a.scale, rb = v[k].y * a.scale;
            if (a.mode == XMODE_C2R) {
                if (va) a.rout[ia] = ra;
                if (vb) a.rout[ib] = rb;
            } else if (a.mode == XMODE_CG_S) {
                const T sa = va ? cgs_point<T>(a, ia, ra) : (T)0;
                const T sb = vb ? cgs_point<T>(a, ib, rb) : (T)0;
                v[k] = cmk<C>(sa, sb);
            } else {  // POT
                if (va) pot_point<T>(a, ia, ra);
                if (vb) pot_point<T>(a, ib, rb);
            }
        }
        if (a.mode != XMODE_CG_S) return;
        // forward transform of s held by rows (x = c + R k): into columns
        rdft<C, R>(v, false);
        twiddle_regs<C, R>(v, c, tw, false);
        xchg_split<C, R, R>(v, v, M, to_rows, from_row);
        rdft<C, R>(v, false);  // v[m] = S[c + R m]
    }
    // natural order, then the (k, N-k) pairs of the split
    C z[2 * NJ];
    xchg_split<C, R, 2 * NJ>(
        v, z, M, [=](int m) { const int p = c + R * m; return p + p / R; },
        [=](int n) {
            const int k = c + R * (n >> 1);
            const int q = (n & 1) ? ((N - k) & (N - 1)) : k;
            return q + q / R;
        });
    const T h = (T)0.5;
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
        const int k = c + R * j;
        if (k < NH) {
            const C zk = z[2 * j], zn = z[2 * j + 1];
            if (va) a.cout[la * a.ldh + k] = cmk<C>(h * (zk.x + zn.x), h * (zk.y - zn.y));
            if (vb) a.cout[lb * a.ldh + k] = cmk<C>(h * (zk.y + zn.y), -h * (zk.x - zn.x));
        }
    }
}

// ------------------------------------------------------------------ y pass
// Block: TX consecutive kx columns of one z plane (tid = c*TX + l).

template <typename T, int R, int TX>
__global__ void __launch_bounds__(R* TX) k_yreg(YPassArgs<T> a) {
    using C = cplx<T>;
    constexpr int N = R * R;
    __shared__ C lds[N * TX];
    __shared__ C tw[N];
    const int tid = threadIdx.x, l = tid % TX, c = tid / TX;
    for (int e = tid; e < N; e += R * TX) tw[e] = twiddle<T>(a.ax, e);
    const int nxh = a.nxh, ldh = a.ldh, kx = blockIdx.x * TX + l, z = blockIdx.y;
    C* buf = a.buf[blockIdx.z];
    const long base = (long)ldh * N * z + kx;
    const bool ok = kx < nxh;
    const bool inv = a.inv != 0;
    C v[R];
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = ok ? buf[base + (long)ldh * (c + R * j)] : cmk<C>((T)0, (T)0);
    __syncthreads();
    col_phase<C, R>(v, c, lds + l, R * TX, TX, tw, inv);
    __syncthreads();
    row_phase<C, R>(v, c, lds + l, R * TX, TX, inv);
    if (ok)
#pragma unroll
        for (int k = 0; k < R; ++k) buf[base + (long)ldh * (c + R * k)] = v[k];
}

// y pass with the split exchange of k_xreg_s (same arithmetic as k_yreg)
template <typename T, int R, int TX>
__global__ void __launch_bounds__(R* TX) k_yreg_s(YPassArgs<T> a) {
    using C = cplx<T>;
    constexpr int N = R * R;
    __shared__ T lds[N * TX];
    __shared__ C tw[N];
    const int tid = threadIdx.x, l = tid % TX, c = tid / TX;
    for (int e = tid; e < N; e += R * TX) tw[e] = twiddle<T>(a.ax, e);
    const int nxh = a.nxh, ldh = a.ldh, kx = blockIdx.x * TX + l, z = blockIdx.y;
    C* buf = a.buf[blockIdx.z];
    const long base = (long)ldh * N * z + kx;
    const bool ok = kx < nxh;
    const bool inv = a.inv != 0;
    C v[R];
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = ok ? buf[base + (long)ldh * (c + R * j)] : cmk<C>((T)0, (T)0);
    __syncthreads();
    rdft<C, R>(v, inv);
    twiddle_regs<C, R>(v, c, tw, inv);
    xchg_split<C, R, R>(v, v, lds, [=](int k) { return (k * R + c) * TX + l; },
                        [=](int n) { return (c * R + n) * TX + l; });
    rdft<C, R>(v, inv);
    if (ok)
#pragma unroll
        for (int k = 0; k < R; ++k) buf[base + (long)ldh * (c + R * k)] = v[k];
}

// ------------------------------------------------------------------ z pass

// NK1: at most one kernel (FWD, INV, or CONV with nk = 1): the product is
// formed in place, which halves the registers held across the inverse
template <typename T, int R, int TX, bool NK1>
__global__ void __launch_bounds__(R* TX) k_zreg(ZPassArgs<T> a) {
    using C = cplx<T>;
    constexpr int N = R * R;
    __shared__ C lds[N * TX];
    __shared__ C tw[N];
    const int tid = threadIdx.x, l = tid % TX, c = tid / TX;
    for (int e = tid; e < N; e += R * TX) tw[e] = twiddle<T>(a.ax, e);
    const int nxh = a.nxh, kx = blockIdx.x * TX + l, ky = blockIdx.y;
    const long zs = (long)a.ldh * a.ny;
    const long base = (long)a.ldh * ky + kx;
    const bool ok = kx < nxh;
    const C* in = a.in[blockIdx.z];
    C* lt = lds + l;
    C v[R];
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = ok ? in[base + zs * (c + R * j)] : cmk<C>((T)0, (T)0);
    __syncthreads();
    if (a.mode == ZMODE_INV) {
        col_phase<C, R>(v, c, lt, R * TX, TX, tw, true);
        __syncthreads();
        row_phase<C, R>(v, c, lt, R * TX, TX, true);
        C* out = a.out[blockIdx.z][0];
        if (ok)
#pragma unroll
            for (int k = 0; k < R; ++k) out[base + zs * (c + R * k)] = v[k];
        return;
    }
    col_phase<C, R>(v, c, lt, R * TX, TX, tw, false);
    __syncthreads();
    row_phase<C, R>(v, c, lt, R * TX, TX, false);  // v[k] = X[c + R k]
    if (a.mode == ZMODE_FWD || a.fwd_out) {
        C* out = a.mode == ZMODE_FWD ? a.out[blockIdx.z][0] : a.fwd_out;
        if (ok)
#pragma unroll
            for (int k = 0; k < R; ++k) out[base + zs * (c + R * k)] = v[k];
        if (a.mode == ZMODE_FWD) return;
    }
    if (a.kout && ok)
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const long o = base + zs * (c + R * k);
            a.kout[o] = cscale(v[k], a.kkern[o]);
        }
    if constexpr (NK1) {
        const T* K = a.kern[0];
#pragma unroll
        for (int k = 0; k < R; ++k) v[k] = cscale(v[k], ok ? K[base + zs * (c + R * k)] : (T)0);
        if (a.addin && ok)
#pragma unroll
            for (int k = 0; k < R; ++k) v[k] = cadd(v[k], a.addin[base + zs * (c + R * k)]);
        __syncthreads();
        rowdata_phase1<C, R>(v, c, lt, R * TX, TX, tw, true);
        __syncthreads();
        row_phase<C, R>(v, c, lt, R * TX, TX, true);
        C* out = a.out[blockIdx.z][0];
        if (ok)
#pragma unroll
            for (int m = 0; m < R; ++m) out[base + zs * (c + R * m)] = v[m];
        return;
    }
    for (int q = 0; q < a.nk; ++q) {
        const T* K = a.kern[q];
        C w[R];
#pragma unroll
        for (int k = 0; k < R; ++k) w[k] = cscale(v[k], ok ? K[base + zs * (c + R * k)] : (T)0);
        if (q == 0 && a.addin && ok)
#pragma unroll
            for (int k = 0; k < R; ++k) w[k] = cadd(w[k], a.addin[base + zs * (c + R * k)]);
        __syncthreads();  // previous readers of the LDS matrix are done
        rowdata_phase1<C, R>(w, c, lt, R * TX, TX, tw, true);
        __syncthreads();
        row_phase<C, R>(w, c, lt, R * TX, TX, true);  // w[m] = x[c + R m]
        C* out = a.out[blockIdx.z][q];
        if (ok)
#pragma unroll
            for (int m = 0; m < R; ++m) out[base + zs * (c + R * m)] = w[m];
    }
}

// --------------------------------------------------------------- launchers

static int reg_radix(int n) {
    switch (n) {
    case 16: return 4;
    case 64: return 8;
    case 256: return 16;
    default: return 0;
    }
}

bool reg_fft_supported(int n) { return reg_radix(n) != 0; }

template <typename T>
hipError_t launch_xreg(XPassArgs<T> a, hipStream_t st) {
    const int R = reg_radix(a.ax.n);
    const long npairs = (a.nlines + 1) / 2;
    if (a.split) {
        switch (R) {
        case 4: hipLaunchKernelGGL((k_xreg_s<T, 4, 32>), dim3((npairs + 31) / 32), dim3(128), 0, st, a); break;
        case 8: hipLaunchKernelGGL((k_xreg_s<T, 8, 16>), dim3((npairs + 15) / 16), dim3(128), 0, st, a); break;
        case 16: hipLaunchKernelGGL((k_xreg_s<T, 16, 8>), dim3((npairs + 7) / 8), dim3(128), 0, st, a); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    switch (R) {
    case 4: hipLaunchKernelGGL((k_xreg<T, 4, 32>), dim3((npairs + 31) / 32), dim3(128), 0, st, a); break;
    case 8: hipLaunchKernelGGL((k_xreg<T, 8, 16>), dim3((npairs + 15) / 16), dim3(128), 0, st, a); break;
    case 16: hipLaunchKernelGGL((k_xreg<T, 16, 8>), dim3((npairs + 7) / 8), dim3(128), 0, st, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename T>
hipError_t launch_yreg(YPassArgs<T> a, int nz, int nbuf, hipStream_t st) {
    const int R = reg_radix(a.ax.n);
    if (a.split) {
        switch (R) {
        case 4: hipLaunchKernelGGL((k_yreg_s<T, 4, 32>), dim3((a.nxh + 31) / 32, nz, nbuf), dim3(128), 0, st, a); break;
        case 8: hipLaunchKernelGGL((k_yreg_s<T, 8, 16>), dim3((a.nxh + 15) / 16, nz, nbuf), dim3(128), 0, st, a); break;
        case 16:
            hipLaunchKernelGGL((k_yreg_s<T, 16, 16>), dim3((a.nxh + 15) / 16, nz, nbuf), dim3(256), 0, st, a);
            break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    switch (R) {
    case 4: hipLaunchKernelGGL((k_yreg<T, 4, 32>), dim3((a.nxh + 31) / 32, nz, nbuf), dim3(128), 0, st, a); break;
    case 8: hipLaunchKernelGGL((k_yreg<T, 8, 16>), dim3((a.nxh + 15) / 16, nz, nbuf), dim3(128), 0, st, a); break;
    case 16: hipLaunchKernelGGL((k_yreg<T, 16, 16>), dim3((a.nxh + 15) / 16, nz, nbuf), dim3(256), 0, st, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename T, int R, int TX>
static hipError_t zreg_go(const ZPassArgs<T>& a, int nbuf, hipStream_t st) {
    const dim3 grid((a.nxh + TX - 1) / TX, a.ny, nbuf);
    if (a.mode != ZMODE_CONV || a.nk == 1)
        hipLaunchKernelGGL((k_zreg<T, R, TX, true>), grid, dim3(R * TX), 0, st, a);
    else
        hipLaunchKernelGGL((k_zreg<T, R, TX, false>), grid, dim3(R * TX), 0, st, a);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_zreg(ZPassArgs<T> a, int nbuf, hipStream_t st) {
    const int R = reg_radix(a.ax.n);
    switch (R) {
    case 4: return zreg_go<T, 4, 32>(a, nbuf, st);
    case 8: return zreg_go<T, 8, 16>(a, nbuf, st);
    case 16: return zreg_go<T, 16, 16>(a, nbuf, st);
    default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_xreg<double>(XPassArgs<double>, hipStream_t);
template hipError_t launch_xreg<float>(XPassArgs<float>, hipStream_t);
template hipError_t launch_yreg<double>(YPassArgs<double>, int, int, hipStream_t);
template hipError_t launch_yreg<float>(YPassArgs<float>, int, int, hipStream_t);
template hipError_t launch_zreg<double>(ZPassArgs<double>, int, hipStream_t);
template hipError_t launch_zreg<float>(ZPassArgs<float>, int, hipStream_t);

}  // namespace hetd
