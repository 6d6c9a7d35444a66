// fft_reg2.hip -- register-resident x, y and z FFT passes for axis lengths
// N = P*Q (P <= Q <= 24, factors 2, 3, 5) that are not squares handled by
// fft_reg.hip: 12 .. 384 (e.g. the 72 .. 192 edges of an ensemble of
// droplets, 128 for a single 128^3 run, 384 for the large one).
//
// One line of N points is held by max(P,Q) threads.  Forward (or inverse)
// transform of data held "by columns" (thread c < Q holds x[c + Q j], j < P):
//   A: DFT_P over j, times W_N^(c k1), to LDS M[k1][c];
//   B: thread r < P reads M[r][.] and does DFT_Q, giving X[r + P k2].
// The inverse of a spectrum held "by rows" (thread r < P holds Y[r + P m2])
// back to columns is the mirror image:
//   C: DFT_Q over m2, times W_N^(r c), to LDS M'[c][r];
//   D: thread c < Q reads M'[c][.] and does DFT_P, giving y[c + Q m1].
// So a z convolution (forward, multiply by the form factor, inverse) keeps
// the spectrum in registers and exchanges through LDS twice, as the square
// passes do.  The DFTs of length P and Q are composed in registers from the
// radix-2/3/4/5/8 butterflies, their inner twiddles read from the per-block
// LDS copy of the exp(-2 pi i e/N) table.
#include "hetd_kernels.h"

namespace hetd {

// In-register DFT of length N (natural order in and out).  tw[e*ts] is
// exp(-2 pi i e/N) (conjugated for the inverse).
template <typename C, int N>
__device__ __forceinline__ void dftn(C* v, bool inv, const C* tw, int ts) {
    if constexpr (N == 1) {
        return;
    } else if constexpr (N == 2 || N == 3 || N == 4 || N == 5 || N == 8) {
        dftR<N>(v, inv);
    } else {
        constexpr int A = (N % 4 == 0) ? 4 : (N % 2 == 0 ? 2 : (N % 3 == 0 ? 3 : 5));
        constexpr int B = N / A;
        static_assert(A * B == N, "length with a prime factor above 5");
        C a[B][A];
#pragma unroll
        for (int j = 0; j < B; ++j) {
#pragma unroll
            for (int l = 0; l < A; ++l) a[j][l] = v[j + B * l];
            dftR<A>(a[j], inv);
        }
#pragma unroll
        for (int j = 1; j < B; ++j)
#pragma unroll
            for (int k = 1; k < A; ++k) {
                C w = tw[((j * k) % N) * ts];
                if (inv) w.y = -w.y;
                a[j][k] = cmul(a[j][k], w);
            }
#pragma unroll
        for (int k = 0; k < A; ++k) {
            C b[B];
#pragma unroll
            for (int j = 0; j < B; ++j) b[j] = a[j][k];
            dftn<C, B>(b, inv, tw, ts * A);
#pragma unroll
            for (int m = 0; m < B; ++m) v[k + A * m] = b[m];
        }
    }
}

// The LDS matrices: M[k1][c] at lt[(k1*SA + c)*TX] (k1 < P, c < Q) and
// M'[c][r] at lt[(c*SC + r)*TX] (c < Q, r < P); SA >= Q, SC >= P (a padded
// row when lines do not interleave, TX = 1).

// phase A (thread c < Q, v[j] = x[c + Q j]) -> M[k1][c]
template <typename C, int P, int Q, int TX, int SA>
__device__ __forceinline__ void phase_a(C* v, int c, C* lt, const C* tw, bool inv) {
    constexpr int N = P * Q;
    dftn<C, P>(v, inv, tw, Q);
#pragma unroll
    for (int k = 0; k < P; ++k) {
        C w = tw[(c * k) % N];
        if (inv) w.y = -w.y;
        lt[(k * SA + c) * TX] = k == 0 ? v[k] : cmul(v[k], w);
    }
}

// phase B (thread r < P): v[k2] = X[r + P k2]
template <typename C, int P, int Q, int TX, int SA>
__device__ __forceinline__ void phase_b(C* v, int r, const C* lt, const C* tw, bool inv) {
#pragma unroll
    for (int n = 0; n < Q; ++n) v[n] = lt[(r * SA + n) * TX];
    dftn<C, Q>(v, inv, tw, P);
}

// phase C (thread r < P, v[m2] = Y[r + P m2]) -> M'[c][r]
template <typename C, int P, int Q, int TX, int SC>
__device__ __forceinline__ void phase_c(C* v, int r, C* lt, const C* tw, bool inv) {
    constexpr int N = P * Q;
    dftn<C, Q>(v, inv, tw, P);
#pragma unroll
    for (int c = 0; c < Q; ++c) {
        C w = tw[(r * c) % N];
        if (inv) w.y = -w.y;
        lt[(c * SC + r) * TX] = c == 0 ? v[c] : cmul(v[c], w);
    }
}

// phase D (thread c < Q): v[m1] = y[c + Q m1]
template <typename C, int P, int Q, int TX, int SC>
__device__ __forceinline__ void phase_d(C* v, int c, const C* lt, const C* tw, bool inv) {
#pragma unroll
    for (int r = 0; r < P; ++r) v[r] = lt[(c * SC + r) * TX];
    dftn<C, P>(v, inv, tw, Q);
}

// ------------------------------------------------------------------ x pass
// Block: TL line pairs x L threads (thread c of pair l: tid = l*L + c), the
// pair packed as one complex line a + i b; LDS per pair: the padded P x Q
// matrices, then the natural-order line for the (k, N-k) split.

__host__ __device__ constexpr int x2_lds(int P, int Q) {
    const int a = P * (Q + 1), b = Q * (P + 1), n = P * Q + (P * Q) / 16 + 1;
    return a > b ? (a > n ? a : n) : (b > n ? b : n);
}

template <typename T, int P, int Q, int TL>
__global__ void __launch_bounds__((P > Q ? P : Q) * TL) k_x2(XPassArgs<T> a) {
    using C = cplx<T>;
    constexpr int N = P * Q, NH = N / 2 + 1, L = P > Q ? P : Q, MS = x2_lds(P, Q);
    constexpr int SA = Q + 1, SC = P + 1;
    __shared__ C lds[TL * MS];
    __shared__ C tw[N];
    const int tid = threadIdx.x, c = tid % L, l = tid / L;
    for (int e = tid; e < N; e += L * TL) tw[e] = twiddle<T>(a.ax, e);
    const long nlines = a.nlines;
    const long pair = (long)blockIdx.x * TL + l;
    const long la = 2 * pair, lb = la + 1;
    const bool va = la < nlines, vb = lb < nlines;
    C* M = lds + l * MS;
    auto pad = [](int i) { return i + (i >> 4); };  // natural order, one pad per 16
    C v[L];
    // line loads first: they overlap the twiddle-table fill (barrier below)
    if (c < Q) {
        if (a.mode == XMODE_R2C) {
#pragma unroll
            for (int j = 0; j < P; ++j) {
                const int x = c + Q * j;
                v[j] = cmk<C>(va ? a.rin[la * N + x] : (T)0, vb ? a.rin[lb * N + x] : (T)0);
            }
        } else {
            // c2r input: Z_k = A_k + i B_k with the Hermitian extension along x;
            // imaginary parts of k = 0 and k = N/2 are not used
#pragma unroll
            for (int j = 0; j < P; ++j) {
                const int k = c + Q * j;
                const int kk = k < NH ? k : N - k;
                C A = va ? a.cin[la * a.ldh + kk] : cmk<C>((T)0, (T)0);
                C B = vb ? a.cin[lb * a.ldh + kk] : cmk<C>((T)0, (T)0);
                if (kk == 0 || 2 * kk == N) { A.y = (T)0; B.y = (T)0; }
                if (k != kk) { A.y = -A.y; B.y = -B.y; }
                v[j] = cmk<C>(A.x - B.y, A.y + B.x);
            }
        }
    }
    __syncthreads();
    const bool inv = a.mode != XMODE_R2C;
    if (c < Q) phase_a<C, P, Q, 1, SA>(v, c, M, tw, inv);
    __syncthreads();
    if (c < P) phase_b<C, P, Q, 1, SA>(v, c, M, tw, inv);  // v[k2] = Z[c + P k2]
    __syncthreads();
    if (inv) {
        if (c < P) {
            // real lines a (re) and b (im) at x = c + P k2, times renor
#pragma unroll
            for (int k = 0; k < Q; ++k) {
                const int x = c + P * k;
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
        }
        if (a.mode != XMODE_CG_S) return;
        // forward transform of s held by rows (x = c + P k): into columns
        if (c < P) phase_c<C, P, Q, 1, SC>(v, c, M, tw, false);
        __syncthreads();
        if (c < Q) phase_d<C, P, Q, 1, SC>(v, c, M, tw, false);  // v[m] = S[c + Q m]
        __syncthreads();
        if (c < Q)
#pragma unroll
            for (int m = 0; m < P; ++m) M[pad(c + Q * m)] = v[m];
    } else {
        if (c < P)
#pragma unroll
            for (int k = 0; k < Q; ++k) M[pad(c + P * k)] = v[k];
    }
    __syncthreads();
    // split Z = A + iB: A_k = (Z_k + conj Z_{N-k})/2, B_k = (Z_k - conj Z_{N-k})/2i
    const T h = (T)0.5;
    for (int k = c; k < NH; k += L) {
        const int km = k == 0 ? 0 : N - k;
        const C zk = M[pad(k)], zn = M[pad(km)];
        if (va) a.cout[la * a.ldh + k] = cmk<C>(h * (zk.x + zn.x), h * (zk.y - zn.y));
        if (vb) a.cout[lb * a.ldh + k] = cmk<C>(h * (zk.y + zn.y), -h * (zk.x - zn.x));
    }
}

// ------------------------------------------------------------------ y pass
// Block: TX consecutive kx columns of one z plane; thread (l, c): tid = c*TX + l.

template <typename T, int P, int Q, int TX>
__global__ void __launch_bounds__((P > Q ? P : Q) * TX) k_y2(YPassArgs<T> a) {
    using C = cplx<T>;
    constexpr int N = P * Q, L = P > Q ? P : Q;
    __shared__ C lds[N * TX];
    __shared__ C tw[N];
    const int tid = threadIdx.x, l = tid % TX, c = tid / TX;
    for (int e = tid; e < N; e += L * TX) tw[e] = twiddle<T>(a.ax, e);
    const int nxh = a.nxh, ldh = a.ldh, kx = blockIdx.x * TX + l, z = blockIdx.y;
    C* buf = a.buf[blockIdx.z];
    const long base = (long)ldh * N * z + kx;
    const bool ok = kx < nxh;
    const bool inv = a.inv != 0;
    C v[L];
    if (c < Q)
#pragma unroll
        for (int j = 0; j < P; ++j) v[j] = ok ? buf[base + (long)ldh * (c + Q * j)] : cmk<C>((T)0, (T)0);
    __syncthreads();
    if (c < Q) phase_a<C, P, Q, TX, Q>(v, c, lds + l, tw, inv);
    __syncthreads();
    if (c < P) {
        phase_b<C, P, Q, TX, Q>(v, c, lds + l, tw, inv);
        if (ok)
#pragma unroll
            for (int k = 0; k < Q; ++k) buf[base + (long)ldh * (c + P * k)] = v[k];
    }
}

// ------------------------------------------------------------------ z pass
// FWD, INV, and CONV with one form factor (the stepping form of poten and
// every other convolution but the OTC first pass, which uses fft.hip).

template <typename T, int P, int Q, int TX>
__global__ void __launch_bounds__((P > Q ? P : Q) * TX) k_z2(ZPassArgs<T> a) {
    using C = cplx<T>;
    constexpr int N = P * Q, L = P > Q ? P : Q;
    __shared__ C lds[N * TX];
    __shared__ C tw[N];
    const int tid = threadIdx.x, l = tid % TX, c = tid / TX;
    for (int e = tid; e < N; e += L * TX) tw[e] = twiddle<T>(a.ax, e);
    const int nxh = a.nxh, kx = blockIdx.x * TX + l, ky = blockIdx.y;
    const long zs = (long)a.ldh * a.ny;
    const long base = (long)a.ldh * ky + kx;
    const bool ok = kx < nxh;
    const C* in = a.in[blockIdx.z];
    C* lt = lds + l;
    const bool inv_only = a.mode == ZMODE_INV;
    C v[L];
    if (c < Q)
#pragma unroll
        for (int j = 0; j < P; ++j) v[j] = ok ? in[base + zs * (c + Q * j)] : cmk<C>((T)0, (T)0);
    __syncthreads();
    if (c < Q) phase_a<C, P, Q, TX, Q>(v, c, lt, tw, inv_only);
    __syncthreads();
    if (c < P) phase_b<C, P, Q, TX, Q>(v, c, lt, tw, inv_only);  // v[k2] = X[c + P k2]
    if (a.mode != ZMODE_CONV) {
        C* out = a.out[blockIdx.z][0];
        if (c < P && ok)
#pragma unroll
            for (int k = 0; k < Q; ++k) out[base + zs * (c + P * k)] = v[k];
        return;
    }
    if (c < P && ok) {
        if (a.fwd_out)
#pragma unroll
            for (int k = 0; k < Q; ++k) a.fwd_out[base + zs * (c + P * k)] = v[k];
        if (a.kout)
#pragma unroll
            for (int k = 0; k < Q; ++k) {
                const long o = base + zs * (c + P * k);
                a.kout[o] = cscale(v[k], a.kkern[o]);
            }
        const T* K = a.kern[0];
#pragma unroll
        for (int k = 0; k < Q; ++k) v[k] = cscale(v[k], K[base + zs * (c + P * k)]);
        if (a.addin)
#pragma unroll
            for (int k = 0; k < Q; ++k) v[k] = cadd(v[k], a.addin[base + zs * (c + P * k)]);
    } else if (c < P) {
#pragma unroll
        for (int k = 0; k < Q; ++k) v[k] = cmk<C>((T)0, (T)0);
    }
    __syncthreads();  // the readers of M are done
    if (c < P) phase_c<C, P, Q, TX, P>(v, c, lt, tw, true);
    __syncthreads();
    if (c < Q) {
        phase_d<C, P, Q, TX, P>(v, c, lt, tw, true);  // v[m1] = x[c + Q m1]
        C* out = a.out[blockIdx.z][0];
        if (ok)
#pragma unroll
            for (int m = 0; m < P; ++m) out[base + zs * (c + Q * m)] = v[m];
    }
}

// --------------------------------------------------------------- launchers

// supported lengths and their split N = P*Q, P <= Q <= 24
#define HETD_REG2_SIZES(X) \
    X(12, 3, 4) X(20, 4, 5) X(24, 4, 6) X(30, 5, 6) X(32, 4, 8) X(36, 6, 6) X(40, 5, 8) X(45, 5, 9) \
    X(48, 6, 8) X(50, 5, 10) X(60, 6, 10) X(72, 8, 9) X(80, 8, 10) X(90, 9, 10) X(96, 8, 12) X(100, 10, 10) \
    X(108, 9, 12) X(120, 10, 12) X(128, 8, 16) X(144, 12, 12) X(150, 10, 15) X(160, 10, 16) X(180, 12, 15) \
    X(192, 12, 16) X(240, 15, 16) X(162, 9, 18) X(216, 12, 18) X(288, 16, 18) X(320, 16, 20) X(384, 16, 24)

bool reg2_supported(int n) {
    switch (n) {
#define HETD_CASE(NN, PP, QQ) case NN: return true;
        HETD_REG2_SIZES(HETD_CASE)
#undef HETD_CASE
    default: return false;
    }
}

// columns per y/z block: 16 (256-byte rows), 8 above N = 256 to keep LDS < 64 KB
template <int N> constexpr int tx2() { return N > 256 ? 8 : 16; }

// line pairs per block: about 256 threads
template <int P, int Q> constexpr int x2_tl() { return 256 / (P > Q ? P : Q) > 0 ? 256 / (P > Q ? P : Q) : 1; }

template <typename T>
hipError_t launch_x2(XPassArgs<T> a, hipStream_t st) {
    const long npairs = (a.nlines + 1) / 2;
    switch (a.ax.n) {
#define HETD_CASE(NN, PP, QQ)                                                                           \
    case NN: {                                                                                          \
        constexpr int TL = x2_tl<PP, QQ>() > 16 ? 16 : x2_tl<PP, QQ>();                                \
        hipLaunchKernelGGL((k_x2<T, PP, QQ, TL>), dim3((unsigned)((npairs + TL - 1) / TL)),            \
                           dim3((PP > QQ ? PP : QQ) * TL), 0, st, a);                                   \
        break;                                                                                          \
    }
        HETD_REG2_SIZES(HETD_CASE)
#undef HETD_CASE
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename T>
hipError_t launch_y2(YPassArgs<T> a, int nz, int nbuf, hipStream_t st) {
    switch (a.ax.n) {
#define HETD_CASE(NN, PP, QQ)                                                                             \
    case NN: {                                                                                            \
        constexpr int TX = tx2<NN>();                                                                     \
        hipLaunchKernelGGL((k_y2<T, PP, QQ, TX>), dim3((a.nxh + TX - 1) / TX, nz, nbuf),                \
                           dim3((PP > QQ ? PP : QQ) * TX), 0, st, a);                                     \
        break;                                                                                            \
    }
        HETD_REG2_SIZES(HETD_CASE)
#undef HETD_CASE
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename T>
hipError_t launch_z2(ZPassArgs<T> a, int nbuf, hipStream_t st) {
    if (a.mode == ZMODE_CONV && a.nk != 1) return hipErrorInvalidValue;
    switch (a.ax.n) {
#define HETD_CASE(NN, PP, QQ)                                                                             \
    case NN: {                                                                                            \
        constexpr int TX = tx2<NN>();                                                                     \
        hipLaunchKernelGGL((k_z2<T, PP, QQ, TX>), dim3((a.nxh + TX - 1) / TX, a.ny, nbuf),              \
                           dim3((PP > QQ ? PP : QQ) * TX), 0, st, a);                                     \
        break;                                                                                            \
    }
        HETD_REG2_SIZES(HETD_CASE)
#undef HETD_CASE
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template hipError_t launch_x2<double>(XPassArgs<double>, hipStream_t);
template hipError_t launch_x2<float>(XPassArgs<float>, hipStream_t);
template hipError_t launch_y2<double>(YPassArgs<double>, int, int, hipStream_t);
template hipError_t launch_y2<float>(YPassArgs<float>, int, int, hipStream_t);
template hipError_t launch_z2<double>(ZPassArgs<double>, int, hipStream_t);
template hipError_t launch_z2<float>(ZPassArgs<float>, int, hipStream_t);

}  // namespace hetd
