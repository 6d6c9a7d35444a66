// hetd_dev.h -- device-side building blocks shared by the gfx950 kernels:
// complex arithmetic, in-register radix-2/3/4/5/8 DFTs, one Stockham stage
// over an LDS tile, wavefront reductions, the impurity-field interpolation,
// and the kernel-argument structures.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hetd {

constexpr int kMaxStages = 14;
constexpr int kMaxNpd = 25;

// ------------------------------------------------------------------ complex

template <typename T> struct cplx_t;
template <> struct cplx_t<double> { using type = double2; };
template <> struct cplx_t<float> { using type = float2; };
template <typename T> using cplx = typename cplx_t<T>::type;

template <typename C> __device__ __forceinline__ C cmk(decltype(C::x) a, decltype(C::x) b) { C r; r.x = a; r.y = b; return r; }
// Solid penalty C_s (1 + t + beta d (1 - t^2)), t = tanh(beta (d - den_m))
// (poten.f90:80-83), through s = exp(-2|x|):
//   1 + t = 2/(1+s) (x >= 0) or 2s/(1+s) (x < 0),   1 - t^2 = 4s/(1+s)^2,
// which has no cancellation where t is near -1 (all of a He density, where
// 1 + t ~ 1e-13) and costs one exp and one division instead of a tanh.
template <typename T> __device__ __forceinline__ T solid_penalty(T d, T csol, T beta, T den_m) {
    const T x = beta * (d - den_m);
    const T sx = exp((T)-2 * fabs(x));
    const T r = (T)1 / ((T)1 + sx);
    const T onept = x >= (T)0 ? (T)2 * r : (T)2 * sx * r;
    const T omt2 = (T)4 * sx * r * r;
    return csol * (onept + beta * d * omt2);
}

// |z|^2 and |z|: x*x + y*y with one rounding fewer than hypot(x,y)^2 (the
// reference's abs(z)**2); the two differ by at most an ulp or two
__device__ __forceinline__ double abs2(double2 v) { return fma(v.x, v.x, v.y * v.y); }
__device__ __forceinline__ double cabs_(double2 v) { return sqrt(abs2(v)); }
// Streams read or written once per kernel: non-temporal (streaming) loads
// and stores, which let the caches keep the data that is read again.
typedef double hetd_d2v __attribute__((ext_vector_type(2)));
typedef float hetd_f2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 ld_nt(const double2* p) {
    const hetd_d2v t = __builtin_nontemporal_load(reinterpret_cast<const hetd_d2v*>(p));
    return make_double2(t[0], t[1]);
}
__device__ __forceinline__ float2 ld_nt(const float2* p) {
    const hetd_f2v t = __builtin_nontemporal_load(reinterpret_cast<const hetd_f2v*>(p));
    return make_float2(t[0], t[1]);
}
__device__ __forceinline__ void st_nt(double2* p, double2 v) {
    hetd_d2v t;
    t[0] = v.x;
    t[1] = v.y;
    __builtin_nontemporal_store(t, reinterpret_cast<hetd_d2v*>(p));
}
__device__ __forceinline__ void st_nt(float2* p, float2 v) {
    hetd_f2v t;
    t[0] = v.x;
    t[1] = v.y;
    __builtin_nontemporal_store(t, reinterpret_cast<hetd_f2v*>(p));
}
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return cmk<C>(a.x + b.x, a.y + b.y); }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return cmk<C>(a.x - b.x, a.y - b.y); }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) { return cmk<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
template <typename C> __device__ __forceinline__ C cscale(C a, decltype(C::x) s) { return cmk<C>(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (inverse)
template <typename C> __device__ __forceinline__ C cmi(C a, bool inv) { return inv ? cmk<C>(-a.y, a.x) : cmk<C>(a.y, -a.x); }

// ------------------------------------------------------------- radix DFTs
// Forward uses exp(-2 pi i/R); inverse the conjugate.  Constants are the
// exact cos/sin of 2 pi/R rounded to double.

template <typename C> __device__ __forceinline__ void dft2(C* v) {
    C a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <typename C> __device__ __forceinline__ void dft3(C* v, bool inv) {
    using R = decltype(C::x);
    const R c = (R)-0.5, s = (R)0.86602540378443864676;
    C t1 = cadd(v[1], v[2]);
    C t2 = cmk<C>(v[0].x + c * t1.x, v[0].y + c * t1.y);
    C d = cscale(csub(v[1], v[2]), s);
    C md = cmi(d, inv);  // -i s (a1 - a2) forward
    v[0] = cadd(v[0], t1);
    v[1] = cadd(t2, md);
    v[2] = csub(t2, md);
}

template <typename C> __device__ __forceinline__ void dft4(C* v, bool inv) {
    C a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    C b0 = cadd(v[1], v[3]), b1 = cmi(csub(v[1], v[3]), inv);
    v[0] = cadd(a0, b0);
    v[2] = csub(a0, b0);
    v[1] = cadd(a1, b1);
    v[3] = csub(a1, b1);
}

template <typename C> __device__ __forceinline__ void dft5(C* v, bool inv) {
    using R = decltype(C::x);
    const R c1 = (R)0.30901699437494742410, c2 = (R)-0.80901699437494742410;
    const R s1 = (R)0.95105651629515357212, s2 = (R)0.58778525229247312917;
    C p1 = cadd(v[1], v[4]), m1 = csub(v[1], v[4]);
    C p2 = cadd(v[2], v[3]), m2 = csub(v[2], v[3]);
    C a = cmk<C>(v[0].x + c1 * p1.x + c2 * p2.x, v[0].y + c1 * p1.y + c2 * p2.y);
    C b = cmk<C>(v[0].x + c2 * p1.x + c1 * p2.x, v[0].y + c2 * p1.y + c1 * p2.y);
    C ta = cmi(cmk<C>(s1 * m1.x + s2 * m2.x, s1 * m1.y + s2 * m2.y), inv);
    C tb = cmi(cmk<C>(s2 * m1.x - s1 * m2.x, s2 * m1.y - s1 * m2.y), inv);
    v[0] = cmk<C>(v[0].x + p1.x + p2.x, v[0].y + p1.y + p2.y);
    v[1] = cadd(a, ta);
    v[4] = csub(a, ta);
    v[2] = cadd(b, tb);
    v[3] = csub(b, tb);
}

template <typename C> __device__ __forceinline__ void dft8(C* v, bool inv) {
    using R = decltype(C::x);
    const R h = (R)0.70710678118654752440;
    C e[4] = {v[0], v[2], v[4], v[6]};
    C o[4] = {v[1], v[3], v[5], v[7]};
    dft4(e, inv);
    dft4(o, inv);
    // o[k] *= w8^k, w8 = exp(-i pi/4) forward
    C o1 = inv ? cmk<C>(h * (o[1].x - o[1].y), h * (o[1].x + o[1].y))
               : cmk<C>(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));
    C o2 = cmi(o[2], inv);
    C o3 = inv ? cmk<C>(-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y))
               : cmk<C>(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y));
    v[0] = cadd(e[0], o[0]); v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);   v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);   v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);   v[7] = csub(e[3], o3);
}

template <int R, typename C> __device__ __forceinline__ void dftR(C* v, bool inv) {
    if constexpr (R == 2) dft2(v);
    else if constexpr (R == 3) dft3(v, inv);
    else if constexpr (R == 4) dft4(v, inv);
    else if constexpr (R == 5) dft5(v, inv);
    else if constexpr (R == 8) dft8(v, inv);
}

// One-dimensional transform plan: n = prod(rad[0..nst)), twiddle table
// tw[m] = exp(-2 pi i m/n) in device memory.
struct FftAxis {
    int n;
    int nst;
    int rad[kMaxStages];
    const double2* tw64;
    const float2* tw32;
};

template <typename T> __device__ __forceinline__ cplx<T> twiddle(const FftAxis& ax, int m);
template <> __device__ __forceinline__ double2 twiddle<double>(const FftAxis& ax, int m) { return ax.tw64[m]; }
template <> __device__ __forceinline__ float2 twiddle<float>(const FftAxis& ax, int m) { return ax.tw32[m]; }

// One Stockham stage (Govindaraju et al., SC'08 formulation) on nlines
// independent lines held in LDS: element (line t, index j) at
// lds[j*ls + t*lt].  Butterfly b of the stage covers line t and position j;
// LINE_FAST selects t = b % nlines (tiles whose lines are the fast LDS
// dimension) or t = b / nb.  PB = butterflies per thread (host-checked).
template <typename T, int R, int PB, bool LINE_FAST>
__device__ __forceinline__ void fft_stage(cplx<T>* lds, const FftAxis& ax, int Ns, int nlines, int ls, int lt,
                                          bool inv, int tid, int nthr) {
    using C = cplx<T>;
    const int n = ax.n;
    const int nb = n / R;
    const int total = nlines * nb;
    const int twstep = n / (Ns * R);
    C v[PB][R];
    int dst[PB];
#pragma unroll
    for (int p = 0; p < PB; ++p) {
        const int b = tid + p * nthr;
        dst[p] = -1;
        if (b < total) {
            const int t = LINE_FAST ? (b % nlines) : (b / nb);
            const int j = LINE_FAST ? (b / nlines) : (b % nb);
            const int k = j % Ns;
#pragma unroll
            for (int r = 0; r < R; ++r) v[p][r] = lds[(j + r * nb) * ls + t * lt];
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    C w = twiddle<T>(ax, r * k * twstep);
                    if (inv) w.y = -w.y;
                    v[p][r] = cmul(v[p][r], w);
                }
            }
            dftR<R>(v[p], inv);
            dst[p] = ((j / Ns) * Ns * R + k) * ls + t * lt;
        }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < PB; ++p) {
        if (dst[p] >= 0) {
#pragma unroll
            for (int r = 0; r < R; ++r) lds[dst[p] + r * Ns * ls] = v[p][r];
        }
    }
    __syncthreads();
}

// Full transform of nlines lines in LDS (forward: exp(-2 pi i ...), unnormalised).
template <typename T, int PB, bool LINE_FAST>
__device__ __forceinline__ void fft_lds(cplx<T>* lds, const FftAxis& ax, int nlines, int ls, int lt, bool inv,
                                        int tid, int nthr) {
    int Ns = 1;
    for (int s = 0; s < ax.nst; ++s) {
        const int R = ax.rad[s];
        switch (R) {
        case 2: fft_stage<T, 2, PB, LINE_FAST>(lds, ax, Ns, nlines, ls, lt, inv, tid, nthr); break;
        case 3: fft_stage<T, 3, PB, LINE_FAST>(lds, ax, Ns, nlines, ls, lt, inv, tid, nthr); break;
        case 4: fft_stage<T, 4, PB, LINE_FAST>(lds, ax, Ns, nlines, ls, lt, inv, tid, nthr); break;
        case 5: fft_stage<T, 5, PB, LINE_FAST>(lds, ax, Ns, nlines, ls, lt, inv, tid, nthr); break;
        default: fft_stage<T, 8, PB, LINE_FAST>(lds, ax, Ns, nlines, ls, lt, inv, tid, nthr); break;
        }
        Ns *= R;
    }
}

// --------------------------------------------------------------- reductions

template <typename T> __device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Block sum of NC values; result valid in thread 0.  scratch: >= NC*(nthr/64) doubles.
template <int NC>
__device__ __forceinline__ void block_sum(double (&v)[NC], double* scratch) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int c = 0; c < NC; ++c) v[c] = wave_sum(v[c]);
    if (lane == 0)
#pragma unroll
        for (int c = 0; c < NC; ++c) scratch[c * nw + wid] = v[c];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            double s = 0.0;
            for (int w = 0; w < nw; ++w) s += scratch[c * nw + w];
            v[c] = s;
        }
    }
    __syncthreads();
}

// ------------------------------------------------------------ impurity state

// Classical impurity (module classicimp, modules.f90:282-320), resident in
// device memory so that the whole step runs without host round trips.
struct ImpState {
    double r[3], v[3], a[3];
    double rold[3][3], vold[3][3], aold[2][3];  // [slot][xyz]
    double qr[3], qv[3], pcr[3], pcv[3];
    double rpot[3];  // position of the last potenimp (u_imp is evaluated there)
    double F[3];
    double errHe, errimp;
    int guard;       // set when den is not > 0 (steprk.f90:89-92)
    int pad;
};

// u_imp(x) = P[i](1-w) + P[i+1] w, r = |x - R|, i = int(r/D) from the
// rounded quotient, w = mod(r,D)/D from the exact remainder (potenimp.f90:
// 82-93).  Where r/D rounds up to an integer the two disagree and the
// reference takes w ~ 1 at the upper node; the same is done here: the exact
// r - i*D is negative exactly then, and mod adds one D back.  The index is
// clamped to the table.
__device__ __forceinline__ double uimp_at(const double* __restrict__ potion, int npot, double D,
                                          double xx, double yt) {
    const double r = sqrt(xx * xx + yt);
    int ir = (int)(r / D);
    ir = ir < 0 ? 0 : (ir > npot - 2 ? npot - 2 : ir);  // also catches a non-finite r
    double rem = fma(-(double)ir, D, r);
    if (rem < 0.0) rem += D;
    const double w = rem / D;
    return potion[ir] * (1.0 - w) + potion[ir + 1] * w;
}

}  // namespace hetd
