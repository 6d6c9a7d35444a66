// fft.hip -- batched 3-D real FFT passes for gfx950 (Stockham in LDS).
//
// The reference calls FFTW 3-D r2c/c2r plans bound to fixed arrays
// (fft.f90:50-64, 81-153): forward unnormalised exp(-2 pi i k.x/n) to the
// (nx/2+1, ny, nz) half spectrum, backward followed by *renor = 1/(nx ny nz).
// Here a 3-D transform is three passes over HBM, one per axis:
//   x: two real lines are packed into one complex line (a + i b), one
//      complex FFT of length nx, then split A_k = (Z_k + conj Z_{n-k})/2,
//      B_k = (Z_k - conj Z_{n-k})/2i (any nx, odd included); the inverse
//      packs A + iB with the Hermitian extension along x, imaginary parts of
//      k = 0 and k = nx/2 dropped (the one-dimensional c2r definition).
//      Pointwise work before / after the x pass is fused into it (modes).
//   y, z: tiles of TX consecutive kx columns, full line length in LDS, so
//      every global access is a TX*16-byte contiguous run.
//   z-conv: forward z FFT, multiply by up to three real form factors
//      (W, V_LJ, K_alpha), inverse z FFT of each product -- the spectral
//      multiply of poten.f90:34-36,46-48,69-71 never leaves LDS.
#include "hetd_kernels.h"

namespace hetd {

// ------------------------------------------------------------------ x pass

template <typename T, int PB>
__global__ void __launch_bounds__(256) k_xpass(XPassArgs<T> a) {
    using C = cplx<T>;
    extern __shared__ unsigned char smem_raw[];
    C* lds = reinterpret_cast<C*>(smem_raw);
    const int nx = a.ax.n, nxh = nx / 2 + 1;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const long nlines = a.nlines;
    const long pair0 = (long)blockIdx.x * a.ppb;
    int npairs = a.ppb;
    if ((pair0 + npairs) * 2 > nlines + 1) npairs = (int)((nlines + 1) / 2 - pair0);
    const bool inv_first = (a.mode != XMODE_R2C);

    // ---- load
    if (!inv_first) {
        for (int idx = tid; idx < npairs * nx; idx += nthr) {
            const int pp = idx / nx, x = idx % nx;
            const long la = 2 * (pair0 + pp), lb = la + 1;
            T va = a.rin[la * nx + x];
            T vb = lb < nlines ? a.rin[lb * nx + x] : (T)0;
            lds[pp * nx + x] = cmk<C>(va, vb);
        }
    } else {
        for (int idx = tid; idx < npairs * nx; idx += nthr) {
            const int pp = idx / nx, k = idx % nx;
            const long la = 2 * (pair0 + pp), lb = la + 1;
            const int kk = k < nxh ? k : nx - k;
            C A = a.cin[la * a.ldh + kk];
            C B = lb < nlines ? a.cin[lb * a.ldh + kk] : cmk<C>((T)0, (T)0);
            if (kk == 0 || (2 * kk == nx)) { A.y = (T)0; B.y = (T)0; }
            if (k != kk) { A.y = -A.y; B.y = -B.y; }
            lds[pp * nx + k] = cmk<C>(A.x - B.y, A.y + B.x);  // A + iB
        }
    }
    __syncthreads();

    if (inv_first) {
        fft_lds<T, PB, false>(lds, a.ax, npairs, 1, nx, true, tid, nthr);
        // ---- pointwise stage on the two real lines (re = line a, im = line b)
        for (int idx = tid; idx < npairs * nx; idx += nthr) {
            const int pp = idx / nx, x = idx % nx;
            const long la = 2 * (pair0 + pp), lb = la + 1;
            const bool hb = lb < nlines;
            C z = lds[pp * nx + x];
            T va = z.x * a.scale, vb = z.y * a.scale;
            const long ia = la * nx + x, ib = lb * nx + x;
            if (a.mode == XMODE_C2R) {
                a.rout[ia] = va;
                if (hb) a.rout[ib] = vb;
            } else if (a.mode == XMODE_CG_S) {
                // dencg (or the folded local part), then the correlation source
                const T sa = cgs_point<T>(a, ia, va);
                const T sb = hb ? cgs_point<T>(a, ib, vb) : (T)0;
                lds[pp * nx + x] = cmk<C>(sa, sb);
            } else {  // XMODE_POT
                pot_point<T>(a, ia, va);
                if (hb) pot_point<T>(a, ib, vb);
            }
        }
        __syncthreads();
        if (a.mode != XMODE_CG_S) return;
    }

    // ---- forward transform of the packed pair and split into two half lines
    fft_lds<T, PB, false>(lds, a.ax, npairs, 1, nx, false, tid, nthr);
    for (int idx = tid; idx < npairs * nxh; idx += nthr) {
        const int pp = idx / nxh, k = idx % nxh;
        const long la = 2 * (pair0 + pp), lb = la + 1;
        const C zk = lds[pp * nx + k];
        const C zn = lds[pp * nx + (k == 0 ? 0 : nx - k)];
        const T h = (T)0.5;
        C A = cmk<C>(h * (zk.x + zn.x), h * (zk.y - zn.y));
        C B = cmk<C>(h * (zk.y + zn.y), -h * (zk.x - zn.x));
        a.cout[la * a.ldh + k] = A;
        if (lb < nlines) a.cout[lb * a.ldh + k] = B;
    }
}

// ---------------------------------------------------------------- y pass

template <typename T, int PB>
__global__ void __launch_bounds__(256) k_ypass(YPassArgs<T> a) {
    using C = cplx<T>;
    extern __shared__ unsigned char smem_raw[];
    C* lds = reinterpret_cast<C*>(smem_raw);
    const int ny = a.ax.n, TX = a.tx, nxh = a.nxh, ldh = a.ldh;
    const int kx0 = blockIdx.x * TX, z = blockIdx.y;
    C* buf = a.buf[blockIdx.z];
    const int tid = threadIdx.x, nthr = blockDim.x;
    const long base = (long)ldh * ny * z;
    for (int idx = tid; idx < ny * TX; idx += nthr) {
        const int j = idx / TX, t = idx % TX, kx = kx0 + t;
        lds[idx] = kx < nxh ? buf[base + (long)ldh * j + kx] : cmk<C>((T)0, (T)0);
    }
    __syncthreads();
    fft_lds<T, PB, true>(lds, a.ax, TX, TX, 1, a.inv != 0, tid, nthr);
    for (int idx = tid; idx < ny * TX; idx += nthr) {
        const int j = idx / TX, t = idx % TX, kx = kx0 + t;
        if (kx < nxh) buf[base + (long)ldh * j + kx] = lds[idx];
    }
}

// ---------------------------------------------------------------- z pass

template <typename T, int PB>
__global__ void __launch_bounds__(256) k_zpass(ZPassArgs<T> a) {
    using C = cplx<T>;
    extern __shared__ unsigned char smem_raw[];
    C* l0 = reinterpret_cast<C*>(smem_raw);
    const int nz = a.ax.n, TX = a.tx, nxh = a.nxh, ny = a.ny;
    const int kx0 = blockIdx.x * TX, ky = blockIdx.y;
    C* l1 = l0 + nz * TX;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const long zs = (long)a.ldh * ny;
    const long base = (long)a.ldh * ky + kx0;
    const C* in = a.in[blockIdx.z];
    for (int idx = tid; idx < nz * TX; idx += nthr) {
        const int j = idx / TX, t = idx % TX;
        l0[idx] = (kx0 + t) < nxh ? in[base + zs * j + t] : cmk<C>((T)0, (T)0);
    }
    __syncthreads();
    if (a.mode == ZMODE_INV) {
        fft_lds<T, PB, true>(l0, a.ax, TX, TX, 1, true, tid, nthr);
        C* out = a.out[blockIdx.z][0];
        for (int idx = tid; idx < nz * TX; idx += nthr) {
            const int j = idx / TX, t = idx % TX;
            if (kx0 + t < nxh) out[base + zs * j + t] = l0[idx];
        }
        return;
    }
    fft_lds<T, PB, true>(l0, a.ax, TX, TX, 1, false, tid, nthr);
    if (a.mode == ZMODE_FWD || a.fwd_out) {
        C* out = a.mode == ZMODE_FWD ? a.out[blockIdx.z][0] : a.fwd_out;
        for (int idx = tid; idx < nz * TX; idx += nthr) {
            const int j = idx / TX, t = idx % TX;
            if (kx0 + t < nxh) out[base + zs * j + t] = l0[idx];
        }
        if (a.mode == ZMODE_FWD) return;
    }
    if (a.kout)
        for (int idx = tid; idx < nz * TX; idx += nthr) {
            const int j = idx / TX, t = idx % TX;
            if (kx0 + t < nxh) a.kout[base + zs * j + t] = cscale(l0[idx], a.kkern[base + zs * j + t]);
        }
    for (int q = 0; q < a.nk; ++q) {
        const T* K = a.kern[q];
        for (int idx = tid; idx < nz * TX; idx += nthr) {
            const int j = idx / TX, t = idx % TX;
            const bool in_ = (kx0 + t) < nxh;
            const T kv = in_ ? K[base + zs * j + t] : (T)0;
            C w = cscale(l0[idx], kv);
            if (q == 0 && a.addin && in_) w = cadd(w, a.addin[base + zs * j + t]);
            l1[idx] = w;
        }
        __syncthreads();
        fft_lds<T, PB, true>(l1, a.ax, TX, TX, 1, true, tid, nthr);
        C* out = a.out[blockIdx.z][q];
        for (int idx = tid; idx < nz * TX; idx += nthr) {
            const int j = idx / TX, t = idx % TX;
            if (kx0 + t < nxh) out[base + zs * j + t] = l1[idx];
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------- launchers

static int pb_for(const FftAxis& ax, int nlines, int nthr) {
    int pb = 1;
    for (int s = 0; s < ax.nst; ++s) {
        const int tot = nlines * (ax.n / ax.rad[s]);
        const int need = (tot + nthr - 1) / nthr;
        if (need > pb) pb = need;
    }
    if (pb <= 1) return 1;
    if (pb <= 2) return 2;
    if (pb <= 4) return 4;
    if (pb <= 8) return 8;
    return -1;
}

// Columns per y/z block: 8 (128-byte runs in FP64), fewer where a long line
// would need more than 8 butterflies per thread (n = 750, 810, 864, 900, 960,
// 972) or more LDS than a gfx950 workgroup holds (the FP64 z convolution,
// two lines in LDS, above n = 640).  0: no width fits.
static constexpr size_t kMaxLdsBytes = 160 * 1024;
static int stockham_tx(const FftAxis& ax, int nl, size_t elt, int nthr) {
    for (int tx = 8; tx >= 1; tx /= 2)
        if (pb_for(ax, tx, nthr) > 0 && (size_t)nl * ax.n * tx * elt <= kMaxLdsBytes) return tx;
    return 0;
}

// Line pairs per x block: about 2048 complex points, fewer where that would
// take more than 4 butterflies per thread (lines with a radix-2 stage: nx =
// 6, 10, 18, 54, 250, 270, 450, 486, 750, 810); the FP64 kernel built for 8
// spills registers.
static int stockham_ppb(const FftAxis& ax, long npairs, int nthr) {
    int ppb = (2048 + ax.n - 1) / ax.n;
    if (ppb < 1) ppb = 1;
    if (ppb > npairs) ppb = (int)npairs;
    while (ppb > 1 && pb_for(ax, ppb, nthr) > 4) --ppb;
    return ppb;
}

bool stockham_plan_ok(const FftAxis& ax) {
    const int nthr = 256;
    const int ppb = stockham_ppb(ax, 1L << 30, nthr);
    const bool x_ok = pb_for(ax, ppb, nthr) > 0 && (size_t)ppb * ax.n * sizeof(cplx<double>) <= kMaxLdsBytes;
    return x_ok && stockham_tx(ax, 2, sizeof(cplx<double>), nthr) > 0;
}

#define HETD_PB_SWITCH(PB, KERN, T, ...)                                                   \
    switch (PB) {                                                                          \
    case 1: hipLaunchKernelGGL((KERN<T, 1>), __VA_ARGS__); break;                          \
    case 2: hipLaunchKernelGGL((KERN<T, 2>), __VA_ARGS__); break;                          \
    case 4: hipLaunchKernelGGL((KERN<T, 4>), __VA_ARGS__); break;                          \
    case 8: hipLaunchKernelGGL((KERN<T, 8>), __VA_ARGS__); break;                          \
    default: return hipErrorInvalidValue;                                                  \
    }

template <typename T>
hipError_t launch_xpass(XPassArgs<T> a, hipStream_t st) {
    const int nthr = 256;
    const int nx = a.ax.n;
    const long npairs = (a.nlines + 1) / 2;
    const int ppb = stockham_ppb(a.ax, npairs, nthr);
    a.ppb = ppb;
    const int pb = pb_for(a.ax, ppb, nthr);
    const size_t lds = (size_t)ppb * nx * sizeof(cplx<T>);
    dim3 grid((unsigned)((npairs + ppb - 1) / ppb));
    HETD_PB_SWITCH(pb, k_xpass, T, grid, dim3(nthr), lds, st, a);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_ypass(YPassArgs<T> a, int nz, int nbuf, hipStream_t st) {
    const int nthr = 256;
    a.tx = stockham_tx(a.ax, 1, sizeof(cplx<T>), nthr);
    if (a.tx == 0) return hipErrorInvalidValue;
    const int pb = pb_for(a.ax, a.tx, nthr);
    const size_t lds = (size_t)a.ax.n * a.tx * sizeof(cplx<T>);
    dim3 grid((a.nxh + a.tx - 1) / a.tx, nz, nbuf);
    HETD_PB_SWITCH(pb, k_ypass, T, grid, dim3(nthr), lds, st, a);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_zpass(ZPassArgs<T> a, int nbuf, hipStream_t st) {
    const int nthr = 256;
    const int nl = (a.mode == ZMODE_CONV) ? 2 : 1;
    a.tx = stockham_tx(a.ax, nl, sizeof(cplx<T>), nthr);
    if (a.tx == 0) return hipErrorInvalidValue;
    const int pb = pb_for(a.ax, a.tx, nthr);
    const size_t lds = (size_t)nl * a.ax.n * a.tx * sizeof(cplx<T>);
    dim3 grid((a.nxh + a.tx - 1) / a.tx, a.ny, nbuf);
    HETD_PB_SWITCH(pb, k_zpass, T, grid, dim3(nthr), lds, st, a);
    return hipGetLastError();
}

template hipError_t launch_xpass<double>(XPassArgs<double>, hipStream_t);
template hipError_t launch_xpass<float>(XPassArgs<float>, hipStream_t);
template hipError_t launch_ypass<double>(YPassArgs<double>, int, int, hipStream_t);
template hipError_t launch_ypass<float>(YPassArgs<float>, int, int, hipStream_t);
template hipError_t launch_zpass<double>(ZPassArgs<double>, int, hipStream_t);
template hipError_t launch_zpass<float>(ZPassArgs<float>, int, hipStream_t);

}  // namespace hetd
