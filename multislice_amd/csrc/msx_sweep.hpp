// msx_sweep.hpp -- the two fused sweep kernels of the multislice slice step.
//
// A wave batch psi[b][ix][iy] (complex, iy contiguous: the layout of the
// reference's cfpix, cfpix.hpp:99-126) is advanced one slice by two sweeps
// over HBM, each reading and writing every pixel once:
//
//   row sweep (contiguous axis y, one line per (b,ix)):
//       [IFFT_y] -> x scale -> x t(ix,iy) -> [FFT_y]
//   col sweep (strided axis x, one line per (b,iy)):
//       [FFT_x] -> x P(kx,ky), band limit -> [detector sums] -> [IFFT_x]
//
// Chained over slices this is exactly the reference loop (mulslice.cpp:
// 706-739; ModelSimulated.cu:52-119): transmit, 2D FFT, propagate with the
// anti-aliasing band limit, inverse 2D FFT scaled by 1/(nx*ny) -- with the
// inverse y transform of slice n and the forward y transform of slice n+1
// done in the same row sweep, so a slice costs 2 HBM round trips instead of
// the 4-6 sweeps of separate transform / pointwise kernels.
//
// The kernels are templated on the complex type (float2: the FP32 engine;
// double2: the FP64 run of the same engine) and on the line length N; the
// FFT direction before (IN) and after (OUT) the pointwise stage is a
// compile-time mode.
#pragma once
#include "msx_fft.hpp"

namespace msx {

constexpr int kMaxDet = 8;

// ----------------------------------------------------------------------------
// Row sweep: lines along the contiguous axis (length N = ny).
// ----------------------------------------------------------------------------
template <typename C> struct RowArgs {
    using Re = typename real_of<C>::type;
    C* psi;                 // out (and in, unless src is set): [nb][nx][N]
    const C* src;           // optional input instead of psi
    long long src_bstride;  // elements between waves of src (0: one input shared by all waves)
    const C* tw;            // twiddles of length N
    const C* trans;         // transmission [nx][N] shared by all waves, or nullptr
    const Re* vreal;        // real potential [nb][nx][N] (replaces the input), or nullptr
    Re vscale;              // phase per unit of vreal (sigma/1000)
    Re scale;               // multiplies the data after the IN transform
    Re* inten;              // optional |psi|^2 output [nb][nx][N] (after the pointwise stage)
    int nx, nb;
};

// IN/OUT: 0 none, +1 forward (exp +i), -1 inverse (exp -i, unscaled)
template <typename C, int N, int IN, int OUT, int LPW>
__global__ __launch_bounds__(Plan<N>::T* LPW) void row_sweep(RowArgs<C> p) {
    using Re = typename real_of<C>::type;
    constexpr int T = Plan<N>::T, E = N / T;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* smem = reinterpret_cast<C*>(smem_raw);
    const int lw = threadIdx.x / T, t = threadIdx.x % T;
    const long long nlines = (long long)p.nx * p.nb;
    long long L = (long long)blockIdx.x * LPW + lw;
    const bool live = L < nlines;
    if (!live) L = nlines - 1;  // keep every thread in the barriers, store nothing
    // line order: waves fastest so consecutive work-groups reuse a t row
    const int ix = (int)(L / p.nb), b = (int)(L % p.nb);
    C* row = p.psi + ((long long)b * p.nx + ix) * N;
    LdsView<C> lds{smem + lw * (N + N / 16), 1, true};

    C a[E];
    if (p.vreal) {
        const Re* v = p.vreal + ((long long)b * p.nx + ix) * N;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            Re ph = p.vscale * v[t + e * T];
            Re s, c;
            if constexpr (sizeof(Re) == 4) sincosf(ph, &s, &c); else sincos(ph, &s, &c);
            a[e] = mk<C>(c, s);
        }
    } else {
        const C* in = p.src ? p.src + (long long)b * p.src_bstride + (long long)ix * N : row;
#pragma unroll
        for (int e = 0; e < E; ++e) a[e] = in[t + e * T];
    }
    if constexpr (IN != 0) fft_line<N, IN>(a, t, lds, p.tw);
    if (p.scale != (Re)1) {
#pragma unroll
        for (int e = 0; e < E; ++e) a[e] = cscale(a[e], p.scale);
    }
    if (p.trans) {
        const C* tr = p.trans + (long long)ix * N;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            // cfpix::operator*= (cfpix.cpp:296-318): (wr*tr - wi*ti, wr*ti + wi*tr)
            a[e] = cmul(a[e], tr[t + e * T]);
        }
    }
    if (p.inten && live) {
        Re* o = p.inten + ((long long)b * p.nx + ix) * N;
#pragma unroll
        for (int e = 0; e < E; ++e) o[t + e * T] = a[e].x * a[e].x + a[e].y * a[e].y;
    }
    if constexpr (OUT != 0) fft_line<N, OUT>(a, t, lds, p.tw);
    if (live) {
#pragma unroll
        for (int e = 0; e < E; ++e) row[t + e * T] = a[e];
    }
}

// ----------------------------------------------------------------------------
// Column sweep: lines along the strided axis (length N = nx), TW adjacent
// columns per work-group so each global access covers TW*sizeof(C) bytes.
// ----------------------------------------------------------------------------
enum ColOp : int {
    kColNone = 0,
    kColProp = 1,    // x py[iy] then x px[ix] inside the band limit, 0 outside
    kColMask = 2,    // band limit only
    kColDetect = 4,  // annular detector sums (after the propagator)
    kColGen = 8,     // generate a STEM probe in k-space (input ignored)
    kColCtf = 16,    // objective lens exp(-i chi) inside the aperture, 0 outside
    kColStore = 32,  // write the result back
};

template <typename C> struct ColArgs {
    using Re = typename real_of<C>::type;
    C* psi;  // [nb][N][ny]
    const C* tw;
    const float* kx2;  // float k^2 tables of freqn (slicelib.cpp:433-450), length N and ny
    const float* ky2;
    float k2max;  // band limit: keep kx2+ky2 < k2max (strict, as slicelib.cpp:755-778)
    const C* px;  // propagator factors of this layer, [N] and [ny]
    const C* py;
    // detectors: keep dlo[d] <= kx2+ky2 <= dhi[d]
    int ndet;
    float dlo[kMaxDet], dhi[kMaxDet];
    double* partial;  // [ndet][nb][ngroups] partial sums of |Psi|^2
    // probe / objective lens: phase(k) = -chi(k) + 2 pi (kx xp + ky yp), FP64
    const double2* pos;  // per-wave probe position (Angstrom), or nullptr
    double ax, by;       // cell size for k = index / a
    double chi1, chi2;   // chi = chi1*k2*(chi2*k2 - df)
    double df;
    float kap_lo, kap_hi;  // aperture: kap_lo <= kx2+ky2 <= kap_hi
    Re amp;                // probe amplitude
    int ny, nb, ops;
};

__device__ __forceinline__ int freq_index(int i, int n) {
    // slicelib.cpp:433-450: imid = (int)(n/2.0 + 0.5); negative for i > imid
    const int imid = (int)(n / 2.0 + 0.5);
    return i > imid ? i - n : i;
}

// exp(i*phase) with the FP64 phase reduced to [-pi, pi] before the
// single-precision evaluation.
template <typename C> __device__ __forceinline__ C expi_reduced(double ph) {
    using Re = typename real_of<C>::type;
    const double twopi = 6.283185307179586476925287;
    ph -= twopi * rint(ph * (1.0 / twopi));
    Re s, c;
    if constexpr (sizeof(Re) == 4) sincosf((float)ph, &s, &c); else sincos(ph, &s, &c);
    return mk<C>(c, s);
}

template <typename C, int N, int IN, int OUT, int TW>
__global__ __launch_bounds__(Plan<N>::T* TW) void col_sweep(ColArgs<C> p) {
    using Re = typename real_of<C>::type;
    constexpr int T = Plan<N>::T, E = N / T;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* smem = reinterpret_cast<C*>(smem_raw);
    const int lw = threadIdx.x % TW, t = threadIdx.x / TW;
    const int groups = (p.ny + TW - 1) / TW;
    const int b = blockIdx.x / groups, g = blockIdx.x % groups;
    int iy = g * TW + lw;
    const bool live = iy < p.ny;
    if (!live) iy = p.ny - 1;
    C* col = p.psi + (long long)b * N * p.ny + iy;
    LdsView<C> lds{smem + lw, TW + 1, false};

    C a[E];
    if (p.ops & kColGen) {
        const double2 pos = p.pos[b];
        const double ky = (double)freq_index(iy, p.ny) / p.by;
        const float ky2 = p.ky2[iy];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int ix = t + e * T;
            const float k2f = p.kx2[ix] + ky2;
            if (k2f <= p.kap_hi && k2f >= p.kap_lo) {
                const double kx = (double)freq_index(ix, N) / p.ax;
                const double k2 = kx * kx + ky * ky;
                const double ph = -p.chi1 * k2 * (p.chi2 * k2 - p.df) +
                                  6.283185307179586476925287 * (kx * pos.x + ky * pos.y);
                a[e] = cscale(expi_reduced<C>(ph), p.amp);
            } else {
                a[e] = mk<C>((Re)0, (Re)0);
            }
        }
    } else {
#pragma unroll
        for (int e = 0; e < E; ++e) a[e] = col[(long long)(t + e * T) * p.ny];
    }
    if constexpr (IN != 0) fft_line<N, IN>(a, t, lds, p.tw);

    if (p.ops & (kColProp | kColMask | kColDetect | kColCtf)) {
        const float ky2 = p.ky2[iy];
        C py = mk<C>((Re)1, (Re)0);
        if (p.ops & kColProp) py = p.py[iy];
        float acc[kMaxDet], cmp[kMaxDet];
#pragma unroll
        for (int d = 0; d < kMaxDet; ++d) acc[d] = cmp[d] = 0.f;
        const double ky = (double)freq_index(iy, p.ny) / p.by;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int ix = t + e * T;
            const float k2 = p.kx2[ix] + ky2;
            if (p.ops & (kColProp | kColMask)) {
                if (k2 < p.k2max) {
                    // slicelib.cpp:765-770: (w * py) * px
                    if (p.ops & kColProp) a[e] = cmul(cmul(a[e], py), p.px[ix]);
                } else {
                    a[e] = mk<C>((Re)0, (Re)0);
                }
            }
            if (p.ops & kColCtf) {
                if (k2 <= p.kap_hi) {
                    const double kx = (double)freq_index(ix, N) / p.ax;
                    const double kk = kx * kx + ky * ky;
                    a[e] = cmul(a[e], expi_reduced<C>(-p.chi1 * kk * (p.chi2 * kk - p.df)));
                } else {
                    a[e] = mk<C>((Re)0, (Re)0);
                }
            }
            if (p.ops & kColDetect) {
                const float v = (float)(a[e].x * a[e].x + a[e].y * a[e].y);
#pragma unroll
                for (int d = 0; d < kMaxDet; ++d) {
                    if (d < p.ndet && k2 >= p.dlo[d] && k2 <= p.dhi[d]) {
                        // Kahan-compensated per-thread accumulation
                        const float y = v - cmp[d];
                        const float s = acc[d] + y;
                        cmp[d] = (s - acc[d]) - y;
                        acc[d] = s;
                    }
                }
            }
        }
        if (p.ops & kColDetect) {
            // deterministic work-group reduction in FP64: wavefront butterfly,
            // then the per-wave results in fixed order through LDS.
            __syncthreads();  // LDS may still be read by the last FFT pass
            double* red = reinterpret_cast<double*>(smem_raw);
            const int nwaves = (T * TW + 63) / 64;
            const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
            for (int d = 0; d < p.ndet; ++d) {
                double v = live ? (double)acc[d] - (double)cmp[d] : 0.0;
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
                if (lane == 0) red[d * 32 + wv] = v;
            }
            __syncthreads();
            if (threadIdx.x < p.ndet) {
                const int d = threadIdx.x;
                double s = 0.0;
                for (int w = 0; w < nwaves; ++w) s += red[d * 32 + w];
                p.partial[((long long)d * p.nb + b) * groups + g] = s;
            }
            __syncthreads();
        }
    }
    if constexpr (OUT != 0) fft_line<N, OUT>(a, t, lds, p.tw);
    if ((p.ops & kColStore) && live) {
#pragma unroll
        for (int e = 0; e < E; ++e) col[(long long)(t + e * T) * p.ny] = a[e];
    }
}

}  // namespace msx
