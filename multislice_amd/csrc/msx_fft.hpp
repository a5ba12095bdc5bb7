// msx_fft.hpp -- register/LDS building blocks of the hand-written Stockham FFT
// used by every sweep kernel of the msx engine (gfx950, 64-wide wavefronts).
//
// A "line" is one 1D transform of length N handled by T cooperating threads.
// Thread t of a line keeps E = N/T complex values in registers, a[e] holding
// element t + e*T.  That layout is what every pass of a Stockham autosort
// radix-R pass reads, and what the LAST pass writes, so:
//   * a line is loaded from / stored to global memory with unit-stride,
//     coalesced accesses across t (or across lines for a strided axis);
//   * an FFT that ends in registers can feed a pointwise multiply and a second
//     FFT (the opposite transform) with no LDS round trip in between.
// Only the P-1 internal exchanges of a P-pass transform go through LDS.
//
// Pass p (radix R, Ns = product of earlier radices) handles butterflies
// j = t + q*T, q < Q = N/(R*T):
//     v[r] = x[j + r*N/R] * W_{Ns*R}^{r*(j mod Ns)}         (r = 0..R-1)
//     v    = DFT_R(v)
//     y[(j div Ns)*Ns*R + (j mod Ns) + r*Ns] = v[r]
// (Stockham autosort formulation; natural order in and out.)
//
// Sign convention: S = +1 is the transform with exponent +2*pi*i*jk/N.
#pragma once
#include <hip/hip_runtime.h>

namespace msx {

template <typename C> struct real_of;
template <> struct real_of<float2> { using type = float; };
template <> struct real_of<double2> { using type = double; };

template <typename C>
__device__ __forceinline__ C mk(typename real_of<C>::type x, typename real_of<C>::type y) {
    C c;
    c.x = x;
    c.y = y;
    return c;
}
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return mk<C>(a.x + b.x, a.y + b.y); }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return mk<C>(a.x - b.x, a.y - b.y); }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
    return mk<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename C> __device__ __forceinline__ C cscale(C a, typename real_of<C>::type s) {
    return mk<C>(a.x * s, a.y * s);
}
// a * (S*i)
template <int S, typename C> __device__ __forceinline__ C mul_si(C a) {
    return (S > 0) ? mk<C>(-a.y, a.x) : mk<C>(a.y, -a.x);
}

// cos/sin(2*pi*j/48): every root of unity needed by radices 2,3,4,6,8,12,16.
struct Roots48 {
    static constexpr double c[48] = {
        1.0, 0.9914448613738104111445575, 0.9659258262890682867497432, 0.9238795325112867561281832,
        0.8660254037844386467637232, 0.793353340291235164579777, 0.7071067811865475244008444,
        0.6087614290087206394160975, 0.5, 0.38268343236508977172846, 0.2588190451025207623488988,
        0.1305261922200515915484062, 0.0, -0.1305261922200515915484062, -0.2588190451025207623488988,
        -0.38268343236508977172846, -0.5, -0.6087614290087206394160975, -0.7071067811865475244008444,
        -0.793353340291235164579777, -0.8660254037844386467637232, -0.9238795325112867561281832,
        -0.9659258262890682867497432, -0.9914448613738104111445575, -1.0, -0.9914448613738104111445575,
        -0.9659258262890682867497432, -0.9238795325112867561281832, -0.8660254037844386467637232,
        -0.793353340291235164579777, -0.7071067811865475244008444, -0.6087614290087206394160975, -0.5,
        -0.38268343236508977172846, -0.2588190451025207623488988, -0.1305261922200515915484062, 0.0,
        0.1305261922200515915484062, 0.2588190451025207623488988, 0.38268343236508977172846, 0.5,
        0.6087614290087206394160975, 0.7071067811865475244008444, 0.793353340291235164579777,
        0.8660254037844386467637232, 0.9238795325112867561281832, 0.9659258262890682867497432,
        0.9914448613738104111445575};
    static constexpr double s(int j) { return c[(j + 36) % 48]; }  // sin(x) = cos(x - pi/2)
};

// omega_R^m = exp(S*2*pi*i*m/R) for R dividing 48
template <int R, int S, typename C> __device__ __forceinline__ C root(int m) {
    static_assert(48 % R == 0, "radix must divide 48");
    using Re = typename real_of<C>::type;
    const int j = ((m % R) * (48 / R)) % 48;
    return mk<C>((Re)Roots48::c[j], (Re)(S * Roots48::s(j)));
}
template <typename C> __device__ __forceinline__ bool is_one(int j) { return j == 0; }

// ---------------------------------------------------------------- small DFTs
// In-place DFT of R values in registers, natural order in and out.
template <int R, int S, typename C> struct Dft;

template <int S, typename C> struct Dft<1, S, C> {
    static __device__ __forceinline__ void run(C*) {}
};
template <int S, typename C> struct Dft<2, S, C> {
    static __device__ __forceinline__ void run(C* a) {
        C t = a[0];
        a[0] = cadd(t, a[1]);
        a[1] = csub(t, a[1]);
    }
};
template <int S, typename C> struct Dft<3, S, C> {
    static __device__ __forceinline__ void run(C* a) {
        using Re = typename real_of<C>::type;
        const Re h = (Re)0.8660254037844386467637232;  // sqrt(3)/2
        C s = cadd(a[1], a[2]);
        C d = csub(a[1], a[2]);
        C m = mk<C>(a[0].x - (Re)0.5 * s.x, a[0].y - (Re)0.5 * s.y);
        C r = mul_si<S>(cscale(d, h));
        a[0] = cadd(a[0], s);
        a[1] = cadd(m, r);
        a[2] = csub(m, r);
    }
};
template <int S, typename C> struct Dft<4, S, C> {
    static __device__ __forceinline__ void run(C* a) {
        C t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
        C t2 = cadd(a[1], a[3]), t3 = mul_si<S>(csub(a[1], a[3]));
        a[0] = cadd(t0, t2);
        a[2] = csub(t0, t2);
        a[1] = cadd(t1, t3);
        a[3] = csub(t1, t3);
    }
};

// Cooley-Tukey split R = A*B inside registers: n = B*n1 + n2, k = k1 + A*k2.
template <int A, int B, int S, typename C> struct DftCT {
    static __device__ __forceinline__ void run(C* x) {
        constexpr int R = A * B;
        C y[R];
#pragma unroll
        for (int n2 = 0; n2 < B; ++n2) {
            C u[A];
#pragma unroll
            for (int n1 = 0; n1 < A; ++n1) u[n1] = x[B * n1 + n2];
            Dft<A, S, C>::run(u);
#pragma unroll
            for (int k1 = 0; k1 < A; ++k1) {
                const int m = n2 * k1;
                y[n2 * A + k1] = (m == 0) ? u[k1] : cmul(u[k1], root<R, S, C>(m));
            }
        }
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) {
            C v[B];
#pragma unroll
            for (int n2 = 0; n2 < B; ++n2) v[n2] = y[n2 * A + k1];
            Dft<B, S, C>::run(v);
#pragma unroll
            for (int k2 = 0; k2 < B; ++k2) x[k1 + A * k2] = v[k2];
        }
    }
};
template <int S, typename C> struct Dft<6, S, C> : DftCT<2, 3, S, C> {};
template <int S, typename C> struct Dft<8, S, C> : DftCT<2, 4, S, C> {};
template <int S, typename C> struct Dft<12, S, C> : DftCT<3, 4, S, C> {};
template <int S, typename C> struct Dft<16, S, C> : DftCT<4, 4, S, C> {};

// ---------------------------------------------------------------- plans
// Radix sequence and threads per line for each supported length.
template <int N> struct Plan;
#define MSX_PLAN(NN, TT, ...)                                  \
    template <> struct Plan<NN> {                              \
        static constexpr int N = NN, T = TT;                   \
        static constexpr int R[] = {__VA_ARGS__};              \
        static constexpr int P = sizeof(R) / sizeof(R[0]);     \
    };
MSX_PLAN(16, 4, 4, 4)
MSX_PLAN(32, 4, 4, 8)
MSX_PLAN(64, 8, 8, 8)
MSX_PLAN(128, 8, 16, 8)
MSX_PLAN(256, 16, 16, 16)
MSX_PLAN(384, 16, 6, 8, 8)
MSX_PLAN(512, 64, 8, 8, 8)
MSX_PLAN(768, 32, 12, 8, 8)
MSX_PLAN(1024, 64, 16, 8, 8)
MSX_PLAN(2048, 128, 16, 16, 8)
#undef MSX_PLAN

// ---------------------------------------------------------------- LDS views
// Element m of the line owned by this thread, in LDS.
//  Contiguous lines: line-major, one padding slot every 16 elements.
//  Strided lines   : element-major with the WG's lines adjacent (+1 pad per row).
template <typename C> struct LdsView {
    C* base;
    int stride;  // distance between consecutive elements (in C units)
    bool padded;
    __device__ __forceinline__ C* at(int m) const { return base + (padded ? (m + (m >> 4)) : m * stride); }
};

// ---------------------------------------------------------------- passes
// Twiddle table tw[m] = exp(+2*pi*i*m/N) (m < N), generated in FP64 on the
// host; the S = -1 transform uses the conjugate.
template <int N, int S, typename C> __device__ __forceinline__ C twiddle(const C* __restrict__ tw, int m) {
    C w = tw[m];
    if (S < 0) w.y = -w.y;
    return w;
}

template <int N, int T, int R, int Ns, int S, bool LAST, typename C>
__device__ __forceinline__ void stockham_pass(C (&a)[N / T], int t, const LdsView<C>& lds,
                                              const C* __restrict__ tw) {
    constexpr int Q = N / (R * T);
    static_assert(Q * R * T == N, "radix/thread split must tile N");
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        C b[R];
#pragma unroll
        for (int r = 0; r < R; ++r) b[r] = a[q + r * Q];
        const int j = t + q * T;
        const int k = j % Ns;
        if (Ns > 1) {
            constexpr int step = N / (Ns * R);
#pragma unroll
            for (int r = 1; r < R; ++r) b[r] = cmul(b[r], twiddle<N, S>(tw, r * k * step));
        }
        Dft<R, S, C>::run(b);
        if (LAST) {
#pragma unroll
            for (int r = 0; r < R; ++r) a[q + r * Q] = b[r];
        } else {
            const int base = (j / Ns) * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) *lds.at(base + r * Ns) = b[r];
        }
    }
    if (!LAST) {
        __syncthreads();
#pragma unroll
        for (int e = 0; e < N / T; ++e) a[e] = *lds.at(t + e * T);
        __syncthreads();
    }
}

template <class PL, int I, int Ns, int S, typename C>
__device__ __forceinline__ void run_passes(C (&a)[PL::N / PL::T], int t, const LdsView<C>& lds,
                                           const C* __restrict__ tw) {
    if constexpr (I < PL::P) {
        constexpr int R = PL::R[I];
        stockham_pass<PL::N, PL::T, R, Ns, S, I == PL::P - 1>(a, t, lds, tw);
        run_passes<PL, I + 1, Ns * R, S>(a, t, lds, tw);
    }
}

// Full length-N transform of the registers a[] (unscaled), sign S.
template <int N, int S, typename C>
__device__ __forceinline__ void fft_line(C (&a)[N / Plan<N>::T], int t, const LdsView<C>& lds,
                                         const C* __restrict__ tw) {
    run_passes<Plan<N>, 0, 1, S>(a, t, lds, tw);
}

}  // namespace msx
