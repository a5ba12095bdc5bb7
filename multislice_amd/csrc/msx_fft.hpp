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
__device__ __forceinline__ float fused(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fused(double a, double b, double c) { return __builtin_fma(a, b, c); }
// (a.x b.x - a.y b.y, a.x b.y + a.y b.x) with the fused multiply-adds spelled
// out: left to the compiler, which product it fuses can differ between
// kernels that inline the same code (and so can the last bit of the result)
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
    return mk<C>(fused(a.x, b.x, -(a.y * b.y)), fused(a.x, b.y, a.y * b.x));
}
// |a|^2 with the same pinned rounding
template <typename C> __device__ __forceinline__ typename real_of<C>::type cnorm2(C a) {
    return fused(a.x, a.x, a.y * a.y);
}
template <typename C> __device__ __forceinline__ C cscale(C a, typename real_of<C>::type s) {
    return mk<C>(a.x * s, a.y * s);
}
// x if keep else 0, as a bitwise AND on the loaded bits: the load stays
// unconditional (a select lets the compiler turn it into a branch-guarded
// load, whose control-flow join then forces a wait on every outstanding load)
__device__ __forceinline__ float2 keep_or_zero(float2 x, unsigned keep) {
    return make_float2(__uint_as_float(__float_as_uint(x.x) & keep), __uint_as_float(__float_as_uint(x.y) & keep));
}
__device__ __forceinline__ double2 keep_or_zero(double2 x, unsigned keep) {
    const unsigned long long k = keep ? ~0ull : 0ull;
    return make_double2(__longlong_as_double(__double_as_longlong(x.x) & k),
                        __longlong_as_double(__double_as_longlong(x.y) & k));
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

// The same roots as float pairs chosen for their modulus: for each angle the
// (cos, sin) pair among the float neighbours of the rounded values whose
// c^2 + s^2 is closest to 1 (then the nearest to the exact root).  Rounding
// cos and sin separately leaves |root|^2 - 1 of up to ~1e-7 with a sign fixed
// per root; the FP32 slice loop applies the same roots to the same strong
// beams every slice, so those errors add up as a drift of the wave's norm
// (about 1e-5 over 100 slices).  These pairs keep |root|^2 - 1 below 4.3e-8
// and their angles within 8.6e-8 rad.
struct RootsF48 {
    static constexpr float cs[48][2] = {
        {1.000000000e+00f, 0.000000000e+00f}, {9.914448857e-01f, 1.305261701e-01f}, {9.659258127e-01f, 2.588190734e-01f},
        {9.238795638e-01f, 3.826833963e-01f}, {8.660253882e-01f, 5.000000000e-01f}, {7.933533788e-01f, 6.087613702e-01f},
        {7.071068287e-01f, 7.071067095e-01f}, {6.087613702e-01f, 7.933533788e-01f}, {5.000000000e-01f, 8.660253882e-01f},
        {3.826833963e-01f, 9.238795638e-01f}, {2.588190734e-01f, 9.659258127e-01f}, {1.305261701e-01f, 9.914448857e-01f},
        {0.000000000e+00f, 1.000000000e+00f}, {-1.305261701e-01f, 9.914448857e-01f}, {-2.588190734e-01f, 9.659258127e-01f},
        {-3.826833963e-01f, 9.238795638e-01f}, {-5.000000000e-01f, 8.660253882e-01f}, {-6.087613702e-01f, 7.933533788e-01f},
        {-7.071067095e-01f, 7.071068287e-01f}, {-7.933533788e-01f, 6.087613702e-01f}, {-8.660253882e-01f, 5.000000000e-01f},
        {-9.238795638e-01f, 3.826833963e-01f}, {-9.659258127e-01f, 2.588190734e-01f}, {-9.914448857e-01f, 1.305261701e-01f},
        {-1.000000000e+00f, 0.000000000e+00f}, {-9.914448857e-01f, -1.305261701e-01f}, {-9.659258127e-01f, -2.588190734e-01f},
        {-9.238795638e-01f, -3.826833963e-01f}, {-8.660253882e-01f, -5.000000000e-01f}, {-7.933533788e-01f, -6.087613702e-01f},
        {-7.071068287e-01f, -7.071067095e-01f}, {-6.087613702e-01f, -7.933533788e-01f}, {-5.000000000e-01f, -8.660253882e-01f},
        {-3.826833963e-01f, -9.238795638e-01f}, {-2.588190734e-01f, -9.659258127e-01f}, {-1.305261701e-01f, -9.914448857e-01f},
        {0.000000000e+00f, -1.000000000e+00f}, {1.305261701e-01f, -9.914448857e-01f}, {2.588190734e-01f, -9.659258127e-01f},
        {3.826833963e-01f, -9.238795638e-01f}, {5.000000000e-01f, -8.660253882e-01f}, {6.087613702e-01f, -7.933533788e-01f},
        {7.071067095e-01f, -7.071068287e-01f}, {7.933533788e-01f, -6.087613702e-01f}, {8.660253882e-01f, -5.000000000e-01f},
        {9.238795638e-01f, -3.826833963e-01f}, {9.659258127e-01f, -2.588190734e-01f}, {9.914448857e-01f, -1.305261701e-01f}};
};

// omega_R^m = exp(S*2*pi*i*m/R) for R dividing 48
template <int R, int S, typename C> __device__ __forceinline__ C root(int m) {
    static_assert(48 % R == 0, "radix must divide 48");
    using Re = typename real_of<C>::type;
    const int j = ((m % R) * (48 / R)) % 48;
    if constexpr (sizeof(Re) == 4) return mk<C>(RootsF48::cs[j][0], (Re)S * RootsF48::cs[j][1]);
    else return mk<C>((Re)Roots48::c[j], (Re)(S * Roots48::s(j)));
}

// v * k for an irrational constant k (1/sqrt2, sqrt3/2) given as its FP64
// value: in FP32, k = k_hi + k_lo with k_hi the float nearest to k, computed
// as fma(k_lo, v, k_hi v).  The float constant alone scales |v|^2 by
// (k_hi / k)^2 (for 1/sqrt2: 1 - 3.4e-8), the same loss on every butterfly of
// every pass, which compounds over the slices of the loop; the compensated
// product is exact to one rounding.
template <typename Re> __device__ __forceinline__ Re mul_const(Re v, double k) {
    if constexpr (sizeof(Re) == 4) {
        const float hi = (float)k, lo = (float)(k - (double)hi);
        return __builtin_fmaf(lo, v, hi * v);
    } else {
        return (Re)k * v;
    }
}
// x * omega_R^m with the trivial roots (+-1, +-i) done as moves/negations and
// the odd multiples of pi/4 as two adds and two multiplies (IEEE rules forbid
// the compiler from folding multiplications by 0 or 1 itself).
template <int R, int S, typename C> __device__ __forceinline__ C mul_root(C x, int m) {
    using Re = typename real_of<C>::type;
    const int j = ((m % R) * (48 / R)) % 48;  // angle in units of 2*pi/48
    const int jj = S > 0 ? j : (48 - j) % 48;
    constexpr double h = 0.7071067811865475244008444;
    auto H = [](Re v) { return mul_const(v, h); };
    switch (jj) {
        case 0: return x;
        case 12: return mk<C>(-x.y, x.x);    // * i
        case 24: return mk<C>(-x.x, -x.y);   // * -1
        case 36: return mk<C>(x.y, -x.x);    // * -i
        case 6: return mk<C>(H(x.x - x.y), H(x.x + x.y));    // * (1+i)/sqrt2
        case 18: return mk<C>(-H(x.x + x.y), H(x.x - x.y));  // * (-1+i)/sqrt2
        case 30: return mk<C>(H(x.y - x.x), -H(x.x + x.y));  // * (-1-i)/sqrt2
        case 42: return mk<C>(H(x.x + x.y), H(x.y - x.x));   // * (1-i)/sqrt2
        default: return cmul(x, root<R, S, C>(m));
    }
}

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
        constexpr double h = 0.8660254037844386467637232;  // sqrt(3)/2
        C s = cadd(a[1], a[2]);
        C d = csub(a[1], a[2]);
        C m = mk<C>(a[0].x - (Re)0.5 * s.x, a[0].y - (Re)0.5 * s.y);
        C r = mul_si<S>(mk<C>(mul_const(d.x, h), mul_const(d.y, h)));
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
            for (int k1 = 0; k1 < A; ++k1) y[n2 * A + k1] = mul_root<R, S>(u[k1], n2 * k1);
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
// A plan is a radix sequence and the number of threads T per line.
template <int NN, int TT, int... RR> struct PlanT {
    static constexpr int N = NN, T = TT;
    static constexpr int R[] = {RR...};
    static constexpr int P = sizeof...(RR);
};
// Row-sweep plans (contiguous lines).  Up to 512 T divides 64, so a line
// stays inside one wavefront and its exchanges need only wavefront barriers;
// at 1024 and 2048 a line spans 2 and 4 wavefronts (work-group barriers),
// which with 8 elements per thread still beats 16 per thread in one
// wavefront (three waves per SIMD instead of two).
template <int N> struct Plan;
template <> struct Plan<16> : PlanT<16, 4, 4, 4> {};
template <> struct Plan<32> : PlanT<32, 4, 4, 8> {};
template <> struct Plan<64> : PlanT<64, 8, 8, 8> {};
template <> struct Plan<128> : PlanT<128, 8, 16, 8> {};
template <> struct Plan<256> : PlanT<256, 32, 8, 8, 4> {};
template <> struct Plan<384> : PlanT<384, 32, 6, 4, 4, 4> {};
template <> struct Plan<512> : PlanT<512, 64, 8, 8, 8> {};
template <> struct Plan<768> : PlanT<768, 64, 12, 4, 4, 4> {};
template <> struct Plan<1024> : PlanT<1024, 128, 8, 8, 4, 4> {};
template <> struct Plan<2048> : PlanT<2048, 256, 8, 8, 8, 4> {};
// Column-sweep plans: fewer threads per line (more elements per thread) so a
// work-group of TW columns stays within 512 threads and its register budget.
#ifndef MSX_COL512_T
#define MSX_COL512_T 64
#endif
template <int N> struct ColPlan : Plan<N> {};
template <> struct ColPlan<256> : PlanT<256, 16, 16, 16> {};
template <> struct ColPlan<512> : PlanT<512, MSX_COL512_T, 8, 8, 8> {};
template <> struct ColPlan<1024> : PlanT<1024, 64, 16, 8, 8> {};
template <> struct ColPlan<2048> : PlanT<2048, 128, 16, 16, 8> {};

// Plan variants the autotuner chooses between (msx_tuning row_plan /
// col_plan): variant 0 is the plan above; variant 1 exists where `alt` is
// true and uses another radix sequence with the same exchange buffers.
template <int N, int V> struct RowPlanSel {
    using type = Plan<N>;
    static constexpr bool alt = false;
};
template <int N, int V> struct ColPlanSel {
    using type = ColPlan<N>;
    static constexpr bool alt = false;
};
// rows: 512 = 16 x 16 x 2 on 32 threads per line: radix-16 butterflies (one
// twiddle stage of 15 factors instead of two of 7) and 16 elements per thread
template <> struct RowPlanSel<512, 1> {
    using type = PlanT<512, 32, 16, 16, 2>;
    static constexpr bool alt = true;
};
// columns (the same threads per line as the built-in plan, which the tile
// widths assume): 512 = 4 x 4 x 4 x 8, one exchange more, radix-4 butterflies
template <> struct ColPlanSel<512, 1> {
    using type = PlanT<512, MSX_COL512_T, 4, 4, 4, 8>;
    static constexpr bool alt = true;
};
// 384 and 768 with half the threads per line and one pass less (6 x 8 x 8
// and 12 x 8 x 8: 24 elements per thread, whose registers allow one wave per
// SIMD; the built-in plans hold 12 and run two)
template <> struct RowPlanSel<384, 1> {
    using type = PlanT<384, 16, 6, 8, 8>;
    static constexpr bool alt = true;
};
template <> struct RowPlanSel<768, 1> {
    using type = PlanT<768, 32, 12, 8, 8>;
    static constexpr bool alt = true;
};
// 256 on half the threads per line (16 elements per thread, two passes):
// the built-in plan's 8 lines per work-group (against 16) spread a few waves
// over twice the work-groups (one wave: 7.7 against 9.5 us per row sweep),
// at 3% cost for 48 waves (23.2 against 22.6 us)
template <> struct RowPlanSel<256, 1> {
    using type = PlanT<256, 16, 16, 16>;
    static constexpr bool alt = true;
};
// 1024 and 2048 on half the threads per line and one pass less (16
// elements per thread, two waves per SIMD; the built-in plans hold 8 and run
// three, their lines spanning 2 and 4 wavefronts)
template <> struct RowPlanSel<1024, 1> {
    using type = PlanT<1024, 64, 16, 8, 8>;
    static constexpr bool alt = true;
};
template <> struct RowPlanSel<2048, 1> {
    using type = PlanT<2048, 128, 16, 16, 8>;
    static constexpr bool alt = true;
};

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

// ---------------------------------------------------------------- twiddles
// Twiddle table tw[m] = exp(+2*pi*i*m/N) (m < N) is generated in FP64 on the
// host.  Each thread needs, per pass with Ns > 1, the factors
// W^(r*k*N/(Ns*R)) of its Q butterflies (k = (t + q*T) mod Ns): they are
// loaded once into registers at kernel start and reused by every transform
// the thread performs (forward ones directly, inverse ones conjugated).
template <class PL> constexpr int tw_offset(int upto) {
    int c = 0, ns = 1;
    for (int i = 0; i < upto; ++i) {
        const int R = PL::R[i], Q = PL::N / (R * PL::T);
        if (ns > 1) c += Q * (R - 1);
        ns *= R;
    }
    return c;
}
template <class PL> constexpr int tw_count() { return tw_offset<PL>(PL::P); }

template <class PL, typename C> struct TwReg {
    static constexpr int K = tw_count<PL>() > 0 ? tw_count<PL>() : 1;
    C w[K];
};

template <class PL, int I, int Ns, typename C>
__device__ __forceinline__ void load_tw_pass(C* w, int t, const C* __restrict__ table) {
    if constexpr (I < PL::P) {
        constexpr int R = PL::R[I], Q = PL::N / (R * PL::T);
        if constexpr (Ns > 1) {
            constexpr int off = tw_offset<PL>(I), step = PL::N / (Ns * R);
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int k = (t + q * PL::T) % Ns;
#pragma unroll
                for (int r = 1; r < R; ++r) w[off + q * (R - 1) + r - 1] = table[r * k * step];
            }
        }
        load_tw_pass<PL, I + 1, Ns * R>(w, t, table);
    }
}
template <class PL, typename C>
__device__ __forceinline__ void load_twiddles(TwReg<PL, C>& tw, int t, const C* __restrict__ table) {
    load_tw_pass<PL, 0, 1>(tw.w, t, table);
}

// Work-group barrier that orders LDS only: the fences are restricted to the
// "local" address space, so outstanding global loads (the prefetch of the
// next tile) and stores stay in flight across it.  __syncthreads() would
// wait for every outstanding global access (vmcnt(0)) at each exchange.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// LDS synchronisation between the write and read halves of an exchange.
// WAVE = true when every line lies inside one wavefront (T divides 64):
// then a wavefront barrier suffices (the __syncwarp pattern, LDS only).
template <bool WAVE> __device__ __forceinline__ void lds_sync() {
    if constexpr (WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
    } else {
        lds_barrier();
    }
}

// ---------------------------------------------------------------- passes
// DB (double buffering): exchange x of a line uses LDS buffer x & 1, so the
// write of exchange x+1 cannot overtake reads of exchange x and one barrier
// per exchange suffices; X is the running exchange index.
template <int N, int T, int R, int Ns, int S, bool LAST, bool WAVE, bool DB, int X, int TWOFF, typename C>
__device__ __forceinline__ void stockham_pass(C (&a)[N / T], int t, const LdsView<C> (&lds)[2], const C* w) {
    constexpr int Q = N / (R * T);
    static_assert(Q * R * T == N, "radix/thread split must tile N");
    const LdsView<C>& L = lds[DB ? (X & 1) : 0];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        C b[R];
#pragma unroll
        for (int r = 0; r < R; ++r) b[r] = a[q + r * Q];
        const int j = t + q * T;
        if constexpr (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                C f = w[TWOFF + q * (R - 1) + r - 1];
                if (S < 0) f.y = -f.y;
                b[r] = cmul(b[r], f);
            }
        }
        Dft<R, S, C>::run(b);
        if constexpr (LAST) {
#pragma unroll
            for (int r = 0; r < R; ++r) a[q + r * Q] = b[r];
        } else {
            const int base = (j / Ns) * Ns * R + (j % Ns);
#pragma unroll
            for (int r = 0; r < R; ++r) *L.at(base + r * Ns) = b[r];
        }
    }
    if constexpr (!LAST) {
        lds_sync<WAVE>();
#pragma unroll
        for (int e = 0; e < N / T; ++e) a[e] = *L.at(t + e * T);
        if constexpr (!DB) lds_sync<WAVE>();
    }
}

template <class PL, int I, int Ns, int S, bool WAVE, bool DB, int X, typename C>
__device__ __forceinline__ void run_passes(C (&a)[PL::N / PL::T], int t, const LdsView<C> (&lds)[2], const C* w) {
    if constexpr (I < PL::P) {
        constexpr int R = PL::R[I];
        constexpr bool LAST = I == PL::P - 1;
        stockham_pass<PL::N, PL::T, R, Ns, S, LAST, WAVE, DB, X, tw_offset<PL>(I)>(a, t, lds, w);
        run_passes<PL, I + 1, Ns * R, S, WAVE, DB, LAST ? X : X + 1>(a, t, lds, w);
    }
}

// Full length-N transform of the registers a[] (unscaled), sign S.  With DB,
// X0 is the index of the transform's first exchange; a transform performs
// PL::P - 1 exchanges.
template <class PL, int S, bool WAVE, bool DB = false, int X0 = 0, typename C>
__device__ __forceinline__ void fft_line(C (&a)[PL::N / PL::T], int t, const LdsView<C> (&lds)[2],
                                         const TwReg<PL, C>& tw) {
    run_passes<PL, 0, 1, S, WAVE, DB, X0>(a, t, lds, tw.w);
}

}  // namespace msx
