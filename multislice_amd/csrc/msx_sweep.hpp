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

// Build-time shape knobs (defaults are the tuned values; the build can
// override them to produce variants for the tuner):
//   MSX_ROW_PREFETCH  prefetch the next wave's row during the transforms
//   MSX_COL_PREFETCH  same for the column tile
//   MSX_COL_CONST     keep per-element k^2 / propagator factors in registers
#ifndef MSX_ROW_PREFETCH
#define MSX_ROW_PREFETCH 1
#endif
#ifndef MSX_COL_PREFETCH
#define MSX_COL_PREFETCH 1
#endif
#ifndef MSX_COL_CONST
#define MSX_COL_CONST 1
#endif
#ifndef MSX_ROW_DRAIN
#define MSX_ROW_DRAIN 1
#endif

namespace msx {

constexpr int kMaxDet = 8;

// Raw buffer stores (gfx9 descriptor word 3: 32-bit data, no swizzle).  A
// store whose byte offset is at or beyond the descriptor's size is dropped
// by the hardware; kBufDrop plus any in-line offset used here stays far
// beyond every image size (at most 2048 x 2048 x 16 B = 64 MiB) and below
// 4 GiB.
constexpr int kBufWord3 = 0x00020000;
constexpr unsigned kBufDrop = 0x80000000u;
__device__ __forceinline__ void buf_store(float2 v, __amdgpu_buffer_rsrc_t r, unsigned off) {
    using u2 = __attribute__((ext_vector_type(2))) unsigned;
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, (int)off, 0, 0);
}
__device__ __forceinline__ void buf_store(double2 v, __amdgpu_buffer_rsrc_t r, unsigned off) {
    using u4 = __attribute__((ext_vector_type(4))) unsigned;
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u4, v), r, (int)off, 0, 0);
}
__device__ __forceinline__ void buf_store(float v, __amdgpu_buffer_rsrc_t r, unsigned off) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)off, 0, 0);
}
__device__ __forceinline__ void buf_store(double v, __amdgpu_buffer_rsrc_t r, unsigned off) {
    using u2 = __attribute__((ext_vector_type(2))) unsigned;
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, (int)off, 0, 0);
}
// Raw buffer loads: an offset at or beyond the descriptor's size reads zero
// without touching memory.
template <typename C> __device__ __forceinline__ C buf_load(__amdgpu_buffer_rsrc_t r, unsigned off);
template <> __device__ __forceinline__ float buf_load<float>(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}
template <> __device__ __forceinline__ double buf_load<double>(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}
template <> __device__ __forceinline__ float2 buf_load<float2>(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}
template <> __device__ __forceinline__ double2 buf_load<double2>(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0));
}

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
    const C* const* trans_tab;  // per-wave transmission [nx][N] (wave b: trans_tab[b]), or nullptr
    const Re* vreal;        // real potential [nb][nx][N] (replaces the input), or nullptr
    const Re* vabs;         // optional absorptive potential, same layout: exp(-vscale*vabs) damping
    Re vscale;              // phase per unit of vreal (sigma/1000)
    Re scale;               // multiplies the data after the IN transform
    Re* inten;              // optional |psi|^2 output [nb][nx][N] (after the pointwise stage)
    int nx, nb;
    int wpb;                // waves handled by one work-group (loop with prefetch)
    // band pruning along y: element iy is "out of band" when ky2[iy] >= k2max.
    // prune & 1: out-of-band inputs are known to be exactly zero (not loaded);
    // prune & 2: out-of-band outputs are never read (not stored).
    const float* ky2;
    float k2max;
    int prune;
    // with inten set: iacc 0 writes |psi|^2, 1 writes iw*|psi|^2, 2 adds
    // iw*|psi|^2 to what inten holds (weighted sums over several launches)
    int iacc;
    Re iw;
    int nostore;  // 1: only the intensity is written, not the wave
    int plan;  // FFT plan variant (RowPlanSel)
};

// IN/OUT: 0 none, +1 forward (exp +i), -1 inverse (exp -i, unscaled)
// Work-group = LPW adjacent rows ix of `wpb` consecutive waves: the
// transmission row and the twiddles are loaded once into registers and the
// next wave's row is prefetched while the current one is transformed.
// VREAL: the input is a real potential turned into exp(i*vscale*V) (the
// transmission build); a compile-time switch so the slice kernels carry none
// of its sincos code.  INTEN: the |psi|^2 output (p.inten) is written; also
// compile-time, so the slice kernels hold no registers for it.
template <typename C, int N, int IN, int OUT, int LPW, bool VREAL, bool INTEN = false, int PV = 0>
__global__ __launch_bounds__((RowPlanSel<N, PV>::type::T * LPW)) void row_sweep(RowArgs<C> p) {
    using Re = typename real_of<C>::type;
    using PL = typename RowPlanSel<N, PV>::type;
    constexpr int T = PL::T, E = N / T;
    constexpr bool WAVE = (64 % T) == 0;  // a line never straddles two wavefronts
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* smem = reinterpret_cast<C*>(smem_raw);
    const int lw = threadIdx.x / T, t = threadIdx.x % T;
    const int nxg = (p.nx + LPW - 1) / LPW;
    const int gx = blockIdx.x % nxg, chunk = blockIdx.x / nxg;
    const int ix_raw = gx * LPW + lw;
    const bool live = ix_raw < p.nx;
    const int ix = live ? ix_raw : p.nx - 1;  // dead rows compute, never store
    const int b0 = chunk * p.wpb, b1 = min(p.nb, b0 + p.wpb);
    const LdsView<C> lds[2] = {{smem + lw * (N + N / 16), 1, true}, {smem + lw * (N + N / 16), 1, true}};
    // element offset of (this line, y) in a wave's image
    auto eo = [&](int y) -> unsigned { return (unsigned)ix * (unsigned)N + (unsigned)y; };

    TwReg<PL, C> tw;
    if constexpr (IN != 0 || OUT != 0) load_twiddles<PL>(tw, t, p.tw);
    // The inverse-transform scale is folded into the transmission factors
    // (a * (scale * t) instead of (a * scale) * t): once per work-group for a
    // shared transmission, per wave for a per-wave one, so both give the same
    // arithmetic.
    const bool fold = (p.trans || p.trans_tab) && p.scale != (Re)1;
    C tr[E];
    if (p.trans) {
        const C* trow = p.trans + (long long)ix * N;
#pragma unroll
        for (int e = 0; e < E; ++e) tr[e] = fold ? cscale(trow[t + e * T], p.scale) : trow[t + e * T];
    }

    unsigned inb = ~0u;  // bit e: element t + e*T lies inside the y band
    if (p.prune) {
        inb = 0u;
#pragma unroll
        for (int e = 0; e < E; ++e) inb |= (p.ky2[t + e * T] < p.k2max ? 1u : 0u) << e;
    }
    const unsigned ld_mask = (p.prune & 1) ? inb : ~0u;
    const unsigned st_mask = (p.prune & 2) ? inb : ~0u;

    auto load = [&](int b, C (&v)[E], bool valid = true) {
        if constexpr (VREAL) {
            const Re* vr = p.vreal + ((long long)b * p.nx + ix) * N;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                Re s, c;
                if constexpr (sizeof(Re) == 4) sincosf(p.vscale * vr[t + e * T], &s, &c);
                else sincos(p.vscale * vr[t + e * T], &s, &c);
                if (p.vabs) {
                    const Re* va = p.vabs + ((long long)b * p.nx + ix) * N;
                    Re d;
                    if constexpr (sizeof(Re) == 4) d = expf(-p.vscale * va[t + e * T]);
                    else d = exp(-p.vscale * va[t + e * T]);
                    c *= d;
                    s *= d;
                }
                v[e] = mk<C>(c, s);
            }
        } else {
            // buffer descriptor of wave b's image (empty when !valid: the
            // loads then read zero and touch no memory); out-of-band
            // elements are loaded from past the descriptor's bound, which
            // reads +0 and touches no memory, so they need no zeroing after
            // the load (unconditional loads keep the prefetch free of
            // branch-induced waits)
            const C* img = p.src ? p.src + (long long)__builtin_amdgcn_readfirstlane(b) * p.src_bstride
                                 : p.psi + (long long)__builtin_amdgcn_readfirstlane(b) * p.nx * N;
            const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                const_cast<C*>(img), 0, valid ? p.nx * N * (int)sizeof(C) : 0, kBufWord3);
#pragma unroll
            for (int e = 0; e < E; ++e)
                v[e] = buf_load<C>(rs, ((ld_mask >> e) & 1u) ? eo(t + e * T) * (unsigned)sizeof(C) : kBufDrop);
        }
    };
    C a[E], nxt[E];
    // the |psi|^2 variant (lens sweeps of many waves, one per work-group)
    // and the transmission build (whose loads compute exp(i vscale V)) do
    // without the prefetch and its registers
    constexpr bool PF = MSX_ROW_PREFETCH && !INTEN && !VREAL;
    if (PF && b0 < b1) {
        load(b0, a);
    }
    // jobs batched in one launch: each wave has its own transmission, whose
    // address is read one wave ahead (after the prefetch, so that using it
    // never waits for this wave's stores)
    const C* tnext = (p.trans_tab && b0 < b1) ? p.trans_tab[b0] : nullptr;
    // Drain the constant loads (and the first row) once: a load still
    // pending at the loop entry makes the compiler wait for everything
    // issued after it in every iteration.
    if (MSX_ROW_DRAIN) __builtin_amdgcn_s_waitcnt(0);
    // |psi|^2 output of wave b: [nb][nx][N] reals
    auto inten_rsrc = [&](int b) {
        return __builtin_amdgcn_make_buffer_rsrc(p.inten + (long long)__builtin_amdgcn_readfirstlane(b) * p.nx * N, 0,
                                                 p.nx * N * (int)sizeof(Re), kBufWord3);
    };
    const unsigned row0 = (unsigned)ix * (unsigned)N;
    // one wave: its row is in cur (with the prefetch) or loaded here; the
    // next wave's row is prefetched into next
    auto wave = [&](int b, C (&cur)[E], C (&next)[E]) {
        if (p.trans_tab) {
            // Loaded before the prefetch, through a buffer descriptor, and
            // scaled only where it is used: the wait counters count in issue
            // order, so consuming a load issued after the prefetch would wait
            // for the whole prefetch.
            const unsigned long long ta = (unsigned long long)tnext;
            const C* tcur = reinterpret_cast<const C*>(
                ((unsigned long long)__builtin_amdgcn_readfirstlane((unsigned)(ta >> 32)) << 32) |
                (unsigned)__builtin_amdgcn_readfirstlane((unsigned)ta));
            const __amdgpu_buffer_rsrc_t rt = __builtin_amdgcn_make_buffer_rsrc(
                const_cast<C*>(tcur), 0, p.nx * N * (int)sizeof(C), kBufWord3);
            const unsigned row = (unsigned)ix * (unsigned)N;
#pragma unroll
            for (int e = 0; e < E; ++e) tr[e] = buf_load<C>(rt, (row + (unsigned)(t + e * T)) * (unsigned)sizeof(C));
        }
        if constexpr (PF) {
            // unconditional: past the last wave of the chunk the descriptor
            // is empty
            load(b + 1 < b1 ? b + 1 : b, next, b + 1 < b1);
        } else {
            load(b, cur);
        }
        Re oi[INTEN ? E : 1];
        if constexpr (INTEN) {
            // the running sums this wave adds to, loaded with its row
            if (p.iacc == 2) {
#pragma unroll
                for (int e = 0; e < E; ++e)
                    oi[e] = buf_load<Re>(inten_rsrc(b), (row0 + (unsigned)(t + e * T)) * (unsigned)sizeof(Re));
            }
        }
        if (p.trans_tab) tnext = p.trans_tab[b + 1 < b1 ? b + 1 : b];
        if constexpr (IN != 0) fft_line<PL, IN, WAVE>(cur, t, lds, tw);
        if (p.scale != (Re)1 && !fold) {
#pragma unroll
            for (int e = 0; e < E; ++e) cur[e] = cscale(cur[e], p.scale);
        }
        if (p.trans || p.trans_tab) {
            // cfpix::operator*= (cfpix.cpp:296-318): (wr*tr - wi*ti, wr*ti + wi*tr)
            const bool scale_here = p.trans_tab && fold;
#pragma unroll
            for (int e = 0; e < E; ++e) cur[e] = cmul(cur[e], scale_here ? cscale(tr[e], p.scale) : tr[e]);
        }
        if constexpr (INTEN) {
            // buffer stores (dead rows aim past the bound)
            const __amdgpu_buffer_rsrc_t ri = inten_rsrc(b);
#pragma unroll
            for (int e = 0; e < E; ++e) {
                Re v = cnorm2(cur[e]);
                if (p.iacc != 0) v = p.iw * v;
                if (p.iacc == 2) v = oi[e] + v;
                buf_store(v, ri, live ? (row0 + (unsigned)(t + e * T)) * (unsigned)sizeof(Re) : kBufDrop);
            }
        }
        if constexpr (OUT != 0) fft_line<PL, OUT, WAVE>(cur, t, lds, tw);
        {
            // buffer stores into wave b's image; dead rows, pruned elements
            // and a sweep that stores nothing aim past the bound (dropped):
            // no branch, so no later wait has to count them conservatively
            C* img = p.psi + (long long)__builtin_amdgcn_readfirstlane(b) * p.nx * N;
            const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                img, 0, p.nostore ? 0 : p.nx * N * (int)sizeof(C), kBufWord3);
#pragma unroll
            for (int e = 0; e < E; ++e)
                buf_store(cur[e], rs, (live && ((st_mask >> e) & 1u)) ? eo(t + e * T) * (unsigned)sizeof(C) : kBufDrop);
        }
    };
    if constexpr (PF) {
        // two waves per trip with the row buffers swapping roles: no copy of
        // the prefetched row into the working registers
        for (int b = b0; b < b1; b += 2) {
            wave(b, a, nxt);
            if (b + 1 < b1) wave(b + 1, nxt, a);
        }
    } else {
        for (int b = b0; b < b1; ++b) wave(b, a, nxt);
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
    kColObl = 64,    // with kColProp: the oblique-cell propagator (FP64 metric per element)
    kColJobs = 128,  // with kColProp: uneven jobs in one launch: per-wave propagator factors
                     // (prop_tab) and units pulled from device-side work counters (queue:
                     // kXcdCount counters kQueueStride apart, zero at launch)
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
    // lens astigmatism (astig != 0): chi gains chi1*k2*(a2 sin(2(phi-phi2)) +
    // a3 |k| sin(3(phi-phi3))), phi = atan2(ky, kx)
    double a2, phi2, a3, phi3;
    int astig;
    C* out;                // store target (nullptr: in place)
    int ny, nb, ops;
    int wpb;               // > 0: waves per work-group (chunked); < 0: balanced schedule (below)
    int tile;              // columns per work-group (0: the built-in width for this length)
    // prune != 0: columns with ky2[iy] >= k2max are skipped entirely (their
    // k-space values are all masked to zero, so nothing downstream reads them)
    int prune;
    // oblique cell (kColObl): x = fx / obl_asg, y = fy / obl_bsg (a sin gam,
    // b sin gam), s = x (x - obl_2c y) + y^2 (obl_2c = 2 cos gam), kept where
    // s < obl_smax with the factor exp(-i (obl_pl s - obl_tx x - obl_ty y))
    double obl_asg, obl_bsg, obl_2c, obl_smax, obl_pl, obl_tx, obl_ty;
    int plan;  // FFT plan variant (ColPlanSel)
    // kColJobs: prop_tab[b] = the factors of wave b's slice, px (N values)
    // then py (ny values), in place of px/py; queue = a zeroed counter from
    // which the balanced schedule's work-groups claim their units after the
    // first (unit = gridDim.x + the counter's previous value)
    const C* const* prop_tab;
    unsigned* queue;
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

// Work-group = TW adjacent columns iy of `wpb` consecutive waves; per-column
// and per-thread constants (k^2, propagator factors, twiddles) are loaded
// once, the next wave's tile is prefetched while the current one is
// transformed.
// OPS: compile-time subset of kColProp | kColMask | kColDetect | kColGen |
// kColCtf (only kColStore is read from p.ops at run time), so the per-slice
// kernel carries no registers for the probe, detector or lens paths.
// DB: double-buffered LDS exchanges (one barrier per exchange); needs twice
// the LDS tile.
#ifndef MSX_COL_MINW
#define MSX_COL_MINW 1
#endif
#ifndef MSX_XCD_REMAP
#define MSX_XCD_REMAP 1
#endif
// Bijection hardware block id -> logical id such that the blocks one XCD
// receives (ids congruent mod 8) cover a contiguous logical range.
constexpr int kXcdCount = 8;
// distance between the per-XCD work counters of one kColJobs launch
// (unsigned; 128 B, one cache line each)
constexpr int kQueueStride = 32;
constexpr int kQueueSlots = kXcdCount * kQueueStride;  // unsigned per launch
__device__ __forceinline__ int xcd_remap(int bid, int nblk) {
    constexpr int kXcd = kXcdCount;
    const int per = nblk / kXcd, rem = nblk % kXcd;
    const int xcd = bid % kXcd, idx = bid / kXcd;
    return xcd * per + min(xcd, rem) + idx;
}
// Byte offset of the balanced schedule's group list in the column sweep's
// dynamic LDS: after the exchange buffers (which also hold the detector
// reduction scratch, at least 2 KiB).
template <typename C, int N, int TW, bool DB> constexpr size_t col_list_offset() {
    constexpr size_t tile = (size_t)N * (TW + 1) * sizeof(C);
    constexpr size_t x = DB ? 2 * tile : tile;
    return x < 2048 ? 2048 : x;
}

// Schedules (p.wpb; the launcher turns 0 into -1):
//   > 0  chunked: work-group = (column group g, waves [b0, b0 + wpb)), the
//        groups of one chunk numbered consecutively;
//   < 0  balanced: the grid is sized to the work-groups the chip holds at
//        once, and work-group w takes the units (live group, wave) u = w,
//        w + G, w + 2G, ... in wave-major order, so every work-group gets
//        the same number of units (to one) and no partly filled last round
//        of work-groups idles most of the chip; at each step the work-groups
//        of one XCD hold consecutive groups of one wave, as in the chunked
//        order.

template <typename C, int N, int IN, int OUT, int TW, int OPS, bool DB, int PV = 0>
__global__ __launch_bounds__((ColPlanSel<N, PV>::type::T * TW), MSX_COL_MINW) void col_sweep(ColArgs<C> p) {
    using Re = typename real_of<C>::type;
    using PL = typename ColPlanSel<N, PV>::type;
    constexpr int T = PL::T, E = N / T;
    constexpr bool WAVE = false;  // a column is spread over several wavefronts
    static_assert(TW <= 32, "live-column masks are 32 bits");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* smem = reinterpret_cast<C*>(smem_raw);
    const int lw = threadIdx.x % TW, t = threadIdx.x / TW;
    // element offset of (row t, column y) of an image (row t + e T adds e T ny)
    auto elem = [&](int y) -> unsigned { return (unsigned)t * (unsigned)p.ny + (unsigned)y; };
    const int groups = (p.ny + TW - 1) / TW;
    // XCD-aware order: work-groups are dispatched round-robin over the 8 XCDs;
    // renumber them so that consecutive logical groups (neighbouring columns,
    // which share L2 lines when TW*sizeof(C) < 128 B) run on the same XCD.
    const int bid = MSX_XCD_REMAP ? xcd_remap((int)blockIdx.x, (int)gridDim.x) : (int)blockIdx.x;
    // bit c set: column gg*TW + c exists and is in band (or pruning is off).
    // Dead lanes of a partly live group load from the group's first live
    // column (same cache lines as a live lane); their column is transformed
    // on its own, never stored and left out of the detector sums, so the
    // values need no zeroing.
    auto live_mask = [&](int gg) -> unsigned {
        unsigned m = 0;
        for (int c = 0; c < TW; ++c) {
            const int y = gg * TW + c;
            if (y < p.ny && (!p.prune || p.ky2[y] < p.k2max)) m |= 1u << c;
        }
        return m;
    };
    constexpr bool jobs = (OPS & kColJobs) != 0;
    const bool bal = jobs || p.wpb < 0;
    int2* glist = reinterpret_cast<int2*>(smem_raw + col_list_offset<C, N, TW, DB>());
    int ngl = 0, gfix = 0, b0 = 0, b1 = 0;
    unsigned mfix = 0;
    if (bal) {
        // stable partition of the groups: the live ones (group, mask) in
        // ascending order at the front, the dead ones after them
        __shared__ int cnt[2][16];
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = (blockDim.x + 63) >> 6;
        int nd = 0;
        for (int base = 0; base < groups; base += blockDim.x) {
            const int gg = base + threadIdx.x;
            const unsigned m = gg < groups ? live_mask(gg) : 0u;
            const bool lv = m != 0, dd = gg < groups && m == 0;
            const unsigned long long bl = __ballot(lv), bd = __ballot(dd);
            if (lane == 0) {
                cnt[0][wv] = __popcll(bl);
                cnt[1][wv] = __popcll(bd);
            }
            __syncthreads();
            int pl = ngl, pd = nd, tl = ngl, td = nd;
            for (int w = 0; w < nwv; ++w) {
                if (w < wv) pl += cnt[0][w], pd += cnt[1][w];
                tl += cnt[0][w], td += cnt[1][w];
            }
            const unsigned long long below = (1ull << lane) - 1ull;
            if (lv) glist[pl + __popcll(bl & below)] = make_int2(gg, (int)m);
            if (dd) glist[groups - 1 - (pd + __popcll(bd & below))] = make_int2(gg, 0);
            __syncthreads();
            ngl = __builtin_amdgcn_readfirstlane(tl);  // uniform (LDS sums of ballots)
            nd = __builtin_amdgcn_readfirstlane(td);
        }
        if constexpr ((OPS & kColDetect) != 0) {
            // the dead groups' detector partials are zero
            const long long V = (long long)nd * p.nb;
            if ((int)threadIdx.x < p.ndet)
                for (long long v = bid; v < V; v += gridDim.x) {
                    const int gd = glist[ngl + (int)(v % nd)].x;
                    p.partial[((long long)threadIdx.x * p.nb + v / nd) * groups + gd] = 0.0;
                }
        }
        if (ngl == 0 || (long long)bid >= (long long)ngl * p.nb) return;
    } else {
        gfix = bid % groups;
        const int chunk = bid / groups;
        b0 = chunk * p.wpb;
        b1 = min(p.nb, b0 + p.wpb);
        mfix = live_mask(gfix);
        if (mfix == 0) {
            // whole group outside the band: its detector partials are zero
            if constexpr ((OPS & kColDetect) != 0) {
                if ((int)threadIdx.x < p.ndet)
                    for (int b = b0; b < b1; ++b) p.partial[((long long)threadIdx.x * p.nb + b) * groups + gfix] = 0.0;
            }
            return;
        }
    }
    // unit i of this work-group: (group, live mask, wave); false past the end
    const long long units = (long long)ngl * p.nb;
    // device-side work queue (kColJobs), one counter per XCD: the units of
    // the balanced schedule's later rounds that fall to an XCD's work-groups
    // (consecutive column groups, whose shared cache lines then meet in that
    // XCD's L2) are handed out by that XCD's counter in round order; after its
    // first unit a work-group takes the units it claims, one claim ahead:
    // unit i+2 is claimed while unit i is transformed and published in
    // qslot[i & 1] at the end of that iteration.  The k-th claim on an XCD
    // is unit (1 + k / W) G + base + k % W (W work-groups from logical index
    // base on that XCD, G in all), which grows with k: the claims end when it
    // passes the unit count, so every work-group exits.
    __shared__ unsigned qslot[2];
    int q_w = 1, q_base = 0;
    unsigned* q_ctr = p.queue;
    if constexpr (jobs) {
        if (MSX_XCD_REMAP && gridDim.x >= kXcdCount) {
            const int x = (int)blockIdx.x % kXcdCount, per = (int)gridDim.x / kXcdCount,
                      rem = (int)gridDim.x % kXcdCount;
            q_w = per + (x < rem ? 1 : 0);
            q_base = x * per + min(x, rem);
            q_ctr = p.queue + x * kQueueStride;
        } else {
            q_w = (int)gridDim.x;
        }
    }
    auto claim_unit = [&](unsigned k) -> unsigned {
        return (1u + k / (unsigned)q_w) * gridDim.x + (unsigned)q_base + k % (unsigned)q_w;
    };
    auto unit = [&](int i, int& g, unsigned& m, int& b) -> bool {
        if (bal) {
            long long u = bid + (long long)i * gridDim.x;
            if constexpr (jobs) {
                if (i > 0) u = (long long)__builtin_amdgcn_readfirstlane(qslot[i & 1]);
            }
            if (u >= units) return false;
            const int2 e = glist[(int)(u % ngl)];
            g = __builtin_amdgcn_readfirstlane(e.x);
            m = (unsigned)__builtin_amdgcn_readfirstlane(e.y);
            b = (int)(u / ngl);
            return true;
        }
        g = gfix;
        m = mfix;
        b = b0 + i;
        return b < b1;
    };
    const LdsView<C> lds[2] = {{smem + lw, TW + 1, false}, {smem + N * (TW + 1) + lw, TW + 1, false}};
    constexpr bool gen = (OPS & kColGen) != 0;
    constexpr bool point = (OPS & (kColProp | kColMask | kColDetect | kColCtf)) != 0;

    TwReg<PL, C> tw;
    if constexpr (IN != 0 || OUT != 0) load_twiddles<PL>(tw, t, p.tw);
    // per-element constants along x (registers)
    float kx2c[MSX_COL_CONST ? E : 1];
    C pxc[MSX_COL_CONST ? E : 1];
    if constexpr (MSX_COL_CONST) {
        if constexpr (point || gen) {
#pragma unroll
            for (int e = 0; e < E; ++e) kx2c[e] = p.kx2[t + e * T];
        }
        if constexpr ((OPS & kColProp) != 0 && !jobs) {
#pragma unroll
            for (int e = 0; e < E; ++e) pxc[e] = p.px[t + e * T];
        }
    }
    auto KX2 = [&](int e) -> float {
        if constexpr (MSX_COL_CONST) return kx2c[e]; else return p.kx2[t + e * T];
    };
    // propagator factors of the current unit's wave (kColJobs: per wave)
    const C* pxp = p.px;
    const C* pyp = p.py;
    auto PX = [&](int e) -> C {
        if constexpr (MSX_COL_CONST) return pxc[e]; else return pxp[t + e * T];
    };
    // per-column constants along y of a unit's group, loaded before the
    // prefetch of the next unit (so that waiting for them never waits for it)
    auto col_y = [&](int g) { return min(g * TW + lw, p.ny - 1); };
    auto load_ky2 = [&](int g) -> float {
        if constexpr (point || gen) return p.ky2[col_y(g)]; else return 0.f;
    };
    auto load_py = [&](int g) -> C {
        if constexpr ((OPS & kColProp) != 0) return pyp[col_y(g)]; else return mk<C>((Re)1, (Re)0);
    };

    // strided accesses: buffer descriptor of the wave's image (uniform) +
    // 32-bit lane byte offset (one image is at most 64 MiB).  With valid
    // false the descriptor is empty: the loads return zeros and read nothing
    // (the prefetch past a work-group's last unit, issued without a branch).
    const unsigned ostep = (unsigned)T * (unsigned)p.ny;
    const int img_bytes = N * p.ny * (int)sizeof(C);
    auto load = [&](int g, unsigned m, int b, bool valid, C (&v)[E]) {
        const int ld_col = ((m >> lw) & 1u) ? lw : __ffs(m) - 1;
        const C* img = p.psi + (long long)__builtin_amdgcn_readfirstlane(b) * N * p.ny;
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(const_cast<C*>(img), 0,
                                                                             valid ? img_bytes : 0, kBufWord3);
        const unsigned off0 = elem(g * TW + ld_col) * (unsigned)sizeof(C);
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = buf_load<C>(rs, off0 + (unsigned)e * ostep * (unsigned)sizeof(C));
    };

    int g, b;
    unsigned m;
    unit(0, g, m, b);  // every work-group that gets here has a unit
    if constexpr (jobs) {
        if (threadIdx.x == 0) qslot[1] = claim_unit(atomicAdd(q_ctr, 1u));
        lds_barrier();
    }
    C a[E], nxt[MSX_COL_PREFETCH ? E : 1];
    // kColJobs: the address of a unit's propagator factors is read one unit
    // ahead (issued before the prefetch and the stores, so that using it at
    // the next unit never waits for them)
    const C* pnext = nullptr;
    if constexpr (jobs) pnext = p.prop_tab[b];
    if (MSX_COL_PREFETCH && !gen) load(g, m, b, true, a);
    // Drain the constant loads (and the first tile) once: the wait counters
    // count in issue order, and a load still pending at the loop entry makes
    // the compiler wait for everything issued after it (the prefetch) in
    // every iteration.
    __builtin_amdgcn_s_waitcnt(0);
    for (int i = 0;; ++i) {
        int g2 = g, b2 = b;
        unsigned m2 = m;
        const bool more = unit(i + 1, g2, m2, b2);
        if (!more) g2 = g, m2 = m, b2 = b;
        // the claim of unit i+2, issued before the prefetch so that reading
        // its result at the end of this unit never waits for the prefetch
        unsigned claim = 0;
        if constexpr (jobs) {
            if (threadIdx.x == 0 && more) claim = atomicAdd(q_ctr, 1u);
        }
        if constexpr (jobs) {
            // this unit's wave's factors, loaded before the prefetch of the
            // next unit (using them then never waits for it)
            const unsigned long long pa = (unsigned long long)pnext;
            pxp = reinterpret_cast<const C*>(
                ((unsigned long long)__builtin_amdgcn_readfirstlane((unsigned)(pa >> 32)) << 32) |
                (unsigned)__builtin_amdgcn_readfirstlane((unsigned)pa));
            pyp = pxp + N;
            if constexpr (MSX_COL_CONST) {
#pragma unroll
                for (int e = 0; e < E; ++e) pxc[e] = pxp[t + e * T];
            }
            pnext = p.prop_tab[b2];
        }
        const float ky2 = load_ky2(g);
        const C py = load_py(g);
        if constexpr (MSX_COL_PREFETCH) {
            // unconditional (past the last unit an empty descriptor reads
            // nothing): a branch around the loads would make the compiler
            // wait for them at the join
            if (!gen) load(g2, m2, b2, more, nxt);
        } else {
            if (!gen) load(g, m, b, true, a);
        }
        const int iy = col_y(g);
        const bool live = (m >> lw) & 1u;
        const double ky = (gen || (OPS & kColCtf)) ? (double)freq_index(iy, p.ny) / p.by : 0.0;
        const double oy = (OPS & kColObl) ? (double)freq_index(iy, p.ny) / p.obl_bsg : 0.0;
        if constexpr (gen) {
            const double2 pos = p.pos[b];
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int ix = t + e * T;
                const float k2f = KX2(e) + ky2;
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
        }
        if constexpr (IN != 0) fft_line<PL, IN, WAVE, DB, 0>(a, t, lds, tw);

        if constexpr (point) {
            Re acc[kMaxDet], cmp[kMaxDet];
#pragma unroll
            for (int d = 0; d < kMaxDet; ++d) acc[d] = cmp[d] = (Re)0;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const float k2 = KX2(e) + ky2;
                if constexpr ((OPS & kColObl) != 0) {
                    // Propagator.cpp:55-107 (this engine's sign), FP64 metric
                    // and phase, the phase reduced mod 2 pi before the
                    // engine-precision sincos
                    const double ox = (double)freq_index(t + e * T, N) / p.obl_asg;
                    const double s2 = ox * (ox - p.obl_2c * oy) + oy * oy;
                    if (s2 < p.obl_smax)
                        a[e] = cmul(a[e], expi_reduced<C>(-(p.obl_pl * s2 - p.obl_tx * ox - p.obl_ty * oy)));
                    else
                        a[e] = mk<C>((Re)0, (Re)0);
                } else if constexpr ((OPS & (kColProp | kColMask)) != 0) {
                    if (k2 < p.k2max) {
                        // slicelib.cpp:765-770: (w * py) * px
                        if constexpr ((OPS & kColProp) != 0) a[e] = cmul(cmul(a[e], py), PX(e));
                    } else {
                        a[e] = mk<C>((Re)0, (Re)0);
                    }
                }
                if constexpr ((OPS & kColCtf) != 0) {
                    if (k2 <= p.kap_hi) {
                        const double kx = (double)freq_index(t + e * T, N) / p.ax;
                        const double kk = kx * kx + ky * ky;
                        double w = p.chi2 * kk - p.df;
                        if (p.astig) {
                            // mulslice.cpp:620-628
                            const double phi = atan2(ky, kx);
                            w += p.a2 * sin(2.0 * (phi - p.phi2)) + p.a3 * sqrt(kk) * sin(3.0 * (phi - p.phi3));
                        }
                        a[e] = cmul(a[e], expi_reduced<C>(-p.chi1 * kk * w));
                    } else {
                        a[e] = mk<C>((Re)0, (Re)0);
                    }
                }
                if constexpr ((OPS & kColDetect) != 0) {
                    const Re v = cnorm2(a[e]);
#pragma unroll
                    for (int d = 0; d < kMaxDet; ++d) {
                        if (d < p.ndet && k2 >= p.dlo[d] && k2 <= p.dhi[d]) {
                            // Kahan-compensated per-thread accumulation
                            const Re y = v - cmp[d];
                            const Re s = acc[d] + y;
                            cmp[d] = (s - acc[d]) - y;
                            acc[d] = s;
                        }
                    }
                }
            }
            if constexpr ((OPS & kColDetect) != 0) {
                // deterministic work-group reduction in FP64: wavefront
                // butterfly, then the per-wave results in fixed order via LDS.
                lds_barrier();  // LDS may still be read by the last FFT pass
                double* red = reinterpret_cast<double*>(smem_raw);
                const int nwaves = (T * TW + 63) / 64;
                const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
                for (int d = 0; d < p.ndet; ++d) {
                    double v = live ? (double)acc[d] - (double)cmp[d] : 0.0;
#pragma unroll
                    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
                    if (lane == 0) red[d * 32 + wv] = v;
                }
                lds_barrier();
                if (threadIdx.x < p.ndet) {
                    const int d = threadIdx.x;
                    double s = 0.0;
                    for (int w = 0; w < nwaves; ++w) s += red[d * 32 + w];
                    p.partial[((long long)d * p.nb + b) * groups + g] = s;
                }
                lds_barrier();
            }
        }
        if constexpr (OUT != 0) fft_line<PL, OUT, WAVE, DB, PL::P - 1>(a, t, lds, tw);
        {
            // Buffer stores bounded by this wave's image: the lanes of dead
            // columns (and every lane when the sweep stores nothing) aim past
            // the bound, where the hardware drops the write.  No branch: a
            // branch around the stores would make every later wait of the
            // loop count them (the wait counters count in issue order and
            // the compiler merges both paths).
            C* img = (p.out ? p.out : p.psi) + (long long)__builtin_amdgcn_readfirstlane(b) * N * p.ny;
            const int bytes = (p.ops & kColStore) ? N * p.ny * (int)sizeof(C) : 0;
            const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(img, 0, bytes, kBufWord3);
            const unsigned so = live ? elem(g * TW + lw) * (unsigned)sizeof(C) : kBufDrop;
#pragma unroll
            for (int e = 0; e < E; ++e) buf_store(a[e], rs, so + (unsigned)e * ostep * (unsigned)sizeof(C));
        }
        if (!more) break;
        if constexpr (jobs) {
            if (threadIdx.x == 0) qslot[i & 1] = claim_unit(claim);
            lds_barrier();
        }
        if constexpr (MSX_COL_PREFETCH) {
#pragma unroll
            for (int e = 0; e < E; ++e) a[e] = nxt[e];
        }
        g = g2;
        m = m2;
        b = b2;
    }
}

}  // namespace msx
