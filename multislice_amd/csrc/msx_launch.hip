// msx_launch.hip -- sweep-kernel instantiations and launchers for one line
// length.  Compiled once per supported length with -DMSX_N=<N> (the build
// runs these compilations in parallel); each object defines
// launch_row_<N>, launch_col_<N> and col_groups_<N> for float2 and double2.
#include <algorithm>
#include <atomic>

#include "msx_internal.h"

#ifndef MSX_N
#error "compile with -DMSX_N=<line length>"
#endif

#define MSX_CAT2(a, b) a##b
#define MSX_CAT(a, b) MSX_CAT2(a, b)

namespace msx {

namespace {
constexpr int kN = MSX_N;
// threads per row / column line of plan variant PV
template <int PV> constexpr int kT = RowPlanSel<kN, PV>::type::T;
template <int PV> constexpr int kTc = ColPlanSel<kN, PV>::type::T;
constexpr bool kRowAlt = RowPlanSel<kN, 1>::alt, kColAlt = ColPlanSel<kN, 1>::alt;
// the column tile widths below are those of variant 0; variant 1 keeps its
// threads per line, so every width fits both
static_assert(!kColAlt || kTc<1> == kTc<0>, "column plan variants share the thread count");

// lines per work-group of the row sweep: 256 threads (4 wavefronts)
template <int PV> constexpr int row_lpw() { return kT<PV> >= 256 ? 1 : 256 / kT<PV>; }

// columns per work-group of the column sweep (coalescing width), bounded by
// 1024 threads and by the LDS tile N*(TW+1) complex values
#ifndef MSX_COL_TW
#define MSX_COL_TW 16
#endif
// widest tile tried at N = 512, measured on gfx950: two 8-column work-groups per
// CU (4 wavefronts each, 74 KB of double-buffered LDS each) run the 256-wave
// slice sweep in 180 us against 198 us for one 16-column group of 8
// wavefronts; the halves of each 128-B line that neighbouring groups read meet
// in L2 because neighbouring groups run on one XCD.  At 256 the 16-column tile
// is faster; from 1024 on the LDS bound picks the width.
#ifndef MSX_COL_TW512
#define MSX_COL_TW512 8
#endif
#ifndef MSX_COL_DB
#define MSX_COL_DB 1
#endif
#ifndef MSX_COL_THREADS
#define MSX_COL_THREADS 512
#endif
template <typename C> constexpr int fit_tw(int tw) {
    while (tw > 1 && (kTc<0> * tw > MSX_COL_THREADS || (long)kN * (tw + 1) * (long)sizeof(C) > 80 * 1024)) tw /= 2;
    return tw;
}
template <typename C> constexpr int col_tw() { return fit_tw<C>(kN == 512 ? MSX_COL_TW512 : MSX_COL_TW); }
// the other width the autotuner may choose (half or double the built-in one;
// equal to it where neither fits or the line is short)
template <typename C> constexpr int col_tw_alt() {
    if (kN < 256) return col_tw<C>();
    const int d = col_tw<C>(), up = fit_tw<C>(2 * d);
    return up != d ? up : (d > 4 ? d / 2 : d);
}
template <typename C> constexpr int tw_of(int tile) { return tile == col_tw_alt<C>() ? col_tw_alt<C>() : col_tw<C>(); }

template <typename K> hipError_t set_smem(K kernel, size_t bytes) {
    if (bytes > 64 * 1024)
        return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    return hipSuccess;
}

// waves per work-group: enough work-groups to fill 256 CUs several times over,
// up to 8 waves per work-group to amortise the per-thread constants
inline int pick_wpb(int requested, long long units, long long target_blocks) {
    if (requested > 0) return requested;
    long long w = units / target_blocks;
    return (int)(w < 1 ? 1 : (w > 8 ? 8 : w));
}

template <typename C, int IN, int OUT, bool VREAL = false, bool INTEN = false, int PV = 0>
hipError_t row_go(const RowArgs<C>& a0, hipStream_t s) {
    constexpr int LPW = row_lpw<PV>();
    RowArgs<C> a = a0;
    const int nxg = (a.nx + LPW - 1) / LPW;
    a.wpb = pick_wpb(a.wpb, (long long)nxg * a.nb, 4096);
    const long long nchunk = (a.nb + a.wpb - 1) / a.wpb;
    const unsigned grid = (unsigned)(nxg * nchunk);
    const size_t smem = (size_t)LPW * (kN + kN / 16) * sizeof(C);
    auto k = row_sweep<C, kN, IN, OUT, LPW, VREAL, INTEN, PV>;
    hipError_t e = set_smem(k, smem);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3(grid), dim3(kT<PV> * LPW), smem, s, a);
    return hipGetLastError();
}

// work-groups of kernel k (block threads, smem bytes) resident on the whole
// device at once, cached per device and LDS size in the caller's
// per-kernel cache ((smem << 24) | (groups + 1); 0: empty)
constexpr int kDev = 16;
template <typename K>
hipError_t resident_groups(K k, int threads, size_t smem, std::atomic<long long> (&cache)[kDev], long long* out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < kDev) {
        const long long c = cache[dev].load(std::memory_order_relaxed);
        if (c != 0 && (size_t)(c >> 24) == smem) {
            *out = (c & 0xffffff) - 1;
            return hipSuccess;
        }
    }
    int per_cu = 0, cus = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(k), threads, smem);
    if (e != hipSuccess) return e;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    *out = std::max(1LL, (long long)per_cu * cus);
    if (dev < kDev && *out < 0xffffff) cache[dev].store(((long long)smem << 24) | (*out + 1), std::memory_order_relaxed);
    return hipSuccess;
}

template <typename C, int IN, int OUT, int OPS, int TW, int PV>
hipError_t col_go_tw(const ColArgs<C>& a0, hipStream_t s) {
    ColArgs<C> a = a0;
    const int groups = (a.ny + TW - 1) / TW;
    constexpr size_t tile = (size_t)kN * (TW + 1) * sizeof(C);
    // double-buffer the exchanges when two tiles fit in the work-group LDS
    // limit (a launch asking for the full 160 KiB is refused)
    constexpr bool DB = MSX_COL_DB && 2 * tile <= 150 * 1024;
    auto k = col_sweep<C, kN, IN, OUT, TW, OPS, DB, PV>;
    size_t smem = col_list_offset<C, kN, TW, DB>();  // exchange tiles, detector reduction scratch
    unsigned grid;
    if (a.wpb <= 0 || (OPS & kColJobs) != 0) {
        // balanced schedule (the built-in one, and the only one of uneven
        // jobs): -wpb work-groups per resident
        // slot (0: one), each with its list of (group, live mask) pairs after
        // the exchange tiles
        a.wpb = a.wpb < 0 ? a.wpb : -1;
        smem += (size_t)groups * sizeof(int2);
        hipError_t e = set_smem(k, smem);
        if (e != hipSuccess) return e;
        static std::atomic<long long> cache[kDev];
        long long slots = 0;
        e = resident_groups(k, kTc<PV> * TW, smem, cache, &slots);
        if (e != hipSuccess) return e;
        const long long g = std::min(slots * std::min(-a.wpb, 8), (long long)groups * a.nb);
        grid = (unsigned)std::max(1LL, g);
    } else {
        const long long nchunk = (a.nb + a.wpb - 1) / a.wpb;
        grid = (unsigned)(groups * nchunk);
        hipError_t e = set_smem(k, smem);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k, dim3(grid), dim3(kTc<PV> * TW), smem, s, a);
    return hipGetLastError();
}
// dispatch on the tile width chosen in the arguments
template <typename C, int IN, int OUT, int OPS, int PV = 0> hipError_t col_go_pv(const ColArgs<C>& a, hipStream_t s) {
    constexpr int D = col_tw<C>(), A = col_tw_alt<C>();
    if constexpr (A != D) {
        if (a.tile == A) return col_go_tw<C, IN, OUT, OPS, A, PV>(a, s);
    }
    return col_go_tw<C, IN, OUT, OPS, D, PV>(a, s);
}
template <typename C, int IN, int OUT, int OPS> hipError_t col_go(const ColArgs<C>& a, hipStream_t s) {
    return col_go_pv<C, IN, OUT, OPS, 0>(a, s);
}
// the slice loop's sweeps also dispatch on the plan variant in the arguments
template <typename C, int IN, int OUT, int OPS> hipError_t col_go_plan(const ColArgs<C>& a, hipStream_t s) {
    if constexpr (kColAlt) {
        if (a.plan == 1) return col_go_pv<C, IN, OUT, OPS, 1>(a, s);
    }
    return col_go_pv<C, IN, OUT, OPS, 0>(a, s);
}
template <typename C, int IN, int OUT> hipError_t row_go_plan(const RowArgs<C>& a, hipStream_t s) {
    if constexpr (kRowAlt) {
        if (a.plan == 1) return row_go<C, IN, OUT, false, false, 1>(a, s);
    }
    return row_go<C, IN, OUT>(a, s);
}
}  // namespace

template <typename C> hipError_t MSX_CAT(launch_row_, MSX_N)(int mode, const RowArgs<C>& a, hipStream_t s) {
    if (a.inten) {  // |psi|^2 output: only after the inverse transform
        if (mode == kRowInv && !a.vreal) return row_go<C, -1, 0, false, true>(a, s);
        return hipErrorInvalidValue;
    }
    if (a.vreal) {
        switch (mode) {
            case kRowFwd: return row_go<C, 0, 1, true>(a, s);
            case kRowNone: return row_go<C, 0, 0, true>(a, s);
            default: return hipErrorInvalidValue;
        }
    }
    switch (mode) {
        case kRowFwd: return row_go_plan<C, 0, 1>(a, s);
        case kRowSlice: return row_go_plan<C, -1, 1>(a, s);
        case kRowInv: return row_go_plan<C, -1, 0>(a, s);
        case kRowInv0: return row_go<C, 0, -1>(a, s);
        case kRowNone: return row_go<C, 0, 0>(a, s);
        default: return hipErrorInvalidValue;
    }
}

template <typename C> hipError_t MSX_CAT(launch_col_, MSX_N)(int mode, const ColArgs<C>& a, hipStream_t s) {
    const int ops = a.ops & ~kColStore;
    constexpr int kPD = kColProp | kColDetect;
    switch (mode) {
        case kColSlice:
            if (ops == kColProp) return col_go_plan<C, 1, -1, kColProp>(a, s);
            if (ops == (kColProp | kColObl)) return col_go<C, 1, -1, kColProp | kColObl>(a, s);
            if (ops == (kColProp | kColJobs)) {
                if (!a.prop_tab || !a.queue) return hipErrorInvalidValue;
                // the plan of the single-job loop, so that a job gives the
                // same bits alone and batched on a tuned grid
                return col_go_plan<C, 1, -1, kColProp | kColJobs>(a, s);
            }
            if (ops == kPD) return col_go_plan<C, 1, -1, kPD>(a, s);
            if (ops == kColMask) return col_go<C, 1, -1, kColMask>(a, s);
            if (ops == kColCtf) return col_go<C, 1, -1, kColCtf>(a, s);
            if (ops == 0) return col_go<C, 1, -1, 0>(a, s);
            break;
        case kColFwd:
            if (ops == kPD) return col_go_plan<C, 1, 0, kPD>(a, s);
            if (ops == 0) return col_go<C, 1, 0, 0>(a, s);
            break;
        case kColGenInv:
            if (ops == kColGen) return col_go_plan<C, 0, -1, kColGen>(a, s);
            break;
        case kColInv:
            if (ops == 0) return col_go<C, -1, 0, 0>(a, s);
            break;
        default:
            break;
    }
    return hipErrorInvalidValue;
}

template <typename C> int MSX_CAT(col_groups_, MSX_N)(int ny, int tile) {
    const int tw = tw_of<C>(tile);
    return (ny + tw - 1) / tw;
}
template <typename C> int MSX_CAT(col_tiles_, MSX_N)(int* def, int* alt) {
    *def = col_tw<C>();
    *alt = col_tw_alt<C>();
    return 0;
}

template hipError_t MSX_CAT(launch_row_, MSX_N)<float2>(int, const RowArgs<float2>&, hipStream_t);
template hipError_t MSX_CAT(launch_row_, MSX_N)<double2>(int, const RowArgs<double2>&, hipStream_t);
template hipError_t MSX_CAT(launch_col_, MSX_N)<float2>(int, const ColArgs<float2>&, hipStream_t);
template hipError_t MSX_CAT(launch_col_, MSX_N)<double2>(int, const ColArgs<double2>&, hipStream_t);
int MSX_CAT(plan_alts_, MSX_N)(int* row, int* col) {
    *row = kRowAlt ? 1 : 0;
    *col = kColAlt ? 1 : 0;
    return 0;
}
template int MSX_CAT(col_groups_, MSX_N)<float2>(int, int);
template int MSX_CAT(col_groups_, MSX_N)<double2>(int, int);
template int MSX_CAT(col_tiles_, MSX_N)<float2>(int*, int*);
template int MSX_CAT(col_tiles_, MSX_N)<double2>(int*, int*);

}  // namespace msx
