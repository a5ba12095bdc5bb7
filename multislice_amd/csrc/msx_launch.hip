// msx_launch.hip -- sweep-kernel instantiations and launchers for one line
// length.  Compiled once per supported length with -DMSX_N=<N> (the build
// runs these compilations in parallel); each object defines
// launch_row_<N>, launch_col_<N> and col_groups_<N> for float2 and double2.
#include "msx_internal.h"

#ifndef MSX_N
#error "compile with -DMSX_N=<line length>"
#endif

#define MSX_CAT2(a, b) a##b
#define MSX_CAT(a, b) MSX_CAT2(a, b)

namespace msx {

namespace {
constexpr int kN = MSX_N;
constexpr int kT = Plan<kN>::T;

// lines per work-group of the row sweep: 256 threads (4 wavefronts)
constexpr int row_lpw() { return kT >= 256 ? 1 : 256 / kT; }

// columns per work-group of the column sweep (coalescing width), bounded by
// 1024 threads and by the LDS tile N*(TW+1) complex values
template <typename C> constexpr int col_tw() {
    int tw = 16;
    while (tw > 1 && (kT * tw > 1024 || (long)kN * (tw + 1) * (long)sizeof(C) > 80 * 1024)) tw /= 2;
    return tw;
}

template <typename K> hipError_t set_smem(K kernel, size_t bytes) {
    if (bytes > 64 * 1024)
        return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    return hipSuccess;
}

template <typename C, int IN, int OUT> hipError_t row_go(const RowArgs<C>& a, hipStream_t s) {
    constexpr int LPW = row_lpw();
    const long long nlines = (long long)a.nx * a.nb;
    const unsigned grid = (unsigned)((nlines + LPW - 1) / LPW);
    const size_t smem = (size_t)LPW * (kN + kN / 16) * sizeof(C);
    auto k = row_sweep<C, kN, IN, OUT, LPW>;
    hipError_t e = set_smem(k, smem);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3(grid), dim3(kT * LPW), smem, s, a);
    return hipGetLastError();
}

template <typename C, int IN, int OUT> hipError_t col_go(const ColArgs<C>& a, hipStream_t s) {
    constexpr int TW = col_tw<C>();
    const int groups = (a.ny + TW - 1) / TW;
    const unsigned grid = (unsigned)groups * (unsigned)a.nb;
    size_t smem = (size_t)kN * (TW + 1) * sizeof(C);
    if (smem < 2048) smem = 2048;  // detector reduction scratch
    auto k = col_sweep<C, kN, IN, OUT, TW>;
    hipError_t e = set_smem(k, smem);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3(grid), dim3(kT * TW), smem, s, a);
    return hipGetLastError();
}
}  // namespace

template <typename C> hipError_t MSX_CAT(launch_row_, MSX_N)(int mode, const RowArgs<C>& a, hipStream_t s) {
    switch (mode) {
        case kRowFwd: return row_go<C, 0, 1>(a, s);
        case kRowSlice: return row_go<C, -1, 1>(a, s);
        case kRowInv: return row_go<C, -1, 0>(a, s);
        case kRowInv0: return row_go<C, 0, -1>(a, s);
        case kRowNone: return row_go<C, 0, 0>(a, s);
        default: return hipErrorInvalidValue;
    }
}

template <typename C> hipError_t MSX_CAT(launch_col_, MSX_N)(int mode, const ColArgs<C>& a, hipStream_t s) {
    switch (mode) {
        case kColSlice: return col_go<C, 1, -1>(a, s);
        case kColGenInv: return col_go<C, 0, -1>(a, s);
        case kColFwd: return col_go<C, 1, 0>(a, s);
        case kColInv: return col_go<C, -1, 0>(a, s);
        case kColPoint: return col_go<C, 0, 0>(a, s);
        default: return hipErrorInvalidValue;
    }
}

template <typename C> int MSX_CAT(col_groups_, MSX_N)(int ny) {
    constexpr int TW = col_tw<C>();
    return (ny + TW - 1) / TW;
}

template hipError_t MSX_CAT(launch_row_, MSX_N)<float2>(int, const RowArgs<float2>&, hipStream_t);
template hipError_t MSX_CAT(launch_row_, MSX_N)<double2>(int, const RowArgs<double2>&, hipStream_t);
template hipError_t MSX_CAT(launch_col_, MSX_N)<float2>(int, const ColArgs<float2>&, hipStream_t);
template hipError_t MSX_CAT(launch_col_, MSX_N)<double2>(int, const ColArgs<double2>&, hipStream_t);
template int MSX_CAT(col_groups_, MSX_N)<float2>(int);
template int MSX_CAT(col_groups_, MSX_N)<double2>(int);

}  // namespace msx
