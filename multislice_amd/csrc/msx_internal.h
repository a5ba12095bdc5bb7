// msx_internal.h -- declarations shared by the engine's translation units
// (not part of the public C ABI, which is include/msx.h).
#pragma once
#include <hip/hip_runtime.h>

#include "msx_sweep.hpp"

namespace msx {

// Row-sweep modes: (IN, OUT) FFT directions around the pointwise stage.
enum RowMode : int {
    kRowFwd = 0,    // (0, +1)
    kRowSlice = 1,  // (-1, +1)
    kRowInv = 2,    // (-1, 0)
    kRowInv0 = 3,   // (0, -1)
    kRowNone = 4,   // (0, 0)
};
// Column-sweep modes.
enum ColMode : int {
    kColSlice = 0,  // (+1, -1)
    kColGenInv = 1, // (0, -1)
    kColFwd = 2,    // (+1, 0)
    kColInv = 3,    // (-1, 0)
    kColPoint = 4,  // (0, 0)
};

// Supported line lengths (both axes).
#define MSX_FOR_EACH_N(X) X(16) X(32) X(64) X(128) X(256) X(384) X(512) X(768) X(1024) X(2048)

#define MSX_DECL_N(NN)                                                                         \
    template <typename C> hipError_t launch_row_##NN(int mode, const RowArgs<C>& a, hipStream_t s); \
    template <typename C> hipError_t launch_col_##NN(int mode, const ColArgs<C>& a, hipStream_t s); \
    template <typename C> int col_groups_##NN(int ny, int tile);                                 \
    template <typename C> int col_tiles_##NN(int* def, int* alt);                               \
    int plan_alts_##NN(int* row, int* col);
MSX_FOR_EACH_N(MSX_DECL_N)
#undef MSX_DECL_N

inline bool supported_n(int n) {
#define MSX_CASE(NN) \
    if (n == NN) return true;
    MSX_FOR_EACH_N(MSX_CASE)
#undef MSX_CASE
    return false;
}

template <typename C> hipError_t launch_row(int n, int mode, const RowArgs<C>& a, hipStream_t s) {
#define MSX_CASE(NN) \
    if (n == NN) return launch_row_##NN<C>(mode, a, s);
    MSX_FOR_EACH_N(MSX_CASE)
#undef MSX_CASE
    return hipErrorInvalidValue;
}
template <typename C> hipError_t launch_col(int n, int mode, const ColArgs<C>& a, hipStream_t s) {
#define MSX_CASE(NN) \
    if (n == NN) return launch_col_##NN<C>(mode, a, s);
    MSX_FOR_EACH_N(MSX_CASE)
#undef MSX_CASE
    return hipErrorInvalidValue;
}
// number of column groups (work-groups per wave) of the column sweep for an
// nx-long line and ny columns
template <typename C> int col_groups(int nx, int ny, int tile = 0) {
#define MSX_CASE(NN) \
    if (nx == NN) return col_groups_##NN<C>(ny, tile);
    MSX_FOR_EACH_N(MSX_CASE)
#undef MSX_CASE
    return 0;
}
// whether an n-long line has an alternative FFT plan for the row sweep
// (RowPlanSel<n, 1>) and for the column sweep (ColPlanSel<n, 1>)
inline void plan_alts(int n, int* row, int* col) {
    *row = *col = 0;
#define MSX_CASE(NN) \
    if (n == NN) plan_alts_##NN(row, col);
    MSX_FOR_EACH_N(MSX_CASE)
#undef MSX_CASE
}
// the built-in column tile width of an nx-long line and the alternative the
// autotuner may pick (equal when there is none)
template <typename C> void col_tiles(int nx, int* def, int* alt) {
    *def = *alt = 0;
#define MSX_CASE(NN) \
    if (nx == NN) col_tiles_##NN<C>(def, alt);
    MSX_FOR_EACH_N(MSX_CASE)
#undef MSX_CASE
}

// potential splat (msx_potential.hip)
struct AtomRec {
    double x, y;   // Angstrom, wrapped into [0,a) x [0,b)
    float occ;
    int16_t Z;     // 1..103
    int16_t slot;  // index of Z among the call's species (radial tables)
};
// Radial tables of the FP32 splat: per species, V_Z(s) at s = r^2 =
// 2^(o + omin) * (1 + (j - 1) / kRadPts), j = 0 .. kRadPts + 2, for octaves
// o = 0 .. noct - 1 (a guard point either side of each octave's kRadPts
// intervals), evaluated by the same FP64 formula as the exact splat without
// its clamp r >= 1e-10 (so the tabulated function stays smooth), followed by
// one entry holding the clamped value V_Z(1e-10) for s < 1e-20.
// The splat reads the cubic through the four points around each interval as
// power coefficients in t (position in the interval, 0..1): per species
// [noct][kRadPts] intervals of (c0, c1) (c2, c3), built from the points.
constexpr int kRadPts = 256;
// lowest octave: s = r^2 from 2^-67 A^2, below 1e-20 = 2^-66.4
constexpr int kRadOmin = -67;
struct SplatArgs {
    const int* cell_start;   // [nslices*nbx*nby + 1] start of each (slice, bin) in atoms
    const AtomRec* atoms;    // sorted by (slice, bin x, bin y)
    double* v64;             // FP64 output [nslices][nx][ny], or nullptr
    float* v32;              // FP32 output [nslices][nx][ny], or nullptr
    double a, b, radius;
    int nx, ny, nslices, nbx, nby, table, formula;
    const double* rtab;      // radial tables [nsp][rt_noct * (kRadPts + 3) + 1] (FP32 output), or nullptr: exact
    const double2* rcoef;    // their interval coefficients [nsp][rt_noct * kRadPts][2]
    int rt_noct;
    // filled by launch_splat: pixel and bin widths (a / nx, b / ny, a / nbx,
    // b / nby; IEEE division on the host gives the kernel's own quotients)
    double dx, dy, bwx, bwy;
};
struct RadArgs {
    double* tab;    // [nsp][noct * (kRadPts + 3) + 1]
    double2* coef;  // [nsp][noct * kRadPts][2]
    int zs[104];  // Z of each species slot
    int nsp, noct, table, formula;
};
// Slab geometry of the splat's binning: slices of thickness dz, nbx x nby
// lateral bins of the cell a x b.
struct BinGeom {
    double a, b, dz;
    int nslices, nbx, nby;
};
// Slice and lateral bin of one atom, with x, y wrapped into the cell; false
// for an atom outside the slab.  Slice rule of ModelPotential.cu:63-66
// (kz*dz <= z < (kz+1)*dz, the floor corrected by one where rounding puts z
// across a boundary).  Evaluated identically on the host and the GPU (no
// contraction), so both bin and wrap every atom the same way; bins are
// clamped into range for any finite input.
__host__ __device__ inline bool bin_atom(const BinGeom& g, double x, double y, double z, long long* cell,
                                         double* xw, double* yw) {
#pragma clang fp contract(off)
    const double fz = floor(z / g.dz);
    if (!(fz >= 0.0 && fz < (double)g.nslices)) return false;
    long long kz = (long long)fz;
    if (!(kz * g.dz <= z && z < (kz + 1) * g.dz)) kz += (z < kz * g.dz) ? -1 : 1;
    if (kz < 0 || kz >= g.nslices) return false;
    x = x - g.a * floor(x / g.a);
    y = y - g.b * floor(y / g.b);
    if (x >= g.a) x -= g.a;
    if (y >= g.b) y -= g.b;
    const double fx = x / (g.a / g.nbx), fy = y / (g.b / g.nby);
    const int bx = fx >= g.nbx - 1 ? g.nbx - 1 : (fx > 0.0 ? (int)fx : 0);
    const int by = fy >= g.nby - 1 ? g.nby - 1 : (fy > 0.0 ? (int)fy : 0);
    *cell = (kz * g.nbx + bx) * g.nby + by;
    *xw = x;
    *yw = y;
    return true;
}
hipError_t launch_splat(const SplatArgs& a, hipStream_t s);
hipError_t launch_radial_tables(const RadArgs& a, hipStream_t s);

// Binning of device-resident atoms (msx_binning.hip): the same records and
// cell starts as the host binning (stable sort by cell, atoms of a cell in
// input order).  flags[0]: index of the first invalid atom (Z outside 1..103
// or a coordinate not finite), INT_MAX if none; flags[Z] = 1 for each Z
// present.  key/val/key2/val2: n ints each.
struct AtomIO;
struct DevBinArgs {
    const AtomIO* atoms;
    int n;
    BinGeom g;
    int ncell;
    unsigned *key, *key2;  // cell keys (sorted into key2)
    int *val, *val2, *flags;
    void* temp;
    size_t temp_bytes;
};
struct SlotMap {
    int16_t s[104];
};
hipError_t bin_sort_temp_bytes(int n, int ncell, size_t* bytes);
hipError_t launch_bin_keys(const DevBinArgs& a, hipStream_t s);
hipError_t launch_bin_sort(const DevBinArgs& a, int* cell_start, AtomRec* recs, const SlotMap& slots,
                           hipStream_t s);
hipError_t upload_fparams(hipStream_t s);

// frozen-phonon displacement (msx_potential.hip); u_rms in device memory [104]
struct AtomIO {
    int32_t Z;
    float occ;
    double x, y, z;
};
hipError_t launch_displace(const AtomIO* in, long long n, const double* u_rms, unsigned long long seed, int config,
                           AtomIO* out, hipStream_t s);

// reciprocal-space potential (msx_potential.hip)
struct RecAtom {
    double x, y;  // fractional coordinates in the unit cell
    double occ;
};
struct RecSeg {            // atoms [start, end) of one slice sharing (Z, B, ads)
    int start, end, Z, pad;
    double dw, ads;
};
struct RecArgs {
    const RecAtom* atoms;
    const RecSeg* segs;
    const int* slice_seg;  // segments of slice kz: [slice_seg[kz], slice_seg[kz+1])
    double2* out;          // V(g) of slices kz0 .. kz0+nslices-1: [slice][nx][ny]
    double A, B, cg, inv_area;  // reciprocal lengths 1/(a sin gam), 1/(b sin gam); cos gam; 1/area
    int nx, ny, nmult, mmult, kz0;
    // tabulated curves (instead of Doyle-Turner): [type][npoints] at |g|/2 = i*dds,
    // already including the Debye-Waller factor and 1/area; seg.Z = type
    const double* curves;
    int npoints;
    double dds;
};
hipError_t launch_rec_sf(const RecArgs& a, int nslices, hipStream_t s);
hipError_t launch_split_potential(const double2* w, long long n, double scale, void* V, void* Vabs, bool f64,
                                  hipStream_t s);

// real-space Doyle-Turner potential over pixel boxes (msx_potential.hip)
struct RsAtom {
    double xf, yf, occ, dw, ads;   // fractional position, occupancy, Debye-Waller B, absorptive fraction
    int Z, ibb, iee, jbb, jee;     // box: pixel index bounds before clipping
    int pad;
};
struct RsArgs {
    const RsAtom* atoms;
    const int2* entries;       // (atom, image 0..8) per tile, grouped by (slice, tile)
    const int* tile_start;     // [nslices * ntx * nty + 1]
    double* v64;               // outputs [nslices][nx][ny] (one of v64 / v32), scaled by `scale`
    float* v32;
    double* a64;               // absorptive part (optional)
    float* a32;
    double a, b, cg, dxa, dyb, scale;
    int nx, ny, ntx, nty, nslices;
    // tabulated curves (instead of Doyle-Turner): [type][npoints] at r = i*dds; atom.Z = type
    const double* curves;
    int npoints;
    double dds;
};
hipError_t launch_rs_dt(const RsArgs& a, hipStream_t s);

// detector finish: out[d][p] += w * sum_g partial[d][b][g] (p = probe0 + b)
// bandwidth probes (msx_copy_probe): contiguous copy of n complex64 values,
// and a copy of nimg 512 x ny images with the column sweep's access pattern
// (tw adjacent columns per work-group, 32 threads x 16 rows per column, XCD
// order)
hipError_t launch_stream_copy(const float2* src, float2* dst, long long n, hipStream_t s);
hipError_t launch_tile_copy(const float2* src, float2* dst, int nimg, int ny, int tw, hipStream_t s);
hipError_t launch_det_finish(const double* partial, int ndet, int nb, int groups, double w, double* out,
                             long long out_stride_d, long long probe0, hipStream_t s);

// partial coherence: tilted incident waves (iq: device [nb] (iqx, iqy)) and
// the FP64 sum of per-wave intensities into one image
template <typename C>
hipError_t launch_tilt_waves(C* psi, int nb, int nx, int ny, const int2* iq, hipStream_t s);
template <typename Re>
hipError_t launch_wave_sum(const Re* inten, int nb, long long img, double scale, double* image, hipStream_t s);

// thickness-series records of the mixed-domain batch m after a column sweep
// (msx_series.hip): beams out[b][i] = scale * sum_x m[b][x][k_i]
// exp(+2 pi i h_i x / nx) for bm[i] = (h_i, k_i, in band, 0) (0 when out of
// band); per-wave partial sums of |m|^2 (columns with ky2 >= k2max skipped
// when ky2 is set) into partial[nb][nparts]
template <typename C>
hipError_t launch_series_beams(const C* m, int nb, int nx, int ny, const int4* bm, int nbeams, double scale, C* out,
                               hipStream_t s);
template <typename C>
hipError_t launch_series_norm(const C* m, int nb, long long img, int ny, const float* ky2, float k2max, int nparts,
                              double* partial, hipStream_t s);

// error reporting shared by the C ABI units: records the message for
// msx_last_error() and returns `code`
int fail(int code, const char* fmt, ...);

}  // namespace msx

// context accessors for C ABI units other than msx_api.hip
struct msx_ctx_s;
int msx_ctx_device(const msx_ctx_s* ctx);
hipStream_t msx_ctx_stream(const msx_ctx_s* ctx);
