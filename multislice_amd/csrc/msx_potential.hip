// msx_potential.hip -- projected-potential splat (the GPU counterpart of
// LMultiSlice calculatePotentialGridGPU, ModelPotential.cu:125-230) and the
// detector finishing reduction.
//
// Differences from the reference kernel, by design:
//  * atoms are binned per (slice, lateral bin) on the host or the GPU, bins
//    at least `radius` wide, so a pixel tile only visits the bins within reach
//    (the reference loops every pixel over every atom of the slice);
//  * the whole slab is one launch over (tile, slice) instead of one launch per
//    4608-atom constant-memory batch per slice;
//  * distances and the fparams sums are FP64 (the reference computes
//    distances in float), which removes the small-r rounding of K0's log.
// The per-pixel contribution rule is the reference's: minimum image per axis,
// r < radius, r clamped >= 1e-10, formula and table selectable.
#include "msx_internal.h"
#include "msx_philox.h"
#include "fparams_tables.inc"
#include "doyle_turner.inc"

namespace msx {

__constant__ double c_fparams[2][103][12];
__constant__ double c_dt[93][8];

hipError_t upload_fparams(hipStream_t s) {
    hipError_t e = hipMemcpyToSymbolAsync(HIP_SYMBOL(c_fparams), MSX_FPARAMS, sizeof(MSX_FPARAMS), 0,
                                          hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    return hipMemcpyToSymbolAsync(HIP_SYMBOL(c_dt), MSX_DOYLE_TURNER, sizeof(MSX_DOYLE_TURNER), 0,
                                  hipMemcpyHostToDevice, s);
}

// Abramowitz & Stegun 9.8.1-9.8.6 polynomials, as slicelib.cpp:168-237 and
// ModelPotential.cu:256-292.
__device__ __forceinline__ double d_bessi0(double x) {
    const double ax = fabs(x);
    if (ax <= 3.75) {
        double t = x / 3.75;
        t = t * t;
        return 1.0 + t * (3.5156229 + t * (3.0899424 + t * (1.2067492 + t * (0.2659732 +
                    t * (0.0360768 + t * 0.0045813)))));
    }
    const double t = 3.75 / ax;
    const double sum = 0.39894228 + t * (0.01328592 + t * (0.00225319 + t * (-0.00157565 +
                       t * (0.00916281 + t * (-0.02057706 + t * (0.02635537 + t * (-0.01647633 +
                       t * 0.00392377)))))));
    return exp(ax) * sum / sqrt(ax);
}
__device__ __forceinline__ double d_bessk0(double x) {
    const double ax = fabs(x);
    if (ax > 0.0 && ax <= 2.0) {
        double x2 = ax / 2.0;
        x2 = x2 * x2;
        const double sum = -0.57721566 + x2 * (0.42278420 + x2 * (0.23069756 + x2 * (0.03488590 +
                           x2 * (0.00262698 + x2 * (0.00010750 + x2 * 0.00000740)))));
        return -log(ax / 2.0) * d_bessi0(x) + sum;
    } else if (ax > 2.0) {
        const double x2 = 2.0 / ax;
        const double sum = 1.25331414 + x2 * (-0.07832358 + x2 * (0.02189568 + x2 * (-0.01062446 +
                           x2 * (0.00587872 + x2 * (-0.00251540 + x2 * 0.00053208)))));
        return exp(-ax) * sum / sqrt(ax);
    }
    return 1.0e20;
}

// vzatom (slicelib.cpp:2666-2699) / calculateProjectedPotential
// (ModelPotential.cu:232-254)
// CLAMP: r < 1e-10 is evaluated at 1e-10, as the reference does; the radial
// tables are filled without it (smooth below the clamp; r > 0 there)
template <bool CLAMP = true>
__device__ __forceinline__ double d_vz(int Z, double r, int table, int formula) {
    const double* p = c_fparams[table][Z - 1];
    if (CLAMP && r < 1.0e-10) r = 1.0e-10;
    double suml = 0.0, sumg = 0.0;
    if (formula == 0) {
        const double pi = 3.141592654;
        const double x = 2.0 * pi * r;
        for (int i = 0; i < 6; i += 2) suml += p[i] * d_bessk0(x * sqrt(p[i + 1]));
        const double x2 = (pi * r) * (pi * r);
        for (int i = 6; i < 12; i += 2) sumg += p[i] * exp(-x2 / p[i + 1]) / p[i + 1];
        return 300.8242834 * suml + 150.4121417 * sumg;
    }
    const double dR1 = 6.2831853071796 * r;
    for (int k = 0; k < 3; k++) suml += p[2 * k] * d_bessk0(dR1 * sqrt(p[2 * k + 1]));
    for (int k = 0; k < 3; k++) sumg += (p[6 + 2 * k] / p[7 + 2 * k]) * exp(-(6.2831853071796 * r * r) / p[7 + 2 * k]);
    return 300.73079394295 * suml + 150.36539697148 * sumg;
}

// V_Z(s = r^2) from the radial table of one species (s in [2^kRadOmin,
// 2^(kRadOmin + noct))): the octave from the exponent of s, the position in
// it from the mantissa (uniform spacing within an octave), the cubic through
// the four surrounding points (Lagrange interpolation, evaluated as
// c0 + t (c1 + t (c2 + t c3)) from the interval's coefficients).  Relative
// error below 3e-9 (99.9th percentile), up to 3e-8 next to the branch
// switches of the reference's Bessel polynomials: below FP32 resolution.
// Below s = 1e-20 (r < 1e-10, where the formula clamps r) the point table's
// last entry holds the clamped value.
// With kRadPts = 2^8 intervals per octave the IEEE fields of s (bits u) give
// both parts: bits 63..44 (sign 0, biased exponent, the mantissa's top 8
// bits) are octave * kRadPts + interval once offset by
// (1023 + kRadOmin) * kRadPts, and the 44 mantissa bits below are t, exactly
// (1.f - 1 with f those bits moved to the top).
static_assert(kRadPts == 256, "the radial interval is read from 8 mantissa bits");
__device__ __forceinline__ unsigned rad_interval(unsigned long long u) {
    return (unsigned)(u >> 44) - (unsigned)((1023 + kRadOmin) * kRadPts);
}
__device__ __forceinline__ double rad_position(unsigned long long u) {
    return __longlong_as_double((long long)(0x3FF0000000000000ull | ((u << 8) & 0x000FFFFFFFFFFFFFull))) - 1.0;
}
__device__ __forceinline__ double rad_cubic(double2 c01, double2 c23, double t) {
    return c01.x + t * (c01.y + t * (c23.x + t * c23.y));
}

// per species: noct * (kRadPts + 3) grid values of the unclamped formula,
// then the value at the clamp r = 1e-10
__global__ __launch_bounds__(256) void radial_table_kernel(RadArgs p) {
    const long long per = (long long)p.noct * (kRadPts + 3);
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (per + 1) * p.nsp) return;
    const int sp = (int)(i / (per + 1));
    const long long rem = i % (per + 1);
    if (rem == per) {
        p.tab[i] = d_vz(p.zs[sp], 1.0e-10, p.table, p.formula);
        return;
    }
    const int o = (int)(rem / (kRadPts + 3)), j = (int)(rem % (kRadPts + 3));
    const double s = ldexp(1.0 + (double)(j - 1) / kRadPts, o + kRadOmin);
    p.tab[i] = d_vz<false>(p.zs[sp], sqrt(s), p.table, p.formula);
}

// power coefficients of the cubic through the points at t = -1, 0, 1, 2
// around each interval (the point table's entries k .. k + 3)
__global__ __launch_bounds__(256) void radial_coef_kernel(RadArgs p) {
    const long long per = (long long)p.noct * kRadPts;
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= per * p.nsp) return;
    const int sp = (int)(i / per);
    const long long rem = i % per;
    const int o = (int)(rem / kRadPts), k = (int)(rem % kRadPts);
    const double* q = p.tab + (long long)sp * ((long long)p.noct * (kRadPts + 3) + 1) + (long long)o * (kRadPts + 3) + k;
    const double p0 = q[0], p1 = q[1], p2 = q[2], p3 = q[3];
    p.coef[2 * i] = make_double2(p1, (-2.0 * p0 - 3.0 * p1 + 6.0 * p2 - p3) * (1.0 / 6.0));
    p.coef[2 * i + 1] = make_double2((p0 - 2.0 * p1 + p2) * 0.5, (-p0 + 3.0 * p1 - 3.0 * p2 + p3) * (1.0 / 6.0));
}

hipError_t launch_radial_tables(const RadArgs& a, hipStream_t s) {
    const long long n = (long long)a.nsp * (a.noct * (kRadPts + 3) + 1);
    hipLaunchKernelGGL(radial_table_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
    const long long nc = (long long)a.nsp * a.noct * kRadPts;
    hipLaunchKernelGGL(radial_coef_kernel, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

// TAB: V_Z from the radial tables (FP32 output); otherwise the exact FP64
// formula per pair (separate instantiations: the table kernel carries none of
// the Bessel code).
// A work-group is a tile of 16 columns (x) by kSplatRows rows (y) of one
// slice, kSplatPpt pixels per thread (two).  The atoms of the bins the tile can reach are
// taken in the order a per-pixel loop would test them: bin column bx, then
// bin rows by, atoms in cell order.  The bin rows of a column form at most
// two runs of contiguous cells (before and after the periodic wrap), so up to
// kSplatCols columns give a table of 2 kSplatCols runs, one per lane, read
// with one vector load per lane and published in LDS as (start in the
// concatenated list, atom offset).  The work-group then loads the
// concatenated runs 256 atoms at a time, one per thread,
// keeps those whose minimum-image distance to the tile centre allows a pixel
// of the tile within the cutoff (a conservative bound: every pixel-atom pair
// it drops fails the per-pixel test), and compacts them in LDS in that order
// (wavefront ballots).  Every thread then sums over that list: the same pairs
// in the same order as testing every atom of the bins, so the potential is
// bit-identical to a per-pixel loop over the bins.
// A kept atom is tagged with its species slot and with whether any pixel of
// the tile can take its periodic image (|x_atom - x_pixel| >= a/2 on an axis,
// bounded by the distance to the tile centre plus the half extent, with a
// margin far above rounding).  Untagged atoms, nearly all of them, skip the
// minimum-image selects: the signed differences square to the same bits.
// The per-work-group set-up (bin bounds, wrapped first bins) runs one value
// per lane, the four quotients and two remainders in one pass each.
struct SplatCand {
    double x, y;
    float occ;
    int16_t Z;
    uint16_t tag;  // slot << 1 | periodic image possible
};
static_assert(sizeof(SplatCand) == 24, "candidate record");
constexpr int kSplatCols = 32;
// pixels per thread, 8 rows apart: a tile is 16 columns (x) by kSplatRows
// rows (y)
#ifndef MSX_SPLAT_PPT
#define MSX_SPLAT_PPT 2
#endif
constexpr int kSplatPpt = MSX_SPLAT_PPT;
constexpr int kSplatRows = 16 * kSplatPpt;

template <bool TAB>
__global__ __launch_bounds__(256) void splat_kernel(SplatArgs p) {
    __shared__ SplatCand cand[256];
    __shared__ int wcnt[4];
    // per run of the current column group: its first position in the
    // concatenated list and (atom index - position)
    __shared__ int2 runs[2 * kSplatCols];
    // each wavefront covers 8 columns by 8 kSplatPpt rows, a quarter of the
    // tile (a compact footprint leaves fewer lanes idle outside the radius of
    // an atom the wavefront evaluates); a thread takes rows iy0 + 8 m
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int iy0 = blockIdx.x * kSplatRows + (wv & 1) * (kSplatRows / 2) + (lane & 7);
    const int ix = blockIdx.y * 16 + (wv >> 1) * 8 + (lane >> 3);
    const int kz = blockIdx.z;
    const double dx = p.dx, dy = p.dy;
    const double xp = ix * dx;
    double yp[kSplatPpt], v[kSplatPpt];
#pragma unroll
    for (int m = 0; m < kSplatPpt; ++m) {
        yp[m] = (iy0 + 8 * m) * dy;
        v[m] = 0.0;
    }
    const double r2max = p.radius * p.radius;
    // bins reached by this tile: floor((t d -/+ radius) / w) for its first and
    // last pixel t on each axis, one bound per lane (lanes 0-3: x low, x high,
    // y low, y high)
    const bool ly = (lane & 2) != 0, lhi = (lane & 1) != 0;
    const int tl = ly ? (int)blockIdx.x * kSplatRows + (lhi ? kSplatRows - 1 : 0) : (int)blockIdx.y * 16 + (lhi ? 15 : 0);
    const double el = (double)tl * (ly ? dy : dx) + (lhi ? p.radius : -p.radius);
    const int ql = (int)floor(el / (ly ? p.bwy : p.bwx));
    int bxa = __builtin_amdgcn_readlane(ql, 0), bxb = __builtin_amdgcn_readlane(ql, 1);
    int bya = __builtin_amdgcn_readlane(ql, 2), byb = __builtin_amdgcn_readlane(ql, 3);
    if (bxb - bxa + 1 >= p.nbx) { bxa = 0; bxb = p.nbx - 1; }
    if (byb - bya + 1 >= p.nby) { bya = 0; byb = p.nby - 1; }
    // the first bin column and row wrapped into the cell (lanes 0 and 1)
    const int vl = (lane & 1) ? bya : bxa, ml = (lane & 1) ? p.nby : p.nbx;
    int wl = vl % ml;
    if (wl < 0) wl += ml;
    const int bxw_a = __builtin_amdgcn_readlane(wl, 0), byw_a = __builtin_amdgcn_readlane(wl, 1);
    // bin rows of the two runs of every column: bya up to the wrap, then
    // from row 0
    const int len1 = min(byb - bya + 1, p.nby - byw_a), len2 = byb - bya + 1 - len1;
    // the tile filter: centre and reach (half extent + cutoff, plus a margin
    // far above rounding) per axis; no filtering along an axis where the
    // reach covers half the period
    constexpr double ry = 0.5 * (kSplatRows - 1);  // half the tile's extent in rows
    const double cx = (double)(blockIdx.y * 16) * dx + 7.5 * dx, cy = (double)(blockIdx.x * kSplatRows) * dy + ry * dy;
    const double hx = 7.5 * dx + p.radius + 1e-6, hy = ry * dy + p.radius + 1e-6;
    const bool fx = hx < 0.5 * p.a, fy = hy < 0.5 * p.b;
    // the image test: no pixel of the tile takes an atom's periodic image on
    // an axis where the atom's distance to the tile centre plus the tile's
    // half extent stays below half the period (margin far above rounding)
    const double gx = 0.5 * p.a * (1.0 - 1e-9) - 7.5 * dx, gy = 0.5 * p.b * (1.0 - 1e-9) - ry * dy;
    const long long cstride = 2ll * p.rt_noct * kRadPts;
    const long long tstride = (long long)p.rt_noct * (kRadPts + 3) + 1;
    for (int g0 = bxa; g0 <= bxb; g0 += kSplatCols) {
        // run table: lane j holds run j (column g0 + j / 2, part j % 2): its
        // first atom and its atom count
        const int nrun = 2 * min(kSplatCols, bxb - g0 + 1);
        int rst = 0, rcnt = 0;
        if (lane < nrun) {
            int bxw = bxw_a + (g0 - bxa) + (lane >> 1);  // < 2 nbx: the range spans at most nbx columns
            if (bxw >= p.nbx) bxw -= p.nbx;
            const int nb = (lane & 1) ? len2 : len1;
            if (nb > 0) {
                const int cell = (kz * p.nbx + bxw) * p.nby + ((lane & 1) ? 0 : byw_a);
                rst = p.cell_start[cell];
                rcnt = p.cell_start[cell + nb] - rst;
            }
        }
        int incl = rcnt;
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(incl, o, 64);
            if (lane >= o) incl += t;
        }
        const int total = __builtin_amdgcn_readlane(incl, 63);
        // every wavefront holds the same table; the first publishes it (the
        // previous table was last read before the previous list's barrier)
        if (wv == 0 && lane < nrun) runs[lane] = make_int2(incl - rcnt, rst - (incl - rcnt));
        for (int base = 0; base < total; base += 256) {
            __syncthreads();  // the run table is written, the previous list has been read
            const int g = base + tid;
            AtomRec at{};
            bool keep = false, img = false;
            if (g < total) {
                // the run holding position g: the last one starting at or
                // before g (an empty run starts where the next one does)
                int delta = 0;
                for (int j = 0; j < nrun; ++j) {
                    const int2 e = runs[j];
                    if (g >= e.x) delta = e.y;
                }
                at = p.atoms[g + delta];
                const double rx = fabs(at.x - cx), ry = fabs(at.y - cy);
                const double ex = rx > 0.5 * p.a ? p.a - rx : rx, ey = ry > 0.5 * p.b ? p.b - ry : ry;
                keep = (!fx || ex <= hx) && (!fy || ey <= hy);
                img = !(rx < gx) || !(ry < gy);
            }
            const unsigned long long bal = __ballot(keep);
            if (lane == 0) wcnt[wv] = __popcll(bal);
            __syncthreads();
            int off = 0;
            for (int w = 0; w < wv; ++w) off += wcnt[w];
            if (keep) {
                cand[off + __popcll(bal & ((1ull << lane) - 1ull))] =
                    SplatCand{at.x, at.y, at.occ, at.Z, (uint16_t)((at.slot << 1) | (img ? 1 : 0))};
            }
            const int n = __builtin_amdgcn_readfirstlane(wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3]);
            __syncthreads();
            for (int k = 0; k < n; ++k) {
                const SplatCand c = cand[k];
                const int tag = __builtin_amdgcn_readfirstlane((int)c.tag);
                double dX = c.x - xp, dY[kSplatPpt];
#pragma unroll
                for (int m = 0; m < kSplatPpt; ++m) dY[m] = c.y - yp[m];
                if (tag & 1) {
                    dX = fabs(dX);
                    if (dX >= p.a / 2.0) dX -= p.a;
#pragma unroll
                    for (int m = 0; m < kSplatPpt; ++m) {
                        dY[m] = fabs(dY[m]);
                        if (dY[m] >= p.b / 2.0) dY[m] -= p.b;
                    }
                }
#pragma unroll
                for (int m = 0; m < kSplatPpt; ++m) {
                    const double r2 = fma(dX, dX, dY[m] * dY[m]);
                    if constexpr (TAB) {
                        const bool in = r2 < r2max, cl = r2 < 1.0e-20;
                        // V_Z(r2) from the species' table: interval and
                        // position from the bits of r2 (every lane reads a
                        // valid interval)
                        const unsigned long long u = (unsigned long long)__double_as_longlong(r2);
                        const double2* cs = p.rcoef + (tag >> 1) * cstride;
                        const unsigned idx = (in && !cl) ? rad_interval(u) : 0u;
                        const double2 c01 = cs[2 * idx], c23 = cs[2 * idx + 1];
                        if (in) {
                            double vz = rad_cubic(c01, c23, rad_position(u));
                            if (cl) vz = p.rtab[(tag >> 1) * tstride + tstride - 1];
                            v[m] += (double)c.occ * vz;
                        }
                    } else {
                        if (r2 < r2max) v[m] += (double)c.occ * d_vz(c.Z, sqrt(r2), p.table, p.formula);
                    }
                }
            }
        }
    }
#pragma unroll
    for (int m = 0; m < kSplatPpt; ++m) {
        const int iy = iy0 + 8 * m;
        if (ix < p.nx && iy < p.ny) {
            const long long o = ((long long)kz * p.nx + ix) * p.ny + iy;
            if (p.v64) p.v64[o] = v[m];
            if (p.v32) p.v32[o] = (float)v[m];
        }
    }
}

hipError_t launch_splat(const SplatArgs& args, hipStream_t s) {
    SplatArgs a = args;
    a.dx = a.a / a.nx;
    a.dy = a.b / a.ny;
    a.bwx = a.a / a.nbx;
    a.bwy = a.b / a.nby;
    dim3 grid((a.ny + kSplatRows - 1) / kSplatRows, (a.nx + 15) / 16, a.nslices);
    if (a.rtab) hipLaunchKernelGGL(splat_kernel<true>, grid, dim3(256), 0, s, a);
    else hipLaunchKernelGGL(splat_kernel<false>, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

// Frozen-phonon displacement: one thread per atom, Philox counter (atom, config)
__global__ __launch_bounds__(256) void displace_kernel(const AtomIO* in, long long n, const double* u_rms,
                                                       unsigned long long seed, int config, AtomIO* out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    AtomIO a = in[i];
    const double u = (a.Z >= 0 && a.Z < 104) ? u_rms[a.Z] : 0.0;
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    const uint32_t c0 = (uint32_t)i, c1 = (uint32_t)((unsigned long long)i >> 32);
    const Philox4 r0 = philox4x32_10(c0, c1, (uint32_t)config, 0u, k0, k1);
    const Philox4 r1 = philox4x32_10(c0, c1, (uint32_t)config, 1u, k0, k1);
    const double twopi = 6.283185307179586476925287;
    const double u1 = uniform53(r0.v[0], r0.v[1]), u2 = uniform53(r0.v[2], r0.v[3]);
    const double u3 = uniform53(r1.v[0], r1.v[1]), u4 = uniform53(r1.v[2], r1.v[3]);
    const double ra = sqrt(-2.0 * log(u1)), rb = sqrt(-2.0 * log(u3));
    a.x += u * ra * cos(twopi * u2);
    a.y += u * ra * sin(twopi * u2);
    a.z += u * rb * cos(twopi * u4);
    out[i] = a;
}

hipError_t launch_displace(const AtomIO* in, long long n, const double* u_rms, unsigned long long seed, int config,
                           AtomIO* out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(displace_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in, n, u_rms, seed,
                       config, out);
    return hipGetLastError();
}

// ------------------------------------------------------------------------
// Reciprocal-space potential (Muslib Potential_rec_DT_DW,
// Pot_rec_DT_DW.cpp:3-132): Fourier coefficients of one slice,
//   V(g) = sum_species F_s(g) sum_{atoms of s} occ exp(-2 pi i (h x_f + k y_f)) (1 + i ads_s),
//   F_s(g) = sum_j a_j exp(-(b_j + B_s) |g|^2/4) / area   (Doyle-Turner, Debye-Waller B),
// at the reflections whose grid index h*nmult (k*mmult) is below n/2 in
// magnitude; every other coefficient (Nyquist, non-multiples) is 0.
// One 16x16 tile of reflections per block.  The phase factor of an atom is
// separable, exp(-2 pi i h x_f) exp(-2 pi i k y_f): per chunk of 8 atoms the
// block evaluates the 16 + 16 factors of its tile row and column once (one
// FP64 sincospi per thread, argument reduced to [-1/2, 1/2]) into LDS, and
// each thread forms the 8 products for its reflection (4 FP64 FMA each).
__device__ __forceinline__ bool refl_index(int i, int n, int mult, int* h) {
    const int half = n / 2;
    if (i < half) {
        *h = i / mult;
        return i % mult == 0;
    }
    const int j = n - i;
    *h = -(j / mult);
    return i > half && j % mult == 0;
}

// linear interpolation in a curve table, nothing past the end (Pot_RS.cpp:40-46,
// Pot_rec_curve_DW.cpp:38-44)
__device__ __forceinline__ bool curve_at(const double* c, int npoints, double u, double* out) {
    const double f = floor(u);
    if (f >= (double)(npoints - 1)) return false;
    const int j = (int)f;
    *out = c[j] + (c[j + 1] - c[j]) * (u - (double)j);
    return true;
}

__global__ __launch_bounds__(256) void rec_sf_kernel(RecArgs p) {
    __shared__ double2 ph[8][32];
    const int ty = threadIdx.x & 15, tx = threadIdx.x >> 4;
    const int iy = blockIdx.x * 16 + ty, ix = blockIdx.y * 16 + tx;
    const int kz = p.kz0 + (int)blockIdx.z;
    int h, k;
    const bool ok = refl_index(ix, p.nx, p.nmult, &h) & refl_index(iy, p.ny, p.mmult, &k);
    const double gx = h * p.A, gy = k * p.B;
    const double s2 = 0.25 * (gx * gx + gy * gy - 2.0 * gx * gy * p.cg);
    // this thread's phase-factor slot: atom (tid / 32) of the chunk, column
    // (tid % 32): 0..15 = h of tile column, 16..31 = k of tile row
    const int slot_a = threadIdx.x >> 5, slot_c = threadIdx.x & 31;
    int slot_q;
    if (slot_c < 16) refl_index(blockIdx.y * 16 + slot_c, p.nx, p.nmult, &slot_q);
    else refl_index(blockIdx.x * 16 + slot_c - 16, p.ny, p.mmult, &slot_q);
    double vr = 0.0, vi = 0.0;
    for (int sg = p.slice_seg[kz]; sg < p.slice_seg[kz + 1]; ++sg) {
        const RecSeg seg = p.segs[sg];
        double sr = 0.0, si = 0.0;
        for (int a0 = seg.start; a0 < seg.end; a0 += 8) {
            const int ai = a0 + slot_a;
            if (ai < seg.end) {
                const RecAtom at = p.atoms[ai];
                double arg = (double)slot_q * (slot_c < 16 ? at.x : at.y);
                arg -= rint(arg);
                double sn, cs;
                sincospi(-2.0 * arg, &sn, &cs);
                if (slot_c < 16) {
                    sn *= at.occ;
                    cs *= at.occ;
                }
                ph[slot_a][slot_c] = make_double2(cs, sn);
            }
            __syncthreads();
            const int na = min(8, seg.end - a0);
            for (int j = 0; j < na; ++j) {
                const double2 a = ph[j][tx], b = ph[j][16 + ty];
                sr = fma(a.x, b.x, fma(-a.y, b.y, sr));
                si = fma(a.x, b.y, fma(a.y, b.x, si));
            }
            __syncthreads();
        }
        double f = 0.0;
        if (p.curves) {
            curve_at(p.curves + (long long)seg.Z * p.npoints, p.npoints, sqrt(s2) / p.dds, &f);
        } else {
            const double* dt = c_dt[seg.Z];
            for (int j = 0; j < 4; ++j) f += dt[2 * j] * exp(-(dt[2 * j + 1] + seg.dw) * s2);
            f *= p.inv_area;
        }
        vr += f * (sr - seg.ads * si);
        vi += f * (si + seg.ads * sr);
    }
    if (!ok) vr = vi = 0.0;
    p.out[((long long)blockIdx.z * p.nx + ix) * p.ny + iy] = make_double2(vr, vi);
}

hipError_t launch_rec_sf(const RecArgs& a, int nslices, hipStream_t s) {
    dim3 grid(a.ny / 16, a.nx / 16, nslices);
    hipLaunchKernelGGL(rec_sf_kernel, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

// V = scale * Re(w), Vabs = scale * Im(w) for n values
template <typename Re>
__global__ void split_potential_kernel(const double2* w, long long n, double scale, Re* V, Re* Vabs) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double2 v = w[i];
    V[i] = (Re)(scale * v.x);
    if (Vabs) Vabs[i] = (Re)(scale * v.y);
}
hipError_t launch_split_potential(const double2* w, long long n, double scale, void* V, void* Vabs, bool f64,
                                  hipStream_t s) {
    const dim3 g((unsigned)((n + 255) / 256));
    if (f64)
        hipLaunchKernelGGL(split_potential_kernel<double>, g, dim3(256), 0, s, w, n, scale, static_cast<double*>(V),
                           static_cast<double*>(Vabs));
    else
        hipLaunchKernelGGL(split_potential_kernel<float>, g, dim3(256), 0, s, w, n, scale, static_cast<float*>(V),
                           static_cast<float*>(Vabs));
    return hipGetLastError();
}

// ------------------------------------------------------------------------
// Real-space Doyle-Turner potential (Muslib Potential_rs_DT_DW_cutpot and
// _periodic, Pot_RS.cpp:111-155, 350-486): every pixel of an atom's box (or of
// a wrapped piece of it, image l) gets occ * 4 pi sum_j a_j/(b_j+B)
// exp(-4 pi^2 r^2/(b_j+B)) (Curves.cpp:23-34), r^2 = x(x + 2y cos gam) + y^2
// to the atom image.  The host lists, per 16x16 pixel tile, the (atom, image)
// pairs whose box meets the tile; each thread tests its pixel against the
// exact integer box of the pair.
__device__ __forceinline__ double dt_rs2(int Z, double r2, double dw) {
    const double* p = c_dt[Z];
    double f = 0.0;
    for (int j = 0; j < 4; ++j)
        if (p[2 * j + 1] > 0) f += p[2 * j] / (p[2 * j + 1] + dw) * exp(-39.47841760435743 * r2 / (p[2 * j + 1] + dw));
    return 12.566370614359172 * f;
}

__global__ __launch_bounds__(256) void rs_dt_kernel(RsArgs p) {
    const int iy = blockIdx.x * 16 + (threadIdx.x & 15);
    const int ix = blockIdx.y * 16 + (threadIdx.x >> 4);
    const int kz = blockIdx.z;
    const int tile = (kz * p.ntx + (int)blockIdx.y) * p.nty + (int)blockIdx.x;
    double v = 0.0, va = 0.0;
    for (int e = p.tile_start[tile]; e < p.tile_start[tile + 1]; ++e) {
        const int2 en = p.entries[e];
        const RsAtom at = p.atoms[en.x];
        const int l = en.y;
        // image l: x part 0 main, +1 wrapped from the low edge (atom at +a),
        // -1 wrapped from the high edge (atom at -a); same for y
        const int sx = (l == 3 || l == 5 || l == 8) ? 1 : ((l == 4 || l == 6 || l == 7) ? -1 : 0);
        const int sy = (l == 1 || l == 5 || l == 7) ? 1 : ((l == 2 || l == 6 || l == 8) ? -1 : 0);
        int ib = max(at.ibb, 0), ie = min(at.iee, p.nx), jb = max(at.jbb, 0), je = min(at.jee, p.ny);
        double x0 = at.xf * p.a, y0 = at.yf * p.b;
        if (sx == 1) { ib = at.ibb + p.nx; ie = p.nx; x0 += p.a; }
        if (sx == -1) { ib = 0; ie = at.iee - p.nx + 1; x0 -= p.a; }
        if (sy == 1) { jb = at.jbb + p.ny; je = p.ny; y0 += p.b; }
        if (sy == -1) { jb = 0; je = at.jee - p.ny + 1; y0 -= p.b; }
        if (ix >= ib && ix < ie && iy >= jb && iy < je) {
            const double x = ix * p.dxa - x0, y = iy * p.dyb - y0;
            const double r2 = x * (x + 2.0 * y * p.cg) + y * y;
            double val = 0.0;
            bool hit = true;
            if (p.curves) hit = curve_at(p.curves + (long long)at.Z * p.npoints, p.npoints, sqrt(r2) / p.dds, &val);
            else val = dt_rs2(at.Z, r2, at.dw);
            if (hit) {
                const double res = val * at.occ;
                v += res;
                va += res * at.ads;
            }
        }
    }
    if (ix < p.nx && iy < p.ny) {
        const long long o = ((long long)kz * p.nx + ix) * p.ny + iy;
        if (p.v64) p.v64[o] = p.scale * v;
        if (p.v32) p.v32[o] = (float)(p.scale * v);
        if (p.a64) p.a64[o] = p.scale * va;
        if (p.a32) p.a32[o] = (float)(p.scale * va);
    }
}

hipError_t launch_rs_dt(const RsArgs& a, hipStream_t s) {
    dim3 grid(a.nty, a.ntx, a.nslices);
    hipLaunchKernelGGL(rs_dt_kernel, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

// each work-group copies 2048 consecutive float4; each thread loads its 8
// (coalesced across the work-group at each step) before storing them, so 8
// loads per thread are in flight.  No bounds test (a branch around the loads
// would serialise them): the caller passes a multiple of 2048 float4.
__global__ __launch_bounds__(256) void stream_copy_kernel(const float4* __restrict__ src, float4* __restrict__ dst) {
    const long long i0 = (long long)blockIdx.x * 2048 + threadIdx.x;
    float4 v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = src[i0 + 256 * k];
#pragma unroll
    for (int k = 0; k < 8; ++k) dst[i0 + 256 * k] = v[k];
}
hipError_t launch_stream_copy(const float2* src, float2* dst, long long n, hipStream_t s) {
    const long long n4 = n / 2;  // two complex64 values per float4; a multiple of 2048 (checked by the caller)
    hipLaunchKernelGGL(stream_copy_kernel, dim3((unsigned)(n4 / 2048)), dim3(256), 0, s,
                       reinterpret_cast<const float4*>(src), reinterpret_cast<float4*>(dst));
    return hipGetLastError();
}
// the column sweep's loads and stores without the transforms: each thread
// loads its 16 rows (stride 32) of one column, then stores them
template <int TW>
__global__ __launch_bounds__(32 * TW) void tile_copy_kernel(const float2* __restrict__ src, float2* __restrict__ dst,
                                                            int ny, int groups) {
    const int bid = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    const int g = bid % groups, b = bid / groups;
    const int lw = threadIdx.x % TW, t = threadIdx.x / TW;
    const long long base = (long long)b * 512 * ny + (long long)g * TW + lw;
    float2 v[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) v[e] = src[base + (long long)(t + 32 * e) * ny];
#pragma unroll
    for (int e = 0; e < 16; ++e) dst[base + (long long)(t + 32 * e) * ny] = v[e];
}
// The column-sweep pattern on images stored in 8-row bands (element (ix, iy)
// at ((ix >> 3) * ny + iy) * 8 + (ix & 7)): a work-group takes 8 adjacent
// columns of 512 rows, wave w and lane l hold column l / 8 and rows
// 8 w + l % 8 + 32 e, so each load or store instruction covers one
// contiguous 512-byte chunk (8 columns x 8 rows of one band).
__global__ __launch_bounds__(256) void band_copy_kernel(const float2* __restrict__ src, float2* __restrict__ dst,
                                                       int ny, int groups) {
    const int bid = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    const int g = bid % groups, b = bid / groups;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const long long base = (long long)b * 512 * ny + (long long)g * 64 + l + (long long)w * ny * 8;
    float2 v[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) v[e] = src[base + (long long)e * ny * 32];
#pragma unroll
    for (int e = 0; e < 16; ++e) dst[base + (long long)e * ny * 32] = v[e];
}
hipError_t launch_tile_copy(const float2* src, float2* dst, int nimg, int ny, int tw, hipStream_t s) {
    const int groups = ny / (tw % 1000);  // ny % (tw % 1000) == 0, checked by the caller
    const dim3 grid((unsigned)(groups * nimg));
    if (tw == 8) hipLaunchKernelGGL(tile_copy_kernel<8>, grid, dim3(256), 0, s, src, dst, ny, groups);
    else if (tw == 16) hipLaunchKernelGGL(tile_copy_kernel<16>, grid, dim3(512), 0, s, src, dst, ny, groups);
    else if (tw == 1008) hipLaunchKernelGGL(band_copy_kernel, grid, dim3(256), 0, s, src, dst, ny, groups);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

__global__ void det_finish_kernel(const double* partial, int ndet, int nb, int groups, double w, double* out,
                                  long long out_stride_d, long long probe0) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ndet * nb) return;
    const int d = idx / nb, b = idx % nb;
    double s = 0.0;
    const double* q = partial + ((long long)d * nb + b) * groups;
    for (int g = 0; g < groups; ++g) s += q[g];
    out[d * out_stride_d + probe0 + b] = s * w;
}

hipError_t launch_det_finish(const double* partial, int ndet, int nb, int groups, double w, double* out,
                             long long out_stride_d, long long probe0, hipStream_t s) {
    const int n = ndet * nb;
    hipLaunchKernelGGL(det_finish_kernel, dim3((n + 255) / 256), dim3(256), 0, s, partial, ndet, nb, groups, w,
                       out, out_stride_d, probe0);
    return hipGetLastError();
}

// Tilted incident waves of the partial-coherence integral (mulslice.cpp:
// 585-589): psi_b(ix, iy) = exp(2 pi i (qx xpos[ix] + qy ypos[iy])) with
// qx = iqx/ax and freqn's xpos[ix] = ix ax/(nx-1), i.e. the phase
// 2 pi (iqx ix/(nx-1) + iqy iy/(ny-1)); the products are reduced modulo
// n-1 in integers, so the FP64 phase is below 4 pi.
template <typename C>
__global__ void tilt_wave_kernel(C* psi, int nb, int nx, int ny, const int2* iq) {
    using Re = typename real_of<C>::type;
    const long long img = (long long)nx * ny;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= img * nb) return;
    const int b = (int)(i / img);
    const long long p = i % img;
    const int ix = (int)(p / ny), iy = (int)(p % ny);
    const int2 q = iq[b];
    const long long mx = (((long long)q.x * ix) % (nx - 1) + (nx - 1)) % (nx - 1);
    const long long my = (((long long)q.y * iy) % (ny - 1) + (ny - 1)) % (ny - 1);
    const double twopi = 6.283185307179586476925287;
    double ph = twopi * ((double)mx / (double)(nx - 1) + (double)my / (double)(ny - 1));
    ph -= twopi * rint(ph * (1.0 / twopi));
    double s, c;
    sincos(ph, &s, &c);
    psi[i] = mk<C>((Re)c, (Re)s);
}

template <typename C>
hipError_t launch_tilt_waves(C* psi, int nb, int nx, int ny, const int2* iq, hipStream_t s) {
    const long long n = (long long)nx * ny * nb;
    hipLaunchKernelGGL(tilt_wave_kernel<C>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, psi, nb, nx, ny, iq);
    return hipGetLastError();
}
template hipError_t launch_tilt_waves<float2>(float2*, int, int, int, const int2*, hipStream_t);
template hipError_t launch_tilt_waves<double2>(double2*, int, int, int, const int2*, hipStream_t);

// image[p] += scale * sum_b inten[b][p], summed in FP64 in wave order
template <typename Re>
__global__ void wave_sum_kernel(const Re* inten, int nb, long long img, double scale, double* image) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= img) return;
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += (double)inten[(long long)b * img + p];
    image[p] += scale * s;
}

template <typename Re>
hipError_t launch_wave_sum(const Re* inten, int nb, long long img, double scale, double* image, hipStream_t s) {
    hipLaunchKernelGGL(wave_sum_kernel<Re>, dim3((unsigned)((img + 255) / 256)), dim3(256), 0, s, inten, nb, img,
                       scale, image);
    return hipGetLastError();
}
template hipError_t launch_wave_sum<float>(const float*, int, long long, double, double*, hipStream_t);
template hipError_t launch_wave_sum<double>(const double*, int, long long, double, double*, hipStream_t);

}  // namespace msx
