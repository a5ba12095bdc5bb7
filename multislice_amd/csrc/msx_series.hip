// msx_series.hip -- what a thickness series records after each slice of the
// slice loop (msx_multislice_series).
//
// Between the column sweep of slice s and the next row sweep the batch holds
// M(x, ky) = IFFT_x(P FFT_x FFT_y(t psi_s)) (unscaled transforms), i.e. nx
// times the y transform of the exit wave psi_{s+1}.  The records read M:
//   beam (h, k) = sum_x M(x, k) exp(+2 pi i h x / nx) / (nx^2 ny): the
//       propagated wave's k-space value at (h, k) over nx ny, the number
//       mulslice.cpp:717-724 prints;
//   intensity = sum |M|^2 / (nx^3 ny^2) = sum |psi_{s+1}|^2 / (nx ny)
//       (Parseval along y; mulslice.cpp:727-737).
// Both sum in FP64 in a fixed order (deterministic for a given launch).
#include "msx_internal.h"

namespace msx {

namespace {
constexpr int kSerThreads = 256;

// Tree sum of the work-group's values in LDS; thread 0 gets the total.
template <typename V> __device__ V block_sum(V v, V* red) {
    red[threadIdx.x] = v;
    __syncthreads();
#pragma unroll
    for (int w = kSerThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + w];
        __syncthreads();
    }
    return red[0];
}

struct D2 {
    double x, y;
    __device__ D2 operator+(const D2& o) const { return {x + o.x, y + o.y}; }
};

// work-group (beam i, wave b): one column k of wave b, DFT at h in FP64.
// bm[i] = (h, k, in band, 0) with h, k already in [0, n-1].
template <typename C>
__global__ __launch_bounds__(kSerThreads) void series_beams_kernel(const C* __restrict__ m, int nx, int ny,
                                                                   const int4* __restrict__ bm, int nbeams,
                                                                   double scale, C* __restrict__ out) {
    using Re = typename real_of<C>::type;
    __shared__ D2 red[kSerThreads];
    const int i = blockIdx.x, b = blockIdx.y;
    const int4 q = bm[i];
    D2 s{0.0, 0.0};
    if (q.z) {
        const C* col = m + (long long)b * nx * ny + q.y;
        for (int x = threadIdx.x; x < nx; x += kSerThreads) {
            // exp(+2 pi i h x / nx) with h x reduced mod nx in integers
            const long long r = ((long long)q.x * x) % nx;
            double sn, cs;
            sincospi(2.0 * (double)r / (double)nx, &sn, &cs);
            const C v = col[(long long)x * ny];
            const double vr = (double)v.x, vi = (double)v.y;
            s.x += vr * cs - vi * sn;
            s.y += vr * sn + vi * cs;
        }
    }
    const D2 t = block_sum(s, red);
    if (threadIdx.x == 0) out[(long long)b * nbeams + i] = mk<C>((Re)(t.x * scale), (Re)(t.y * scale));
}

// work-group (part j, wave b): sum |M|^2 over the j-th of nparts contiguous
// pieces of wave b; with ky2 set, columns with ky2 >= k2max are skipped (the
// column sweep did not write them).
template <typename C>
__global__ __launch_bounds__(kSerThreads) void series_norm_kernel(const C* __restrict__ m, long long img, int ny,
                                                                  const float* __restrict__ ky2, float k2max,
                                                                  int nparts, double* __restrict__ partial) {
    __shared__ double red[kSerThreads];
    const int j = blockIdx.x, b = blockIdx.y;
    const long long per = (img + nparts - 1) / nparts;
    const long long p0 = (long long)j * per, p1 = p0 + per < img ? p0 + per : img;
    const C* w = m + (long long)b * img;
    double s = 0.0;
    for (long long p = p0 + threadIdx.x; p < p1; p += kSerThreads) {
        if (ky2 && !(ky2[p % ny] < k2max)) continue;
        const C v = w[p];
        s += (double)v.x * (double)v.x + (double)v.y * (double)v.y;
    }
    const double t = block_sum(s, red);
    if (threadIdx.x == 0) partial[(long long)b * nparts + j] = t;
}
}  // namespace

template <typename C>
hipError_t launch_series_beams(const C* m, int nb, int nx, int ny, const int4* bm, int nbeams, double scale, C* out,
                               hipStream_t s) {
    if (nb <= 0 || nbeams <= 0 || nbeams > 65535 || nb > 65535) return hipErrorInvalidValue;
    hipLaunchKernelGGL(series_beams_kernel<C>, dim3((unsigned)nbeams, (unsigned)nb), dim3(kSerThreads), 0, s, m, nx,
                       ny, bm, nbeams, scale, out);
    return hipGetLastError();
}

template <typename C>
hipError_t launch_series_norm(const C* m, int nb, long long img, int ny, const float* ky2, float k2max, int nparts,
                              double* partial, hipStream_t s) {
    if (nb <= 0 || nb > 65535 || nparts <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(series_norm_kernel<C>, dim3((unsigned)nparts, (unsigned)nb), dim3(kSerThreads), 0, s, m, img,
                       ny, ky2, k2max, nparts, partial);
    return hipGetLastError();
}

template hipError_t launch_series_beams<float2>(const float2*, int, int, int, const int4*, int, double, float2*,
                                                hipStream_t);
template hipError_t launch_series_beams<double2>(const double2*, int, int, int, const int4*, int, double, double2*,
                                                 hipStream_t);
template hipError_t launch_series_norm<float2>(const float2*, int, long long, int, const float*, float, int, double*,
                                               hipStream_t);
template hipError_t launch_series_norm<double2>(const double2*, int, long long, int, const float*, float, int,
                                                double*, hipStream_t);

}  // namespace msx
