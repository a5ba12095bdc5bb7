// msx_unet_kernels.hip -- GPU kernels of the frozen-phonon surrogate (a small
// U-Net that maps a 4-configuration mean ADF image to the converged
// 64-configuration one).
//
// Every convolution (3x3 or 1x1, stride 1, zero padding) is an implicit GEMM on
// the gfx950 matrix cores: v_mfma_f32_32x32x16_bf16, BF16 operands converted
// from the FP32 tensors while they are staged in LDS, FP32 accumulation.
// The same kernel runs the three products of training:
//   FWD : Y[n,co,p]   = sum_{ci,tap} W[co,ci,tap] X[n,ci,p+off(tap)]    (+bias, ReLU)
//   BWDD: dX[n,ci,p]  = sum_{co,tap} W[co,ci,flip(tap)] G[n,co,p+off(tap)]
//   BWDW: dW[co,ci,tap] += sum_{n,p} G[n,co,p] X[n,ci,p+off(tap)]       (split over p)
// where G = dY masked by the ReLU of the layer's output.  Tensors are NCHW FP32.
#include <hip/hip_runtime.h>

#include "msx_unet.h"

namespace msx {
namespace unet {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// Block tile: BM output rows (32 or 64) x BN = 4096/BM columns, K step 32;
// 4 wavefronts, each owning one 32x32 output tile (two MFMAs per K step).
constexpr int BK = 32, KP = BK + 8;  // +8 bf16: 16-byte rows, spread banks

// Operand staging.  A rows are the GEMM's M rows, B rows its N columns; each
// thread stages a run of consecutive k of one row:
//   A: row tid / (BK/KPA), k run (tid % (BK/KPA)) * KPA, KPA = BM*BK/256
//   B: row tid % BN,       k run (tid / BN) * KPB,        KPB = BN*BK/256
// so pixel rows are gathered by consecutive threads (coalesced) and each
// thread writes whole runs of its LDS row.  Indices are decoded once per run
// and stepped, not divided per element.
template <int MODE, int KS, int BM>
__global__ __launch_bounds__(256) void conv_gemm(ConvArgs p) {
    constexpr int BN = 4096 / BM, KPA = BM * BK / 256, KPB = BN * BK / 256, TAPS = KS * KS, R = KS / 2;
    __shared__ __align__(16) __bf16 As[2][BM][KP];
    __shared__ __align__(16) __bf16 Bs[2][BN][KP];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = BM == 64 ? (wave & 1) : 0, wn = BM == 64 ? (wave >> 1) : wave;
    const int HW = p.H * p.W;
    long long M, N, K;
    if (MODE == kFwd) { M = p.Cout; N = (long long)p.N * HW; K = (long long)p.Cin * TAPS; }
    else if (MODE == kBwdData) { M = p.Cin; N = (long long)p.N * HW; K = (long long)p.Cout * TAPS; }
    else { M = p.Cout; N = (long long)p.Cin * TAPS; K = (long long)p.N * HW; }
    const long long m0 = (long long)blockIdx.y * BM, n0 = (long long)blockIdx.x * BN;
    long long k_beg = 0, k_end = K;
    if (MODE == kBwdWeight) {  // split over pixels
        const long long chunk = ((K + gridDim.z - 1) / gridDim.z + BK - 1) / BK * BK;
        k_beg = (long long)blockIdx.z * chunk;
        k_end = k_beg + chunk < K ? k_beg + chunk : K;
    }
    const int a_row = tid / (BK / KPA), a_k = (tid % (BK / KPA)) * KPA;
    const int b_row = tid % BN, b_k = (tid / BN) * KPB;
    const long long am = m0 + a_row, bn = n0 + b_row;
    // B row: a pixel (forward, input gradient) or a (channel, tap) (weight gradient)
    int b_img = 0, b_y = 0, b_x = 0, b_ci = 0, b_dy = 0, b_dx = 0;
    if (bn < N) {
        if (MODE != kBwdWeight) {
            b_img = (int)(bn / HW);
            const int pix = (int)(bn % HW);
            b_y = pix / p.W;
            b_x = pix % p.W;
        } else {
            b_ci = (int)(bn / TAPS);
            const int tap = (int)(bn % TAPS);
            b_dy = tap / KS - R;
            b_dx = tap % KS - R;
        }
    }

    float ra[KPA], rb[KPB];
    auto fetch = [&](long long k0) {
        const long long ka = k0 + a_k;
        if (MODE == kFwd) {
#pragma unroll
            for (int j = 0; j < KPA; ++j) ra[j] = (am < M && ka + j < k_end) ? p.w[am * K + ka + j] : 0.f;
        } else if (MODE == kBwdData) {
            int co = (int)(ka / TAPS), tap = (int)(ka % TAPS);
#pragma unroll
            for (int j = 0; j < KPA; ++j) {
                ra[j] = (am < M && ka + j < k_end) ? p.w[((long long)co * p.Cin + am) * TAPS + (TAPS - 1 - tap)] : 0.f;
                if (++tap == TAPS) { tap = 0; ++co; }
            }
        } else {
            int img = (int)(ka / HW), pix = (int)(ka % HW);
#pragma unroll
            for (int j = 0; j < KPA; ++j) {
                float v = 0.f;
                if (am < M && ka + j < k_end) {
                    const long long o = ((long long)img * p.Cout + am) * HW + pix;
                    v = p.g[o];
                    if (p.relu && p.y[o] <= 0.f) v = 0.f;
                }
                ra[j] = v;
                if (++pix == HW) { pix = 0; ++img; }
            }
        }
        const long long kb = k0 + b_k;
        if (MODE == kFwd || MODE == kBwdData) {
            int c = (int)(kb / TAPS), tap = (int)(kb % TAPS);
            const int C = MODE == kFwd ? p.Cin : p.Cout;
            const float* src = MODE == kFwd ? p.x : p.g;
#pragma unroll
            for (int j = 0; j < KPB; ++j) {
                float v = 0.f;
                if (bn < N && kb + j < k_end) {
                    const int yy = b_y + tap / KS - R, xx = b_x + tap % KS - R;
                    if ((unsigned)yy < (unsigned)p.H && (unsigned)xx < (unsigned)p.W) {
                        const long long o = (((long long)b_img * C + c) * p.H + yy) * p.W + xx;
                        v = src[o];
                        if (MODE == kBwdData && p.relu && p.y[o] <= 0.f) v = 0.f;
                    }
                }
                rb[j] = v;
                if (++tap == TAPS) { tap = 0; ++c; }
            }
        } else {
            int img = (int)(kb / HW);
            const int pix = (int)(kb % HW);
            int y = pix / p.W, x = pix % p.W;
#pragma unroll
            for (int j = 0; j < KPB; ++j) {
                float v = 0.f;
                if (bn < N && kb + j < k_end) {
                    const int yy = y + b_dy, xx = x + b_dx;
                    if ((unsigned)yy < (unsigned)p.H && (unsigned)xx < (unsigned)p.W)
                        v = p.x[(((long long)img * p.Cin + b_ci) * p.H + yy) * p.W + xx];
                }
                rb[j] = v;
                if (++x == p.W) { x = 0; if (++y == p.H) { y = 0; ++img; } }
            }
        }
    };

    f32x16 acc;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = 0.f;
    const int r = lane & 31, kh = (lane >> 5) * 8;
    int buf = 0;
    if (k_beg < k_end) fetch(k_beg);
    for (long long k0 = k_beg; k0 < k_end; k0 += BK) {
#pragma unroll
        for (int j = 0; j < KPA; ++j) As[buf][a_row][a_k + j] = (__bf16)ra[j];
#pragma unroll
        for (int j = 0; j < KPB; ++j) Bs[buf][b_row][b_k + j] = (__bf16)rb[j];
        __syncthreads();
        if (k0 + BK < k_end) fetch(k0 + BK);  // the next step's loads overlap this step's MFMAs
#pragma unroll
        for (int s = 0; s < BK; s += 16) {
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(&As[buf][wm * 32 + r][s + kh]);
            const bf16x8 b = *reinterpret_cast<const bf16x8*>(&Bs[buf][wn * 32 + r][s + kh]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        buf ^= 1;  // double buffer: the next step writes the other buffer
    }
    // epilogue: acc[v] is C[row][col], col = lane % 32,
    // row = 8*(v/4) + 4*(lane/32) + v%4 (32x32 MFMA output layout)
    const int col = lane & 31;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int row = 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        const long long m = m0 + wm * 32 + row, n = n0 + wn * 32 + col;
        if (m >= M || n >= N) continue;
        float c = acc[v];
        if (MODE == kFwd) {
            if (p.bias) c += p.bias[m];
            if (p.relu && c < 0.f) c = 0.f;
        }
        if (MODE == kFwd || MODE == kBwdData) {
            const int img = (int)(n / HW), pix = (int)(n % HW);
            float* dst = p.out + ((long long)img * p.out_c + p.out_c0 + m) * HW + pix;
            if (p.accumulate) *dst += c; else *dst = c;
        } else if (gridDim.z > 1) {
            // this reduction slice's share of dW[co][ci][tap]; wgrad_reduce
            // adds the slices in order (no atomics: the same sums every run)
            p.partial[(long long)blockIdx.z * M * N + m * N + n] = c;
        } else {
            p.out[m * N + n] += c;  // dW[co][ci][tap], one writer per element
        }
    }
}

__global__ void wgrad_reduce(const float* __restrict__ part, float* out, long long mn, int splits) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= mn) return;
    float s = 0.f;
    for (int z = 0; z < splits; ++z) s += part[(long long)z * mn + i];
    out[i] += s;
}

namespace {
void conv_gemm_shape(int mode, int ks, const ConvArgs& a, long long* M, long long* N, long long* K) {
    const int HW = a.H * a.W;
    if (mode == kFwd) { *M = a.Cout; *N = (long long)a.N * HW; *K = (long long)a.Cin * ks * ks; }
    else if (mode == kBwdData) { *M = a.Cin; *N = (long long)a.N * HW; *K = (long long)a.Cout * ks * ks; }
    else { *M = a.Cout; *N = (long long)a.Cin * ks * ks; *K = (long long)a.N * HW; }
}
int conv_bm(long long M) { return M <= 32 ? 32 : 64; }
// reduction slices of the weight gradient: enough work-groups to fill the
// chip, each slice at least 4 K-steps long
int wgrad_splits(long long M, long long N, long long K) {
    const int bm = conv_bm(M), bn = 4096 / bm;
    const long long tiles = ((M + bm - 1) / bm) * ((N + bn - 1) / bn);
    long long want = 2048 / (tiles > 0 ? tiles : 1);
    long long maxs = (K + 4 * BK - 1) / (4 * BK);
    int splits = (int)(want < 1 ? 1 : (want > maxs ? maxs : want));
    return splits < 1 ? 1 : splits;
}
}  // namespace

size_t conv_wgrad_scratch(int ks, const ConvArgs& a) {
    long long M, N, K;
    conv_gemm_shape(kBwdWeight, ks, a, &M, &N, &K);
    const int splits = wgrad_splits(M, N, K);
    return splits > 1 ? (size_t)splits * (size_t)(M * N) : 0;
}

hipError_t launch_conv(int mode, int ks, const ConvArgs& a, hipStream_t s) {
    long long M, N, K;
    conv_gemm_shape(mode, ks, a, &M, &N, &K);
    const int bm = conv_bm(M), bn = 4096 / bm;
    const int splits = mode == kBwdWeight ? wgrad_splits(M, N, K) : 1;
    if (splits > 1 && !a.partial) return hipErrorInvalidValue;
    dim3 grid((unsigned)((N + bn - 1) / bn), (unsigned)((M + bm - 1) / bm), (unsigned)splits);
#define MSX_CONV(MODE_, KS_, BM_) hipLaunchKernelGGL((conv_gemm<MODE_, KS_, BM_>), grid, dim3(256), 0, s, a)
#define MSX_CONV_BM(MODE_, KS_)                    \
    do {                                           \
        if (bm == 32) MSX_CONV(MODE_, KS_, 32);    \
        else MSX_CONV(MODE_, KS_, 64);             \
    } while (0)
    if (ks == 3) {
        if (mode == kFwd) MSX_CONV_BM(kFwd, 3);
        else if (mode == kBwdData) MSX_CONV_BM(kBwdData, 3);
        else MSX_CONV_BM(kBwdWeight, 3);
    } else if (ks == 1) {
        if (mode == kFwd) MSX_CONV_BM(kFwd, 1);
        else if (mode == kBwdData) MSX_CONV_BM(kBwdData, 1);
        else MSX_CONV_BM(kBwdWeight, 1);
    } else {
        return hipErrorInvalidValue;
    }
#undef MSX_CONV_BM
#undef MSX_CONV
    if (splits > 1) {
        const long long mn = M * N;
        hipLaunchKernelGGL(wgrad_reduce, dim3((unsigned)((mn + 255) / 256)), dim3(256), 0, s, a.partial, a.out, mn,
                           splits);
    }
    return hipGetLastError();
}

// ------------------------------------------------------------------ small ops
__global__ void avgpool2_fwd(const float* x, float* y, long long nc, int H, int W) {
    const int Ho = H / 2, Wo = W / 2;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc * Ho * Wo) return;
    const long long c = i / (Ho * Wo);
    const int pix = (int)(i % (Ho * Wo)), yo = pix / Wo, xo = pix % Wo;
    const float* s = x + c * H * W + (2 * yo) * W + 2 * xo;
    y[i] = 0.25f * (s[0] + s[1] + s[W] + s[W + 1]);
}
__global__ void avgpool2_bwd(const float* gy, float* gx, long long nc, int H, int W) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc * H * W) return;
    const long long c = i / (H * W);
    const int pix = (int)(i % (H * W)), y = pix / W, x = pix % W;
    gx[i] += 0.25f * gy[(c * (H / 2) + y / 2) * (W / 2) + x / 2];
}
// nearest-neighbour x2 upsampling into channels [c0, c0+C) of a tensor with Ct channels
__global__ void upsample2_fwd(const float* x, float* y, int N, int C, int H, int W, int Ct, int c0) {
    const long long tot = (long long)N * C * 4 * H * W;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= tot) return;
    const int Wo = 2 * W, Ho = 2 * H;
    const int xo = (int)(i % Wo), yo = (int)((i / Wo) % Ho);
    const long long nc = i / ((long long)Wo * Ho);
    const int n = (int)(nc / C), c = (int)(nc % C);
    y[(((long long)n * Ct + c0 + c) * Ho + yo) * Wo + xo] = x[((long long)nc * H + yo / 2) * W + xo / 2];
}
__global__ void upsample2_bwd(const float* gy, float* gx, int N, int C, int H, int W, int Ct, int c0) {
    const long long tot = (long long)N * C * H * W;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= tot) return;
    const int x = (int)(i % W), y = (int)((i / W) % H);
    const long long nc = i / ((long long)W * H);
    const int n = (int)(nc / C), c = (int)(nc % C);
    const int Wo = 2 * W;
    const float* s = gy + (((long long)n * Ct + c0 + c) * 2 * H + 2 * y) * Wo + 2 * x;
    gx[i] += s[0] + s[1] + s[Wo] + s[Wo + 1];
}
// copy channels: dst[n, d0 + c] (+)= src[n, s0 + c], c < C
__global__ void copy_channels(const float* src, int Cs, int s0, float* dst, int Cd, int d0, int N, int C, int HW,
                              int accumulate) {
    const long long tot = (long long)N * C * HW;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= tot) return;
    const int pix = (int)(i % HW);
    const long long nc = i / HW;
    const int n = (int)(nc / C), c = (int)(nc % C);
    const float v = src[((long long)n * Cs + s0 + c) * HW + pix];
    float* d = dst + ((long long)n * Cd + d0 + c) * HW + pix;
    if (accumulate) *d += v; else *d = v;
}
// bias gradient: one block per channel, deterministic tree reduction
__global__ __launch_bounds__(256) void bias_grad(const float* g, const float* y, int relu, int N, int C, int HW,
                                                 float* db) {
    __shared__ float red[256];
    const int c = blockIdx.x;
    float s = 0.f;
    for (long long i = threadIdx.x; i < (long long)N * HW; i += 256) {
        const int n = (int)(i / HW), pix = (int)(i % HW);
        const long long o = ((long long)n * C + c) * HW + pix;
        const float v = g[o];
        s += (relu && y[o] <= 0.f) ? 0.f : v;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[c] = red[0];
}
// residual output out = in + z; MSE loss gradient g = 2 (out - t) / count
__global__ void residual_mse(const float* in, const float* z, const float* t, float* out, float* g, long long n,
                             float scale) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float o = in[i] + z[i];
    out[i] = o;
    if (g) g[i] = scale * (o - t[i]);
}
__global__ void channel_mean(const float* x, float* y, int C, long long HW, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long b = i / HW, p = i % HW;
    const float* q = x + b * C * HW + p;
    float s = 0.f;
    for (int c = 0; c < C; ++c) s += q[c * HW];
    y[i] = s * (1.0f / (float)C);
}
// per-block partial sums of (a - b)^2 in FP64 (fixed order: deterministic loss)
__global__ __launch_bounds__(256) void sq_diff(const float* a, const float* b, long long n, double* partial) {
    __shared__ double red[256];
    double s = 0.0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const double d = (double)a[i] - (double)b[i];
        s += d * d;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}
// Adam (Kingma & Ba 2015), bias-corrected
__global__ void adam_step(float* w, const float* gw, float* m, float* v, long long n, float lr, float b1, float b2,
                          float eps, float c1, float c2) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = gw[i];
    const float mi = b1 * m[i] + (1.f - b1) * g;
    const float vi = b2 * v[i] + (1.f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

#define MSX_GRID(n) dim3((unsigned)(((n) + 255) / 256)), dim3(256), 0, s
hipError_t launch_avgpool_fwd(const float* x, float* y, long long nc, int H, int W, hipStream_t s) {
    hipLaunchKernelGGL(avgpool2_fwd, MSX_GRID(nc * (H / 2) * (W / 2)), x, y, nc, H, W);
    return hipGetLastError();
}
hipError_t launch_avgpool_bwd(const float* gy, float* gx, long long nc, int H, int W, hipStream_t s) {
    hipLaunchKernelGGL(avgpool2_bwd, MSX_GRID(nc * H * W), gy, gx, nc, H, W);
    return hipGetLastError();
}
hipError_t launch_upsample_fwd(const float* x, float* y, int N, int C, int H, int W, int Ct, int c0, hipStream_t s) {
    hipLaunchKernelGGL(upsample2_fwd, MSX_GRID((long long)N * C * 4 * H * W), x, y, N, C, H, W, Ct, c0);
    return hipGetLastError();
}
hipError_t launch_upsample_bwd(const float* gy, float* gx, int N, int C, int H, int W, int Ct, int c0, hipStream_t s) {
    hipLaunchKernelGGL(upsample2_bwd, MSX_GRID((long long)N * C * H * W), gy, gx, N, C, H, W, Ct, c0);
    return hipGetLastError();
}
hipError_t launch_copy_channels(const float* src, int Cs, int s0, float* dst, int Cd, int d0, int N, int C, int HW,
                                int accumulate, hipStream_t s) {
    hipLaunchKernelGGL(copy_channels, MSX_GRID((long long)N * C * HW), src, Cs, s0, dst, Cd, d0, N, C, HW, accumulate);
    return hipGetLastError();
}
hipError_t launch_bias_grad(const float* g, const float* y, int relu, int N, int C, int HW, float* db, hipStream_t s) {
    hipLaunchKernelGGL(bias_grad, dim3(C), dim3(256), 0, s, g, y, relu, N, C, HW, db);
    return hipGetLastError();
}
hipError_t launch_residual_mse(const float* in, const float* z, const float* t, float* out, float* g, long long n,
                               float scale, hipStream_t s) {
    hipLaunchKernelGGL(residual_mse, MSX_GRID(n), in, z, t, out, g, n, scale);
    return hipGetLastError();
}
hipError_t launch_channel_mean(const float* x, float* y, int N, int C, long long HW, hipStream_t s) {
    const long long n = (long long)N * HW;
    hipLaunchKernelGGL(channel_mean, MSX_GRID(n), x, y, C, HW, n);
    return hipGetLastError();
}
hipError_t launch_sq_diff(const float* a, const float* b, long long n, double* partial, int nblocks, hipStream_t s) {
    hipLaunchKernelGGL(sq_diff, dim3(nblocks), dim3(256), 0, s, a, b, n, partial);
    return hipGetLastError();
}
hipError_t launch_adam(float* w, const float* gw, float* m, float* v, long long n, float lr, float b1, float b2,
                       float eps, float c1, float c2, hipStream_t s) {
    hipLaunchKernelGGL(adam_step, MSX_GRID(n), w, gw, m, v, n, lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}
#undef MSX_GRID

}  // namespace unet
}  // namespace msx
