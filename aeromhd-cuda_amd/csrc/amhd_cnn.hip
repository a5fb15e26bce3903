// amhd_cnn.hip -- 3x3 convolution layers of the learned flux closure on
// gfx950 matrix cores: BF16 operands, FP32 accumulation
// (v_mfma_f32_16x16x32_bf16), written as implicit GEMMs.
//
// Tensors are NHWC over the coarse grid: (n, nx, ny, C), x rows slow, y
// fast, channels innermost.  Activations are BF16 with C in {16, 32}
// (channels zero-padded); gradients are FP32.  Padding: zero outside
// [0, nx) in x, periodic in y (the solver's y boundary) -- the transposed
// convolution with the same rule is the exact adjoint.
//
// Weights are stored GEMM-ready: W[co][k] with k = tap * Cin + ci,
// tap = (di + 1) * 3 + (dj + 1) for the source offset (di, dj), and k
// zero-padded to a multiple of 32.  The backward-data pass takes the
// transposed layout WT[ci][tap * Cout + co].
//
// MFMA 16x16x32 (one wavefront): A is 16 x 32, B is 32 x 16, D is 16 x 16.
// Lane l supplies A[l % 16][8 (l / 16) .. 8 (l / 16) + 7] and
// B[8 (l / 16) .. + 7][l % 16], and receives D[4 (l / 16) + i][l % 16],
// i = 0..3 (checked against a host matmul by the GPU tests).
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <cstdarg>
#include <cstdio>
#include <string>

#include "aeromhd.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16;

__device__ __forceinline__ u16 f2bf(float x) {
    // round to nearest even
    unsigned u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) return (u16)(u >> 16);   // inf / nan
    u += 0x7fffu + ((u >> 16) & 1u);
    return (u16)(u >> 16);
}
__device__ __forceinline__ float bf2f(u16 h) { return __uint_as_float((unsigned)h << 16); }

union Frag {
    bf16x8 v;
    u16 h[8];
    uint4 q;
};

__device__ __forceinline__ f32x4 mfma(const Frag& a, const Frag& b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, b.v, c, 0, 0, 0);
}

// source pixel of (i, j) shifted by (di, dj): -1 if outside in x, wrap in y
__device__ __forceinline__ long src_pixel(int nimg, int i, int j, int di, int dj, int nx, int ny) {
    const int si = i + di;
    if (si < 0 || si >= nx) return -1;
    int sj = j + dj;
    sj = sj < 0 ? sj + ny : (sj >= ny ? sj - ny : sj);
    return ((long)nimg * nx + si) * ny + sj;
}

// A fragment of the implicit GEMM: pixel p (row of A), 8 consecutive k
// starting at k0 (k = tap * C + c).  `sign` = +1 forward (gather at p + off),
// -1 backward-data (gather at p - off).  BF16 source.
template <int C>
__device__ __forceinline__ Frag gather_bf16(const u16* __restrict__ x, long p, int k0, int sign, int nx, int ny) {
    Frag f;
    f.q = make_uint4(0, 0, 0, 0);
    const int tap = k0 / C, c0 = k0 % C;
    if (p < 0 || tap >= 9) return f;
    const long plane = (long)nx * ny;
    const int nimg = (int)(p / plane);
    const long r = p - (long)nimg * plane;
    const int i = (int)(r / ny), j = (int)(r - (long)i * ny);
    const long s = src_pixel(nimg, i, j, sign * (tap / 3 - 1), sign * (tap % 3 - 1), nx, ny);
    if (s < 0) return f;
    f.q = *(const uint4*)(x + s * C + c0);
    return f;
}

// FP32 source converted to BF16 on load.
template <int C>
__device__ __forceinline__ Frag gather_f32(const float* __restrict__ x, long p, int k0, int sign, int nx, int ny) {
    Frag f;
    f.q = make_uint4(0, 0, 0, 0);
    const int tap = k0 / C, c0 = k0 % C;
    if (p < 0 || tap >= 9) return f;
    const long plane = (long)nx * ny;
    const int nimg = (int)(p / plane);
    const long r = p - (long)nimg * plane;
    const int i = (int)(r / ny), j = (int)(r - (long)i * ny);
    const long s = src_pixel(nimg, i, j, sign * (tap / 3 - 1), sign * (tap % 3 - 1), nx, ny);
    if (s < 0) return f;
    const float4 a = *(const float4*)(x + s * C + c0);
    const float4 b = *(const float4*)(x + s * C + c0 + 4);
    f.h[0] = f2bf(a.x); f.h[1] = f2bf(a.y); f.h[2] = f2bf(a.z); f.h[3] = f2bf(a.w);
    f.h[4] = f2bf(b.x); f.h[5] = f2bf(b.y); f.h[6] = f2bf(b.z); f.h[7] = f2bf(b.w);
    return f;
}

// ------------------------------------------------------------- forward
// y[p][co] = act(bias[co] + sum_k A[p][k] W[co][k]); one wavefront = 16
// pixels x COUT channels.  Output BF16 (hidden layers) or FP32 (last).
template <int CIN, int COUT, bool RELU, bool OUT_F32>
__global__ __launch_bounds__(256) void conv_fwd_kernel(const u16* __restrict__ x, const u16* __restrict__ w,
                                                      const float* __restrict__ bias, void* __restrict__ y,
                                                      long npix, int nx, int ny) {
    constexpr int KPAD = (9 * CIN + 31) / 32 * 32;
    constexpr int NT = COUT / 16;
    const int lane = threadIdx.x & 63;
    const long tile = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const long p0 = tile * 16;
    if (p0 >= npix) return;
    const int r = lane & 15, q = lane >> 4;
    const long pa = p0 + r < npix ? p0 + r : -1;
    f32x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = (f32x4){0.f, 0.f, 0.f, 0.f};
    for (int kc = 0; kc < KPAD; kc += 32) {
        const int k0 = kc + 8 * q;
        const Frag a = gather_bf16<CIN>(x, pa, k0, +1, nx, ny);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            Frag b;
            b.q = *(const uint4*)(w + (long)(t * 16 + r) * KPAD + k0);
            acc[t] = mfma(a, b, acc[t]);
        }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int co = t * 16 + r;
        const float bv = bias[co];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const long p = p0 + 4 * q + i;
            if (p >= npix) continue;
            float v = acc[t][i] + bv;
            if (RELU) v = v > 0.f ? v : 0.f;
            if (OUT_F32) ((float*)y)[p * COUT + co] = v;
            else ((u16*)y)[p * COUT + co] = f2bf(v);
        }
    }
}

// -------------------------------------------------------- backward data
// dx[p][ci] = mask[p][ci] * sum_{tap, co} dy[p - off(tap)][co] WT[ci][tap*COUT+co]
// mask: the layer input's ReLU (x > 0), or none for the network input.
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void conv_bwd_data_kernel(const float* __restrict__ dy, const u16* __restrict__ wt,
                                                           const u16* __restrict__ mask, float* __restrict__ dx,
                                                           long npix, int nx, int ny) {
    constexpr int KPAD = 9 * COUT;   // COUT in {16, 32}: a multiple of 16; padded below
    constexpr int KP = (KPAD + 31) / 32 * 32;
    constexpr int NT = CIN / 16;
    const int lane = threadIdx.x & 63;
    const long tile = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const long p0 = tile * 16;
    if (p0 >= npix) return;
    const int r = lane & 15, q = lane >> 4;
    const long pa = p0 + r < npix ? p0 + r : -1;
    f32x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = (f32x4){0.f, 0.f, 0.f, 0.f};
    for (int kc = 0; kc < KP; kc += 32) {
        const int k0 = kc + 8 * q;
        const Frag a = gather_f32<COUT>(dy, pa, k0, -1, nx, ny);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            Frag b;
            b.q = *(const uint4*)(wt + (long)(t * 16 + r) * KP + k0);
            acc[t] = mfma(a, b, acc[t]);
        }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int ci = t * 16 + r;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const long p = p0 + 4 * q + i;
            if (p >= npix) continue;
            float v = acc[t][i];
            if (mask && !(bf2f(mask[p * CIN + ci]) > 0.f)) v = 0.f;
            dx[p * CIN + ci] = v;
        }
    }
}

// ------------------------------------------------------ backward weights
// Partial dW over a pixel range: dWp[s][co][k] = sum_p dy[p][co] A[p][k],
// with A the forward gather of x.  One wavefront per (16 co x 16 k) tile
// and pixel split s; the reduction over pixels runs 32 pixels per MFMA
// (A operand = dy^T, B operand = gathered x).
template <int CIN, int COUT>
__global__ __launch_bounds__(64) void conv_bwd_weight_kernel(const u16* __restrict__ x, const float* __restrict__ dy,
                                                            float* __restrict__ part, long npix, int nx, int ny,
                                                            long pix_per_split) {
    constexpr int KPAD = (9 * CIN + 31) / 32 * 32;
    constexpr int KT = KPAD / 16;
    const int lane = threadIdx.x;
    const int tile = blockIdx.x;               // co_t * KT + k_t
    const int co_t = tile / KT, k_t = tile % KT;
    const int split = blockIdx.y;
    const int r = lane & 15, q = lane >> 4;
    const long pbeg = (long)split * pix_per_split;
    const long pend = pbeg + pix_per_split < npix ? pbeg + pix_per_split : npix;
    const long plane = (long)nx * ny;
    const int k = k_t * 16 + r;                // this lane's B column
    const int tap = k / CIN, ci = k % CIN;
    f32x4 acc = (f32x4){0.f, 0.f, 0.f, 0.f};
    for (long pc = pbeg; pc < pend; pc += 32) {
        Frag a, b;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const long p = pc + 8 * q + e;
            float av = 0.f, bv = 0.f;
            if (p < pend) {
                av = dy[p * COUT + co_t * 16 + r];
                if (tap < 9) {
                    const int nimg = (int)(p / plane);
                    const long rr = p - (long)nimg * plane;
                    const int i = (int)(rr / ny), j = (int)(rr - (long)i * ny);
                    const long s = src_pixel(nimg, i, j, tap / 3 - 1, tap % 3 - 1, nx, ny);
                    if (s >= 0) bv = bf2f(x[s * CIN + ci]);
                }
            }
            a.h[e] = f2bf(av);
            b.h[e] = f2bf(bv);
        }
        acc = mfma(a, b, acc);
    }
    // D[row = co][col = k]: lane holds rows 4q..4q+3, column r
    float* dst = part + (long)split * COUT * KPAD;
#pragma unroll
    for (int i = 0; i < 4; ++i) dst[(long)(co_t * 16 + 4 * q + i) * KPAD + k_t * 16 + r] = acc[i];
}

// Sum the per-split partials (fixed order) and the bias gradient.
__global__ void reduce_splits_kernel(const float* __restrict__ part, int n_split, int count, float* __restrict__ out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    float s = 0.f;
    for (int k = 0; k < n_split; ++k) s += part[(long)k * count + e];
    out[e] = s;
}

__global__ void bias_grad_kernel(const float* __restrict__ dy, long npix, int cout, float* __restrict__ db) {
    // one workgroup per output channel, fixed-order tree
    __shared__ float sh[256];
    const int co = blockIdx.x;
    float s = 0.f;
    for (long p = threadIdx.x; p < npix; p += blockDim.x) s += dy[p * cout + co];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[co] = sh[0];
}

// MFMA lane-layout probe: D = A (16x32) * B (32x16), row-major FP32 in/out
// (operands rounded to BF16).
__global__ void mfma_probe_kernel(const float* __restrict__ A, const float* __restrict__ B, float* __restrict__ D) {
    const int lane = threadIdx.x, r = lane & 15, q = lane >> 4;
    Frag a, b;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        a.h[e] = f2bf(A[r * 32 + 8 * q + e]);
        b.h[e] = f2bf(B[(8 * q + e) * 16 + r]);
    }
    f32x4 acc = mfma(a, b, (f32x4){0.f, 0.f, 0.f, 0.f});
#pragma unroll
    for (int i = 0; i < 4; ++i) D[(4 * q + i) * 16 + r] = acc[i];
}

bool ok_c(int c) { return c == 16 || c == 32; }

}  // namespace

// shared error channel of the library (amhd_api.hip)
int amhd_set_error(int code, const char* text);

namespace {
int cfail(int code, const char* fmt, ...) {
    char buf[256];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    return amhd_set_error(code, buf);
}
}  // namespace

extern "C" {

int amhd_conv3x3_kpad(int cin) { return (9 * cin + 31) / 32 * 32; }

int amhd_conv3x3_fwd(const void* x, int n, int nx, int ny, int cin, const void* w, const float* bias, int cout,
                     int relu, void* y, int y_f32, void* stream) {
    if (!x || !w || !bias || !y || n <= 0 || nx <= 0 || ny <= 0) return cfail(-1, "amhd_conv3x3_fwd: bad argument");
    if (!ok_c(cin) || !ok_c(cout)) return cfail(-4, "amhd_conv3x3_fwd: channels must be 16 or 32");
    const long npix = (long)n * nx * ny;
    const dim3 grid((unsigned)((npix + 63) / 64)), block(256);
    hipStream_t s = (hipStream_t)stream;
    const u16* X = (const u16*)x;
    const u16* Wt = (const u16*)w;
#define FWD(CI, CO, R, F) conv_fwd_kernel<CI, CO, R, F><<<grid, block, 0, s>>>(X, Wt, bias, y, npix, nx, ny)
    if (cin == 16 && cout == 32 && relu && !y_f32) FWD(16, 32, true, false);
    else if (cin == 32 && cout == 32 && relu && !y_f32) FWD(32, 32, true, false);
    else if (cin == 32 && cout == 32 && !relu && y_f32) FWD(32, 32, false, true);
    else if (cin == 32 && cout == 16 && !relu && y_f32) FWD(32, 16, false, true);
    else if (cin == 16 && cout == 16 && !relu && y_f32) FWD(16, 16, false, true);
    else if (cin == 16 && cout == 16 && relu && !y_f32) FWD(16, 16, true, false);
    else return cfail(-4, "amhd_conv3x3_fwd: unsupported (cin %d, cout %d, relu %d, f32 %d)", cin, cout, relu, y_f32);
#undef FWD
    return hipGetLastError() == hipSuccess ? 0 : cfail(-2, "amhd_conv3x3_fwd: launch failed");
}

int amhd_conv3x3_bwd_data(const float* dy, int n, int nx, int ny, int cout, const void* wt, int cin, const void* mask,
                          float* dx, void* stream) {
    if (!dy || !wt || !dx || n <= 0) return cfail(-1, "amhd_conv3x3_bwd_data: bad argument");
    if (!ok_c(cin) || !ok_c(cout)) return cfail(-4, "amhd_conv3x3_bwd_data: channels must be 16 or 32");
    const long npix = (long)n * nx * ny;
    const dim3 grid((unsigned)((npix + 63) / 64)), block(256);
    hipStream_t s = (hipStream_t)stream;
    const u16* W = (const u16*)wt;
    const u16* M = (const u16*)mask;
#define BWD(CI, CO) conv_bwd_data_kernel<CI, CO><<<grid, block, 0, s>>>(dy, W, M, dx, npix, nx, ny)
    if (cin == 32 && cout == 32) BWD(32, 32);
    else if (cin == 32 && cout == 16) BWD(32, 16);
    else if (cin == 16 && cout == 32) BWD(16, 32);
    else BWD(16, 16);
#undef BWD
    return hipGetLastError() == hipSuccess ? 0 : cfail(-2, "amhd_conv3x3_bwd_data: launch failed");
}

// workspace: n_split * cout * kpad floats; n_split chosen by the caller.
int amhd_conv3x3_bwd_weight(const void* x, const float* dy, int n, int nx, int ny, int cin, int cout, float* dw,
                            float* dbias, float* workspace, int n_split, void* stream) {
    if (!x || !dy || !dw || !dbias || !workspace || n <= 0 || n_split <= 0)
        return cfail(-1, "amhd_conv3x3_bwd_weight: bad argument");
    if (!ok_c(cin) || !ok_c(cout)) return cfail(-4, "amhd_conv3x3_bwd_weight: channels must be 16 or 32");
    const long npix = (long)n * nx * ny;
    long per = (npix + n_split - 1) / n_split;
    per = (per + 31) / 32 * 32;
    const int kpad = amhd_conv3x3_kpad(cin);
    hipStream_t s = (hipStream_t)stream;
    const dim3 grid((unsigned)((cout / 16) * (kpad / 16)), (unsigned)n_split);
    const u16* X = (const u16*)x;
#define BWW(CI, CO) conv_bwd_weight_kernel<CI, CO><<<grid, 64, 0, s>>>(X, dy, workspace, npix, nx, ny, per)
    if (cin == 32 && cout == 32) BWW(32, 32);
    else if (cin == 32 && cout == 16) BWW(32, 16);
    else if (cin == 16 && cout == 32) BWW(16, 32);
    else BWW(16, 16);
#undef BWW
    if (hipGetLastError() != hipSuccess) return cfail(-2, "amhd_conv3x3_bwd_weight: launch failed");
    const int count = cout * kpad;
    reduce_splits_kernel<<<(count + 255) / 256, 256, 0, s>>>(workspace, n_split, count, dw);
    bias_grad_kernel<<<cout, 256, 0, s>>>(dy, npix, cout, dbias);
    return hipGetLastError() == hipSuccess ? 0 : cfail(-2, "amhd_conv3x3_bwd_weight: reduce launch failed");
}

int amhd_mfma_probe(const float* A, const float* B, float* D, void* stream) {
    if (!A || !B || !D) return cfail(-1, "amhd_mfma_probe: null argument");
    mfma_probe_kernel<<<1, 64, 0, (hipStream_t)stream>>>(A, B, D);
    return hipGetLastError() == hipSuccess ? 0 : cfail(-2, "amhd_mfma_probe: launch failed");
}

}  // extern "C"
