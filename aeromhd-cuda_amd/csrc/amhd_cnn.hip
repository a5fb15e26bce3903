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
#include <vector>
#include <cmath>
#include <mutex>

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

// The same fragment for a pixel whose (image, row, column) the caller has
// already split off p once (the forward pass gathers 9 taps and every K
// chunk of one pixel, so the two 64-bit divisions leave the K loop).
struct PixelPos {
    int nimg, i, j;
};
__device__ __forceinline__ PixelPos pixel_pos(long p, int nx, int ny) {
    const long plane = (long)nx * ny;
    PixelPos q;
    q.nimg = (int)(p / plane);
    const long r = p - (long)q.nimg * plane;
    q.i = (int)(r / ny);
    q.j = (int)(r - (long)q.i * ny);
    return q;
}
template <int C>
__device__ __forceinline__ Frag gather_bf16_at(const u16* __restrict__ x, bool valid, const PixelPos& q, int k0,
                                               int sign, int nx, int ny) {
    Frag f;
    f.q = make_uint4(0, 0, 0, 0);
    const int tap = k0 / C, c0 = k0 % C;
    if (!valid || tap >= 9) return f;
    const long s = src_pixel(q.nimg, q.i, q.j, sign * (tap / 3 - 1), sign * (tap % 3 - 1), nx, ny);
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
    const PixelPos pos = pixel_pos(pa >= 0 ? pa : 0, nx, ny);
    for (int kc = 0; kc < KPAD; kc += 32) {
        const int k0 = kc + 8 * q;
        const Frag a = gather_bf16_at<CIN>(x, pa >= 0, pos, k0, +1, nx, ny);
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

// ------------------------------------------- closure inference on the state
// The network sees NF = 9 conserved fields and returns, per cell, the
// correction of the 9 x-face fluxes at i-1/2 (channels 0..8) and of the 9
// y-face fluxes at j-1/2 (channels 9..17) in units of fscale.
constexpr int NF = 9;
constexpr int CHUNK = 32;     // instances per launch (pointers and step sizes travel as kernel arguments)

struct ConstPtrs { const void* p[CHUNK]; };
struct MutPtrs { void* p[CHUNK]; };
struct Feat {
    int field[NF];            // plane of feature f in the instance's array
    float mean[NF], std[NF];
};
struct FluxScale { float s[NF]; int field[NF]; };
struct StepDt { double dt[CHUNK], dtdx[CHUNK], dtdy[CHUNK]; };   // per instance: dt, dt/dx, dt/dy

// x[(m, i, j)][f] = bf16((U_m[field[f]][i, j] - mean[f]) / std[f]), f < 9;
// channels 9..15 zero.
template <class real>
__global__ __launch_bounds__(256) void closure_features_kernel(ConstPtrs st, Feat ft, int n, long plane,
                                                               u16* __restrict__ x) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long)n * plane) return;
    const int m = (int)(k / plane);
    const long c = k - (long)m * plane;
    const real* u = (const real*)st.p[m];
    Frag lo, hi;
    lo.q = make_uint4(0, 0, 0, 0);
    hi.q = make_uint4(0, 0, 0, 0);
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        const u16 h = f2bf((float(u[ft.field[f] * plane + c]) - ft.mean[f]) / ft.std[f]);
        if (f < 8) lo.h[f] = h; else hi.h[f - 8] = h;
    }
    uint4* o = (uint4*)(x + k * 16);
    o[0] = lo.q;
    o[1] = hi.q;
}

// correction of field f through the four faces of cell (m, i, j), each
// scaled by its face limiter; x faces on the domain boundary carry none.
struct FaceCorr {
    float xl[NF], xr[NF], yl[NF], yr[NF];
};
// A pixel's 32 channels are 128 contiguous bytes and neighbouring lanes
// hold neighbouring pixels, so every load touches one cache line per lane:
// the channels are read 16 bytes at a time (11 loads of 4 channels instead
// of 36 of one), channels 0..19 of the cell, 0..11 of the x neighbour and
// 8..19 of the y neighbour.
__device__ __forceinline__ void load_faces(const float* __restrict__ dF, long p, int i, int j, int nx, int ny,
                                           FaceCorr& fc) {
    const long pr = p + ny;
    const long pu = j + 1 < ny ? p + 1 : p + 1 - ny;
    const float4* own = (const float4*)(dF + p * 32);
    const float4* up = (const float4*)(dF + pu * 32 + 8);
    float a[20], u[12], r[12];
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        const float4 v = own[q];
        a[4 * q] = v.x; a[4 * q + 1] = v.y; a[4 * q + 2] = v.z; a[4 * q + 3] = v.w;
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const float4 v = up[q];
        u[4 * q] = v.x; u[4 * q + 1] = v.y; u[4 * q + 2] = v.z; u[4 * q + 3] = v.w;
    }
    if (i + 1 < nx) {
        const float4* rt = (const float4*)(dF + pr * 32);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const float4 v = rt[q];
            r[4 * q] = v.x; r[4 * q + 1] = v.y; r[4 * q + 2] = v.z; r[4 * q + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int q = 0; q < 12; ++q) r[q] = 0.f;
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        fc.xl[f] = i > 0 ? a[f] : 0.f;
        fc.xr[f] = r[f];
        fc.yl[f] = a[NF + f];
        fc.yr[f] = u[NF + f - 8];
    }
}

// Per-cell limiter.  The correction of cell c is a sum over its four faces,
// D = sum_f w_f D_f, with face weights w_f = min(theta_left, theta_right).
// theta_c is the smallest over the faces of the largest t in
// {1, 1/2, ..., 1/32, 0} for which U + 4 t D_f alone is admissible
// (rho > 0, rho >= keep * rho(U), p >= keep * p(U); the bounds are rho(U),
// p(U) themselves where those are not positive).  The admissible set is
// convex (rho linear, p concave in U) and contains U, so for any weights
// 0 <= w_f <= theta_c, U + D = (1/4) sum_f (U + 4 w_f D_f) is admissible:
// the limited correction keeps the bound in every cell.
template <class real>
__global__ __launch_bounds__(256) void closure_theta_kernel(const float* __restrict__ dF, ConstPtrs st, FluxScale fs,
                                                            int n, int nx, int ny, StepDt sd, double gamma,
                                                            double keep, float* __restrict__ theta) {
    const long plane = (long)nx * ny;
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long)n * plane) return;
    const int m = (int)(k / plane);
    const long c = k - (long)m * plane;
    const int i = (int)(c / ny), j = (int)(c - (long)i * ny);
    FaceCorr fc;
    load_faces(dF, k, i, j, nx, ny, fc);
    const double dtdx = sd.dtdx[m], dtdy = sd.dtdy[m];
    const real* u = (const real*)st.p[m];
    double u0[10];
#pragma unroll
    for (int v = 0; v < 10; ++v) u0[v] = double(u[v * plane + c]);
    auto pres = [&](const double* w) {
        return (gamma - 1.0) * (w[7] - 0.5 * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]) / w[0] -
                                0.5 * (w[4] * w[4] + w[5] * w[5] + w[6] * w[6]));
    };
    const double rmin = u0[0] > 0.0 ? keep * u0[0] : u0[0];
    const double p0 = pres(u0);
    const double pmin = p0 > 0.0 ? keep * p0 : p0;
    // largest grid t with U + 4 t D_f admissible; D_f = sign * scale * face[f] * fscale[f]
    auto largest = [&](const float* face, double sc) -> double {
        double t = 1.0;
        for (int s = 0; s < 6; ++s, t *= 0.5) {
            double w[10];
#pragma unroll
            for (int v = 0; v < 10; ++v) w[v] = u0[v];
#pragma unroll
            for (int f = 0; f < NF; ++f) w[fs.field[f]] += 4.0 * t * sc * (double)face[f] * fs.s[f];
            if (w[0] > 0.0 && w[0] >= rmin && pres(w) >= pmin) return t;
        }
        return 0.0;
    };
    double th = largest(fc.xl, dtdx);              // S gains +dF(i-1/2) / dx
    th = fmin(th, largest(fc.xr, -dtdx));          //          -dF(i+1/2) / dx
    th = fmin(th, largest(fc.yl, dtdy));
    th = fmin(th, largest(fc.yr, -dtdy));
    theta[k] = (float)th;
}

// Position of rho and rhoC among the network fields (the host checks
// fields[0] == 0 and fields[8] == 9).  The rhoC face flux is rebuilt as
// donor-cell transport of the mass flux plus the network's mixing flux G:
// F_rhoC = C_donor * F_rho + G, the donor being the cell upwind of F_rho.
constexpr int FRHO = 0, FRC = 8;

template <class real>
__device__ __forceinline__ double conc(const real* __restrict__ u, long plane, long c) {
    const double r = double(u[c]);
    return r > 0.0 ? double(u[9 * plane + c]) / r : 0.0;
}

// The cell's four face weights: min of theta over the face's two cells
// (x faces on the domain boundary 0; 1 everywhere without a limiter).
struct FaceW { float xl, xr, yl, yr; };
__device__ __forceinline__ FaceW face_weights(const float* __restrict__ th, long k, int i, int j, int nx, int ny) {
    FaceW w{1.f, 1.f, 1.f, 1.f};
    if (!th) {
        w.xl = i > 0 ? 1.f : 0.f;
        w.xr = i + 1 < nx ? 1.f : 0.f;
        return w;
    }
    const float t0 = th[k];
    w.xl = i > 0 ? fminf(t0, th[k - ny]) : 0.f;
    w.xr = i + 1 < nx ? fminf(t0, th[k + ny]) : 0.f;
    w.yl = fminf(t0, th[j > 0 ? k - 1 : k - 1 + ny]);
    w.yr = fminf(t0, th[j + 1 < ny ? k + 1 : k + 1 - ny]);
    return w;
}

// Concentration of the cell and of its four face neighbours (the cell's
// own value stands in beyond an x boundary, where the face carries nothing).
struct Nbr { double c, xl, xr, yl, yr; };
template <class real>
__device__ __forceinline__ Nbr neighbours(const real* __restrict__ u, long plane, long c, int i, int j, int nx, int ny) {
    Nbr q;
    q.c = conc(u, plane, c);
    q.xl = i > 0 ? conc(u, plane, c - ny) : q.c;
    q.xr = i + 1 < nx ? conc(u, plane, c + ny) : q.c;
    q.yl = conc(u, plane, j > 0 ? c - 1 : c - 1 + ny);
    q.yr = conc(u, plane, j + 1 < ny ? c + 1 : c + 1 - ny);
    return q;
}

// Limited physical face fluxes of cell (i, j): non-mixing fields weighted
// by w, G by wm, F_rhoC = C_donor * F_rho + G.
struct Fluxes { double xl[NF], xr[NF], yl[NF], yr[NF]; };
__device__ __forceinline__ void limited_fluxes(const FaceCorr& fc, const FluxScale& fs, const FaceW& w,
                                               const FaceW& wm, const Nbr& q, Fluxes& F) {
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        const double sc = fs.s[f];
        const bool g = f == FRC;
        F.xl[f] = (g ? wm.xl : w.xl) * (double)fc.xl[f] * sc;
        F.xr[f] = (g ? wm.xr : w.xr) * (double)fc.xr[f] * sc;
        F.yl[f] = (g ? wm.yl : w.yl) * (double)fc.yl[f] * sc;
        F.yr[f] = (g ? wm.yr : w.yr) * (double)fc.yr[f] * sc;
    }
    F.xl[FRC] += (F.xl[FRHO] > 0.0 ? q.xl : q.c) * F.xl[FRHO];
    F.xr[FRC] += (F.xr[FRHO] > 0.0 ? q.c : q.xr) * F.xr[FRHO];
    F.yl[FRC] += (F.yl[FRHO] > 0.0 ? q.yl : q.c) * F.yl[FRHO];
    F.yr[FRC] += (F.yr[FRHO] > 0.0 ? q.c : q.yr) * F.yr[FRHO];
}

// Limiter of the mixing flux: with the mass flux limited by theta and its
// donor-cell rhoC transport applied (rho1, rc1), the smallest over the
// faces of the largest t in {1, ..., 1/32, 0} for which rc1 + 4 t G_face
// alone stays within [C_lo rho1, C_hi rho1], C_lo / C_hi the min / max of C
// over the cell and its face neighbours -- in the state and, when `bst`
// is given, in a second state too (the state before the step: Zalesak's
// bounds).  The set is convex and, by the donor-cell form, holds rc1
// itself, so any face weights up to theta_mix keep C within the bounds.
template <class real>
__global__ __launch_bounds__(256) void closure_theta_mix_kernel(const float* __restrict__ dF, ConstPtrs st,
                                                                ConstPtrs bst, bool has_b, FluxScale fs, int n,
                                                                int nx, int ny, StepDt sd,
                                                                const float* __restrict__ theta,
                                                                float* __restrict__ theta_mix) {
    const long plane = (long)nx * ny;
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long)n * plane) return;
    const int m = (int)(k / plane);
    const long c = k - (long)m * plane;
    const int i = (int)(c / ny), j = (int)(c - (long)i * ny);
    FaceCorr fc;
    load_faces(dF, k, i, j, nx, ny, fc);
    const real* u = (const real*)st.p[m];
    const double dtdx = sd.dtdx[m], dtdy = sd.dtdy[m];
    const FaceW w = face_weights(theta, k, i, j, nx, ny);
    const FaceW zero{0.f, 0.f, 0.f, 0.f};
    const Nbr q = neighbours(u, plane, c, i, j, nx, ny);
    Fluxes F;
    limited_fluxes(fc, fs, w, zero, q, F);
    const double rho1 = double(u[c]) + dtdx * (F.xl[FRHO] - F.xr[FRHO]) + dtdy * (F.yl[FRHO] - F.yr[FRHO]);
    const double rc1 = double(u[9 * plane + c]) + dtdx * (F.xl[FRC] - F.xr[FRC]) + dtdy * (F.yl[FRC] - F.yr[FRC]);
    double clo = fmin(q.c, fmin(fmin(q.xl, q.xr), fmin(q.yl, q.yr)));
    double chi = fmax(q.c, fmax(fmax(q.xl, q.xr), fmax(q.yl, q.yr)));
    if (has_b) {
        const Nbr b = neighbours((const real*)bst.p[m], plane, c, i, j, nx, ny);
        clo = fmin(clo, fmin(b.c, fmin(fmin(b.xl, b.xr), fmin(b.yl, b.yr))));
        chi = fmax(chi, fmax(b.c, fmax(fmax(b.xl, b.xr), fmax(b.yl, b.yr))));
    }
    const double lo = clo * rho1, hi = chi * rho1, sc = fs.s[FRC];
    auto largest = [&](double g) -> double {
        double t = 1.0;
        for (int s = 0; s < 6; ++s, t *= 0.5) {
            const double v = rc1 + 4.0 * t * g;
            if (v >= lo && v <= hi) return t;
        }
        return 0.0;
    };
    double th = largest(dtdx * (double)fc.xl[FRC] * sc);
    th = fmin(th, largest(-dtdx * (double)fc.xr[FRC] * sc));
    th = fmin(th, largest(dtdy * (double)fc.yl[FRC] * sc));
    th = fmin(th, largest(-dtdy * (double)fc.yr[FRC] * sc));
    theta_mix[k] = (float)th;
}

// S[field[f]] = -((F_xr - F_xl) / dx + (F_yr - F_yl) / dy) of the limited
// fluxes.  ACC = false: out = S, and the planes that no feature names
// (psi) get 0; ACC = true: out += dt * S on the named planes (out is the
// state).  theta / theta_mix null: no limiter.
template <class real, bool ACC>
__global__ __launch_bounds__(256) void closure_source_kernel(const float* __restrict__ dF, ConstPtrs st,
                                                             const float* __restrict__ theta,
                                                             const float* __restrict__ theta_mix, MutPtrs outs,
                                                             FluxScale fs, int n, int nx, int ny, double rdx,
                                                             double rdy, StepDt sd) {
    const long plane = (long)nx * ny;
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long)n * plane) return;
    const int m = (int)(k / plane);
    const long c = k - (long)m * plane;
    const int i = (int)(c / ny), j = (int)(c - (long)i * ny);
    FaceCorr fc;
    load_faces(dF, k, i, j, nx, ny, fc);
    const Nbr q = neighbours((const real*)st.p[m], plane, c, i, j, nx, ny);
    Fluxes F;
    limited_fluxes(fc, fs, face_weights(theta, k, i, j, nx, ny), face_weights(theta_mix, k, i, j, nx, ny), q, F);
    real* s = (real*)outs.p[m];
    unsigned named = 0u;   // variables written, one bit each (a register, not a private array)
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        const double v = -((F.xr[f] - F.xl[f]) * rdx + (F.yr[f] - F.yl[f]) * rdy);
        real& o = s[fs.field[f] * plane + c];
        if (ACC) o = real(double(o) + sd.dt[m] * v);
        else o = real(v);
        named |= 1u << fs.field[f];
    }
    if (!ACC) {
#pragma unroll
        for (int v = 0; v < 10; ++v)
            if (!((named >> v) & 1u)) s[v * plane + c] = real(0);
    }
}

// Workgroups per instance (the partials the caller's workspace holds) and
// their size: 512 lanes give a 256x128 instance one cell per lane, and
// twice the wavefronts in flight of 256-lane workgroups for this
// latency-bound pass.  The results are minima, maxima and a count, the same
// for any split of the cells.
constexpr int CHECK_BLOCKS = 64;
constexpr int CHECK_THREADS = 512;

// Guard quantities of U + dt S without writing it: per instance (blockIdx.y)
// and workgroup, {min rho, min p, max |divB|, non-finite cells} with
// p = (gamma-1)(E - |m|^2/(2 rho) - |B|^2/2) and divB as amhd_state_check
// (central differences on rows and columns 1 .. n-2), all in FP64.
template <class real>
__global__ __launch_bounds__(CHECK_THREADS) void closure_check_kernel(ConstPtrs st, ConstPtrs src, StepDt sd, int nx, int ny,
                                                            double gamma, double dx, double dy,
                                                            double* __restrict__ part) {
    const int m = blockIdx.y;
    const real* u = (const real*)st.p[m];
    const real* s = (const real*)src.p[m];
    const double dt = sd.dt[m];
    const long plane = (long)nx * ny;
    auto val = [&](int v, long c) {
        const double x = double(u[v * plane + c]);
        return dt != 0.0 ? x + dt * double(s[v * plane + c]) : x;
    };
    double mr = 1e300, mp = 1e300, md = 0.0, nf = 0.0;
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < plane; k += (long)gridDim.x * blockDim.x) {
        const int i = (int)(k / ny), j = (int)(k - (long)i * ny);
        double w[10];
        bool fin = true;
#pragma unroll
        for (int v = 0; v < 10; ++v) { w[v] = val(v, k); fin = fin && isfinite(w[v]); }
        if (!fin) nf += 1.0;
        const double ke = 0.5 * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]) / w[0];
        const double me = 0.5 * (w[4] * w[4] + w[5] * w[5] + w[6] * w[6]);
        mr = fmin(mr, w[0]);
        mp = fmin(mp, (gamma - 1.0) * (w[7] - ke - me));
        if (i >= 1 && i < nx - 1 && j >= 1 && j < ny - 1) {
            const double d = (0.5 * (w[4] + val(4, k + ny)) - 0.5 * (val(4, k - ny) + w[4])) / dx +
                             (0.5 * (w[5] + val(5, k + 1)) - 0.5 * (val(5, k - 1) + w[5])) / dy;
            md = fmax(md, fabs(d));
        }
    }
    __shared__ double sh[4][CHECK_THREADS / 64];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int off = 32; off > 0; off >>= 1) {
        mr = fmin(mr, __shfl_xor(mr, off, 64));
        mp = fmin(mp, __shfl_xor(mp, off, 64));
        md = fmax(md, __shfl_xor(md, off, 64));
        nf += __shfl_xor(nf, off, 64);
    }
    if (lane == 0) { sh[0][wv] = mr; sh[1][wv] = mp; sh[2][wv] = md; sh[3][wv] = nf; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q) {
            sh[0][0] = fmin(sh[0][0], sh[0][q]); sh[1][0] = fmin(sh[1][0], sh[1][q]);
            sh[2][0] = fmax(sh[2][0], sh[2][q]); sh[3][0] += sh[3][q];
        }
        double* o = part + ((long)m * gridDim.x + blockIdx.x) * 4;
        o[0] = sh[0][0]; o[1] = sh[1][0]; o[2] = sh[2][0]; o[3] = sh[3][0];
    }
}

// U += dt S on all 10 planes of each instance.
template <class real>
__global__ __launch_bounds__(256) void closure_apply_kernel(MutPtrs st, ConstPtrs src, StepDt sd, int n, long plane) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long)n * plane) return;
    const int m = (int)(k / plane);
    const long c = k - (long)m * plane;
    real* u = (real*)st.p[m];
    const real* s = (const real*)src.p[m];
#pragma unroll
    for (int v = 0; v < 10; ++v) u[v * plane + c] = real(double(u[v * plane + c]) + sd.dt[m] * double(s[v * plane + c]));
}

// Limited anti-diffusion of C across x faces (flux-corrected transport,
// Zalesak 1979), pass 1 of 2: per cell, the anti-diffusive rhoC flux of its
// face i+1/2, A = kappa (rho_i + rho_{i+1}) / 2 (C_{i+1} - C_i) (0 on the
// last column), and the ratios R+ = min(1, Q+ / P+), R- = min(1, Q- / P-)
// of the room left below C_hi and above C_lo (the min / max of C over the
// cell and its face neighbours, in the state and, when has_b, in the state
// before the step) to the cell's inflows and outflows over its two x faces.
// work holds three planes of cnt * nx * ny doubles: R+, R-, A.
template <class real>
__global__ __launch_bounds__(256) void sharpen_ratios_kernel(ConstPtrs st, ConstPtrs bst, bool has_b, int n, int nx,
                                                             int ny, double kappa, double* __restrict__ work) {
    const long plane = (long)nx * ny;
    const long total = (long)n * plane;
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= total) return;
    const int m = (int)(k / plane);
    const long c = k - (long)m * plane;
    const int i = (int)(c / ny), j = (int)(c - (long)i * ny);
    const real* u = (const real*)st.p[m];
    const Nbr q = neighbours(u, plane, c, i, j, nx, ny);
    const double r = double(u[c]);
    const double al = i > 0 ? kappa * 0.5 * (double(u[c - ny]) + r) * (q.c - q.xl) : 0.0;
    const double ar = i + 1 < nx ? kappa * 0.5 * (r + double(u[c + ny])) * (q.xr - q.c) : 0.0;
    const double gin = fmax(al, 0.0) + fmax(-ar, 0.0);
    const double gout = fmax(ar, 0.0) + fmax(-al, 0.0);
    double lo = fmin(q.c, fmin(fmin(q.xl, q.xr), fmin(q.yl, q.yr)));
    double hi = fmax(q.c, fmax(fmax(q.xl, q.xr), fmax(q.yl, q.yr)));
    if (has_b) {
        const Nbr b = neighbours((const real*)bst.p[m], plane, c, i, j, nx, ny);
        lo = fmin(lo, fmin(b.c, fmin(fmin(b.xl, b.xr), fmin(b.yl, b.yr))));
        hi = fmax(hi, fmax(b.c, fmax(fmax(b.xl, b.xr), fmax(b.yl, b.yr))));
    }
    const double rc = double(u[9 * plane + c]);
    const double qp = fmax(r * hi - rc, 0.0), qm = fmax(rc - r * lo, 0.0);
    work[k] = gin > 0.0 ? fmin(1.0, qp / gin) : 0.0;
    work[total + k] = gout > 0.0 ? fmin(1.0, qm / gout) : 0.0;
    work[2 * total + k] = ar;
}

// Pass 2: rhoC_i += L(i-1/2) - L(i+1/2), L = A * min(R+ of the receiving
// cell, R- of the giving one).  Reads only the work planes and the cell's
// own rhoC, so no cell reads what another writes.
template <class real>
__global__ __launch_bounds__(256) void sharpen_apply_kernel(MutPtrs st, int n, int nx, int ny,
                                                            const double* __restrict__ work) {
    const long plane = (long)nx * ny;
    const long total = (long)n * plane;
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= total) return;
    const int m = (int)(k / plane);
    const long c = k - (long)m * plane;
    const int i = (int)(c / ny);
    const double* rp = work;
    const double* rm = work + total;
    const double* a = work + 2 * total;
    auto limited = [&](long kl, long kr) {   // face between cell kl and cell kr = kl + ny
        const double f = a[kl];
        return (f > 0.0 ? fmin(rp[kr], rm[kl]) : fmin(rp[kl], rm[kr])) * f;
    };
    const double fl = i > 0 ? limited(k - ny, k) : 0.0;
    const double fr = i + 1 < (int)nx ? limited(k, k + ny) : 0.0;
    real* u = (real*)st.p[m];
    u[9 * plane + c] = real(double(u[9 * plane + c]) + (fl - fr));
}



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

int amhd_closure_features(const void* const* states, int dtype, const int* fields, int n, int nx, int ny,
                          const float* mean, const float* std, void* x, void* stream) {
    if (!states || !fields || !mean || !std || !x || n <= 0 || nx <= 0 || ny <= 0)
        return cfail(-1, "amhd_closure_features: bad argument");
    if (dtype != AMHD_F64 && dtype != AMHD_F32) return cfail(-1, "amhd_closure_features: bad dtype %d", dtype);
    Feat ft;
    for (int f = 0; f < NF; ++f) {
        if (fields[f] < 0 || fields[f] >= 10) return cfail(-1, "amhd_closure_features: field %d out of range", fields[f]);
        if (!(std[f] > 0.f)) return cfail(-1, "amhd_closure_features: std[%d] must be positive", f);
        ft.field[f] = fields[f]; ft.mean[f] = mean[f]; ft.std[f] = std[f];
    }
    const long plane = (long)nx * ny;
    hipStream_t s = (hipStream_t)stream;
    for (int m0 = 0; m0 < n; m0 += CHUNK) {
        const int cnt = n - m0 < CHUNK ? n - m0 : CHUNK;
        ConstPtrs st{};
        for (int m = 0; m < cnt; ++m) {
            if (!states[m0 + m]) return cfail(-1, "amhd_closure_features: null state %d", m0 + m);
            st.p[m] = states[m0 + m];
        }
        const long tot = (long)cnt * plane;
        u16* xo = (u16*)x + (long)m0 * plane * 16;
        const unsigned g = (unsigned)((tot + 255) / 256);
        if (dtype == AMHD_F64) closure_features_kernel<double><<<g, 256, 0, s>>>(st, ft, cnt, plane, xo);
        else closure_features_kernel<float><<<g, 256, 0, s>>>(st, ft, cnt, plane, xo);
        if (hipGetLastError() != hipSuccess) return cfail(-2, "amhd_closure_features: launch failed");
    }
    return 0;
}

int amhd_closure_source(const float* dF, const void* const* states, void* const* srcs, int dtype, const int* fields,
                        int n, int nx, int ny, const float* fscale, double dx, double dy, const double* dt,
                        double gamma, double keep, float* theta, const void* const* bounds, int accumulate,
                        void* stream) {
    if (!dF || !states || !srcs || !fields || !fscale || n <= 0 || nx <= 0 || ny <= 0 || !(dx > 0) || !(dy > 0))
        return cfail(-1, "amhd_closure_source: bad argument");
    if (dtype != AMHD_F64 && dtype != AMHD_F32) return cfail(-1, "amhd_closure_source: bad dtype %d", dtype);
    if ((uintptr_t)dF % 16) return cfail(-1, "amhd_closure_source: dF must be 16-byte aligned");
    if (fields[FRHO] != 0 || fields[FRC] != 9)
        return cfail(-1, "amhd_closure_source: fields[0] must be rho (0) and fields[8] rhoC (9)");
    if ((theta || accumulate) && !dt) return cfail(-1, "amhd_closure_source: the limiter and accumulate need dt");
    for (int m = 0; dt && m < n; ++m)
        if (!(dt[m] >= 0)) return cfail(-1, "amhd_closure_source: dt[%d] must be >= 0", m);
    if (theta && !(keep >= 0 && keep < 1)) return cfail(-1, "amhd_closure_source: keep must be in [0, 1)");
    FluxScale fs;
    for (int f = 0; f < NF; ++f) {
        if (fields[f] < 0 || fields[f] >= 10) return cfail(-1, "amhd_closure_source: field %d out of range", fields[f]);
        fs.s[f] = fscale[f]; fs.field[f] = fields[f];
    }
    const long plane = (long)nx * ny;
    const long total = (long)n * plane;
    hipStream_t s = (hipStream_t)stream;
    for (int m0 = 0; m0 < n; m0 += CHUNK) {
        const int cnt = n - m0 < CHUNK ? n - m0 : CHUNK;
        ConstPtrs st{}, bs{};
        MutPtrs sp{};
        for (int m = 0; m < cnt; ++m) {
            if (!srcs[m0 + m] || !states[m0 + m] || (bounds && !bounds[m0 + m]))
                return cfail(-1, "amhd_closure_source: null pointer %d", m0 + m);
            sp.p[m] = srcs[m0 + m];
            st.p[m] = states[m0 + m];
            bs.p[m] = bounds ? bounds[m0 + m] : nullptr;
        }
        const bool hb = bounds != nullptr;
        StepDt sd{};
        for (int m = 0; m < cnt; ++m) {
            const double t = dt ? dt[m0 + m] : 0.0;
            sd.dt[m] = t; sd.dtdx[m] = t / dx; sd.dtdy[m] = t / dy;
        }
        const long tot = (long)cnt * plane;
        const float* d = dF + (long)m0 * plane * 32;
        float* th = theta ? theta + (long)m0 * plane : nullptr;
        float* tm = theta ? theta + total + (long)m0 * plane : nullptr;
        const unsigned g = (unsigned)((tot + 255) / 256);
        if (th) {
            if (dtype == AMHD_F64) {
                closure_theta_kernel<double><<<g, 256, 0, s>>>(d, st, fs, cnt, nx, ny, sd, gamma, keep, th);
                closure_theta_mix_kernel<double><<<g, 256, 0, s>>>(d, st, bs, hb, fs, cnt, nx, ny, sd, th, tm);
            } else {
                closure_theta_kernel<float><<<g, 256, 0, s>>>(d, st, fs, cnt, nx, ny, sd, gamma, keep, th);
                closure_theta_mix_kernel<float><<<g, 256, 0, s>>>(d, st, bs, hb, fs, cnt, nx, ny, sd, th, tm);
            }
            if (hipGetLastError() != hipSuccess) return cfail(-2, "amhd_closure_source: limiter launch failed");
        }
#define SRC(R, A) closure_source_kernel<R, A><<<g, 256, 0, s>>>(d, st, th, tm, sp, fs, cnt, nx, ny, 1.0 / dx, 1.0 / dy, sd)
        if (dtype == AMHD_F64) { if (accumulate) SRC(double, true); else SRC(double, false); }
        else { if (accumulate) SRC(float, true); else SRC(float, false); }
#undef SRC
        if (hipGetLastError() != hipSuccess) return cfail(-2, "amhd_closure_source: launch failed");
    }
    return 0;
}

int amhd_closure_check(const void* const* states, const void* const* srcs, const double* dt, int dtype, int n,
                       int nx, int ny, double gamma, double dx, double dy, double* work, double* out, void* stream) {
    if (!states || !srcs || !dt || !work || !out || n <= 0 || nx <= 0 || ny <= 0 || !(dx > 0) || !(dy > 0))
        return cfail(-1, "amhd_closure_check: bad argument");
    if (dtype != AMHD_F64 && dtype != AMHD_F32) return cfail(-1, "amhd_closure_check: bad dtype %d", dtype);
    hipStream_t s = (hipStream_t)stream;
    // The per-workgroup results come back through one pinned host buffer of
    // the process (64 KiB, allocated at the first call and kept), one call
    // at a time.
    static std::mutex pin_mu;
    static double* pin = nullptr;
    std::lock_guard<std::mutex> lock(pin_mu);
    if (!pin && hipHostMalloc((void**)&pin, sizeof(double) * CHUNK * CHECK_BLOCKS * 4, hipHostMallocPortable) != hipSuccess) {
        pin = nullptr;
        return cfail(-3, "amhd_closure_check: pinned staging allocation failed");
    }
    const double* h = pin;
    for (int m0 = 0; m0 < n; m0 += CHUNK) {
        const int cnt = n - m0 < CHUNK ? n - m0 : CHUNK;
        ConstPtrs st{}, sr{};
        StepDt sd{};
        for (int m = 0; m < cnt; ++m) {
            if (!states[m0 + m] || !srcs[m0 + m]) return cfail(-1, "amhd_closure_check: null pointer %d", m0 + m);
            st.p[m] = states[m0 + m];
            sr.p[m] = srcs[m0 + m];
            sd.dt[m] = dt[m0 + m];
        }
        const dim3 grid(CHECK_BLOCKS, cnt);
        if (dtype == AMHD_F64)
            closure_check_kernel<double><<<grid, CHECK_THREADS, 0, s>>>(st, sr, sd, nx, ny, gamma, dx, dy, work);
        else closure_check_kernel<float><<<grid, CHECK_THREADS, 0, s>>>(st, sr, sd, nx, ny, gamma, dx, dy, work);
        if (hipGetLastError() != hipSuccess) return cfail(-2, "amhd_closure_check: launch failed");
        const size_t bytes = sizeof(double) * (size_t)cnt * CHECK_BLOCKS * 4;
        if (hipMemcpyAsync(pin, work, bytes, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return cfail(-2, "amhd_closure_check: copy failed");
        for (int m = 0; m < cnt; ++m) {
            double* o = out + 4 * (size_t)(m0 + m);
            const double* p = h + (size_t)m * CHECK_BLOCKS * 4;
            o[0] = p[0]; o[1] = p[1]; o[2] = p[2]; o[3] = p[3];
            for (int b = 1; b < CHECK_BLOCKS; ++b) {
                o[0] = std::fmin(o[0], p[4 * b]); o[1] = std::fmin(o[1], p[4 * b + 1]);
                o[2] = std::fmax(o[2], p[4 * b + 2]); o[3] += p[4 * b + 3];
            }
        }
    }
    return 0;
}

int amhd_closure_apply(void* const* states, const void* const* srcs, const double* dt, int dtype, int n, int nx,
                       int ny, void* stream) {
    if (!states || !srcs || !dt || n < 0 || nx <= 0 || ny <= 0) return cfail(-1, "amhd_closure_apply: bad argument");
    if (dtype != AMHD_F64 && dtype != AMHD_F32) return cfail(-1, "amhd_closure_apply: bad dtype %d", dtype);
    const long plane = (long)nx * ny;
    hipStream_t s = (hipStream_t)stream;
    for (int m0 = 0; m0 < n; m0 += CHUNK) {
        const int cnt = n - m0 < CHUNK ? n - m0 : CHUNK;
        MutPtrs st{};
        ConstPtrs sr{};
        StepDt sd{};
        for (int m = 0; m < cnt; ++m) {
            if (!states[m0 + m] || !srcs[m0 + m]) return cfail(-1, "amhd_closure_apply: null pointer %d", m0 + m);
            st.p[m] = states[m0 + m];
            sr.p[m] = srcs[m0 + m];
            sd.dt[m] = dt[m0 + m];
        }
        const unsigned g = (unsigned)(((long)cnt * plane + 255) / 256);
        if (dtype == AMHD_F64) closure_apply_kernel<double><<<g, 256, 0, s>>>(st, sr, sd, cnt, plane);
        else closure_apply_kernel<float><<<g, 256, 0, s>>>(st, sr, sd, cnt, plane);
        if (hipGetLastError() != hipSuccess) return cfail(-2, "amhd_closure_apply: launch failed");
    }
    return 0;
}

int amhd_closure_sharpen(void* const* states, const void* const* bounds, int dtype, int n, int nx, int ny,
                         double kappa, double* work, void* stream) {
    if (!states || n < 0 || nx <= 0 || ny <= 0 || !(kappa >= 0.0) || !std::isfinite(kappa))
        return cfail(-1, "amhd_closure_sharpen: bad argument");
    if (dtype != AMHD_F64 && dtype != AMHD_F32) return cfail(-1, "amhd_closure_sharpen: bad dtype %d", dtype);
    if (kappa == 0.0 || n == 0) return 0;
    if (!work) return cfail(-1, "amhd_closure_sharpen: work is null");
    const long plane = (long)nx * ny;
    hipStream_t s = (hipStream_t)stream;
    for (int m0 = 0; m0 < n; m0 += CHUNK) {
        const int cnt = n - m0 < CHUNK ? n - m0 : CHUNK;
        MutPtrs st{};
        ConstPtrs cs{}, bs{};
        for (int m = 0; m < cnt; ++m) {
            if (!states[m0 + m] || (bounds && !bounds[m0 + m]))
                return cfail(-1, "amhd_closure_sharpen: null pointer %d", m0 + m);
            st.p[m] = states[m0 + m];
            cs.p[m] = states[m0 + m];
            bs.p[m] = bounds ? bounds[m0 + m] : nullptr;
        }
        const unsigned g = (unsigned)(((long)cnt * plane + 255) / 256);
        if (dtype == AMHD_F64) {
            sharpen_ratios_kernel<double><<<g, 256, 0, s>>>(cs, bs, bounds != nullptr, cnt, nx, ny, kappa, work);
            sharpen_apply_kernel<double><<<g, 256, 0, s>>>(st, cnt, nx, ny, work);
        } else {
            sharpen_ratios_kernel<float><<<g, 256, 0, s>>>(cs, bs, bounds != nullptr, cnt, nx, ny, kappa, work);
            sharpen_apply_kernel<float><<<g, 256, 0, s>>>(st, cnt, nx, ny, work);
        }
        if (hipGetLastError() != hipSuccess) return cfail(-2, "amhd_closure_sharpen: launch failed");
    }
    return 0;
}

}  // extern "C"
