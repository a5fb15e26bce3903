// amhd_api.hip -- the C ABI of libaeromhd_hip.so (declared in
// include/aeromhd.h).  Built with -ffp-contract=off: the standalone FP64
// operators here are the reference's IEEE arithmetic.
#include <algorithm>
#include <cstdarg>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>
#include <cstdio>
#include <cstring>
#include <cmath>

#include "aeromhd.h"
#define AMHD_API_TU
#include "amhd_launch.hpp"

using namespace amhd;

namespace {

// The last error's text, per thread, in a fixed buffer (reporting an error
// allocates nothing).
thread_local char g_err[512];

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

}  // namespace

int amhd_set_error(int code, const char* text) {
    snprintf(g_err, sizeof g_err, "%s", text ? text : "");
    return code;
}

namespace {

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) return fail(-2, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                                          __FILE__, __LINE__);                                 \
    } while (0)

// Per-device scratch for the context-free operators (smax bits, CFL
// partials, a StepCtl).  Allocated once per device, freed at exit.
constexpr int SCRATCH_PARTS = 1024;
struct Scratch {
    double* part = nullptr;               // SCRATCH_PARTS * NPART_DT
    StepCtl* ctl = nullptr;
    unsigned long long* bits = nullptr;
};
std::mutex g_scratch_mu;
std::unordered_map<int, Scratch> g_scratch;

int get_scratch(Scratch** out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch.find(dev);
    if (it == g_scratch.end()) {
        Scratch s;
        HIP_TRY(hipMalloc(&s.part, sizeof(double) * SCRATCH_PARTS * NPART_DT));
        HIP_TRY(hipMalloc(&s.ctl, sizeof(StepCtl)));
        HIP_TRY(hipMalloc(&s.bits, sizeof(unsigned long long)));
        it = g_scratch.emplace(dev, s).first;
    }
    *out = &it->second;
    return 0;
}

// Scratch of amhd_compute_dt_w_noalloc: a file-scope device array instead of
// allocated memory, for callers that cannot allocate.  NOALLOC_PARTS CFL
// partials (NPART_DT doubles each), then one StepCtl: 32 KiB + 128 B per
// device, held until the process ends.  In host code `scratch` only names
// the symbol; its device address comes from hipGetSymbolAddress, once per
// device (noalloc_scratch), and only that address is passed to kernels.
constexpr int NOALLOC_PARTS = 1024;
constexpr int NOALLOC_CTL = NOALLOC_PARTS * NPART_DT;   // offset of the StepCtl, in doubles
constexpr int NOALLOC_DOUBLES = NOALLOC_CTL + (int)((sizeof(StepCtl) + sizeof(double) - 1) / sizeof(double));
static_assert(NOALLOC_CTL * sizeof(double) % alignof(StepCtl) == 0, "StepCtl alignment");
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu;
double* g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

int noalloc_scratch(double** out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(-4, "amhd_compute_dt_w_noalloc: device %d beyond %d", dev,
                                                           NOALLOC_MAX_DEVICES);
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_noalloc_ptr[dev] = (double*)p;
    }
    *out = g_noalloc_ptr[dev];
    return 0;
}

// Most segments of a narrow last strip per wavefront (default 0: off; a
// tuning switch).  Each lane group of a packed wavefront is on its own row,
// so its loads and stores touch a cache line per group: measured on the
// 2048 x 1024 FP32 sieve (4 valid cells in the last strip), packed
// wavefronts ran about 5% longer than the others with 2 groups, 8% with 4
// and 25% with 8 (stages 2 and 3, which also read U0).  With equal-row
// segments 2 groups still gained 1%; with the cost-sized segments the packed
// wavefronts (whose x faces take the uniform-face path only where both
// groups' rows are uniform, and which carry no row flags) ended last in
// every launch, and the sieve ran 3% faster without packing.  FP64 (one
// wavefront per SIMD) lost 10% with 2 groups.
#ifndef AMHD_PACK_MAX_GROUPS
#define AMHD_PACK_MAX_GROUPS 0
#endif

struct Ctx {
    int nx = 0, ny = 0, dtype = AMHD_F64;
    amhd_params p{};
    int rows = 0, n_strips = 0, n_segs = 0;   // rows: requested rows per wave (0 = auto)
    int pk_groups = 0, pk_shift = 0;           // packing of a narrow last strip (0 = none)
    int pack_max = AMHD_PACK_MAX_GROUPS;       // most packed groups (amhd_set_packing; 0 = off)
    int slots = 2048;                          // resident stage-kernel wavefronts on the device
    int cus = 0;                               // compute units of the device (0 = unknown)
    long plane = 0;
    int pitch = 0;
    void* buf[3] = {nullptr, nullptr, nullptr};
    int cur = 0;
    double* staging = nullptr;   // ghosted f64, (10, nx+4, ny+4)
    double* part_dt = nullptr;
    int part_cap = 0, n_part = 0;
    bool part_valid = false;
    double* part_bf = nullptr;
    double* gsum = nullptr;      // global-sum partials
    StepCtl* ctl = nullptr;
    StepCtl* ctl_next = nullptr; // the device loop's current step (the fused stage kernels)
    hipStream_t stream = nullptr;
    hipGraphExec_t graph[3] = {nullptr, nullptr, nullptr};
    const void* halo[2] = {nullptr, nullptr};   // x-slab halo buffers (lo, hi), caller-owned
    int* seg_i0 = nullptr;       // device [nx + 1]: first row of each segment, then nx
    int* seg_next = nullptr;     // device [nx + 1]: the table the device loop's stages 2 and 3 read
    int* seg_first = nullptr;    // device [nx + 1]: the device loop's next first-stage table
    int balance = 0;             // size the segments by cost each step (amhd_set_balance)
    unsigned char* rowflags = nullptr;   // device [nx][RB_FLAG_PITCH]: rows equal to the next, per strip
};


size_t esize(int dtype) { return dtype == AMHD_F32 ? sizeof(float) : sizeof(double); }

// Strips of STRIP y cells; x rows split into n_segs balanced segments.
// A last strip with few valid columns (r = ny - (n_strips - 1) STRIP <= 28)
// is packed: groups of 64 / G lanes (r + 4 <= 64 / G; G = 2, 4 or 8, at
// most the context's limit: 2 by default with two wavefronts per SIMD, else
// 0) take G of its segments per wavefront, so it costs n_segs / G
// wavefronts instead of n_segs.  rows == 0: as many segments as fill every
// resident wavefront slot but one workgroup's once, so that every wave of a
// stage launch runs concurrently with equal work.
void set_tiling(Ctx* c, int rows) {
    c->rows = rows;
    c->n_strips = (c->ny + STRIP - 1) / STRIP;
    const int r = c->ny - (c->n_strips - 1) * STRIP;
    int g = 0;
    if (c->pack_max) g = r + 2 * HALO <= 8 ? 8 : r + 2 * HALO <= 16 ? 4 : r + 2 * HALO <= 32 ? 2 : 0;
    g = std::min(g, c->pack_max);
    c->pk_groups = g;
    c->pk_shift = g == 8 ? 3 : g == 4 ? 4 : g == 2 ? 5 : 0;
    const long wide = g ? c->n_strips - 1 : c->n_strips;
    // the device loop's first and third stages add one side workgroup: keep
    // its slots free so that it runs beside the strips
    const long slots = std::max(1, c->slots - WPB);
    long segs;
    if (rows > 0) {
        segs = (c->nx + rows - 1) / rows;
    } else if (!g) {
        segs = slots / c->n_strips;
    } else {
        // largest S with wide * S + ceil(S / g) <= slots
        segs = (long)slots * g / (wide * g + 1);
        while (segs > 1 && wide * segs + (segs + g - 1) / g > slots) --segs;
    }
    c->n_segs = (int)std::max(1L, std::min(segs, (long)c->nx));
}

int n_waves(const Ctx* c) {
    if (!c->pk_groups) return c->n_strips * c->n_segs;
    return (c->n_strips - 1) * c->n_segs + (c->n_segs + c->pk_groups - 1) / c->pk_groups;
}

// Segment boundaries of the tiling.  nx rows in n_segs segments are
// nx % n_segs segments of one row more than the rest; the longer ones go to
// the segments whose wavefronts belong to the workgroups dispatched first
// onto each CU.  Those lead their CU's second workgroup (the SIMD issues
// from the older wavefront on ties), so the longer segments there end with
// the rest.  The mapping followed here is the stage kernel's: workgroup b
// (dealt round-robin over the 8 XCDs, the first CU-count of them one per CU)
// runs logical block xcd_logical_block(), whose WPB wavefronts w take
// strip w % ns and segment w / ns of the ns wide strips, and after those the
// G segments of a packed wavefront.  Only the balance depends on the
// dispatch order; any boundaries give the same result.
std::vector<int> segment_rows(const Ctx* c) {
    const int ng = c->n_segs, nx = c->nx, g = c->pk_groups;
    const int ns = g ? c->n_strips - 1 : c->n_strips;
    const int base = nx / ng, n_long = nx - base * ng;
    std::vector<int> first_waves(ng, 0);
    const long wide = (long)ns * ng, waves = n_waves(c);
    const long blocks = (waves + WPB - 1) / WPB;
    const long nb8 = blocks & ~7L, per_xcd = nb8 / 8;
    for (long b = 0; b < blocks; ++b) {
        if (b >= c->cus) break;   // only the first-dispatched blocks count
        const long lb = b < nb8 ? (b & 7) * per_xcd + (b >> 3) : b;
        for (int k = 0; k < WPB; ++k) {
            const long w = lb * WPB + k;
            if (w < wide) first_waves[w / ns] += 1;
            else if (w < waves)
                for (long s = (w - wide) * g; s < (w - wide + 1) * g && s < ng; ++s) first_waves[s] += 1;
        }
    }
    std::vector<int> order(ng);
    for (int g = 0; g < ng; ++g) order[g] = g;
    // most first-dispatched wavefronts first; ties spread over the segments
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        if (first_waves[x] != first_waves[y]) return first_waves[x] > first_waves[y];
        return (long)x * 7919 % ng < (long)y * 7919 % ng;
    });
    std::vector<int> len(ng, base);
    for (int k = 0; k < n_long; ++k) len[order[k]] += 1;
    std::vector<int> i0(ng + 1, 0);
    for (int g = 0; g < ng; ++g) i0[g + 1] = i0[g] + len[g];
    return i0;
}

hipError_t upload_segments(Ctx* c) {
    const std::vector<int> i0 = segment_rows(c);
    hipError_t e = hipMemcpyAsync(c->seg_i0, i0.data(), sizeof(int) * i0.size(), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->seg_next, i0.data(), sizeof(int) * i0.size(), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->seg_first, i0.data(), sizeof(int) * i0.size(), hipMemcpyHostToDevice, c->stream);
    return e == hipSuccess ? hipStreamSynchronize(c->stream) : e;
}

template <class real>
StageParams<real> stage_args(const Ctx* c, int stage, const real* u0, const real* us, real* uo) {
    StageParams<real> a{};
    a.u0 = u0; a.us = us; a.uo = uo;
    a.plane = c->plane; a.oplane = c->plane;
    a.pitch = c->pitch; a.opitch = c->pitch;
    a.nx = c->nx; a.ny = c->ny;
    a.bcx_periodic = c->p.bc_x_periodic; a.bcy_periodic = c->p.bc_y_periodic;
    a.n_strips = c->n_strips; a.n_segs = c->n_segs;
    a.in_bytes = a.out_bytes = (unsigned)(esize(c->dtype) * NV * (size_t)c->plane);
    a.gamma = c->p.gamma; a.dx = c->p.dx; a.dy = c->p.dy;
    a.ctl = nullptr;
    a.part_dt = stage == 3 ? c->part_dt : nullptr;
    a.part_bf = stage == 1 ? c->part_bf : nullptr;
    a.halo_lo = (const real*)c->halo[0];
    a.halo_hi = (const real*)c->halo[1];
    a.first_blocks = c->cus;
    a.seg_i0 = c->seg_i0;
    a.pk_groups = c->pk_groups;
    a.pk_shift = c->pk_shift;
    a.rowflags = (c->balance && stage == 3 && (c->pk_groups ? c->n_strips - 1 : c->n_strips) <= RB_FLAG_PITCH)
                     ? c->rowflags : nullptr;
    return a;
}

hipError_t launch_stage_any(const Ctx* c, int stage, const void* u0, const void* us, void* uo, double dt, double ch,
                            double decay, const StepCtl* ctl) {
    if (c->dtype == AMHD_F64) {
        auto a = stage_args<double>(c, stage, (const double*)u0, (const double*)us, (double*)uo);
        a.dt = dt; a.ch = ch; a.decay = decay; a.ctl = ctl;
        return launch_stage_f64(stage, c->p.powell != 0, ctl != nullptr || ch > 0.0, a, c->stream);
    }
    auto a = stage_args<float>(c, stage, (const float*)u0, (const float*)us, (float*)uo);
    a.dt = dt; a.ch = ch; a.decay = decay; a.ctl = ctl;
    return launch_stage_f32(stage, c->p.powell != 0, ctl != nullptr || ch > 0.0, a, c->stream);
}

// One SSP-RK3 step on buffers rotating cur -> cur+1 (mod 3).
hipError_t enqueue_step(Ctx* c, int cur, double dt, double ch, double decay, const StepCtl* ctl) {
    void* u0 = c->buf[cur];
    void* u1 = c->buf[(cur + 1) % 3];
    void* u2 = c->buf[(cur + 2) % 3];
    hipError_t e;
    if ((e = launch_stage_any(c, 1, u0, u0, u1, dt, ch, decay, ctl)) != hipSuccess) return e;
    if ((e = launch_stage_any(c, 2, u0, u1, u2, dt, ch, decay, ctl)) != hipSuccess) return e;
    return launch_stage_any(c, 3, u0, u2, u1, dt, ch, decay, ctl);
}

// One SSP-RK3 step of the device loop with the step control fused into the
// stage kernels (no control launch).  Stage 1 reduces the CFL partials of
// the previous step (or of the state upload) in every workgroup and begins
// the step from the committed block `ctl` into `ctl_next`; with cost-sized
// segments its side workgroup sizes two tables from the row flags, with
// the weights of each: seg_next for this step's stages 2 and 3 and
// seg_first for the next step's stage 1.  Stages 2 and 3 read ctl_next and
// seg_next; the side workgroup of stage 3 closes the step into `ctl` and
// copies seg_first to seg_i0, the table stage 1 reads.  No table or block
// is written while a launch reads it.
// CFL partials the fused third stage leaves: one per strip workgroup.
int loop_parts(const Ctx* c) { return (n_waves(c) + WPB - 1) / WPB; }

template <class real>
hipError_t launch_loop_stage(Ctx* c, int stage, const void* u0, const void* us, void* uo, hipEvent_t ev0 = nullptr,
                             hipEvent_t ev1 = nullptr) {
    StageParams<real> a = stage_args<real>(c, stage, (const real*)u0, (const real*)us, (real*)uo);
    a.fuse = 1;
    const bool bal = c->balance != 0;
    if (stage == 1) {
        a.ctl = c->ctl;
        a.ctl_out = c->ctl_next;
        a.part_in = c->part_dt;
        a.n_part_in = loop_parts(c);
        a.seg_i0 = c->seg_i0;
        if (bal) {
            // this step's table for stages 2 and 3, the next step's for stage 1
            a.side = 1;
            const int nw = c->pk_groups ? c->n_strips - 1 : c->n_strips;
            const bool f32 = sizeof(real) == 4;
            a.rb = Rebalance{c->rowflags, nw, c->nx, c->n_segs, c->seg_next, f32 ? RB_W23 : RB_W23_F64};
            a.rb1 = Rebalance{c->rowflags, nw, c->nx, c->n_segs, c->seg_first, f32 ? RB_W1 : RB_W1_F64};
        }
    } else {
        a.ctl = c->ctl_next;
        a.seg_i0 = bal ? c->seg_next : c->seg_i0;
        if (stage == 3) {
            a.side = 1;
            a.ctl_out = c->ctl;
            a.part_bf_in = c->p.bc_x_periodic ? nullptr : c->part_bf;
            a.n_part_next = loop_parts(c);
            a.seg_copy = bal ? c->seg_i0 : nullptr;
            a.seg_copy_src = c->seg_first;
        }
    }
    if constexpr (sizeof(real) == 8) return launch_stage_f64(stage, c->p.powell != 0, true, a, c->stream, ev0, ev1);
    else return launch_stage_f32(stage, c->p.powell != 0, true, a, c->stream, ev0, ev1);
}

hipError_t enqueue_loop_step(Ctx* c, int cur) {
    void* u0 = c->buf[cur];
    void* u1 = c->buf[(cur + 1) % 3];
    void* u2 = c->buf[(cur + 2) % 3];
    const bool f64 = c->dtype == AMHD_F64;
    hipError_t e = f64 ? launch_loop_stage<double>(c, 1, u0, u0, u1) : launch_loop_stage<float>(c, 1, u0, u0, u1);
    if (e == hipSuccess) e = f64 ? launch_loop_stage<double>(c, 2, u0, u1, u2) : launch_loop_stage<float>(c, 2, u0, u1, u2);
    if (e == hipSuccess) e = f64 ? launch_loop_stage<double>(c, 3, u0, u2, u1) : launch_loop_stage<float>(c, 3, u0, u2, u1);
    return e;
}

hipError_t enqueue_ctl(Ctx* c, int n_part, double host_dt, double cfl, int mode, StepCtl* ctl) {
    const double* bf = c->p.bc_x_periodic ? nullptr : c->part_bf;
    if (c->balance && mode != CTL_DT_ONLY) {
        // the step's last stage has recorded its row flags: size the next
        // step's segments (second workgroup, beside the step control)
        Rebalance rb{c->rowflags, c->pk_groups ? c->n_strips - 1 : c->n_strips, c->nx, c->n_segs, c->seg_i0,
                     c->dtype == AMHD_F32 ? RB_W23 : RB_W23_F64};
        step_ctl_balance_kernel<<<2, 256, 0, c->stream>>>(ctl, c->part_dt, n_part, bf, c->n_strips, c->p.dx, c->p.dy,
                                                          host_dt, cfl, mode, rb);
    } else {
        step_ctl_kernel<<<1, 256, 0, c->stream>>>(ctl, c->part_dt, n_part, bf, c->n_strips, c->p.dx, c->p.dy, host_dt,
                                                  cfl, mode);
    }
    return hipGetLastError();
}

int ensure_partials(Ctx* c) {
    if (c->part_valid) return 0;
    int g;
    if (c->dtype == AMHD_F64)
        g = launch_dt_partials_f64((const double*)c->buf[c->cur], c->plane, c->pitch, c->nx, c->ny, c->p.gamma,
                                   c->part_dt, c->part_cap, c->stream);
    else
        g = launch_dt_partials_f32((const float*)c->buf[c->cur], c->plane, c->pitch, c->nx, c->ny, c->p.gamma,
                                   c->part_dt, c->part_cap, c->stream);
    if (g < 0) return fail(-2, "dt partials launch failed: %s", hipGetErrorString(hipGetLastError()));
    c->n_part = g;
    c->part_valid = true;
    return 0;
}

// The device loop's first stage reduces loop_parts(c) partials (a launch
// constant, so that no wavefront waits on a count before loading them):
// the partials of the current state in exactly that many slots, the
// unused ones 0 (neutral for maxima of non-negative terms).
int ensure_loop_partials(Ctx* c) {
    const int nw = loop_parts(c);
    if (c->part_valid && c->n_part == nw) return 0;
    int g;
    if (c->dtype == AMHD_F64)
        g = launch_dt_partials_f64((const double*)c->buf[c->cur], c->plane, c->pitch, c->nx, c->ny, c->p.gamma,
                                   c->part_dt, nw, c->stream);
    else
        g = launch_dt_partials_f32((const float*)c->buf[c->cur], c->plane, c->pitch, c->nx, c->ny, c->p.gamma,
                                   c->part_dt, nw, c->stream);
    if (g < 0) return fail(-2, "dt partials launch failed: %s", hipGetErrorString(hipGetLastError()));
    if (g < nw) {
        const hipError_t e = hipMemsetAsync(c->part_dt + (size_t)g * NPART_DT, 0, sizeof(double) * NPART_DT * (nw - g),
                                            c->stream);
        if (e != hipSuccess) return fail(-2, "partials fill failed: %s", hipGetErrorString(e));
    }
    c->n_part = nw;
    c->part_valid = true;
    return 0;
}

void drop_graphs(Ctx* c) {
    for (auto& g : c->graph) {
        if (g) hipGraphExecDestroy(g);
        g = nullptr;
    }
}

// FP64 global sums of the interior state (one partial per workgroup).
template <class real>
__global__ __launch_bounds__(256) void global_sums_kernel(const real* __restrict__ u, long plane, int pitch, int nx,
                                                          int ny, double dx, double dy, double* part) {
    double s_rho = 0.0, s_en = 0.0, m_div = 0.0, s_div2 = 0.0;
    const long n = (long)nx * ny;
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
        const int i = (int)(k / ny), j = (int)(k - (long)i * ny);
        const long o = (long)i * pitch + j;
        s_rho += double(u[RHO * plane + o]);
        s_en += double(u[EN * plane + o]);
        if (i >= 1 && i < nx - 1 && j >= 1 && j < ny - 1) {
            const real* bx = u + BX * plane;
            const real* by = u + BY * plane;
            const double d = (0.5 * (double(bx[o]) + double(bx[o + pitch])) -
                              0.5 * (double(bx[o - pitch]) + double(bx[o]))) / dx +
                             (0.5 * (double(by[o]) + double(by[o + 1])) - 0.5 * (double(by[o - 1]) + double(by[o]))) /
                                 dy;
            m_div = fmax(m_div, fabs(d));
            s_div2 += d * d;
        }
    }
    __shared__ double sh[4][4];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    s_rho = wave_sum(s_rho);
    s_en = wave_sum(s_en);
    s_div2 = wave_sum(s_div2);
    m_div = wave_max(m_div);
    if (lane == 0) { sh[0][w] = s_rho; sh[1][w] = s_en; sh[2][w] = m_div; sh[3][w] = s_div2; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0, b = 0, c = 0, d = 0;
        for (int q = 0; q < (int)(blockDim.x >> 6); ++q) { a += sh[0][q]; b += sh[1][q]; c = fmax(c, sh[2][q]); d += sh[3][q]; }
        part[blockIdx.x * 4 + 0] = a;
        part[blockIdx.x * 4 + 1] = b;
        part[blockIdx.x * 4 + 2] = c;
        part[blockIdx.x * 4 + 3] = d;
    }
}

constexpr int GSUM_BLOCKS = 512;

// y-profile of the concentration (solver.py:340, 352-364): for interior
// row i, C = clip(rhoC * (1 / max(rho, 1e-12)), 0, 1) (cons_to_prim's C),
// out[i] = mean_j C and out[nx + i] = mean_j (C - mean)^2, both in FP64 with
// a fixed summation order.  One workgroup per row.
template <class real>
__global__ __launch_bounds__(256) void conc_profile_kernel(const real* __restrict__ u, long plane, int pitch, int nx,
                                                           int ny, double* __restrict__ out) {
    const int i = blockIdx.x;
    const real* r = u + RHO * plane + (long)i * pitch;
    const real* rc = u + RHOC * plane + (long)i * pitch;
    auto conc = [&](int j) {
        const double rho = fmax(double(r[j]), 1e-12);
        const double c = double(rc[j]) * (1.0 / rho);
        return c < 0.0 ? 0.0 : (c > 1.0 ? 1.0 : c);
    };
    __shared__ double sh[4];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    auto block_sum = [&](double v) {
        v = wave_sum(v);
        __syncthreads();
        if (lane == 0) sh[w] = v;
        __syncthreads();
        return (sh[0] + sh[1]) + (sh[2] + sh[3]);
    };
    double s = 0.0;
    for (int j = threadIdx.x; j < ny; j += blockDim.x) s += conc(j);
    const double mean = block_sum(s) / ny;
    double s2 = 0.0;
    for (int j = threadIdx.x; j < ny; j += blockDim.x) {
        const double d = conc(j) - mean;
        s2 += d * d;
    }
    const double var = block_sum(s2) / ny;
    if (threadIdx.x == 0) {
        out[i] = mean;
        out[nx + i] = var;
    }
}

// Primitive values the flow diagnostics read (solver.py:338-347 through
// cons_to_prim, physics.py:284-301, with its 1e-12 floors), in FP64.
template <class real>
struct DiagCell {
    const real* u;
    long plane;
    int pitch;
    __device__ double rho(long o) const { return fmax(double(u[RHO * plane + o]), 1e-12); }
    __device__ double vel(int comp, long o) const { return double(u[comp * plane + o]) * (1.0 / rho(o)); }
    __device__ double conc(long o) const {
        const double c = double(u[RHOC * plane + o]) * (1.0 / rho(o));
        return c < 0.0 ? 0.0 : (c > 1.0 ? 1.0 : c);   // np.clip, NaN passes
    }
};

// Per interior row i: out[i] = sum over y of rho omega^2 with omega =
// dvy/dx - dvx/dy as np.gradient forms it (central differences inside,
// one-sided at the first and last row / column; solver.py:396-399), and
// out[nx + i] = max over y of p + rho |v|^2 / 2 + |B|^2 / 2 (solver.py:412-414).
// One workgroup per row; FP64 with a fixed summation order.
template <class real>
__global__ __launch_bounds__(256) void flow_rows_kernel(const real* __restrict__ u, long plane, int pitch, int nx,
                                                        int ny, double gamma, double dx, double dy,
                                                        double* __restrict__ out) {
    const int i = blockIdx.x;
    const DiagCell<real> q{u, plane, pitch};
    const long row = (long)i * pitch;
    double ens = 0.0, stag = -INFINITY;
    bool nan = false;
    for (int j = threadIdx.x; j < ny; j += blockDim.x) {
        const long o = row + j;
        double dvydx, dvxdy;
        if (nx < 2) dvydx = 0.0;
        else if (i == 0) dvydx = (q.vel(MY, o + pitch) - q.vel(MY, o)) / dx;
        else if (i == nx - 1) dvydx = (q.vel(MY, o) - q.vel(MY, o - pitch)) / dx;
        else dvydx = (q.vel(MY, o + pitch) - q.vel(MY, o - pitch)) / (2.0 * dx);
        if (ny < 2) dvxdy = 0.0;
        else if (j == 0) dvxdy = (q.vel(MX, o + 1) - q.vel(MX, o)) / dy;
        else if (j == ny - 1) dvxdy = (q.vel(MX, o) - q.vel(MX, o - 1)) / dy;
        else dvxdy = (q.vel(MX, o + 1) - q.vel(MX, o - 1)) / (2.0 * dy);
        const double om = (nx > 1 && ny > 1) ? dvydx - dvxdy : 0.0;
        const double rho = q.rho(o);
        ens += rho * (om * om);
        const double vx = q.vel(MX, o), vy = q.vel(MY, o), vz = q.vel(MZ, o);
        const double bx = double(u[BX * plane + o]), by = double(u[BY * plane + o]), bz = double(u[BZ * plane + o]);
        const double ke = 0.5 * rho * (vx * vx + vy * vy + vz * vz);
        const double me = 0.5 * (bx * bx + by * by + bz * bz);
        const double p = fmax((gamma - 1.0) * (double(u[EN * plane + o]) - ke - me), 1e-12);
        const double s = p + 0.5 * rho * (vx * vx + vy * vy) + 0.5 * (bx * bx + by * by + bz * bz);
        nan = nan || s != s;
        stag = s > stag ? s : stag;
    }
    __shared__ double sh[2][4];
    __shared__ int shn[4];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    ens = wave_sum(ens);
    stag = wave_max(stag);
    const bool wnan = __builtin_amdgcn_ballot_w64(nan) != 0;
    if (lane == 0) { sh[0][w] = ens; sh[1][w] = stag; shn[w] = wnan; }
    __syncthreads();
    if (threadIdx.x == 0) {
        out[i] = (sh[0][0] + sh[0][1]) + (sh[0][2] + sh[0][3]);
        const double m = fmax(fmax(sh[1][0], sh[1][1]), fmax(sh[1][2], sh[1][3]));
        out[nx + i] = (shn[0] | shn[1] | shn[2] | shn[3]) ? NAN : m;   // np.max propagates NaN
    }
}

// Interface x position of each interior column j (solver.py:372-384): with
// g(i) = |dC/dx| as np.gradient forms it on C = clip(C, 0, 1), if max g >
// 1e-6, k = the first row holding the maximum, hw = min(10, k, nx - k - 1),
// and for hw > 0 the g-weighted mean of x over rows k - hw .. k + hw when
// the weights sum above 1e-12; else NaN.  A column holding a NaN gradient
// gives NaN (np.max is NaN there).  Workgroup: 64 columns (x of the thread
// index) by IFACE_CHUNKS row chunks (y), combined in LDS.
constexpr int IFACE_CHUNKS = 16;
template <class real>
__global__ __launch_bounds__(64 * IFACE_CHUNKS) void iface_cols_kernel(const real* __restrict__ u, long plane,
                                                                       int pitch, int nx, int ny, double dx,
                                                                       double x_min, double* __restrict__ pos) {
    const int jl = threadIdx.x, ch = threadIdx.y;
    const int j = blockIdx.x * 64 + jl;
    const DiagCell<real> q{u, plane, pitch};
    auto grad = [&](int i) {
        const long o = (long)i * pitch + j;
        double g;
        if (i == 0) g = (q.conc(o + pitch) - q.conc(o)) / dx;
        else if (i == nx - 1) g = (q.conc(o) - q.conc(o - pitch)) / dx;
        else g = (q.conc(o + pitch) - q.conc(o - pitch)) / (2.0 * dx);
        return fabs(g);
    };
    double best = -INFINITY;
    int bi = 0;
    bool nan = false;
    if (j < ny && nx > 1) {
        const int per = (nx + IFACE_CHUNKS - 1) / IFACE_CHUNKS;
        const int ia = ch * per, ib = min(nx, ia + per);
        for (int i = ia; i < ib; ++i) {
            const double g = grad(i);
            nan = nan || g != g;
            if (g > best) { best = g; bi = i; }
        }
    }
    __shared__ double sb[IFACE_CHUNKS][64];
    __shared__ int si[IFACE_CHUNKS][64];
    __shared__ int sn[IFACE_CHUNKS][64];
    sb[ch][jl] = best; si[ch][jl] = bi; sn[ch][jl] = nan;
    __syncthreads();
    if (ch != 0 || j >= ny) return;
    // chunks in row order: a later chunk wins only with a larger value, so
    // the first row holding the maximum is kept (np.argmax)
    for (int c = 1; c < IFACE_CHUNKS; ++c) {
        if (sb[c][jl] > best) { best = sb[c][jl]; bi = si[c][jl]; }
        nan = nan || sn[c][jl];
    }
    double r = NAN;
    if (!nan && nx > 1 && best > 1e-6) {
        const int k = bi;
        const int hw = min(10, min(k, nx - k - 1));
        if (hw > 0) {
            double ws = 0.0, xs = 0.0;
            for (int t = k - hw; t <= k + hw; ++t) {
                const double g = grad(t);
                ws += g;
                xs += (x_min + (t + 0.5) * dx) * g;
            }
            if (ws > 1e-12) r = xs / ws;
        }
    }
    pos[j] = r;
}

}  // namespace

extern "C" {

const char* amhd_last_error(void) { return g_err; }
int amhd_abi_version(void) { return AMHD_ABI_VERSION; }

int amhd_create(void** out, int nx, int ny, int dtype, const amhd_params* p) {
    if (!out || !p) return fail(-1, "amhd_create: null argument");
    *out = nullptr;
    if (nx < 2 || ny < 2) return fail(-1, "amhd_create: grid %dx%d too small (need nx, ny >= 2)", nx, ny);
    if (dtype != AMHD_F64 && dtype != AMHD_F32) return fail(-1, "amhd_create: unknown dtype %d", dtype);
    if (!(p->gamma > 1.0) || !(p->dx > 0.0) || !(p->dy > 0.0))
        return fail(-1, "amhd_create: need gamma > 1, dx > 0, dy > 0");
    Ctx* c = new Ctx();
    c->nx = nx; c->ny = ny; c->dtype = dtype; c->p = *p;
    c->pitch = ny;
    c->plane = (long)nx * ny;
    if ((double)esize(dtype) * NV * (double)c->plane > 4294967295.0 ||
        (double)sizeof(double) * NV * (double)(nx + 4) * (ny + 4) > 4294967295.0) {
        delete c;
        return fail(-4, "amhd_create: %dx%d state exceeds the 4 GiB reach of 32-bit buffer offsets", nx, ny);
    }
    {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0) {
            const int wpc = dtype == AMHD_F64 ? stage_waves_per_cu_f64(p->powell != 0)
                                             : stage_waves_per_cu_f32(p->powell != 0);
            c->slots = cus * wpc;
            c->cus = cus;
        }
    }
    set_tiling(c, 0);
    const size_t sb = esize(dtype) * NV * (size_t)c->plane;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    for (int k = 0; k < 3 && e == hipSuccess; ++k) e = hipMalloc(&c->buf[k], sb);
    if (e == hipSuccess) e = hipMalloc(&c->staging, sizeof(double) * NV * (size_t)(nx + 4) * (ny + 4));
    // CFL partials: one per wavefront at any tiling (rows >= 1), or one per
    // workgroup of the standalone partials kernel.
    c->part_cap = std::max((int)((long)nx * ((ny + STRIP - 1) / STRIP)), 4096);   // >= n_strips * n_segs
    if (e == hipSuccess) e = hipMalloc(&c->part_dt, sizeof(double) * NPART_DT * (size_t)c->part_cap);
    if (e == hipSuccess) e = hipMalloc(&c->part_bf, sizeof(double) * 2 * (size_t)c->n_strips);
    if (e == hipSuccess) e = hipMalloc(&c->gsum, sizeof(double) * 4 * GSUM_BLOCKS);
    if (e == hipSuccess) e = hipMalloc(&c->ctl, sizeof(StepCtl));
    if (e == hipSuccess) e = hipMalloc(&c->ctl_next, sizeof(StepCtl));
    if (e == hipSuccess) e = hipMalloc(&c->seg_i0, sizeof(int) * (size_t)(nx + 1));
    if (e == hipSuccess) e = hipMalloc(&c->seg_next, sizeof(int) * (size_t)(nx + 1));
    if (e == hipSuccess) e = hipMalloc(&c->seg_first, sizeof(int) * (size_t)(nx + 1));
    // cost-sized segments: rows that repeat the previous one are cheaper
    c->balance = 1;
    if (e == hipSuccess) e = hipMalloc(&c->rowflags, (size_t)RB_FLAG_PITCH * nx);
    if (e == hipSuccess) e = hipMemsetAsync(c->rowflags, 0, (size_t)RB_FLAG_PITCH * nx, c->stream);
    if (e == hipSuccess) e = upload_segments(c);
    if (e == hipSuccess) {
        StepCtl z{};
        e = hipMemcpyAsync(c->ctl, &z, sizeof z, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(c->ctl_next, &z, sizeof z, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = hipMemsetAsync(c->part_bf, 0, sizeof(double) * 2 * c->n_strips, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    }
    if (e != hipSuccess) {
        int rc = fail(e == hipErrorOutOfMemory ? -3 : -2, "amhd_create: %s", hipGetErrorString(e));
        amhd_destroy(c);
        return rc;
    }
    *out = c;
    return 0;
}

int amhd_destroy(void* ctx) {
    if (!ctx) return 0;
    Ctx* c = (Ctx*)ctx;
    if (c->stream) hipStreamSynchronize(c->stream);
    drop_graphs(c);
    for (auto& b : c->buf) if (b) hipFree(b);
    if (c->staging) hipFree(c->staging);
    if (c->part_dt) hipFree(c->part_dt);
    if (c->part_bf) hipFree(c->part_bf);
    if (c->gsum) hipFree(c->gsum);
    if (c->ctl) hipFree(c->ctl);
    if (c->ctl_next) hipFree(c->ctl_next);
    if (c->seg_i0) hipFree(c->seg_i0);
    if (c->seg_next) hipFree(c->seg_next);
    if (c->seg_first) hipFree(c->seg_first);
    if (c->rowflags) hipFree(c->rowflags);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

int amhd_set_rows(void* ctx, int rows) {
    if (!ctx) return fail(-1, "amhd_set_rows: null context");
    if (rows < 0) return fail(-1, "amhd_set_rows: rows must be >= 0");
    Ctx* c = (Ctx*)ctx;
    HIP_TRY(hipStreamSynchronize(c->stream));
    set_tiling(c, rows);
    HIP_TRY(upload_segments(c));
    drop_graphs(c);
    c->part_valid = false;   // per-wave partials follow the tiling
    return 0;
}

int amhd_set_balance(void* ctx, int on) {
    if (!ctx) return fail(-1, "amhd_set_balance: null context");
    if (on != 0 && on != 1) return fail(-1, "amhd_set_balance: on must be 0 or 1 (got %d)", on);
    Ctx* c = (Ctx*)ctx;
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->balance = on;
    HIP_TRY(upload_segments(c));
    drop_graphs(c);
    return 0;
}

#ifdef AMHD_WAVE_TIMES
// Diagnostic builds only: the context's segment tables (each n_segs + 1
// ints: the one stage 1 reads, the one the device loop's stages 2 and 3
// read, the next step's first-stage table) and row flags (nx *
// RB_FLAG_PITCH bytes) copied to host memory.  Not part of the ABI of
// include/aeromhd.h.
int amhd_debug_tables(void* ctx, int* seg_i0, int* seg_next, int* seg_first, unsigned char* flags, int* n_segs) {
    if (!ctx || !seg_i0 || !seg_next || !seg_first || !flags || !n_segs)
        return fail(-1, "amhd_debug_tables: null argument");
    Ctx* c = (Ctx*)ctx;
    HIP_TRY(hipStreamSynchronize(c->stream));
    *n_segs = c->n_segs;
    const size_t tb = sizeof(int) * (size_t)(c->n_segs + 1);
    HIP_TRY(hipMemcpy(seg_i0, c->seg_i0, tb, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(seg_next, c->seg_next, tb, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(seg_first, c->seg_first, tb, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(flags, c->rowflags, (size_t)RB_FLAG_PITCH * c->nx, hipMemcpyDeviceToHost));
    return 0;
}
#endif

int amhd_set_packing(void* ctx, int max_groups) {
    if (!ctx) return fail(-1, "amhd_set_packing: null context");
    if (max_groups != 0 && max_groups != 2 && max_groups != 4 && max_groups != 8)
        return fail(-1, "amhd_set_packing: max_groups must be 0, 2, 4 or 8 (got %d)", max_groups);
    Ctx* c = (Ctx*)ctx;
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->pack_max = max_groups;
    set_tiling(c, c->rows);
    HIP_TRY(upload_segments(c));
    drop_graphs(c);
    c->part_valid = false;
    return 0;
}

int amhd_get_rows(void* ctx) {
    if (!ctx) return -1;
    Ctx* c = (Ctx*)ctx;
    return (c->nx + c->n_segs - 1) / c->n_segs;
}

int amhd_upload_state(void* ctx, const double* U, int on_device) {
    if (!ctx || !U) return fail(-1, "amhd_upload_state: null argument");
    Ctx* c = (Ctx*)ctx;
    const size_t gbytes = sizeof(double) * NV * (size_t)(c->nx + 4) * (c->ny + 4);
    HIP_TRY(hipMemcpyAsync(c->staging, U, gbytes, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                           c->stream));
    hipError_t e = c->dtype == AMHD_F64
                       ? launch_ingest_f64(c->staging, (double*)c->buf[c->cur], c->plane, c->pitch, c->nx, c->ny,
                                           c->stream)
                       : launch_ingest_f32(c->staging, (float*)c->buf[c->cur], c->plane, c->pitch, c->nx, c->ny,
                                           c->stream);
    HIP_TRY(e);
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->part_valid = false;
    return 0;
}

int amhd_download_state(void* ctx, double* U, int on_device) {
    if (!ctx || !U) return fail(-1, "amhd_download_state: null argument");
    Ctx* c = (Ctx*)ctx;
    hipError_t e = c->dtype == AMHD_F64
                       ? launch_materialize_f64((const double*)c->buf[c->cur], c->plane, c->pitch, c->nx, c->ny,
                                                c->p.bc_x_periodic, c->p.bc_y_periodic, c->staging, c->stream)
                       : launch_materialize_f32((const float*)c->buf[c->cur], c->plane, c->pitch, c->nx, c->ny,
                                                c->p.bc_x_periodic, c->p.bc_y_periodic, c->staging, c->stream);
    HIP_TRY(e);
    const size_t gbytes = sizeof(double) * NV * (size_t)(c->nx + 4) * (c->ny + 4);
    HIP_TRY(hipMemcpyAsync(U, c->staging, gbytes, on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int amhd_compute_dt(void* ctx, double cfl, double* dt, double* ch) {
    if (!ctx || !dt || !ch) return fail(-1, "amhd_compute_dt: null argument");
    Ctx* c = (Ctx*)ctx;
    int rc = ensure_partials(c);
    if (rc) return rc;
    HIP_TRY(enqueue_ctl(c, c->n_part, 0.0, cfl, CTL_DT_ONLY, c->ctl));
    StepCtl h;
    HIP_TRY(hipMemcpyAsync(&h, c->ctl, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    *dt = h.dt;
    *ch = h.ch;
    return 0;
}

int amhd_step_ssprk3(void* ctx, double dt, double ch, double psi_decay) {
    if (!ctx) return fail(-1, "amhd_step_ssprk3: null context");
    Ctx* c = (Ctx*)ctx;
    HIP_TRY(enqueue_step(c, c->cur, dt, ch, psi_decay, nullptr));
    HIP_TRY(enqueue_ctl(c, 0, dt, 0.0, CTL_HOST_STEP_END, c->ctl));
    c->cur = (c->cur + 1) % 3;
    c->n_part = n_waves(c);
    c->part_valid = true;
    return 0;
}

int amhd_set_halo(void* ctx, const void* lo, const void* hi) {
    if (!ctx) return fail(-1, "amhd_set_halo: null context");
    Ctx* c = (Ctx*)ctx;
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (c->halo[0] != lo || c->halo[1] != hi) drop_graphs(c);
    c->halo[0] = lo;
    c->halo[1] = hi;
    return 0;
}

int amhd_stage(void* ctx, int stage, double dt, double ch, double psi_decay) {
    if (!ctx) return fail(-1, "amhd_stage: null context");
    if (stage < 1 || stage > 3) return fail(-1, "amhd_stage: stage must be 1, 2 or 3 (got %d)", stage);
    Ctx* c = (Ctx*)ctx;
    void* u0 = c->buf[c->cur];
    void* u1 = c->buf[(c->cur + 1) % 3];
    void* u2 = c->buf[(c->cur + 2) % 3];
    if (stage == 1) HIP_TRY(launch_stage_any(c, 1, u0, u0, u1, dt, ch, psi_decay, nullptr));
    else if (stage == 2) HIP_TRY(launch_stage_any(c, 2, u0, u1, u2, dt, ch, psi_decay, nullptr));
    else {
        HIP_TRY(launch_stage_any(c, 3, u0, u2, u1, dt, ch, psi_decay, nullptr));
        HIP_TRY(enqueue_ctl(c, 0, dt, 0.0, CTL_HOST_STEP_END, c->ctl));
        c->cur = (c->cur + 1) % 3;
        c->n_part = n_waves(c);
        c->part_valid = true;
    }
    return 0;
}

int amhd_edge_rows(void* ctx, int stage, int side, void* dst) {
    if (!ctx || !dst) return fail(-1, "amhd_edge_rows: null argument");
    if (stage < 1 || stage > 3 || (side != 0 && side != 1)) return fail(-1, "amhd_edge_rows: bad stage or side");
    Ctx* c = (Ctx*)ctx;
    if (c->nx < 2) return fail(-1, "amhd_edge_rows: the slab needs at least 2 rows");
    const void* src = c->buf[(c->cur + stage - 1) % 3];   // the stage's input
    const size_t es = esize(c->dtype);
    const size_t row0 = side == 0 ? 0 : (size_t)(c->nx - 2);
    HIP_TRY(hipMemcpy2DAsync(dst, 2 * (size_t)c->pitch * es, (const char*)src + row0 * c->pitch * es,
                             (size_t)c->plane * es, 2 * (size_t)c->pitch * es, NV, hipMemcpyDeviceToDevice, c->stream));
    return 0;
}

int amhd_cfl_maxima(void* ctx, double m[4]) {
    if (!ctx || !m) return fail(-1, "amhd_cfl_maxima: null argument");
    Ctx* c = (Ctx*)ctx;
    int rc = ensure_partials(c);
    if (rc) return rc;
    std::vector<double> h((size_t)c->n_part * NPART_DT);
    HIP_TRY(hipMemcpyAsync(h.data(), c->part_dt, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int q = 0; q < NPART_DT; ++q) m[q] = 0.0;
    for (int k = 0; k < c->n_part; ++k)
        for (int q = 0; q < NPART_DT; ++q) m[q] = std::max(m[q], h[(size_t)k * NPART_DT + q]);
    return 0;
}

int amhd_dt_from_maxima(const double m[4], double cfl, double dx, double dy, double* dt, double* ch) {
    if (!m || !dt || !ch || !(dx > 0) || !(dy > 0)) return fail(-1, "amhd_dt_from_maxima: bad argument");
    // the expressions of the device step control (solver.py:211-223)
    const double cfmax = std::sqrt(m[0]);
    const double c_h = std::fmax(cfmax + m[1], 1.0) * 1.5;
    const double sx = m[2] + c_h, sy = m[3] + c_h;
    const double sm = std::fmax(sx / dx, sy / dy);
    *dt = (sm < 1e-14) ? cfl * std::fmin(dx, dy) : cfl / sm;
    *ch = c_h;
    return 0;
}

int amhd_synchronize(void* ctx) {
    if (!ctx) return fail(-1, "amhd_synchronize: null context");
    HIP_TRY(hipStreamSynchronize(((Ctx*)ctx)->stream));
    return 0;
}

int amhd_run(void* ctx, int n_steps, double cfl, double t_end, int max_steps, int use_graph, double* t_inout,
             int* step_inout, double* last_dt) {
    if (!ctx || !t_inout || !step_inout) return fail(-1, "amhd_run: null argument");
    if (n_steps < 0) return fail(-1, "amhd_run: n_steps < 0");
    Ctx* c = (Ctx*)ctx;
    StepCtl h;
    HIP_TRY(hipMemcpyAsync(&h, c->ctl, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    const int start = *step_inout;
    h.t = *t_inout;
    h.steps = start;
    h.t_end = t_end;
    h.cfl = cfl;
    h.max_steps = (int)std::min<long>((long)max_steps, (long)start + n_steps);
    h.glm_alpha = c->p.glm_alpha;
    h.min_dxdy = std::min(c->p.dx, c->p.dy);
    h.done = 0;
    h.pending = 0;
    int rc = ensure_loop_partials(c);
    if (rc) return rc;
    h.n_part = c->n_part;   // the first stage reduces these partials
    HIP_TRY(hipMemcpyAsync(c->ctl, &h, sizeof h, hipMemcpyHostToDevice, c->stream));
    const int nw = loop_parts(c);
    if (use_graph && n_steps > 0) {

        hipGraphExec_t& ge = c->graph[c->cur];
        if (!ge) {
            hipGraph_t g;
            HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            hipError_t e = hipSuccess;
            for (int s = 0; s < 3 && e == hipSuccess; ++s) e = enqueue_loop_step(c, (c->cur + s) % 3);
            hipError_t e2 = hipStreamEndCapture(c->stream, &g);
            HIP_TRY(e);
            HIP_TRY(e2);
            HIP_TRY(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
            hipGraphDestroy(g);
        }
        for (int r = 0; r < (n_steps + 2) / 3; ++r) HIP_TRY(hipGraphLaunch(ge, c->stream));
    } else {
        for (int s = 0; s < n_steps; ++s) HIP_TRY(enqueue_loop_step(c, (c->cur + s) % 3));
    }
    HIP_TRY(hipMemcpyAsync(&h, c->ctl, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    const int ran = h.steps - start;
    c->cur = (c->cur + ran) % 3;
    if (ran > 0) {
        c->n_part = nw;
        c->part_valid = true;
    }
    *t_inout = h.t;
    *step_inout = h.steps;
    if (last_dt) *last_dt = h.dt;
    return 0;
}

int amhd_run_profiled(void* ctx, int n_steps, double cfl, double t_end, int max_steps, double* t_inout,
                      int* step_inout, double* last_dt, double stage_ms[3], double* ctl_ms) {
    if (!ctx || !t_inout || !step_inout || !stage_ms) return fail(-1, "amhd_run_profiled: null argument");
    if (n_steps < 0) return fail(-1, "amhd_run_profiled: n_steps < 0");
    Ctx* c = (Ctx*)ctx;
    StepCtl h;
    HIP_TRY(hipMemcpyAsync(&h, c->ctl, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    const int start = *step_inout;
    h.t = *t_inout; h.steps = start; h.t_end = t_end; h.cfl = cfl;
    h.max_steps = (int)std::min<long>((long)max_steps, (long)start + n_steps);
    h.glm_alpha = c->p.glm_alpha; h.min_dxdy = std::min(c->p.dx, c->p.dy);
    h.done = 0; h.pending = 0;
    int rc = ensure_loop_partials(c);
    if (rc) return rc;
    h.n_part = c->n_part;
    HIP_TRY(hipMemcpyAsync(c->ctl, &h, sizeof h, hipMemcpyHostToDevice, c->stream));
    const int nw = loop_parts(c);
    // the start and end of each of the 3 launches of a step (the stage
    // kernels carry the step control), recorded by the dispatches themselves
    std::vector<hipEvent_t> ev((size_t)n_steps * 6);
    for (auto& e : ev) HIP_TRY(hipEventCreate(&e));
    hipError_t err = hipSuccess;
    for (int st = 0; st < n_steps && err == hipSuccess; ++st) {
        const int cur = (c->cur + st) % 3;
        void* u0 = c->buf[cur];
        void* u1 = c->buf[(cur + 1) % 3];
        void* u2 = c->buf[(cur + 2) % 3];
        hipEvent_t* e = &ev[(size_t)st * 6];
        const bool f64 = c->dtype == AMHD_F64;
        if (err == hipSuccess)
            err = f64 ? launch_loop_stage<double>(c, 1, u0, u0, u1, e[0], e[1])
                      : launch_loop_stage<float>(c, 1, u0, u0, u1, e[0], e[1]);
        if (err == hipSuccess)
            err = f64 ? launch_loop_stage<double>(c, 2, u0, u1, u2, e[2], e[3])
                      : launch_loop_stage<float>(c, 2, u0, u1, u2, e[2], e[3]);
        if (err == hipSuccess)
            err = f64 ? launch_loop_stage<double>(c, 3, u0, u2, u1, e[4], e[5])
                      : launch_loop_stage<float>(c, 3, u0, u2, u1, e[4], e[5]);
    }
    if (err == hipSuccess) err = hipMemcpyAsync(&h, c->ctl, sizeof h, hipMemcpyDeviceToHost, c->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(c->stream);
    double acc[3] = {0, 0, 0};
    for (int st = 0; st < n_steps && err == hipSuccess; ++st) {
        for (int k = 0; k < 3 && err == hipSuccess; ++k) {
            float ms = 0.f;
            err = hipEventElapsedTime(&ms, ev[(size_t)st * 6 + 2 * k], ev[(size_t)st * 6 + 2 * k + 1]);
            acc[k] += ms;
        }
    }
    for (auto& e : ev) hipEventDestroy(e);
    HIP_TRY(err);
    const int ran = h.steps - start;
    c->cur = (c->cur + ran) % 3;
    if (ran > 0) { c->n_part = nw; c->part_valid = true; }
    *t_inout = h.t;
    *step_inout = h.steps;
    if (last_dt) *last_dt = h.dt;
    for (int k = 0; k < 3; ++k) stage_ms[k] = acc[k];
    if (ctl_ms) *ctl_ms = 0.0;   // no step-control launch in the device loop
    return 0;
}

int amhd_boundary_energy(void* ctx, double* cumulative, double* last) {
    if (!ctx) return fail(-1, "amhd_boundary_energy: null context");
    Ctx* c = (Ctx*)ctx;
    StepCtl h;
    HIP_TRY(hipMemcpyAsync(&h, c->ctl, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (cumulative) *cumulative = h.bflux_cum;
    if (last) *last = h.bflux;
    return 0;
}

int amhd_reset_boundary_energy(void* ctx, double value) {
    if (!ctx) return fail(-1, "amhd_reset_boundary_energy: null context");
    Ctx* c = (Ctx*)ctx;
    StepCtl h;
    HIP_TRY(hipMemcpyAsync(&h, c->ctl, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    h.bflux_cum = value;
    h.bflux = 0.0;
    HIP_TRY(hipMemcpyAsync(c->ctl, &h, sizeof h, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int amhd_compute_rhs(void* ctx, double* U, double* R, double ch) {
    if (!ctx || !U || !R) return fail(-1, "amhd_compute_rhs: null argument");
    Ctx* c = (Ctx*)ctx;
    const int nxt = c->nx + 4, nyt = c->ny + 4;
    const long gplane = (long)nxt * nyt;
    apply_bc_kernel<<<grid_for(gplane, 256), 256, 0, c->stream>>>(U, c->nx, c->ny, c->p.bc_x_periodic,
                                                                   c->p.bc_y_periodic);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemsetAsync(R, 0, sizeof(double) * NV * gplane, c->stream));
    StageParams<double> a{};
    a.u0 = a.us = U + 2 * nyt + 2;
    a.uo = R + 2 * nyt + 2;
    a.plane = a.oplane = gplane;
    a.pitch = a.opitch = nyt;
    a.nx = c->nx; a.ny = c->ny;
    a.bcx_periodic = c->p.bc_x_periodic; a.bcy_periodic = c->p.bc_y_periodic;
    a.n_strips = c->n_strips; a.n_segs = c->n_segs;
    a.in_bytes = a.out_bytes = (unsigned)(sizeof(double) * ((size_t)NV * gplane - (size_t)(2 * nyt + 2)));
    a.gamma = c->p.gamma; a.dx = c->p.dx; a.dy = c->p.dy;
    a.dt = 0.0; a.ch = ch; a.decay = 1.0;
    HIP_TRY(launch_stage_f64(0, c->p.powell != 0, ch > 0.0, a, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int amhd_state_check(void* ctx, double out[4]) {
    if (!ctx || !out) return fail(-1, "amhd_state_check: null argument");
    Ctx* c = (Ctx*)ctx;
    const int g = grid_for((long)c->nx * c->ny, 256, GSUM_BLOCKS);
    if (c->dtype == AMHD_F64)
        state_check_kernel<double><<<g, 256, 0, c->stream>>>((const double*)c->buf[c->cur], c->plane, c->pitch, c->nx,
                                                             c->ny, c->p.gamma, c->p.dx, c->p.dy, c->gsum);
    else
        state_check_kernel<float><<<g, 256, 0, c->stream>>>((const float*)c->buf[c->cur], c->plane, c->pitch, c->nx,
                                                            c->ny, c->p.gamma, c->p.dx, c->p.dy, c->gsum);
    HIP_TRY(hipGetLastError());
    double h[4 * GSUM_BLOCKS];
    HIP_TRY(hipMemcpyAsync(h, c->gsum, sizeof(double) * 4 * g, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    out[0] = h[0]; out[1] = h[1]; out[2] = h[2]; out[3] = h[3];
    for (int k = 1; k < g; ++k) {
        out[0] = std::min(out[0], h[4 * k]); out[1] = std::min(out[1], h[4 * k + 1]);
        out[2] = std::max(out[2], h[4 * k + 2]); out[3] += h[4 * k + 3];
    }
    return 0;
}

int amhd_state_ptr(void* ctx, void** state) {
    if (!ctx || !state) return fail(-1, "amhd_state_ptr: null argument");
    Ctx* c = (Ctx*)ctx;
    HIP_TRY(hipStreamSynchronize(c->stream));
    *state = c->buf[c->cur];
    c->part_valid = false;   // the caller may edit the state: recompute the CFL partials
    return 0;
}

int amhd_concentration_profile(void* ctx, double* mean, double* var) {
    if (!ctx || !mean || !var) return fail(-1, "amhd_concentration_profile: null argument");
    Ctx* c = (Ctx*)ctx;
    double* d = c->staging;   // 2 nx doubles of scratch (the staging buffer holds far more)
    if (c->dtype == AMHD_F64)
        conc_profile_kernel<double><<<c->nx, 256, 0, c->stream>>>((const double*)c->buf[c->cur], c->plane, c->pitch,
                                                                 c->nx, c->ny, d);
    else
        conc_profile_kernel<float><<<c->nx, 256, 0, c->stream>>>((const float*)c->buf[c->cur], c->plane, c->pitch,
                                                                c->nx, c->ny, d);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(mean, d, sizeof(double) * c->nx, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(var, d + c->nx, sizeof(double) * c->nx, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int amhd_flow_diagnostics(void* ctx, double* row_ens, double* row_stag, double* col_pos) {
    if (!ctx || !row_ens || !row_stag || !col_pos) return fail(-1, "amhd_flow_diagnostics: null argument");
    Ctx* c = (Ctx*)ctx;
    double* d = c->staging;   // 2 nx + ny doubles of scratch
    const dim3 ib(64, IFACE_CHUNKS);
    const int ig = (c->ny + 63) / 64;
    if (c->dtype == AMHD_F64) {
        const double* u = (const double*)c->buf[c->cur];
        flow_rows_kernel<double><<<c->nx, 256, 0, c->stream>>>(u, c->plane, c->pitch, c->nx, c->ny, c->p.gamma,
                                                               c->p.dx, c->p.dy, d);
        iface_cols_kernel<double><<<ig, ib, 0, c->stream>>>(u, c->plane, c->pitch, c->nx, c->ny, c->p.dx, c->p.x_min,
                                                            d + 2 * (long)c->nx);
    } else {
        const float* u = (const float*)c->buf[c->cur];
        flow_rows_kernel<float><<<c->nx, 256, 0, c->stream>>>(u, c->plane, c->pitch, c->nx, c->ny, c->p.gamma,
                                                              c->p.dx, c->p.dy, d);
        iface_cols_kernel<float><<<ig, ib, 0, c->stream>>>(u, c->plane, c->pitch, c->nx, c->ny, c->p.dx, c->p.x_min,
                                                           d + 2 * (long)c->nx);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(row_ens, d, sizeof(double) * c->nx, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(row_stag, d + c->nx, sizeof(double) * c->nx, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(col_pos, d + 2 * (long)c->nx, sizeof(double) * c->ny, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int amhd_get_stream(void* ctx, void** stream) {
    if (!ctx || !stream) return fail(-1, "amhd_get_stream: null argument");
    *stream = (void*)((Ctx*)ctx)->stream;
    return 0;
}

int amhd_global_sums(void* ctx, double out[4]) {
    if (!ctx || !out) return fail(-1, "amhd_global_sums: null argument");
    Ctx* c = (Ctx*)ctx;
    const int g = grid_for((long)c->nx * c->ny, 256, GSUM_BLOCKS);
    if (c->dtype == AMHD_F64)
        global_sums_kernel<double><<<g, 256, 0, c->stream>>>((const double*)c->buf[c->cur], c->plane, c->pitch, c->nx,
                                                             c->ny, c->p.dx, c->p.dy, c->gsum);
    else
        global_sums_kernel<float><<<g, 256, 0, c->stream>>>((const float*)c->buf[c->cur], c->plane, c->pitch, c->nx,
                                                            c->ny, c->p.dx, c->p.dy, c->gsum);
    HIP_TRY(hipGetLastError());
    double h[4 * GSUM_BLOCKS];
    HIP_TRY(hipMemcpyAsync(h, c->gsum, sizeof(double) * 4 * g, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    double a = 0, b = 0, m = 0, d2 = 0;
    for (int k = 0; k < g; ++k) { a += h[4 * k]; b += h[4 * k + 1]; m = std::max(m, h[4 * k + 2]); d2 += h[4 * k + 3]; }
    out[0] = a * c->p.dx * c->p.dy;
    out[1] = b * c->p.dx * c->p.dy;
    out[2] = m;
    out[3] = std::sqrt(d2 / ((double)c->nx * c->ny));
    return 0;
}

// ------------------------------------------------------------- standalone
int amhd_cons_to_prim(const double* U, double* W, long n, double gamma, void* stream) {
    if (!U || !W || n < 0) return fail(-1, "amhd_cons_to_prim: bad argument");
    if (n == 0) return 0;
    cons_to_prim_kernel<<<grid_for(n, 256), 256, 0, (hipStream_t)stream>>>(U, W, n, gamma);
    HIP_TRY(hipGetLastError());
    return 0;
}

int amhd_prim_to_cons(const double* W, double* U, long n, double gamma, void* stream) {
    if (!U || !W || n < 0) return fail(-1, "amhd_prim_to_cons: bad argument");
    if (n == 0) return 0;
    prim_to_cons_kernel<<<grid_for(n, 256), 256, 0, (hipStream_t)stream>>>(W, U, n, gamma);
    HIP_TRY(hipGetLastError());
    return 0;
}

int amhd_muscl(const double* W, int nxt, int nyt, int dir, double* WL, double* WR, void* stream) {
    if (!W || !WL || !WR) return fail(-1, "amhd_muscl: null argument");
    if (dir != 0 && dir != 1) return fail(-1, "amhd_muscl: dir must be 0 or 1");
    if ((dir == 0 && nxt < 4) || (dir == 1 && nyt < 4) || nxt < 1 || nyt < 1)
        return fail(-1, "amhd_muscl: need at least 4 cells along the reconstruction axis");
    const long nf = dir == 0 ? (long)(nxt - 3) * nyt : (long)nxt * (nyt - 3);
    muscl_kernel<<<grid_for(nf, 256), 256, 0, (hipStream_t)stream>>>(W, nxt, nyt, dir, WL, WR);
    HIP_TRY(hipGetLastError());
    return 0;
}

int amhd_hlld_flux(const void* WL, const void* WR, void* F, void* smax_arr, double* smax, long n, int dir,
                   double gamma, double ch, double floor_rho, double floor_pr, int dtype, void* stream) {
    if (!WL || !WR || !F || n < 0) return fail(-1, "amhd_hlld_flux: bad argument");
    if (dir != 0 && dir != 1) return fail(-1, "amhd_hlld_flux: dir must be 0 or 1");
    hipStream_t s = (hipStream_t)stream;
    unsigned long long* bits = nullptr;
    if (smax) {
        Scratch* sc;
        int rc = get_scratch(&sc);
        if (rc) return rc;
        bits = sc->bits;
        HIP_TRY(hipMemsetAsync(bits, 0, sizeof *bits, s));
    }
    if (n > 0) {
        if (dtype == AMHD_F64) {
            const FluxConst<double> kc = make_flux_const<double>(gamma, ch, floor_rho, floor_pr);
            hlld_kernel<double, ExactMath<double>><<<grid_for(n, 256), 256, 0, s>>>(
                (const double*)WL, (const double*)WR, (double*)F, (double*)smax_arr, n, dir, kc, bits);
            HIP_TRY(hipGetLastError());
        } else if (dtype == AMHD_F32) {
            HIP_TRY(launch_hlld_f32((const float*)WL, (const float*)WR, (float*)F, (float*)smax_arr, n, dir, gamma, ch, floor_rho,
                                    floor_pr, bits, s));
        } else {
            return fail(-1, "amhd_hlld_flux: unknown dtype %d", dtype);
        }
    }
    if (smax) {
        unsigned long long h = 0;
        HIP_TRY(hipMemcpyAsync(&h, bits, sizeof h, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        double v;
        std::memcpy(&v, &h, sizeof v);
        *smax = v;
    }
    return 0;
}

int amhd_apply_bc(double* U, int nx, int ny, int bcx, int bcy, void* stream) {
    if (!U || nx < 2 || ny < 2) return fail(-1, "amhd_apply_bc: bad argument");
    const long gplane = (long)(nx + 4) * (ny + 4);
    apply_bc_kernel<<<grid_for(gplane, 256), 256, 0, (hipStream_t)stream>>>(U, nx, ny, bcx, bcy);
    HIP_TRY(hipGetLastError());
    return 0;
}

int amhd_enforce_scalar_bounds(double* U, long n, void* stream) {
    if (!U || n < 0) return fail(-1, "amhd_enforce_scalar_bounds: bad argument");
    if (n == 0) return 0;
    enforce_bounds_kernel<<<grid_for(n, 256), 256, 0, (hipStream_t)stream>>>(U, n);
    HIP_TRY(hipGetLastError());
    return 0;
}

int amhd_compute_dt_w(const double* W, long n, double gamma, double cfl, double dx, double dy, double* dt,
                      double* ch, void* stream) {
    if (!W || n <= 0 || !dt || !ch) return fail(-1, "amhd_compute_dt_w: bad argument");
    Scratch* sc;
    int rc = get_scratch(&sc);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    const int g = grid_for(n, 256, SCRATCH_PARTS);
    dt_partials_w_kernel<<<g, 256, 0, s>>>(W, n, gamma, sc->part);
    HIP_TRY(hipGetLastError());
    step_ctl_kernel<<<1, 256, 0, s>>>(sc->ctl, sc->part, g, nullptr, 0, dx, dy, 0.0, cfl, CTL_DT_ONLY);
    HIP_TRY(hipGetLastError());
    StepCtl h;
    HIP_TRY(hipMemcpyAsync(&h, sc->ctl, sizeof h, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    *dt = h.dt;
    *ch = h.ch;
    return 0;
}

// amhd_compute_dt_w with its CFL partials and step-control block in the
// library's file-scope device array (no allocation).  Calls on one device
// share that array: overlapping calls on one device race.  Every field the
// step control reads back is written by this call (the partials by the
// partial kernel, dt / ch / cf / v maxima by the control kernel), so the
// previous call's contents never reach the result.
int amhd_compute_dt_w_noalloc(const double* W, long n, double gamma, double cfl, double dx, double dy, double* dt,
                              double* ch, void* stream) {
    if (!W || n <= 0 || !dt || !ch) return fail(-1, "amhd_compute_dt_w_noalloc: bad argument");
    const int g = grid_for(n, 256, NOALLOC_PARTS);   // grid-stride: any n fits in NOALLOC_PARTS partials
    if (g < 1 || g > NOALLOC_PARTS)
        return fail(-1, "amhd_compute_dt_w_noalloc: %d partials exceed the scratch of %d", g, NOALLOC_PARTS);
    double* sp = nullptr;
    int rc = noalloc_scratch(&sp);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    StepCtl* ctl = (StepCtl*)(sp + NOALLOC_CTL);
    dt_partials_w_kernel<<<g, 256, 0, s>>>(W, n, gamma, sp);
    HIP_TRY(hipGetLastError());
    step_ctl_kernel<<<1, 256, 0, s>>>(ctl, sp, g, nullptr, 0, dx, dy, 0.0, cfl, CTL_DT_ONLY);
    HIP_TRY(hipGetLastError());
    StepCtl h;
    HIP_TRY(hipMemcpyAsync(&h, ctl, sizeof h, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    *dt = h.dt;
    *ch = h.ch;
    return 0;
}

}  // extern "C"

// ------------------------------------------------------------------ batch
namespace {

struct Batch {
    int n = 0, dtype = AMHD_F32;
    std::vector<int> nx, ny;
    std::vector<amhd_params> p;
    std::vector<BatchInst> inst;       // host copy of the device descriptors
    std::vector<int> n_part;           // CFL partial count when they come from the standalone kernel
    std::vector<char> part_valid, part_from_stage3;
    std::vector<int> part_cap;
    BatchInst* dinst = nullptr;
    StepCtl* ctl = nullptr;            // [n]
    StepCtl* ctl_pin = nullptr;        // [n] pinned host staging of ctl's transfers
    std::vector<StepCtl> hctl;         // host copy of ctl from the last read
    bool hctl_valid = false;           // false from a run's first launch to its final read
    double* staging = nullptr;
    size_t staging_bytes = 0;
    int rows = 0, n_tasks = 0, resident = 0;
    hipStream_t stream = nullptr;
    hipGraphExec_t graph = nullptr;
};

// Tasks of the batch with r rows per segment.
long batch_tasks(const Batch* b, int r) {
    long t = 0;
    for (int k = 0; k < b->n; ++k)
        t += (long)((b->ny[k] + STRIP - 1) / STRIP) * std::max(1, std::min((b->nx[k] + r - 1) / r, b->nx[k]));
    return t;
}

int batch_layout(Batch* b, int rows) {
    b->rows = rows;
    int r = rows;
    if (r <= 0) {
        // The resident wavefronts take the tasks round by round, so a step
        // lasts about ceil(tasks / resident) tasks of r rows plus the
        // segment prologue (about 0.6 of a row: two extra primitive rows
        // and one extra x face).  Take the r in [4, 64] that minimises it.
        double best = 1e300;
        for (int c = 4; c <= 64; ++c) {
            const long rounds = (batch_tasks(b, c) + b->resident - 1) / b->resident;
            const double cost = (double)rounds * (c + 0.6);
            if (cost < best) { best = cost; r = c; }
        }
    }
    int base = 0;
    for (int k = 0; k < b->n; ++k) {
        BatchInst& d = b->inst[k];
        d.n_strips = (b->ny[k] + STRIP - 1) / STRIP;
        d.n_segs = std::max(1, std::min((b->nx[k] + r - 1) / r, b->nx[k]));
        d.task_base = base;
        base += d.n_strips * d.n_segs;
    }
    b->n_tasks = base;
    HIP_TRY(hipMemcpyAsync(b->dinst, b->inst.data(), sizeof(BatchInst) * b->n, hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (b->graph) { hipGraphExecDestroy(b->graph); b->graph = nullptr; }
    for (int k = 0; k < b->n; ++k) b->part_from_stage3[k] = 0, b->part_valid[k] = 0;
    return 0;
}

// The control blocks as the device holds them.  Only amhd_batch_run's
// launches write them, and every entry point returns with the batch's
// stream drained, so between runs the copy from the last read is exact and
// no transfer (or wait) is needed.
int batch_read_ctl(Batch* b, std::vector<StepCtl>& h) {
    if (b->hctl_valid) {
        h = b->hctl;
        return 0;
    }
    HIP_TRY(hipMemcpyAsync(b->ctl_pin, b->ctl, sizeof(StepCtl) * b->n, hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    h.assign(b->ctl_pin, b->ctl_pin + b->n);
    b->hctl = h;
    b->hctl_valid = true;
    return 0;
}

hipError_t batch_enqueue_step(Batch* b) {
    hipError_t e = hipSuccess;
    for (int st = 1; st <= 3 && e == hipSuccess; ++st)
        e = b->dtype == AMHD_F64 ? launch_batch_stage_f64(st, b->dinst, b->n, b->n_tasks, b->resident, b->stream)
                                 : launch_batch_stage_f32(st, b->dinst, b->n, b->n_tasks, b->resident, b->stream);
    if (e != hipSuccess) return e;
    batch_ctl_kernel<<<b->n, 256, 0, b->stream>>>(b->dinst, 0);
    return hipGetLastError();
}

}  // namespace

extern "C" {

int amhd_batch_destroy(void* handle) {
    if (!handle) return 0;
    Batch* b = (Batch*)handle;
    if (b->stream) hipStreamSynchronize(b->stream);
    if (b->graph) hipGraphExecDestroy(b->graph);
    for (auto& d : b->inst) {
        for (auto& x : d.buf) if (x) hipFree(x);
        if (d.part_dt) hipFree(d.part_dt);
        if (d.part_bf) hipFree(d.part_bf);
    }
    if (b->dinst) hipFree(b->dinst);
    if (b->ctl) hipFree(b->ctl);
    if (b->ctl_pin) hipHostFree(b->ctl_pin);
    if (b->staging) hipFree(b->staging);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
    return 0;
}

int amhd_batch_create(void** out, int n, const int* nx, const int* ny, int dtype, const amhd_params* params) {
    if (!out || !nx || !ny || !params || n <= 0) return fail(-1, "amhd_batch_create: bad argument");
    *out = nullptr;
    if (dtype != AMHD_F64 && dtype != AMHD_F32) return fail(-1, "amhd_batch_create: unknown dtype %d", dtype);
    Batch* b = new Batch();
    b->n = n; b->dtype = dtype;
    b->nx.assign(nx, nx + n); b->ny.assign(ny, ny + n); b->p.assign(params, params + n);
    b->inst.resize(n); b->n_part.assign(n, 0); b->part_valid.assign(n, 0); b->part_from_stage3.assign(n, 0);
    b->part_cap.assign(n, 0);
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    int dev = 0, cus = 0;
    if (e == hipSuccess) e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    b->resident = std::max(1, cus) * (dtype == AMHD_F64 ? batch_waves_per_cu_f64() : batch_waves_per_cu_f32());
    for (int k = 0; k < n && e == hipSuccess; ++k) {
        const amhd_params& p = params[k];
        if (nx[k] < 2 || ny[k] < 2 || !(p.gamma > 1.0) || !(p.dx > 0.0) || !(p.dy > 0.0)) {
            amhd_batch_destroy(b);
            return fail(-1, "amhd_batch_create: instance %d: need nx, ny >= 2, gamma > 1, dx, dy > 0", k);
        }
        BatchInst& d = b->inst[k];
        std::memset(&d, 0, sizeof d);
        d.nx = nx[k]; d.ny = ny[k]; d.pitch = ny[k]; d.plane = (long)nx[k] * ny[k];
        const double bytes = (double)esize(dtype) * NV * (double)d.plane;
        if (bytes > 4294967295.0) {
            amhd_batch_destroy(b);
            return fail(-4, "amhd_batch_create: instance %d exceeds 4 GiB", k);
        }
        d.bytes = (unsigned)bytes;
        d.bcx_periodic = p.bc_x_periodic; d.bcy_periodic = p.bc_y_periodic; d.powell = p.powell;
        d.gamma = p.gamma; d.dx = p.dx; d.dy = p.dy;
        const int nstr = (ny[k] + STRIP - 1) / STRIP;
        b->part_cap[k] = std::max(nx[k] * nstr, SCRATCH_PARTS);
        for (int q = 0; q < 3 && e == hipSuccess; ++q) e = hipMalloc(&d.buf[q], (size_t)bytes);
        if (e == hipSuccess) e = hipMalloc(&d.part_dt, sizeof(double) * NPART_DT * (size_t)b->part_cap[k]);
        if (e == hipSuccess) e = hipMalloc(&d.part_bf, sizeof(double) * 2 * (size_t)nstr);
        if (e == hipSuccess) e = hipMemsetAsync(d.part_bf, 0, sizeof(double) * 2 * (size_t)nstr, b->stream);
        b->staging_bytes = std::max(b->staging_bytes, sizeof(double) * NV * (size_t)(nx[k] + 4) * (ny[k] + 4));
    }
    if (e == hipSuccess) e = hipMalloc(&b->staging, b->staging_bytes);
    if (e == hipSuccess) e = hipMalloc(&b->ctl, sizeof(StepCtl) * n);
    if (e == hipSuccess) e = hipHostMalloc((void**)&b->ctl_pin, sizeof(StepCtl) * n, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc(&b->dinst, sizeof(BatchInst) * n);
    if (e == hipSuccess) {
        std::vector<StepCtl> z(n);
        std::memset(z.data(), 0, sizeof(StepCtl) * n);
        for (int k = 0; k < n; ++k) b->inst[k].ctl = b->ctl + k;
        e = hipMemcpyAsync(b->ctl, z.data(), sizeof(StepCtl) * n, hipMemcpyHostToDevice, b->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    }
    if (e != hipSuccess) {
        int rc = fail(e == hipErrorOutOfMemory ? -3 : -2, "amhd_batch_create: %s", hipGetErrorString(e));
        amhd_batch_destroy(b);
        return rc;
    }
    int rc = batch_layout(b, 0);
    if (rc) { amhd_batch_destroy(b); return rc; }
    *out = b;
    return 0;
}

int amhd_batch_set_rows(void* handle, int rows) {
    if (!handle || rows < 0) return fail(-1, "amhd_batch_set_rows: bad argument");
    Batch* b = (Batch*)handle;
    HIP_TRY(hipStreamSynchronize(b->stream));
    return batch_layout(b, rows);
}

int amhd_batch_info(void* handle, int* n_tasks, int* resident, int* rows) {
    if (!handle) return fail(-1, "amhd_batch_info: null handle");
    Batch* b = (Batch*)handle;
    if (n_tasks) *n_tasks = b->n_tasks;
    if (resident) *resident = b->resident;
    if (rows) {
        int r = 0;
        for (int k = 0; k < b->n; ++k) r = std::max(r, (b->nx[k] + b->inst[k].n_segs - 1) / b->inst[k].n_segs);
        *rows = r;
    }
    return 0;
}

int amhd_batch_upload(void* handle, int k, const double* U, int on_device) {
    if (!handle || !U) return fail(-1, "amhd_batch_upload: null argument");
    Batch* b = (Batch*)handle;
    if (k < 0 || k >= b->n) return fail(-1, "amhd_batch_upload: instance %d out of range", k);
    std::vector<StepCtl> h;
    int rc = batch_read_ctl(b, h);
    if (rc) return rc;
    const BatchInst& d = b->inst[k];
    const size_t gb = sizeof(double) * NV * (size_t)(b->nx[k] + 4) * (b->ny[k] + 4);
    HIP_TRY(hipMemcpyAsync(b->staging, U, gb, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, b->stream));
    HIP_TRY(b->dtype == AMHD_F64 ? launch_ingest_f64(b->staging, (double*)d.buf[h[k].cur], d.plane, d.pitch, d.nx, d.ny, b->stream)
                                 : launch_ingest_f32(b->staging, (float*)d.buf[h[k].cur], d.plane, d.pitch, d.nx, d.ny, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->part_valid[k] = 0;
    b->part_from_stage3[k] = 0;
    return 0;
}

int amhd_batch_download(void* handle, int k, double* U, int on_device) {
    if (!handle || !U) return fail(-1, "amhd_batch_download: null argument");
    Batch* b = (Batch*)handle;
    if (k < 0 || k >= b->n) return fail(-1, "amhd_batch_download: instance %d out of range", k);
    std::vector<StepCtl> h;
    int rc = batch_read_ctl(b, h);
    if (rc) return rc;
    const BatchInst& d = b->inst[k];
    HIP_TRY(b->dtype == AMHD_F64
                ? launch_materialize_f64((const double*)d.buf[h[k].cur], d.plane, d.pitch, d.nx, d.ny, d.bcx_periodic,
                                         d.bcy_periodic, b->staging, b->stream)
                : launch_materialize_f32((const float*)d.buf[h[k].cur], d.plane, d.pitch, d.nx, d.ny, d.bcx_periodic,
                                         d.bcy_periodic, b->staging, b->stream));
    const size_t gb = sizeof(double) * NV * (size_t)(b->nx[k] + 4) * (b->ny[k] + 4);
    HIP_TRY(hipMemcpyAsync(U, b->staging, gb, on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return 0;
}

int amhd_batch_last_dt(void* handle, double* dt) {
    if (!handle || !dt) return fail(-1, "amhd_batch_last_dt: null argument");
    Batch* b = (Batch*)handle;
    std::vector<StepCtl> h;
    int rc = batch_read_ctl(b, h);
    if (rc) return rc;
    for (int k = 0; k < b->n; ++k) dt[k] = h[k].dt;   // a finished run leaves the last step's dt
    return 0;
}

int amhd_batch_state_ptrs(void* handle, void** cur, void** prev) {
    if (!handle || !cur) return fail(-1, "amhd_batch_state_ptrs: null argument");
    Batch* b = (Batch*)handle;
    std::vector<StepCtl> h;
    int rc = batch_read_ctl(b, h);
    if (rc) return rc;
    for (int k = 0; k < b->n; ++k) {
        cur[k] = b->inst[k].buf[h[k].cur];
        if (prev) prev[k] = b->inst[k].buf[(h[k].cur + 2) % 3];
        b->part_valid[k] = 0;   // the caller may edit the state: recompute the CFL partials
        b->part_from_stage3[k] = 0;
    }
    return 0;
}

int amhd_batch_run(void* handle, int n_steps, const double* cfl, const double* t_end, const int* max_steps,
                   int use_graph, double* t_inout, int* step_inout) {
    if (!handle || !cfl || !t_end || !max_steps || !t_inout || !step_inout || n_steps < 0)
        return fail(-1, "amhd_batch_run: bad argument");
    Batch* b = (Batch*)handle;
    std::vector<StepCtl> h;
    int rc = batch_read_ctl(b, h);
    if (rc) return rc;
    std::vector<int> start(b->n);
    // CFL partials of the instances whose stage-3 partials are not current,
    // DT_LIST instances per launch
    DtList list{};
    int cnt = 0, gmax = 0;
    auto flush = [&]() -> int {
        if (cnt == 0) return 0;
        const hipError_t e = b->dtype == AMHD_F64 ? launch_batch_dt_partials_f64(b->dinst, list, cnt, gmax, b->stream)
                                                  : launch_batch_dt_partials_f32(b->dinst, list, cnt, gmax, b->stream);
        if (e != hipSuccess) return fail(-2, "amhd_batch_run: CFL partials launch failed: %s", hipGetErrorString(e));
        cnt = 0;
        gmax = 0;
        return 0;
    };
    for (int k = 0; k < b->n; ++k) {
        const BatchInst& d = b->inst[k];
        start[k] = step_inout[k];
        h[k].t = t_inout[k]; h[k].steps = step_inout[k];
        h[k].t_end = t_end[k]; h[k].cfl = cfl[k];
        h[k].max_steps = (int)std::min<long>((long)max_steps[k], (long)step_inout[k] + n_steps);
        h[k].glm_alpha = b->p[k].glm_alpha;
        h[k].min_dxdy = std::min(b->p[k].dx, b->p[k].dy);
        h[k].done = 0; h[k].pending = 0;
        if (!b->part_valid[k]) {
            // as many workgroups as a launch for this instance alone
            const int g = grid_for((long)d.nx * d.ny, 256, b->part_cap[k]);
            list.k[cnt] = k;
            list.cur[cnt] = h[k].cur;
            list.g[cnt] = g;
            gmax = std::max(gmax, g);
            if (++cnt == DT_LIST && (rc = flush())) return rc;
            b->n_part[k] = g;
            b->part_valid[k] = 1;
            b->part_from_stage3[k] = 0;
        }
        h[k].n_part = b->part_from_stage3[k] ? d.n_strips * d.n_segs : b->n_part[k];
    }
    if ((rc = flush())) return rc;
    b->hctl_valid = false;   // the launches below change the control blocks
    // through the pinned staging, whose last use (a read) has completed; the
    // final read below reuses it after this copy in stream order
    std::memcpy(b->ctl_pin, h.data(), sizeof(StepCtl) * b->n);
    HIP_TRY(hipMemcpyAsync(b->ctl, b->ctl_pin, sizeof(StepCtl) * b->n, hipMemcpyHostToDevice, b->stream));
    batch_ctl_kernel<<<b->n, 256, 0, b->stream>>>(b->dinst, 1);
    HIP_TRY(hipGetLastError());
    if (use_graph && n_steps > 0) {
        if (!b->graph) {
            hipGraph_t g;
            HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            hipError_t e = batch_enqueue_step(b);
            hipError_t e2 = hipStreamEndCapture(b->stream, &g);
            HIP_TRY(e);
            HIP_TRY(e2);
            HIP_TRY(hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0));
            hipGraphDestroy(g);
        }
        for (int s = 0; s < n_steps; ++s) HIP_TRY(hipGraphLaunch(b->graph, b->stream));
    } else {
        for (int s = 0; s < n_steps; ++s) HIP_TRY(batch_enqueue_step(b));
    }
    rc = batch_read_ctl(b, h);
    if (rc) return rc;
    for (int k = 0; k < b->n; ++k) {
        if (h[k].steps > start[k]) b->part_from_stage3[k] = 1;
        t_inout[k] = h[k].t;
        step_inout[k] = h[k].steps;
    }
    return 0;
}

}  // extern "C"
