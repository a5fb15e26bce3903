// amhd_kernels.hpp -- gfx950 kernels of the MHD update.
//
// State layout in HBM ("interior layout"): NV planes, each nx rows of ny
// cells (x slow, y fast), row pitch `pitch` elements, plane stride `plane`
// elements.  Ghost cells are not stored: every read of a cell outside the
// interior goes through remap() (periodic wrap or zero-gradient clamp, and
// psi = 0 on the x-ghosts of a non-periodic x boundary), which is what
// solver.py:158-197 leaves in the ghost cells before every RHS evaluation.
//
// The fused stage kernel (one launch per SSP-RK3 stage) is a strip march:
// each wavefront owns a strip of STRIP = 60 consecutive y cells (64 lanes:
// 2 halo lanes on each side) and walks down a segment of `rows` x rows.
// Per row it loads one new conserved row, converts it to primitives once,
// and keeps a 3-row primitive window in registers: the x-face MUSCL + HLLD
// uses the window, the y-face MUSCL + HLLD uses DPP lane shifts, the x-flux
// of the previous row is carried in registers, so each face is solved once
// (apart from one extra x-face per segment and the 4 halo lanes).  The
// epilogue applies the Powell source, the RK stage combine, the scalar
// bounds, the psi decay (stage 3) and writes per-wave partial maxima for the
// next CFL step.
#pragma once
#include "amhd_device.hpp"

// Optional occupancy request for the stage kernels: -DAMHD_STAGE_WAVES_PER_EU=n asks
// the compiler to fit n waves per SIMD (register budget 512/n per lane).
#if defined(AMHD_STAGE_WAVES_PER_EU) && AMHD_STAGE_WAVES_PER_EU > 0
#define AMHD_STAGE_OCC __attribute__((amdgpu_waves_per_eu(AMHD_STAGE_WAVES_PER_EU)))
#else
#define AMHD_STAGE_OCC
#endif

namespace amhd {

constexpr int WAVE = 64;
constexpr int HALO = 2;
constexpr int STRIP = WAVE - 2 * HALO;  // output columns per wavefront
constexpr int WPB = 4;                  // wavefronts per workgroup
constexpr int NPART_DT = 4;             // {max cf, max(|vx|+|vy|), max|vx|, max|vy|}

#ifdef AMHD_WAVE_TIMES
// Diagnostic builds only (-DAMHD_WAVE_TIMES): times of every stage-kernel
// wavefront of the latest launch of each stage, in ticks of the 100 MHz
// s_memrealtime clock, [stage][wave][start, end of the segment prologue,
// end of the first row, end], and its hardware place [5] = XCC_ID << 32 |
// HW_ID (wave slot, SIMD, CU, shader array, shader engine).
constexpr int WT_WAVES = 8192;
constexpr int WT_FIELDS = 5;
__device__ unsigned long long amhd_wave_times[4][WT_WAVES][WT_FIELDS];
#define AMHD_WT_MARK(STG, W, K)                                                       \
    do {                                                                              \
        const unsigned long long t_ = __builtin_amdgcn_s_memrealtime();               \
        if (lane == 0 && (W) >= 0 && (W) < WT_WAVES) amhd_wave_times[STG][W][K] = t_; \
    } while (0)
#else
#define AMHD_WT_MARK(STG, W, K) do { } while (0)
#endif

// Device-resident step control for the host-free time loop.
struct StepCtl {
    double t, dt, ch, decay;
    double t_end, cfl, glm_alpha, min_dxdy;
    double bflux, bflux_cum;
    double cfmax, vmax;
    int steps, max_steps, done, pending;
    int cur, n_part;   // batched solver: current state buffer, CFL partial count
};

// The two halves of the device loop's step control, shared by the step
// control kernel and the fused stage kernels (contraction off: both
// translation units round them alike).
// The end of a step that ran: boundary energy (solver.py:316-317), t and
// step advance (solver.py:481-482).
__device__ __forceinline__ void ctl_close(StepCtl& h, double bflux) {
#pragma clang fp contract(off)
    h.bflux = bflux;
    h.bflux_cum += bflux * h.dt;
    h.t += h.dt;
    h.steps += 1;
    h.cur = (h.cur + 1) % 3;
    h.pending = 0;
}
// (ch, dt) from the four CFL maxima, exactly as solver.py:206-223 (all
// reductions are maxima; since ch >= 1.5 max cf, max(cf, ch) == ch in every
// cell and max_k (|v_k| + ch)/dx == (max_k |v_k| + ch)/dx, so the four maxima
// determine dt bit for bit); records ch and the maxima in h.
__device__ __forceinline__ double ctl_dt(StepCtl& h, const double (&r)[4], double dx, double dy, double cfl) {
#pragma clang fp contract(off)
    const double cfmax = sqrt(r[0]);   // partials hold max cf^2
    const double ch = fmax(cfmax + r[1], 1.0) * 1.5;
    const double sx = r[2] + ch, sy = r[3] + ch;
    const double sm = fmax(sx / dx, sy / dy);
    const double dt = (sm < 1e-14) ? cfl * fmin(dx, dy) : cfl / sm;
    h.cfmax = cfmax;
    h.vmax = r[1];
    h.ch = ch;
    return dt;
}
// The loop test of solver.py:475-479 and the step's dt (clipped to t_end)
// and psi decay; pending marks a step that is to run.
__device__ __forceinline__ void ctl_begin(StepCtl& h, double dt) {
#pragma clang fp contract(off)
    if (!(h.t < h.t_end && h.steps < h.max_steps)) { h.done = 1; return; }
    const double rem = h.t_end - h.t;
    if (rem < dt) dt = rem;
    if (dt <= 1e-16) { h.done = 1; return; }
    h.dt = dt;
    h.decay = h.ch > 0.0 ? exp(-h.glm_alpha * h.ch * dt / h.min_dxdy) : 1.0;
    h.pending = 1;
}

// Segment re-sizing (rebalance_block below): row flags in, boundary table
// out.  A row's cost is wb, plus wx where its x faces were solved, plus wy
// where its y faces were (the previous row differs).
struct RbWeights {
    int wb, wx, wy;
};
struct Rebalance {
    const unsigned char* flags;   // [nx][RB_FLAG_PITCH]
    int n_wide, nx, n_segs;
    int* seg_i0;                  // [n_segs + 1]
    RbWeights w;
};
// The weights of the first stage's table and of the table of stages 2 and
// 3 (and of the host-driven step, whose three stages share one table).
// FP32: stage 1 reads and writes two states per cell, so a row whose x and
// y faces repeat the previous row's costs little; stages 2 and 3 move
// three states per cell, and such a row costs about what its memory
// traffic does.  FP64 is bound by its arithmetic in every stage.
#ifndef AMHD_RB_W1
#define AMHD_RB_W1 3, 2, 3
#endif
#ifndef AMHD_RB_W23
#define AMHD_RB_W23 4, 2, 1
#endif
#ifndef AMHD_RB_W1_F64
#define AMHD_RB_W1_F64 4, 3, 3
#endif
#ifndef AMHD_RB_W23_F64
#define AMHD_RB_W23_F64 4, 3, 3
#endif
constexpr RbWeights RB_W1{AMHD_RB_W1}, RB_W23{AMHD_RB_W23};
constexpr RbWeights RB_W1_F64{AMHD_RB_W1_F64}, RB_W23_F64{AMHD_RB_W23_F64};

template <class real> struct StageParams {
    const real* u0;   // U^n (interior layout)
    const real* us;   // stage input U^(s)
    real* uo;         // stage output (interior layout), or R (STAGE 0)
    long plane;       // plane stride of u0/us (elements)
    long oplane;      // plane stride of uo (R output may be ghosted)
    int pitch, opitch;
    unsigned in_bytes, out_bytes;   // addressable bytes from u0/us and from uo
    int nx, ny, bcx_periodic, bcy_periodic;
    int n_strips, n_segs;   // wave w: strip w % n_strips, segment w / n_strips (rows from seg_i0)
    // Packing of the last strip when it is narrow (pk_groups = G in {2, 4, 8},
    // 0 = none): the first (n_strips - 1) * n_segs wavefronts cover the other
    // strips as above; each of the ceil(n_segs / G) after them covers G
    // segments of the last strip, one per group of 64 / G lanes (2 halo lanes
    // on each side), all marching in step.
    int pk_groups, pk_shift;   // G, log2(64 / G)
    double gamma, dx, dy;
    double dt, ch, decay;   // used when ctl == nullptr
    const StepCtl* ctl;     // device-resident dt/ch/decay
    double* part_dt;        // STAGE 3: [n_waves][4]
    double* part_bf;        // STAGE 1: [n_strips][2] (non-periodic x only)
    // x-slab halos: rows -2, -1 (halo_lo) and nx, nx+1 (halo_hi) of the stage
    // input taken from neighbouring slabs, 10 planes of 2 rows of `pitch`
    // cells; null = the domain boundary (wrap / clamp, psi = 0).
    const real* halo_lo;
    const real* halo_hi;
    int first_blocks;       // workgroups dispatched first onto each CU (the CU count); 0 = not known
    const int* seg_i0;      // [n_segs + 1] first row of each segment (null: seg*nx/n_segs)
    // STAGE 3, wide strips: rowflags[i * RB_FLAG_PITCH + strip] = 1 where
    // rows i and i+1 hold the same bits in every lane of the strip's
    // wavefront (FP32: where face i+1/2 took the uniform-face path there)
    // (null = none); the next step's segments are sized from them
    unsigned char* rowflags;
    // Fused device loop (fuse = 1): STAGE 1 reduces the previous step's CFL
    // partials (part_in, ctl->n_part of them) in every workgroup, reads the
    // committed control block `ctl` and writes the step's block to ctl_out
    // (workgroup 0); STAGE 2 and 3 read the step's block through `ctl`.
    // side = 1: one workgroup after the strips does side work instead:
    // STAGE 1 re-sizes the segments into rb.seg_i0 and, when rb1.seg_i0 is
    // set, into rb1.seg_i0 with rb1's weights (a copy of seg_i0 where it
    // cannot), STAGE 3 closes the step from `ctl` and part_bf_in into
    // ctl_out (the committed block) and copies seg_copy_src to seg_copy.
    int fuse, side;
    const double* part_in;
    int n_part_in;   // STAGE 1 fused: the CFL partials in part_in, at least 1 (a launch constant of the device loop)
    StepCtl* ctl_out;
    const double* part_bf_in;
    int n_part_next;   // STAGE 3 side: the CFL partials the stage writes (one per strip workgroup)
    Rebalance rb, rb1;
    int* seg_copy;
    const int* seg_copy_src;
};

// Step values a fused stage kernel hands to its wavefronts.
struct CtlVals {
    double dt, ch, decay;
};

// Row flags of the cost-sized segments: one 16-byte-aligned record per row
// with a byte per wide strip (at most RB_FLAG_PITCH strips; more: no flags)
constexpr int RB_FLAG_PITCH = 64;

// Wavefronts of one stage launch: the wide strips' segments, then the
// packed wavefronts of a narrow last strip.
template <class real>
__host__ __device__ inline int stage_wave_count(const StageParams<real>& a) {
    if (!a.pk_groups) return a.n_strips * a.n_segs;
    return (a.n_strips - 1) * a.n_segs + (a.n_segs + a.pk_groups - 1) / a.pk_groups;
}

// ---------------------------------------------------------- buffer access
// Raw buffer loads/stores: one 32-bit per-lane offset (the column), the row
// and variable plane go into the scalar offset, so a row of 10 variables
// costs no vector address arithmetic (a packed wavefront, whose lane groups
// are on different rows, adds the row to the per-lane offset instead).
// Offsets past `bytes` read as 0; stores there are dropped.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float bload(float, __amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, vo, so, 0));
}
typedef unsigned int u32x2 __attribute__((__vector_size__(8)));
__device__ __forceinline__ double bload(double, __amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0));
}
__device__ __forceinline__ void bstore(float x, __amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(x), r, vo, so, 0);
}
__device__ __forceinline__ void bstore(double x, __amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, x), r, vo, so, 0);
}

// XCD-aware workgroup order: the hardware deals workgroups round-robin over
// the 8 XCDs; renumber so that consecutive logical workgroups (neighbouring
// strips, which share halo columns) land on the same XCD and its L2.
__device__ __forceinline__ int xcd_logical_block(int nb) {
    const int b = blockIdx.x;
    const int nb8 = nb & ~7;
    return b < nb8 ? (b & 7) * (nb8 >> 3) + (b >> 3) : b;
}
__device__ __forceinline__ int xcd_logical_block() { return xcd_logical_block((int)gridDim.x); }

// Floors of the MUSCL face states (physics.py:329-336).  The exact policy
// keeps the reference's comparison (a NaN passes through); the fast one
// uses max, one instruction per value.
template <class real, class P>
__device__ __forceinline__ void floor_face(real (&w)[NV]) {
    if constexpr (P::exact) {
        if (w[iRHO] < real(FLOOR_RHO)) w[iRHO] = real(FLOOR_RHO);
        if (w[iPR] < real(FLOOR_PR)) w[iPR] = real(FLOOR_PR);
    } else {
        w[iRHO] = fmax(w[iRHO], real(FLOOR_RHO));
        w[iPR] = fmax(w[iPR], real(FLOOR_PR));
    }
}

template <class real>
__device__ __forceinline__ void load_cell(const real* __restrict__ base, long plane, real (&u)[NV]) {
#pragma unroll
    for (int v = 0; v < NV; ++v) u[v] = base[v * plane];
}

// CFL contributions of one cell (solver.py:209-214), from conserved u, in
// the state's precision, as the reference computes them on its arrays:
// m[0] collects cf^2 = gamma p/rho + B^2/rho (sqrt is monotone, so
// sqrt(max cf^2) == max cf; the square root is taken once per step),
// m[1] |vx|+|vy|, m[2] |vx|, m[3] |vy|.  FP64 state: the reference's
// expression; FP32 state: (gamma p + B^2) * (1/rho) with one v_rcp_f32.
template <class real>
__device__ __forceinline__ void dt_terms(const real (&u)[NV], double gamma, real (&m)[NPART_DT]) {
    real cf2, ax, ay;
    if constexpr (sizeof(real) == sizeof(double)) {
        // the reference's expressions (solver.py:208-214)
        double w[NV];
        cons_to_prim<double, ExactMath<double>>(u, w, gamma, false);
        const double b2 = w[iBX] * w[iBX] + w[iBY] * w[iBY] + w[iBZ] * w[iBZ];
        cf2 = gamma * w[iPR] / w[iRHO] + b2 / w[iRHO];
        ax = fabs(w[iVX]);
        ay = fabs(w[iVY]);
    } else {
        const float g = float(gamma);
        const float rho = fmaxf(u[RHO], float(FLOOR_RHO));
        const float ri = __builtin_amdgcn_rcpf(rho);
        const float vx = u[MX] * ri, vy = u[MY] * ri, vz = u[MZ] * ri;
        const float b2 = u[BX] * u[BX] + u[BY] * u[BY] + u[BZ] * u[BZ];
        const float rv2 = u[MX] * vx + u[MY] * vy + u[MZ] * vz;
        const float p = fmaxf((g - 1.0f) * (u[EN] - 0.5f * (rv2 + b2)), float(FLOOR_PR));
        cf2 = (g * p + b2) * ri;
        ax = fabsf(vx);
        ay = fabsf(vy);
    }
    m[0] = fmax(m[0], cf2);
    m[1] = fmax(m[1], ax + ay);
    m[2] = fmax(m[2], ax);
    m[3] = fmax(m[3], ay);
}

// Boundary energy-flux density of one cell (solver.py:233-239), FP64.
template <class real>
__device__ __forceinline__ double bflux_term(const real (&u)[NV], double gamma, double dy) {
    double ud[NV], w[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) ud[v] = double(u[v]);
    cons_to_prim<double, ExactMath<double>>(ud, w, gamma, false);
    const double b2 = w[iBX] * w[iBX] + w[iBY] * w[iBY] + w[iBZ] * w[iBZ];
    const double pt = w[iPR] + 0.5 * b2;
    const double vb = w[iVX] * w[iBX] + w[iVY] * w[iBY] + w[iVZ] * w[iBZ];
    return ((ud[EN] + pt) * w[iVX] - w[iBX] * vb) * dy;
}

// Segment boundaries sized by cost.  Row r of a strip costs a base wb, plus
// wx where its x faces were solved (flag 0: not every face of the
// wavefront took the uniform-face path), plus wy where row r-1's were
// (then rows r-1 and r differ somewhere and row r solves its y faces; a
// flag 1 on row r-1 means the two rows hold the same bits).  Row r of the
// wide strips costs the sum over strips, and segment g (0 < g < n_segs)
// starts after the row whose cost interval (P(r), P(r + 1)] holds the
// target g total / n_segs.  When the cheapest row times the rows per segment
// exceeds the dearest row (nx wb > n_segs (wb + wx + wy)), the
// targets lie more than one row's cost apart, so the boundaries increase
// strictly and every segment keeps at least one row.  One workgroup; larger
// grids than the LDS limits below, or fewer rows per segment, keep the
// table as it is (returns false).
constexpr int RB_MAX_ROWS = 8192;
__device__ __forceinline__ bool rebalance_block(const Rebalance& rb) {
    __shared__ int rcost[RB_MAX_ROWS];
    __shared__ long wtot[4];
    const int T = (int)blockDim.x, t = (int)threadIdx.x;
    const int nx = rb.nx, ns = rb.n_wide, ng = rb.n_segs;
    const int wb = rb.w.wb, wx = rb.w.wx, wy = rb.w.wy;
    if (nx > RB_MAX_ROWS || wb < 1 || wx < 0 || wy < 0 || (long)nx * wb <= (long)ng * (wb + wx + wy) || ng < 2 ||
        ns < 1 || ns > RB_FLAG_PITCH || T != 256)
        return false;
    // each thread's run of rows [r0, r1)
    const int R = (nx + T - 1) / T;
    const int r0 = t * R < nx ? t * R : nx, r1 = r0 + R < nx ? r0 + R : nx;
    // uniform strips per row: the flag bytes (0 or 1, padding 0) of a row
    // counted by popcount, 16 bytes per load, four rows' loads in flight at
    // a time
    const int words = (ns + 15) / 16;
    for (int rb0 = r0; rb0 < r1; rb0 += 4) {
        uint4 v[4][RB_FLAG_PITCH / 16];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
#pragma unroll
            for (int w = 0; w < RB_FLAG_PITCH / 16; ++w) {
                const bool ok = rb0 + u < r1 && w < words;
                v[u][w] = ok ? *(const uint4*)(rb.flags + (long)(rb0 + u) * RB_FLAG_PITCH + 16 * w) : make_uint4(0, 0, 0, 0);
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            int uni = 0;
#pragma unroll
            for (int w = 0; w < RB_FLAG_PITCH / 16; ++w)
                uni += __popc(v[u][w].x) + __popc(v[u][w].y) + __popc(v[u][w].z) + __popc(v[u][w].w);
            if (rb0 + u < r1) rcost[rb0 + u] = uni;
        }
    }
    // row costs in place, each from its own count and the previous row's
    // (read before any thread overwrites it; no row before row 0)
    __syncthreads();
    int prev = r0 > 0 && r0 < r1 ? rcost[r0 - 1] : 0;
    __syncthreads();
    long loc = 0;
    for (int r = r0; r < r1; ++r) {
        const int uni = rcost[r];
        const int c = ns * wb + (ns - uni) * wx + (ns - prev) * wy;
        prev = uni;
        rcost[r] = c;
        loc += c;
    }
    // exclusive prefix of the runs: inclusive scan in each wavefront, then
    // the wavefront totals
    const int lane = t & 63, wv = t >> 6;
    long inc = loc;
    for (int d = 1; d < 64; d <<= 1) {
        const long o = __shfl_up(inc, d, 64);
        if (lane >= d) inc += o;
    }
    if (lane == 63) wtot[wv] = inc;
    __syncthreads();
    long P = inc - loc, total = 0;
    for (int u = 0; u < 4; ++u) {
        if (u < wv) P += wtot[u];
        total += wtot[u];
    }
    for (int r = r0; r < r1; ++r) {
        const long Pn = P + rcost[r];
        for (long g = P * ng / total + 1; g <= Pn * ng / total && g < ng; ++g) rb.seg_i0[g] = r + 1;
        P = Pn;
    }
    if (t == 0) {
        rb.seg_i0[0] = 0;
        rb.seg_i0[ng] = nx;
    }
    return true;
}

// Net boundary energy flux of a step from the per-strip sums of its first
// stage (one wavefront, a fixed order; valid in every lane): the left
// boundary's sum minus the right one's.
__device__ __forceinline__ double bflux_wave(const double* part_bf, int n_bf, int lane) {
#pragma clang fp contract(off)
    double b0 = 0.0, b1 = 0.0;
    for (int k = lane; k < n_bf; k += 64) { b0 += part_bf[2 * k]; b1 += part_bf[2 * k + 1]; }
    b0 = wave_sum(b0);
    b1 = wave_sum(b1);
    return b0 - b1;
}

// The CFL maxima of the wavefronts of a workgroup (fused third stage).
__device__ __forceinline__ double (&block_red())[WPB][NPART_DT] {
    __shared__ double r[WPB][NPART_DT];
    return r;
}

// A double known to be equal in every lane, moved to scalar registers.
__device__ __forceinline__ double uniform_d(double x) {
    const int lo = __builtin_amdgcn_readfirstlane(__double2loint(x));
    const int hi = __builtin_amdgcn_readfirstlane(__double2hiint(x));
    return __hiloint2double(hi, lo);
}

// Fused device loop, STAGE 1: every wavefront reduces the previous step's
// CFL partials itself and begins the step from the committed control block
// the way the step control kernel does (the same maxima and expressions, so
// every wavefront holds the same values); wavefront 0 of workgroup 0 stores
// the step's block.  First half, in two parts so that the partials' loads
// and the segment's first rows are in flight together: cfl_issue loads this
// lane's partials lane, lane + 64, ..., lane + 448 (an index past the last
// reads the last partial again, which leaves the maxima unchanged, so the
// loads need no test), and cfl_reduce, after the rows are requested, takes
// their maxima and those of any partials past the first 512.
constexpr int CFL_U = 8;
__device__ __forceinline__ void cfl_issue(const double* __restrict__ part, int n_part, int lane,
                                          double (&v)[CFL_U][NPART_DT]) {
#pragma unroll
    for (int u = 0; u < CFL_U; ++u) {
        const int k = min(lane + u * WAVE, n_part - 1);
#pragma unroll
        for (int q = 0; q < NPART_DT; ++q) v[u][q] = part[(long)k * NPART_DT + q];
    }
}
__device__ __forceinline__ void cfl_reduce(const double* __restrict__ part, int n_part, int lane,
                                           double (&v)[CFL_U][NPART_DT], double (&m)[NPART_DT]) {
    // an empty statement that takes the partials and clobbers memory: the
    // maxima below cannot be placed before the loads issued ahead of this
    // point (the compiler otherwise waits on the partials before issuing them)
#pragma unroll
    for (int u = 0; u < CFL_U; ++u) {
        asm volatile("" : "+v"(v[u][0]), "+v"(v[u][1]), "+v"(v[u][2]), "+v"(v[u][3])::"memory");
    }
#pragma unroll
    for (int q = 0; q < NPART_DT; ++q) m[q] = 0.0;
#pragma unroll
    for (int u = 0; u < CFL_U; ++u) {
#pragma unroll
        for (int q = 0; q < NPART_DT; ++q) m[q] = fmax(m[q], v[u][q]);
    }
    for (int k = lane + CFL_U * WAVE; k < n_part; k += WAVE) {
#pragma unroll
        for (int q = 0; q < NPART_DT; ++q) m[q] = fmax(m[q], part[(long)k * NPART_DT + q]);
    }
}
// Second half: the wave's maxima and the step values.  Returns false when
// the loop is done (no step runs).
template <class real>
__device__ __forceinline__ bool fused_begin(const StageParams<real>& a, const double (&m)[NPART_DT], const StepCtl& h0,
                                           CtlVals& cv) {
    double r[NPART_DT];
#pragma unroll
    for (int q = 0; q < NPART_DT; ++q) r[q] = wave_max(m[q]);
    StepCtl h = h0;
    if (!h.done) ctl_begin(h, ctl_dt(h, r, a.dx, a.dy, h.cfl));
    if (blockIdx.x == 0 && threadIdx.x == 0) *a.ctl_out = h;
    cv.dt = uniform_d(h.dt);
    cv.ch = uniform_d(h.ch);
    cv.decay = uniform_d(h.decay);
    return !h.done;
}

// The side workgroup of a fused launch (see StageParams::side).
template <class real, int STAGE>
__device__ __forceinline__ void stage_side(const StageParams<real>& a) {
    if constexpr (STAGE == 1) {
        // the next segment table from the row flags of the previous step's
        // third stage; a copy of the current one where it cannot be re-sized
        if (!rebalance_block(a.rb)) {
            for (int k = threadIdx.x; k <= a.n_segs; k += blockDim.x) a.rb.seg_i0[k] = a.seg_i0[k];
        }
        if (a.rb1.seg_i0) {
            __syncthreads();   // the second table reuses the first one's shared arrays
            if (!rebalance_block(a.rb1)) {
                for (int k = threadIdx.x; k <= a.n_segs; k += blockDim.x) a.rb1.seg_i0[k] = a.seg_i0[k];
            }
        }
    } else if constexpr (STAGE == 3) {
        // every stage of the step has read the step's block: close the step
        if (threadIdx.x < 64) {
            const double bflux = a.part_bf_in ? bflux_wave(a.part_bf_in, a.n_strips, (int)threadIdx.x) : 0.0;
            if (threadIdx.x == 0) {
                StepCtl h = *a.ctl;
                if (h.pending) {
                    ctl_close(h, bflux);
                    h.n_part = a.n_part_next;
                }
                *a.ctl_out = h;
            }
        }
        if (a.seg_copy) {
            for (int k = threadIdx.x; k <= a.n_segs; k += blockDim.x) a.seg_copy[k] = a.seg_copy_src[k];
        }
    }
}

// ------------------------------------------------------------ stage kernel
// STAGE 0: R = L(U^s) only (compute_rhs, solver.py:268-308)
// STAGE 1: U1 = U0 + dt L(U0)                              (solver.py:319)
// STAGE 2: U2 = 0.75 U0 + 0.25 (U1 + dt L(U1))             (solver.py:323)
// STAGE 3: U  = (1/3) U0 + (2/3) (U2 + dt L(U2)), psi decay (solver.py:327-333)
// followed by the scalar bounds of solver.py:199-204 for stages 1-3.
// The work of one wavefront: strip `wave % n_strips`, segment
// `wave / n_strips` of the grid described by `a`.  `wave` is wave-uniform.
//
// PK = false: strip `wave % n_wide`, segment `wave / n_wide`, with n_wide the
// strips not packed (n_strips, or n_strips - 1); the row index is
// wave-uniform.  PK = true: packed wavefront `wave` of the last
// strip (StageParams::pk_groups); lane group g takes segment wave * G + g, so
// the row index differs between groups and enters the buffer accesses
// through the per-lane offset.  `part` is the wavefront's slot in the
// per-wave partials.
template <class real, class P, int STAGE, bool POWELL, bool GLM, bool PK = false>
__device__ __forceinline__ void stage_wave(const StageParams<real>& a, const int wave, const int lane,
                                           const bool older = false, const int part = -1) {
    const int nx = a.nx, ny = a.ny;
    int strip, seg, gl;
    bool live = true;   // the lane's group has a segment (PK)
    if constexpr (PK) {
        strip = a.n_strips - 1;
        gl = lane & ((1 << a.pk_shift) - 1);
        seg = wave * a.pk_groups + (lane >> a.pk_shift);
        live = seg < a.n_segs;
        if (!live) seg = a.n_segs - 1;
    } else {
        // the wide strips: all of them, or all but a packed last one
        const int n_wide = a.pk_groups ? a.n_strips - 1 : a.n_strips;
        strip = wave % n_wide;
        seg = wave / n_wide;
        gl = lane;
    }
    const int gw = PK ? (1 << a.pk_shift) : WAVE;   // lanes per strip piece
    const int i0 = a.seg_i0 ? a.seg_i0[seg] : (int)(((long)seg * nx) / a.n_segs);
    const int i1 = a.seg_i0 ? a.seg_i0[seg + 1] : (int)(((long)(seg + 1) * nx) / a.n_segs);
    // rows this wavefront steps through (wave-uniform); a PK lane whose
    // segment is shorter idles through the last step
    int n_rows = i1 - i0;
    if constexpr (PK) n_rows = wave_max_int(live ? n_rows : 0);

    // The step's dt, ch and psi decay: launch arguments, the device loop's
    // control block, or (fused first stage) reduced by the wavefront beside
    // the segment's first rows (cfl_issue, cfl_reduce, fused_begin in the
    // prologue).
    const bool fused = STAGE == 1 && a.fuse;
    double dt_d = a.dt, ch_d = a.ch, decay_d = a.decay;
    if (a.ctl && !fused) {
        if (a.ctl->done) return;
        dt_d = a.ctl->dt; ch_d = a.ctl->ch; decay_d = a.ctl->decay;
    }
    FluxConst<real> kc;
    // uniform x faces take the physical flux (FP32 with GLM only: the FP64
    // path keeps the reference's arithmetic bit for bit)
    constexpr bool XSHORT = !P::exact && GLM;
    real dtr, cx, cy, pcx, pcy;
    const real dxr = real(a.dx), dyr = real(a.dy);
    const real idx = real(1.0 / a.dx), idy = real(1.0 / a.dy);
    // fast path: weight of dt L(U) in the stage's combine, times 1/dx, 1/dy, 1
    constexpr double wst = STAGE == 1 ? 1.0 : (STAGE == 2 ? 0.25 : 2.0 / 3.0);
    auto set_step = [&](double dt, double ch, double decay) {
        dt_d = dt; ch_d = ch; decay_d = decay;
        kc = make_flux_const<real>(a.gamma, ch_d);
        dtr = real(dt_d);
        cx = real(wst * dt_d / a.dx); cy = real(wst * dt_d / a.dy);
        pcx = real(0.5 * wst * dt_d / a.dx); pcy = real(0.5 * wst * dt_d / a.dy);
    };
    if (!fused) set_step(dt_d, ch_d, decay_d);

    constexpr unsigned ES = sizeof(real);
    const unsigned pb = (unsigned)a.plane * ES;             // plane stride, bytes
    const unsigned rb = (unsigned)a.pitch * ES;             // row stride, bytes
    const int j = strip * STRIP - HALO + gl;
    const int sj = remap(j, ny, a.bcy_periodic);
    const unsigned vo = (unsigned)sj * ES;
    const bool out_lane = live && gl >= HALO && gl < gw - HALO && j < ny;
    // buffer offsets of row si (row stride rs bytes): per-lane part, wave-uniform part
    auto voff = [&](int si, unsigned rs) -> unsigned { return PK ? vo + (unsigned)si * rs : vo; };
    auto soff = [&](int si, unsigned rs) -> unsigned { return PK ? 0u : (unsigned)si * rs; };
    const __amdgpu_buffer_rsrc_t r_us = make_rsrc(a.us, a.in_bytes);
    const __amdgpu_buffer_rsrc_t r_u0 = make_rsrc(a.u0, a.in_bytes);
    const __amdgpu_buffer_rsrc_t r_uo = make_rsrc(a.uo, a.out_bytes);
    const bool xper = a.bcx_periodic != 0;

    // x ghost rows: a neighbouring slab's halo rows, else wrap (periodic) or
    // clamp with psi = 0 on non-periodic x ghosts
    const bool hlo = a.halo_lo != nullptr, hhi = a.halo_hi != nullptr;
    const unsigned hbytes = 2u * rb * NV;
    const __amdgpu_buffer_rsrc_t r_hlo = make_rsrc(hlo ? (const void*)a.halo_lo : a.us, hlo ? hbytes : 0u);
    const __amdgpu_buffer_rsrc_t r_hhi = make_rsrc(hhi ? (const void*)a.halo_hi : a.us, hhi ? hbytes : 0u);
    auto row_src = [&](int i) -> int {
        if (i < 0) return xper ? i + nx : 0;
        if (i >= nx) return xper ? i - nx : nx - 1;
        return i;
    };
    auto load_row = [&](__amdgpu_buffer_rsrc_t r, int si, real (&u)[NV]) {
        const unsigned lo = voff(si, rb), so = soff(si, rb);
#pragma unroll
        for (int v = 0; v < NV; ++v) u[v] = bload(real(0), r, lo, so + (unsigned)v * pb);
    };
    auto load_halo = [&](__amdgpu_buffer_rsrc_t r, int hi, real (&u)[NV]) {
        const unsigned lo = voff(hi, rb), so = soff(hi, rb);
#pragma unroll
        for (int v = 0; v < NV; ++v) u[v] = bload(real(0), r, lo, so + (unsigned)v * 2u * rb);
    };
    // stage-input row i for any i in [-2, nx + 2) (wave-uniform branch; per
    // lane group with PK, each buffer resource staying wave-uniform)
    auto load_in = [&](int i, real (&u)[NV]) {
        if constexpr (PK) {
            if (i < 0 && hlo) load_halo(r_hlo, i + 2, u);
            else if (i >= nx && hhi) load_halo(r_hhi, i - nx, u);
            else load_row(r_us, row_src(i), u);
        } else {
            if ((i < 0 && hlo) || (i >= nx && hhi)) load_halo(i < 0 ? r_hlo : r_hhi, i < 0 ? i + 2 : i - nx, u);
            else load_row(r_us, row_src(i), u);
        }
    };
    auto psi_ghost = [&](int i) { return !xper && ((i < 0 && !hlo) || (i >= nx && !hhi)); };
    auto prim_row = [&](int i, real (&w)[NV]) {
        real u[NV];
        load_in(i, u);
        cons_to_prim<real, P>(u, w, kc.gamma, psi_ghost(i));
    };

    // prologue: W at rows i0, i0+1, the x-differences and the flux through
    // face i0-1/2.  The row loop keeps three primitive rows (A, B, C), their
    // conserved values (QA, QB, QC: the stage input of the row the combine
    // needs, loaded once) and three x-flux rows (F0, F1, F2), whose roles
    // rotate with period 3, so the window slides without register copies.
    real A[NV], B[NV], C[NV];
    real QA[NV], QB[NV], QC[NV];
    real F0[NV], F1[NV], F2[NV];
    real xl[NV];   // left state of face i+1/2
    real bxm;      // Bx of row i-1 (Powell)
    {
        real wm2[NV], wm1[NV], fa[NV], fb[NV];
        real um2[NV], um1[NV];
        // fused first stage: the CFL partials (cache-resident) are loaded
        // first and reduced while the segment's rows, loaded next, are on
        // their way from memory
        double pv[CFL_U][NPART_DT], pm[NPART_DT];
        if (fused) cfl_issue(a.part_in, a.n_part_in, lane, pv);
        load_in(i0 - 2, um2);
        load_in(i0 - 1, um1);
        load_in(i0, QA);
        load_in(i0 + 1, QB);
        if (fused) cfl_reduce(a.part_in, a.n_part_in, lane, pv, pm);
        if (fused) {
            // (scalar loads return out of order, so a wait on any of them
            // waits on all: the control block is loaded after the rows are
            // requested)
            const StepCtl hc = *a.ctl;
            CtlVals v;
            if (!fused_begin(a, pm, hc, v)) return;   // the loop is done
            set_step(v.dt, v.ch, v.decay);
        }
        cons_to_prim<real, P>(um2, wm2, kc.gamma, psi_ghost(i0 - 2));
        cons_to_prim<real, P>(um1, wm1, kc.gamma, psi_ghost(i0 - 1));
        cons_to_prim<real, P>(QA, A, kc.gamma, psi_ghost(i0));
        cons_to_prim<real, P>(QB, B, kc.gamma, psi_ghost(i0 + 1));
        // the uniform-face rule of the row loop, for face i0-1/2: a lane whose
        // face is uniform gets F(W_i0) and the left state W_i0 of the next
        // face, whichever branch its wavefront takes
        bool uni0 = false;
        if constexpr (XSHORT) uni0 = same_bits(A, wm1);
        if (XSHORT && __builtin_amdgcn_ballot_w64(!uni0) == 0) {
            phys_flux_x(A, F0, kc);
#pragma unroll
            for (int v = 0; v < NV; ++v) xl[v] = A[v];
        } else {
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const real dmm = wm1[v] - wm2[v];
                const real dm0 = A[v] - wm1[v];
                const real h_m1 = vl_half_d<real, P>(dmm, dm0);
                const real h_0 = vl_half_d<real, P>(dm0, B[v] - A[v]);
                fa[v] = wm1[v] + h_m1;
                fb[v] = A[v] - h_0;
                xl[v] = A[v] + h_0;
            }
            floor_face<real, P>(fa); floor_face<real, P>(fb); floor_face<real, P>(xl);
            hlld_face<real, P, GLM, true>(fa, fb, F0, kc);
            if constexpr (XSHORT) {
                if (uni0) {
                    phys_flux_x(A, F0, kc);
#pragma unroll
                    for (int v = 0; v < NV; ++v) xl[v] = A[v];
                }
            }
        }
        bxm = wm1[iBX];
    }
    if constexpr (!PK) AMHD_WT_MARK(STAGE, wave, 1);

    real mx[NPART_DT] = {0, 0, 0, 0};
    double bf_sum[2] = {0.0, 0.0};
    real upre[NV];
    load_in(i0 + 2, upre);

    // Rows that repeat.  Where rows i-1 and i hold the same bits in every
    // lane of the wavefront (y_same, wave-uniform), the y faces of row i
    // have the inputs of those of row i-1 and so their fluxes: fy, by_prev
    // and by_next keep row i-1's.  A flow that depends on y alone keeps its
    // rows identical bit for bit (each gets the same update), so this skips
    // the y solves of every row but the first of such a run.
    real fy[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) fy[v] = real(0);
    real by_prev = real(0), by_next = real(0);
    bool y_same = false;
    // Without the uniform-face path (FP64), the x faces repeat too: where
    // rows i-2 .. i+2 hold the same bits in every lane, face i+1/2 has the
    // inputs of face i-1/2 (the same states and slopes), and the left state
    // of face i+3/2 is that of face i+1/2.  idh bit k: rows i-k and i-k+1
    // are the same (row i = the row about to be computed); none is known
    // before the segment's first row but rows i0, i0+1.
    unsigned idh = 0u;
    if constexpr (!XSHORT) idh = __builtin_amdgcn_ballot_w64(!same_bits(A, B)) == 0 ? 1u : 0u;

    // One output row i.  w0 = W_i, wp1 = W_{i+1}; wp2 receives W_{i+2};
    // ucur = U_i (stage input); qp2 receives U_{i+2}; fxl = F_{i-1/2}; fxr
    // receives F_{i+1/2}.
    auto row = [&](int i, real (&w0)[NV], real (&wp1)[NV], real (&wp2)[NV], const real (&ucur)[NV],
                   real (&qp2)[NV], const real (&fxl)[NV], real (&fxr)[NV]) __attribute__((always_inline)) {
        cons_to_prim<real, P>(upre, wp2, kc.gamma, psi_ghost(i + 2));
#pragma unroll
        for (int v = 0; v < NV; ++v) qp2[v] = upre[v];
        if (i + 1 < i1) load_in(i + 3, upre);

        // base at this row (halo lanes read a valid remapped column and
        // never store)
        real ubase[NV];
        if constexpr (STAGE >= 2) load_row(r_u0, i, ubase);

        // FP64: which faces of row i repeat those of row i-1
        bool xrep = false;
        unsigned row_flag = 0u;
        if constexpr (!XSHORT) {
            const bool id1 = __builtin_amdgcn_ballot_w64(!same_bits(wp1, wp2)) == 0;   // rows i+1, i+2
            row_flag = idh & 1u;             // rows i, i+1
            y_same = (idh & 2u) != 0u;       // rows i-1, i
            xrep = id1 && (idh & 7u) == 7u;  // rows i-2 .. i+2
            idh = ((idh << 1) | (id1 ? 1u : 0u)) & 7u;
        }

        // ---- x face i+1/2: slope at row i+1 from the differences
        // Uniform faces (FP32): where W_i and W_{i+1} hold the same bits in
        // all ten components, both van Leer slopes next to the face are 0, so
        // its two states equal that cell and its flux is the physical flux
        // F(W_{i+1}); the left state of face i+3/2 is W_{i+1}.  Every lane
        // whose face is uniform gets exactly that, whichever branch its
        // wavefront takes; a wavefront whose faces are all uniform skips the
        // slopes and the HLLD solve.  If face i-1/2 was uniform in every lane
        // too (y_same), F(W_{i+1}) is F(W_i), the flux of that face.
        bool uni = false;
        if constexpr (XSHORT) uni = same_bits(w0, wp1);
        const bool all_uni = XSHORT && __builtin_amdgcn_ballot_w64(!uni) == 0;
        if (all_uni) {
#pragma unroll
            for (int v = 0; v < NV; ++v) xl[v] = wp1[v];
            if (y_same) {
#pragma unroll
                for (int v = 0; v < NV; ++v) fxr[v] = fxl[v];
            } else {
                phys_flux_x(wp1, fxr, kc);
            }
        } else if (xrep) {
#pragma unroll
            for (int v = 0; v < NV; ++v) fxr[v] = fxl[v];
        } else {
            real xr[NV], xn[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const real dn = wp2[v] - wp1[v];
                const real h1 = vl_half_d<real, P>(wp1[v] - w0[v], dn);
                xr[v] = wp1[v] - h1;
                xn[v] = wp1[v] + h1;
            }
            floor_face<real, P>(xr);
            hlld_face<real, P, GLM, true>(xl, xr, fxr, kc);
#pragma unroll
            for (int v = 0; v < NV; ++v) xl[v] = xn[v];
            floor_face<real, P>(xl);
            if constexpr (XSHORT) {
                if (uni) {
                    phys_flux_x(wp1, fxr, kc);
#pragma unroll
                    for (int v = 0; v < NV; ++v) xl[v] = wp1[v];
                }
            }
        }

        if constexpr (STAGE == 3 && !PK) {
            // rows i and i+1 the same: the uniform-face path in every lane (FP32)
            if (a.rowflags && lane == 0)
                a.rowflags[(long)i * RB_FLAG_PITCH + strip] = XSHORT ? (all_uni ? 1 : 0) : (unsigned char)row_flag;
        }

        // ---- y face j+1/2 of row i (neighbours through DPP lane shifts),
        // unless row i repeats row i-1
        if (!y_same) {
            real yl[NV], yr[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const real wnv = from_next(w0[v]);
                const real dp = wnv - w0[v];
                const real dm = from_prev(dp);
                if (v == iBY) { by_next = wnv; by_prev = from_prev(w0[v]); }
                const real h = vl_half_d<real, P>(dm, dp);
                yl[v] = w0[v] + h;
                yr[v] = from_next(w0[v] - h);
            }
            floor_face<real, P>(yl); floor_face<real, P>(yr);
            hlld_face_y<real, P, GLM, true>(yl, yr, fy, kc);
        }
        // does row i+1 repeat row i?
        if constexpr (XSHORT) y_same = all_uni;

        // ---- RHS: ((0 - dFx/dx) - dFy/dy) + S   (solver.py:297-306)
        real divb = real(0);
        if constexpr (POWELL && (P::exact || STAGE == 0)) {
            // solver.py:249-254, on the BC-filled primitive field
            const real half = real(0.5);
            divb = P::divi(half * (w0[iBX] + wp1[iBX]) - half * (bxm + w0[iBX]), dxr, idx) +
                   P::divi(half * (w0[iBY] + by_next) - half * (by_prev + w0[iBY]), dyr, idy);
        }
        // fast path: divB c dt, with c the stage's weight on dt L(U); the
        // centre cell of the two half-sums cancels: divB = (Bx_{i+1} -
        // Bx_{i-1}) / (2 dx) + (By_{j+1} - By_{j-1}) / (2 dy)
        real pw = real(0);
        if constexpr (POWELL && !(P::exact || STAGE == 0)) pw = (wp1[iBX] - bxm) * pcx + (by_next - by_prev) * pcy;
        bxm = w0[iBX];
        real unew[NV];
        if constexpr (P::exact || STAGE == 0) {
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const real fyl = from_prev(fy[v]);
                real r = real(0) - P::divi(fxr[v] - fxl[v], dxr, idx);
                r = r - P::divi(fy[v] - fyl, dyr, idy);
                if constexpr (POWELL) {
                    // S = -divB (0, B, v, v.B, 0, 0)   (solver.py:256-264)
                    real s = real(0);
                    if (v == MX) s = -divb * w0[iBX];
                    if (v == MY) s = -divb * w0[iBY];
                    if (v == MZ) s = -divb * w0[iBZ];
                    if (v == BX) s = -divb * w0[iVX];
                    if (v == BY) s = -divb * w0[iVY];
                    if (v == BZ) s = -divb * w0[iVZ];
                    if (v == EN) s = -divb * (w0[iVX] * w0[iBX] + w0[iVY] * w0[iBY] + w0[iVZ] * w0[iBZ]);
                    r = r + s;
                }
                if constexpr (STAGE == 0) unew[v] = r;
                else if constexpr (STAGE == 1) unew[v] = ucur[v] + dtr * r;
                else if constexpr (STAGE == 2) unew[v] = real(0.75) * ubase[v] + real(0.25) * (ucur[v] + dtr * r);
                else unew[v] = real(1.0 / 3.0) * ubase[v] + real(2.0 / 3.0) * (ucur[v] + dtr * r);
            }
        } else {
            // The same combine with the stage's weight on dt L(U) folded
            // into per-launch constants: U1 = U + dt L, U2 = 3/4 U0 + 1/4 U1
            // + (dt/4) L, U3 = 1/3 U0 + 2/3 U2 + (2 dt/3) L, L = -dFx/dx -
            // dFy/dy + S, each flux difference entering with one FMA.
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const real fyl = from_prev(fy[v]);
                real acc;
                if constexpr (STAGE == 1) acc = ucur[v];
                else if constexpr (STAGE == 2) acc = real(0.75) * ubase[v] + real(0.25) * ucur[v];
                else acc = real(1.0 / 3.0) * ubase[v] + real(2.0 / 3.0) * ucur[v];
                acc = acc + (fxl[v] - fxr[v]) * cx;
                acc = acc + (fyl - fy[v]) * cy;
                if constexpr (POWELL) {
                    if (v == MX) acc = acc - pw * w0[iBX];
                    if (v == MY) acc = acc - pw * w0[iBY];
                    if (v == MZ) acc = acc - pw * w0[iBZ];
                    if (v == BX) acc = acc - pw * w0[iVX];
                    if (v == BY) acc = acc - pw * w0[iVY];
                    if (v == BZ) acc = acc - pw * w0[iVZ];
                    if (v == EN) acc = acc - pw * (w0[iVX] * w0[iBX] + w0[iVY] * w0[iBY] + w0[iVZ] * w0[iBZ]);
                }
                unew[v] = acc;
            }
        }
        if constexpr (STAGE >= 1) {
            // enforce_scalar_bounds (solver.py:199-204)
            if constexpr (P::exact) {
                const real rho = unew[RHO] > real(FLOOR_RHO) ? unew[RHO] : real(FLOOR_RHO);
                real c = P::div(unew[RHOC], rho);
                c = c < real(0) ? real(0) : (c > real(1) ? real(1) : c);
                unew[RHOC] = rho * c;
            } else {
                const real rho = fmax(unew[RHO], real(FLOOR_RHO));
                const real c = fmin(fmax(P::div(unew[RHOC], rho), real(0)), real(1));
                unew[RHOC] = rho * c;
            }
        }
        if constexpr (STAGE == 3) {
            if (ch_d > 0.0) unew[PSI] = unew[PSI] * real(decay_d);
        }

        if (out_lane && (!PK || i < i1)) {
            if constexpr (STAGE == 0) {
                const unsigned ors = (unsigned)a.opitch * ES;
                const unsigned lo = voff(i, ors), so = soff(i, ors);
                const unsigned ob = (unsigned)a.oplane * ES;
#pragma unroll
                for (int v = 0; v < NV; ++v) bstore(unew[v], r_uo, lo, so + (unsigned)v * ob);
            } else {
                const unsigned lo = voff(i, rb), so = soff(i, rb);
#pragma unroll
                for (int v = 0; v < NV; ++v) bstore(unew[v], r_uo, lo, so + (unsigned)v * pb);
            }
            if constexpr (STAGE == 3) dt_terms(unew, a.gamma, mx);
            if constexpr (STAGE == 1) {
                if (!xper) {   // physical x boundaries only (not slab edges)
                    if (i == 0 && !hlo) bf_sum[0] += bflux_term(ucur, a.gamma, a.dy);
                    if (i == nx - 1 && !hhi) bf_sum[1] += bflux_term(ucur, a.gamma, a.dy);
                }
            }
        }
    };

    // Issue priority by progress (FP32, two waves per SIMD; the FP64 kernels
    // hold one).  Among equal priorities the SIMD issues from the older wave:
    // left alone, the wave dispatched first runs ahead and the other
    // finishes its last rows alone, with nothing to hide its latencies
    // (measured: half the waves ended 20 us after the other half).  Each
    // wave lowers its priority at half its rows and with two and one rows
    // left: the one behind is preferred until it reaches the same mark, and
    // the wave that finishes first leads the other by at most about a row.
    constexpr bool PRIO = sizeof(real) == 4;
    // The wave dispatched first onto its CU (older: it wins ties) steps down
    // one row earlier than the later one, which ends the two closer together
    // (measured +1.6% against equal thresholds; two rows earlier was slower).
    const int sh = older ? 1 : 0;
    const int q1 = n_rows / 2 - sh, q2 = n_rows - 2 - sh, q3 = n_rows - 1 - sh;
    if constexpr (PRIO) __builtin_amdgcn_s_setprio(3);
    auto progress = [&](int done) __attribute__((always_inline)) {
        if constexpr (PRIO) {
            if (done == q1) __builtin_amdgcn_s_setprio(2);
            else if (done == q2) __builtin_amdgcn_s_setprio(1);
            else if (done == q3) __builtin_amdgcn_s_setprio(0);
        }
    };
    // t: rows done (wave-uniform); the row itself is i0 + t
    for (int t = 0;;) {
        if (t >= n_rows) break;
        row(i0 + t, A, B, C, QA, QC, F0, F1);
        progress(++t);
        if (!PK && t == 1) AMHD_WT_MARK(STAGE, wave, 2);
        if (t >= n_rows) break;
        row(i0 + t, B, C, A, QB, QA, F1, F2);
        progress(++t);
        if (t >= n_rows) break;
        row(i0 + t, C, A, B, QC, QB, F2, F0);
        progress(++t);
    }

    if constexpr (STAGE == 3) {
        if (a.part_dt || a.fuse) {
            double r[NPART_DT];
#pragma unroll
            for (int q = 0; q < NPART_DT; ++q) r[q] = wave_max(double(mx[q]));
            const int slot = part >= 0 ? part : wave;
            if (lane == 0) {
#pragma unroll
                for (int q = 0; q < NPART_DT; ++q) {
                    // fused loop: one partial per workgroup, combined by the caller
                    if (a.fuse) block_red()[threadIdx.x >> 6][q] = r[q];
                    else a.part_dt[(long)slot * NPART_DT + q] = r[q];
                }
            }
        }
    }
    if constexpr (STAGE == 1) {
        if (a.part_bf && !xper) {
            const double s0 = wave_sum(bf_sum[0]);
            const double s1 = wave_sum(bf_sum[1]);
            // the wavefront holding the strip's first (last) segment
            const bool first_seg = PK ? __builtin_amdgcn_ballot_w64(live && i0 == 0) != 0 : i0 == 0;
            const bool last_seg = PK ? __builtin_amdgcn_ballot_w64(live && i1 == nx) != 0 : i1 == nx;
            if (lane == 0) {
                if (first_seg) a.part_bf[strip * 2 + 0] = s0;
                if (last_seg) a.part_bf[strip * 2 + 1] = s1;
            }
        }
    }
}

template <class real, class P, int STAGE, bool POWELL, bool GLM>
__global__ __launch_bounds__(WAVE* WPB) AMHD_STAGE_OCC void stage_kernel(StageParams<real> a) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int nb = (int)gridDim.x - a.side;   // the strip workgroups
    if ((int)blockIdx.x >= nb) {
        if constexpr (STAGE == 1 || STAGE == 3) stage_side<real, STAGE>(a);
        return;
    }
    // wave-uniform quantities live in scalar registers
#ifdef AMHD_WAVE_TIMES
    const unsigned long long t_start = __builtin_amdgcn_s_memrealtime();
#endif
    const int wave = __builtin_amdgcn_readfirstlane(xcd_logical_block(nb) * WPB + (int)(threadIdx.x >> 6));
    const bool older = (int)blockIdx.x < a.first_blocks;
    const int n_wide = (a.pk_groups ? a.n_strips - 1 : a.n_strips) * a.n_segs;
    if (wave < n_wide) {
        stage_wave<real, P, STAGE, POWELL, GLM>(a, wave, lane, older);
    } else if (STAGE >= 1 && a.pk_groups && wave < stage_wave_count(a)) {
        if constexpr (STAGE >= 1) stage_wave<real, P, STAGE, POWELL, GLM, true>(a, wave - n_wide, lane, older, wave);
    } else if (STAGE == 3 && a.fuse && lane == 0) {
#pragma unroll
        for (int q = 0; q < NPART_DT; ++q) block_red()[threadIdx.x >> 6][q] = 0.0;
    }
    if constexpr (STAGE == 3) {
        // fused loop: the workgroup's CFL partial, the maxima over its
        // wavefronts (none when the loop is done: the last step's stay)
        if (a.fuse) {
            __syncthreads();
            if (threadIdx.x < NPART_DT && !a.ctl->done) {
                double m = block_red()[0][threadIdx.x];
                for (int w = 1; w < WPB; ++w) m = fmax(m, block_red()[w][threadIdx.x]);
                a.part_dt[(long)xcd_logical_block(nb) * NPART_DT + threadIdx.x] = m;
            }
        }
    }
#ifdef AMHD_WAVE_TIMES
    const unsigned long long t_end = __builtin_amdgcn_s_memrealtime();
    if (lane == 0 && wave < WT_WAVES) {
        amhd_wave_times[STAGE][wave][0] = t_start;
        amhd_wave_times[STAGE][wave][3] = t_end;
        const unsigned hw = __builtin_amdgcn_s_getreg((31 << 11) | 4);
        const unsigned xcc = __builtin_amdgcn_s_getreg((31 << 11) | 20);
        amhd_wave_times[STAGE][wave][4] = ((unsigned long long)xcc << 32) | hw;
    }
#endif
}

// ------------------------------------------------- CFL partials from state
// One partial per workgroup over the interior layout (used when no stage-3
// partials are current, e.g. right after a state upload).
// Workgroup blockIdx.x of `nblocks` covers cells blockIdx.x * 256 + t,
// stepping by nblocks * 256.
template <class real>
__device__ __forceinline__ void dt_partials_block(const real* __restrict__ u, long plane, int pitch, int nx, int ny,
                                                  double gamma, double* part, int nblocks) {
    real m[NPART_DT] = {0, 0, 0, 0};
    const long n = (long)nx * ny;
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)nblocks * blockDim.x) {
        const long i = k / ny, jj = k - i * ny;
        real c[NV];
        load_cell(u + i * pitch + jj, plane, c);
        dt_terms(c, gamma, m);
    }
    __shared__ double sh[NPART_DT][4];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int q = 0; q < NPART_DT; ++q) {
        const double r = wave_max(double(m[q]));
        if (lane == 0) sh[q][w] = r;
    }
    __syncthreads();
    if (threadIdx.x < NPART_DT) {
        double r = sh[threadIdx.x][0];
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q) r = fmax(r, sh[threadIdx.x][q]);
        part[(long)blockIdx.x * NPART_DT + threadIdx.x] = r;
    }
}

template <class real>
__global__ __launch_bounds__(256) void dt_partials_kernel(const real* __restrict__ u, long plane, int pitch,
                                                          int nx, int ny, double gamma, double* part) {
    dt_partials_block(u, plane, pitch, nx, ny, gamma, part, (int)gridDim.x);
}

// ------------------------------------------------------------- step control
enum : int { CTL_DT_ONLY = 0, CTL_HOST_STEP_END = 1, CTL_DEVICE_LOOP = 2 };

#ifdef AMHD_API_TU
// CFL partials from a primitive array W (NV planes of n cells), as
// compute_dt(W) of solver.py:206-223 reads it (floors re-applied).
__global__ __launch_bounds__(256) void dt_partials_w_kernel(const double* __restrict__ W, long n, double gamma,
                                                            double* part) {
    double m[NPART_DT] = {0.0, 0.0, 0.0, 0.0};
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
        const double rho = fmax(W[iRHO * n + k], FLOOR_RHO);
        const double p = fmax(W[iPR * n + k], FLOOR_PR);
        const double bx = W[iBX * n + k], by = W[iBY * n + k], bz = W[iBZ * n + k];
        const double b2 = bx * bx + by * by + bz * bz;
        const double cf2 = gamma * p / rho + b2 / rho;
        const double ax = fabs(W[iVX * n + k]), ay = fabs(W[iVY * n + k]);
        m[0] = fmax(m[0], cf2);
        m[1] = fmax(m[1], ax + ay);
        m[2] = fmax(m[2], ax);
        m[3] = fmax(m[3], ay);
    }
    __shared__ double sh[NPART_DT][4];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int q = 0; q < NPART_DT; ++q) {
        const double r = wave_max(m[q]);
        if (lane == 0) sh[q][w] = r;
    }
    __syncthreads();
    if (threadIdx.x < NPART_DT) {
        double r = sh[threadIdx.x][0];
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q) r = fmax(r, sh[threadIdx.x][q]);
        part[(long)blockIdx.x * NPART_DT + threadIdx.x] = r;
    }
}


// Reduces CFL partials to (ch, dt) (ctl_dt) and, in the device loop, also
// closes the previous step (ctl_close) and begins the next (ctl_begin).
// One workgroup of 256 threads.  CTL_DEVICE_LOOP reads cfl from ctl.
// The step-control update of step_ctl_block on a register copy of StepCtl.
__device__ __forceinline__ void step_ctl_update(StepCtl& h, const double (&r)[NPART_DT], double bflux, double dx,
                                                double dy, double host_dt, double cfl, int mode) {
    if (mode == CTL_HOST_STEP_END) {
        h.bflux = bflux;
        h.bflux_cum += bflux * host_dt;
        return;
    }
    if (mode == CTL_DEVICE_LOOP && h.pending) ctl_close(h, bflux);
    if (mode == CTL_DEVICE_LOOP) {
        if (h.done) return;
        cfl = h.cfl;
    }
    const double dt = ctl_dt(h, r, dx, dy, cfl);
    if (mode == CTL_DT_ONLY) {
        h.dt = dt;
        return;
    }
    ctl_begin(h, dt);   // device loop: solver.py:475-479
}

__device__ __forceinline__ void step_ctl_block(StepCtl* ctl, const double* part_dt, int n_part, const double* part_bf,
                                               int n_bf, double dx, double dy, double host_dt, double cfl, int mode) {
    __shared__ double sh[NPART_DT][4];
    __shared__ double shb[2][4];
    // thread 0 works on a register copy of the control block, loaded here so
    // that its latency overlaps the partials' and stored back once
    StepCtl h;
    if (threadIdx.x == 0) h = *ctl;
    // Eight partials per thread in flight at once (one load latency for 2048
    // partials with 256 threads); the maxima are exact in any order, and
    // every partial is >= 0, so a missing one counts as 0.
    constexpr int U = 8;
    double m[NPART_DT] = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = threadIdx.x; k0 < n_part; k0 += blockDim.x * U) {
        double v[U][NPART_DT];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = k0 + u * (int)blockDim.x;
#pragma unroll
            for (int q = 0; q < NPART_DT; ++q) v[u][q] = k < n_part ? part_dt[(long)k * NPART_DT + q] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int q = 0; q < NPART_DT; ++q) m[q] = fmax(m[q], v[u][q]);
        }
    }
    // boundary flux: strips summed in a fixed order by one wavefront
    double b0 = 0.0, b1 = 0.0;
    if (threadIdx.x < 64 && part_bf) {
        for (int k = threadIdx.x; k < n_bf; k += 64) { b0 += part_bf[2 * k]; b1 += part_bf[2 * k + 1]; }
    }
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int q = 0; q < NPART_DT; ++q) {
        const double r = wave_max(m[q]);
        if (lane == 0) sh[q][w] = r;
    }
    b0 = wave_sum(b0);
    b1 = wave_sum(b1);
    if (lane == 0) { shb[0][w] = b0; shb[1][w] = b1; }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double r[NPART_DT];
    for (int q = 0; q < NPART_DT; ++q) {
        r[q] = sh[q][0];
        for (int x = 1; x < (int)(blockDim.x >> 6); ++x) r[q] = fmax(r[q], sh[q][x]);
    }
    const double bflux = part_bf ? (shb[0][0] - shb[1][0]) : 0.0;

    step_ctl_update(h, r, bflux, dx, dy, host_dt, cfl, mode);
    *ctl = h;
}

__global__ __launch_bounds__(256) void step_ctl_kernel(StepCtl* ctl, const double* part_dt, int n_part,
                                                       const double* part_bf, int n_bf, double dx, double dy,
                                                       double host_dt, double cfl, int mode) {
    step_ctl_block(ctl, part_dt, n_part, part_bf, n_bf, dx, dy, host_dt, cfl, mode);
}

// Step control (workgroup 0) and, beside it, the segment boundaries of the
// next step re-sized from this step's row flags (workgroup 1).  The two
// touch disjoint memory, so they run at once on two CUs.
__global__ __launch_bounds__(256) void step_ctl_balance_kernel(StepCtl* ctl, const double* part_dt, int n_part,
                                                               const double* part_bf, int n_bf, double dx, double dy,
                                                               double host_dt, double cfl, int mode, Rebalance rb) {
    if (blockIdx.x == 1) {
        rebalance_block(rb);
        return;
    }
    step_ctl_block(ctl, part_dt, n_part, part_bf, n_bf, dx, dy, host_dt, cfl, mode);
}

#endif  // AMHD_API_TU

// -------------------------------------------------------- layout transfers
// Ghosted f64 (NV, nx+4, ny+4) -> interior layout (any real).
template <class real>
__global__ void ingest_kernel(const double* __restrict__ ug, real* __restrict__ u, long plane, int pitch, int nx,
                              int ny) {
    const int nyt = ny + 4;
    const long gplane = (long)(nx + 4) * nyt;
    const long n = (long)nx * ny;
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
        const long i = k / ny, jj = k - i * ny;
        const long src = (i + 2) * nyt + (jj + 2);
#pragma unroll
        for (int v = 0; v < NV; ++v) u[v * plane + i * pitch + jj] = real(ug[v * gplane + src]);
    }
}

// Interior layout -> ghosted f64 with ghosts filled as solver.py:158-197
// leaves them (images of interior cells; psi = 0 on non-periodic x ghosts).
template <class real>
__global__ void materialize_kernel(const real* __restrict__ u, long plane, int pitch, int nx, int ny,
                                   int bcx_periodic, int bcy_periodic, double* __restrict__ ug) {
    const int nxt = nx + 4, nyt = ny + 4;
    const long gplane = (long)nxt * nyt;
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < gplane; k += (long)gridDim.x * blockDim.x) {
        const int gi = (int)(k / nyt), gj = (int)(k - (long)gi * nyt);
        const int i = gi - 2, jj = gj - 2;
        const int si = remap(i, nx, bcx_periodic), sj = remap(jj, ny, bcy_periodic);
        const bool xg = (i < 0 || i >= nx);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            double x = double(u[v * plane + (long)si * pitch + sj]);
            if (v == PSI && xg && !bcx_periodic) x = 0.0;
            ug[v * gplane + k] = x;
        }
    }
}

#ifdef AMHD_API_TU
// ------------------------------------------------ standalone (drop-in API)
// physics.py:284-301 over (NV, n)
__global__ void cons_to_prim_kernel(const double* __restrict__ U, double* __restrict__ W, long n, double gamma) {
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
        double u[NV], w[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) u[v] = U[v * n + k];
        cons_to_prim<double, ExactMath<double>>(u, w, gamma, false);
#pragma unroll
        for (int v = 0; v < NV; ++v) W[v * n + k] = w[v];
    }
}

// physics.py:304-320 over (NV, n)
__global__ void prim_to_cons_kernel(const double* __restrict__ W, double* __restrict__ U, long n, double gamma) {
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
        const double rho = fmax(W[iRHO * n + k], FLOOR_RHO);
        const double vx = W[iVX * n + k], vy = W[iVY * n + k], vz = W[iVZ * n + k];
        const double bx = W[iBX * n + k], by = W[iBY * n + k], bz = W[iBZ * n + k];
        U[RHO * n + k] = rho;
        U[MX * n + k] = rho * vx;
        U[MY * n + k] = rho * vy;
        U[MZ * n + k] = rho * vz;
        U[BX * n + k] = bx;
        U[BY * n + k] = by;
        U[BZ * n + k] = bz;
        U[PSI * n + k] = W[iPSI * n + k];
        U[RHOC * n + k] = rho * W[iCLR * n + k];
        const double ke = 0.5 * rho * (vx * vx + vy * vy + vz * vz);
        const double me = 0.5 * (bx * bx + by * by + bz * bz);
        U[EN * n + k] = W[iPR * n + k] / (gamma - 1.0) + ke + me;
    }
}

// physics.py:323-350.  dir 0: faces along x, output (NV, nxt-3, nyt);
// dir 1: along y, output (NV, nxt, nyt-3).
__global__ void muscl_kernel(const double* __restrict__ W, int nxt, int nyt, int dir, double* __restrict__ WL,
                             double* __restrict__ WR) {
    const int fx = dir == 0 ? nxt - 3 : nxt;
    const int fy = dir == 0 ? nyt : nyt - 3;
    const long nf = (long)fx * fy, plane = (long)nxt * nyt;
    const long st = dir == 0 ? nyt : 1;
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < nf; k += (long)gridDim.x * blockDim.x) {
        const long a = k / fy, b = k - a * fy;
        const long c0 = dir == 0 ? a * nyt + b : a * nyt + b;  // first of the 4 cells
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const double* w = W + v * plane + c0;
            const double d1 = vl_slope<double, ExactMath<double>>(w[0], w[st], w[2 * st]);
            const double d2 = vl_slope<double, ExactMath<double>>(w[st], w[2 * st], w[3 * st]);
            double l = w[st] + 0.5 * d1;
            double r = w[2 * st] - 0.5 * d2;
            if (v == iRHO || v == iPR) {
                const double fl = v == iRHO ? FLOOR_RHO : FLOOR_PR;
                l = fmax(l, fl);
                r = fmax(r, fl);
            }
            WL[v * nf + k] = l;
            WR[v * nf + k] = r;
        }
    }
}

#endif  // AMHD_API_TU

// physics.py:43-275 + 353-379 over (NV, n) face arrays; smax_arr (may be
// null) receives every face's max(|SL|, |SR|) (physics.py:108-109), smax
// the maximum as the bit pattern of a non-negative double (ordered like the
// value) for atomicMax.
template <class real, class P>
__global__ __launch_bounds__(256) void hlld_kernel(const real* __restrict__ WL, const real* __restrict__ WR,
                                                   real* __restrict__ F, real* __restrict__ smax_arr, long n, int dir,
                                                   FluxConst<real> kc, unsigned long long* smax_bits) {
    double smax = 0.0;
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
        real wl[NV], wr[NV], f[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) { wl[v] = WL[v * n + k]; wr[v] = WR[v * n + k]; }
        const real s = dir == 0 ? hlld_face<real, P>(wl, wr, f, kc) : hlld_face_y<real, P>(wl, wr, f, kc);
        if (smax_arr) smax_arr[k] = s;
        smax = fmax(smax, double(s));
#pragma unroll
        for (int v = 0; v < NV; ++v) F[v * n + k] = f[v];
    }
    if (smax_bits) {
        smax = wave_max(smax);
        if ((threadIdx.x & 63) == 0) atomicMax(smax_bits, (unsigned long long)__double_as_longlong(smax));
    }
}

#ifdef AMHD_API_TU
// min rho, min p, max |divB| (interior stencil of solver.py:419-428) and a
// non-finite count over an interior-layout state; one partial per workgroup.
template <class real>
__global__ __launch_bounds__(256) void state_check_kernel(const real* __restrict__ u, long plane, int pitch, int nx,
                                                          int ny, double gamma, double dx, double dy, double* part) {
    double mr = 1e300, mp = 1e300, md = 0.0, nf = 0.0;
    const long n = (long)nx * ny;
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
        const int i = (int)(k / ny), j = (int)(k - (long)i * ny);
        const long o = (long)i * pitch + j;
        double c[NV];
        bool fin = true;
#pragma unroll
        for (int v = 0; v < NV; ++v) { c[v] = double(u[v * plane + o]); fin = fin && isfinite(c[v]); }
        if (!fin) nf += 1.0;
        const double rho = c[RHO];
        const double ke = 0.5 * (c[MX] * c[MX] + c[MY] * c[MY] + c[MZ] * c[MZ]) / rho;
        const double me = 0.5 * (c[BX] * c[BX] + c[BY] * c[BY] + c[BZ] * c[BZ]);
        const double p = (gamma - 1.0) * (c[EN] - ke - me);
        mr = fmin(mr, rho);
        mp = fmin(mp, p);
        if (i >= 1 && i < nx - 1 && j >= 1 && j < ny - 1) {
            const real* bx = u + BX * plane;
            const real* by = u + BY * plane;
            const double d = (0.5 * (double(bx[o]) + double(bx[o + pitch])) - 0.5 * (double(bx[o - pitch]) + double(bx[o]))) / dx +
                             (0.5 * (double(by[o]) + double(by[o + 1])) - 0.5 * (double(by[o - 1]) + double(by[o]))) / dy;
            md = fmax(md, fabs(d));
        }
    }
    __shared__ double sh[4][4];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int off = 32; off > 0; off >>= 1) {
        mr = fmin(mr, __shfl_xor(mr, off, 64));
        mp = fmin(mp, __shfl_xor(mp, off, 64));
        md = fmax(md, __shfl_xor(md, off, 64));
        nf += __shfl_xor(nf, off, 64);
    }
    if (lane == 0) { sh[0][w] = mr; sh[1][w] = mp; sh[2][w] = md; sh[3][w] = nf; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q) {
            sh[0][0] = fmin(sh[0][0], sh[0][q]); sh[1][0] = fmin(sh[1][0], sh[1][q]);
            sh[2][0] = fmax(sh[2][0], sh[2][q]); sh[3][0] += sh[3][q];
        }
        for (int q = 0; q < 4; ++q) part[blockIdx.x * 4 + q] = sh[q][0];
    }
}

// solver.py:158-197 in place on a ghosted f64 array.  Every ghost value is
// an image of an interior cell, so one pass over the ghost cells suffices.
__global__ void apply_bc_kernel(double* __restrict__ U, int nx, int ny, int bcx_periodic, int bcy_periodic) {
    const int nxt = nx + 4, nyt = ny + 4;
    const long gplane = (long)nxt * nyt;
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < gplane; k += (long)gridDim.x * blockDim.x) {
        const int gi = (int)(k / nyt), gj = (int)(k - (long)gi * nyt);
        const int i = gi - 2, jj = gj - 2;
        const bool xg = (i < 0 || i >= nx), yg = (jj < 0 || jj >= ny);
        if (!xg && !yg) continue;
        const int si = remap(i, nx, bcx_periodic) + 2, sj = remap(jj, ny, bcy_periodic) + 2;
        const long src = (long)si * nyt + sj;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            double x = U[v * gplane + src];
            if (v == PSI && xg && !bcx_periodic) x = 0.0;
            U[v * gplane + k] = x;
        }
    }
}

// solver.py:199-204 over (NV, n)
__global__ void enforce_bounds_kernel(double* __restrict__ U, long n) {
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
        const double rho = fmax(U[RHO * n + k], FLOOR_RHO);
        double c = U[RHOC * n + k] / rho;
        c = c < 0.0 ? 0.0 : (c > 1.0 ? 1.0 : c);
        U[RHOC * n + k] = rho * c;
    }
}

// psi *= decay over a plane (solver.py:333)
__global__ void scale_plane_kernel(double* __restrict__ x, long n, double s) {
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) x[k] *= s;
}

#endif  // AMHD_API_TU


// ------------------------------------------------------------ batched form
// Many independent instances advanced by one persistent launch per stage.
// A task is one (instance, strip, segment) wave of work; the persistent grid
// (one wave per resident slot) walks the task list with a fixed stride.
// Each instance keeps its own StepCtl (dt, ch, psi decay, t, step, done,
// current state buffer), so the captured step replays for every instance.
struct BatchInst {
    void* buf[3];          // state buffers (interior layout)
    double* part_dt;       // [n_strips * n_segs][4]
    double* part_bf;       // [n_strips][2]
    StepCtl* ctl;
    long plane;
    int pitch;
    unsigned bytes;
    int nx, ny, bcx_periodic, bcy_periodic, powell;
    int n_strips, n_segs, task_base, pad;
    double gamma, dx, dy;
};

// CFL partials of up to DT_LIST batch instances in one launch: row y of
// the grid is instance k[y], its state in buffer cur[y], over g[y]
// workgroups -- the cells and partials that dt_partials_kernel with g[y]
// workgroups gives that instance alone.
constexpr int DT_LIST = 32;
struct DtList {
    int k[DT_LIST], cur[DT_LIST], g[DT_LIST];
};

template <class real>
__global__ __launch_bounds__(256) void batch_dt_partials_kernel(const BatchInst* __restrict__ insts, DtList list) {
    const int y = blockIdx.y;
    const int g = list.g[y];
    if ((int)blockIdx.x >= g) return;   // uniform over the workgroup
    const BatchInst& d = insts[list.k[y]];
    dt_partials_block((const real*)d.buf[list.cur[y]], d.plane, d.pitch, d.nx, d.ny, d.gamma, d.part_dt, g);
}

template <class real, class P, int STAGE>
__global__ __launch_bounds__(WAVE* WPB) AMHD_STAGE_OCC void batch_stage_kernel(const BatchInst* __restrict__ insts, int n_inst,
                                                                int n_tasks) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int first = __builtin_amdgcn_readfirstlane(xcd_logical_block() * WPB + (int)(threadIdx.x >> 6));
    const int stride = gridDim.x * WPB;
    for (int task = first; task < n_tasks; task += stride) {
        // instance holding this task: last k with task_base <= task
        int lo = 0, hi = n_inst - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (insts[mid].task_base <= task) lo = mid; else hi = mid - 1;
        }
        const BatchInst& d = insts[lo];
        const int cur = d.ctl->cur;
        StageParams<real> a{};
        a.u0 = (const real*)d.buf[cur];
        a.us = STAGE == 1 ? (const real*)d.buf[cur] : STAGE == 2 ? (const real*)d.buf[(cur + 1) % 3]
                                                                : (const real*)d.buf[(cur + 2) % 3];
        a.uo = STAGE == 2 ? (real*)d.buf[(cur + 2) % 3] : (real*)d.buf[(cur + 1) % 3];
        a.plane = a.oplane = d.plane;
        a.pitch = a.opitch = d.pitch;
        a.in_bytes = a.out_bytes = d.bytes;
        a.nx = d.nx; a.ny = d.ny;
        a.bcx_periodic = d.bcx_periodic; a.bcy_periodic = d.bcy_periodic;
        a.n_strips = d.n_strips; a.n_segs = d.n_segs;
        a.gamma = d.gamma; a.dx = d.dx; a.dy = d.dy;
        a.dt = 0.0; a.ch = 0.0; a.decay = 1.0;
        a.ctl = d.ctl;
        a.part_dt = STAGE == 3 ? d.part_dt : nullptr;
        a.part_bf = STAGE == 1 ? d.part_bf : nullptr;
        a.halo_lo = nullptr;
        a.halo_hi = nullptr;
        a.first_blocks = 0;
        a.seg_i0 = nullptr;
        a.pk_groups = 0;
        a.pk_shift = 0;
        const int wave = task - d.task_base;
        if (d.powell) stage_wave<real, P, STAGE, true, true>(a, wave, lane);
        else stage_wave<real, P, STAGE, false, true>(a, wave, lane);
    }
}

#ifdef AMHD_API_TU
// One workgroup per instance: close the step and compute the next dt.
// first != 0: the partials come from the standalone CFL kernel
// (ctl->n_part of them); otherwise from stage 3 (n_strips * n_segs).
__global__ __launch_bounds__(256) void batch_ctl_kernel(const BatchInst* __restrict__ insts, int first) {
    const BatchInst& d = insts[blockIdx.x];
    const int n_part = first ? d.ctl->n_part : d.n_strips * d.n_segs;
    step_ctl_block(d.ctl, d.part_dt, n_part, d.bcx_periodic ? nullptr : d.part_bf, d.n_strips, d.dx, d.dy, 0.0,
                   0.0, CTL_DEVICE_LOOP);
}
#endif  // AMHD_API_TU

}  // namespace amhd
