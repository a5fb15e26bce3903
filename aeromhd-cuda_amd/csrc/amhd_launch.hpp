// amhd_launch.hpp -- host-side launchers shared by the precision TUs.
//
// amhd_f64.hip instantiates them for double with ExactMath (built with
// -ffp-contract=off: the reference's IEEE FP64 arithmetic), amhd_f32.hip for
// float with FastMath (contraction allowed).  amhd_api.hip only sees the
// declarations below.
#pragma once
#include <algorithm>
#include <hip/hip_ext.h>
#include "amhd_kernels.hpp"

namespace amhd {

// One SSP-RK3 stage (0 = RHS only) of the fused strip-march kernel.  With
// ev0 and ev1 (created events), the launch records the kernel's own start
// and end in them (the dispatch's timestamps).
hipError_t launch_stage_f64(int stage, bool powell, bool glm, const StageParams<double>& a, hipStream_t s,
                            hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
hipError_t launch_stage_f32(int stage, bool powell, bool glm, const StageParams<float>& a, hipStream_t s,
                            hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);

// Batched stage launch over a device array of BatchInst (persistent grid of
// `resident` wavefronts).
hipError_t launch_batch_stage_f64(int stage, const BatchInst* insts, int n_inst, int n_tasks, int resident,
                                  hipStream_t s);
hipError_t launch_batch_stage_f32(int stage, const BatchInst* insts, int n_inst, int n_tasks, int resident,
                                  hipStream_t s);
int batch_waves_per_cu_f64();
int batch_waves_per_cu_f32();

// Resident wavefronts per CU of the stage kernel (occupancy query).
int stage_waves_per_cu_f64(bool powell);
int stage_waves_per_cu_f32(bool powell);

// CFL partial maxima over an interior-layout state; returns #partials.
int launch_dt_partials_f64(const double* u, long plane, int pitch, int nx, int ny, double gamma, double* part,
                           int max_parts, hipStream_t s);
int launch_dt_partials_f32(const float* u, long plane, int pitch, int nx, int ny, double gamma, double* part,
                           int max_parts, hipStream_t s);
// The same for list entries 0 .. cnt-1 of batch instances (gmax = the
// largest list.g).
hipError_t launch_batch_dt_partials_f64(const BatchInst* insts, const DtList& list, int cnt, int gmax, hipStream_t s);
hipError_t launch_batch_dt_partials_f32(const BatchInst* insts, const DtList& list, int cnt, int gmax, hipStream_t s);

hipError_t launch_ingest_f64(const double* ug, double* u, long plane, int pitch, int nx, int ny, hipStream_t s);
hipError_t launch_ingest_f32(const double* ug, float* u, long plane, int pitch, int nx, int ny, hipStream_t s);
hipError_t launch_materialize_f64(const double* u, long plane, int pitch, int nx, int ny, int bcx, int bcy,
                                  double* ug, hipStream_t s);
hipError_t launch_materialize_f32(const float* u, long plane, int pitch, int nx, int ny, int bcx, int bcy,
                                  double* ug, hipStream_t s);

hipError_t launch_hlld_f32(const float* wl, const float* wr, float* f, float* smax_arr, long n, int dir, double gamma,
                           double ch, double floor_rho, double floor_pr, unsigned long long* smax_bits, hipStream_t s);

inline int grid_for(long n, int block, int cap = 4096) {
    long g = (n + block - 1) / block;
    if (g < 1) g = 1;
    return (int)(g > cap ? cap : g);
}

// Template bodies, included by the two precision TUs only.
#ifdef AMHD_LAUNCH_IMPL
template <class real, class P, int STAGE>
void launch_stage_t(bool powell, bool glm, const StageParams<real>& a, dim3 grid, dim3 block, hipStream_t s,
                    hipEvent_t ev0, hipEvent_t ev1) {
    auto go = [&](auto kernel) {
        if (ev0 || ev1) hipExtLaunchKernelGGL(kernel, grid, block, 0, s, ev0, ev1, 0, a);
        else kernel<<<grid, block, 0, s>>>(a);
    };
    if (powell) {
        if (glm) go(stage_kernel<real, P, STAGE, true, true>);
        else go(stage_kernel<real, P, STAGE, true, false>);
    } else {
        if (glm) go(stage_kernel<real, P, STAGE, false, true>);
        else go(stage_kernel<real, P, STAGE, false, false>);
    }
}

// glm: the caller guarantees ch > 0 for this launch (the device loop always
// has ch >= 1.5; host-driven steps pass their ch).
template <class real, class P>
hipError_t launch_stage_impl(int stage, bool powell, bool glm, const StageParams<real>& a, hipStream_t s,
                             hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr) {
    const int n_waves = stage_wave_count(a);
    const dim3 grid((n_waves + WPB - 1) / WPB + (a.side ? 1 : 0)), block(WAVE * WPB);
    switch (stage) {
        case 0: launch_stage_t<real, P, 0>(powell, glm, a, grid, block, s, ev0, ev1); break;
        case 1: launch_stage_t<real, P, 1>(powell, glm, a, grid, block, s, ev0, ev1); break;
        case 2: launch_stage_t<real, P, 2>(powell, glm, a, grid, block, s, ev0, ev1); break;
        case 3: launch_stage_t<real, P, 3>(powell, glm, a, grid, block, s, ev0, ev1); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class real, class P>
hipError_t launch_batch_stage_impl(int stage, const BatchInst* insts, int n_inst, int n_tasks, int resident,
                                   hipStream_t s) {
    const dim3 grid((resident + WPB - 1) / WPB), block(WAVE * WPB);
    switch (stage) {
        case 1: batch_stage_kernel<real, P, 1><<<grid, block, 0, s>>>(insts, n_inst, n_tasks); break;
        case 2: batch_stage_kernel<real, P, 2><<<grid, block, 0, s>>>(insts, n_inst, n_tasks); break;
        case 3: batch_stage_kernel<real, P, 3><<<grid, block, 0, s>>>(insts, n_inst, n_tasks); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class real, class P>
int batch_waves_per_cu_impl() {
    // resident wavefronts of the persistent launch: the fewest over its stages
    auto blocks = [](auto kernel) {
        int nb = 0;
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, WAVE * WPB, 0) == hipSuccess && nb > 0 ? nb : 1;
    };
    return std::min({blocks(batch_stage_kernel<real, P, 1>), blocks(batch_stage_kernel<real, P, 2>),
                     blocks(batch_stage_kernel<real, P, 3>)}) * WPB;
}

template <class real, class P>
int stage_waves_per_cu_impl(bool powell) {
    // the three stages share one tiling: size it for the one with the fewest
    // resident wavefronts (their register counts differ)
    auto blocks = [](auto kernel) {
        int nb = 0;
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, WAVE * WPB, 0) == hipSuccess && nb > 0 ? nb : 1;
    };
    const int nb = powell ? std::min({blocks(stage_kernel<real, P, 1, true, true>), blocks(stage_kernel<real, P, 2, true, true>),
                                      blocks(stage_kernel<real, P, 3, true, true>)})
                          : std::min({blocks(stage_kernel<real, P, 1, false, true>),
                                      blocks(stage_kernel<real, P, 2, false, true>),
                                      blocks(stage_kernel<real, P, 3, false, true>)});
    return nb * WPB;
}

template <class real>
int launch_dt_partials_impl(const real* u, long plane, int pitch, int nx, int ny, double gamma, double* part,
                            int max_parts, hipStream_t s) {
    const int g = grid_for((long)nx * ny, 256, max_parts);
    dt_partials_kernel<real><<<g, 256, 0, s>>>(u, plane, pitch, nx, ny, gamma, part);
    return hipGetLastError() == hipSuccess ? g : -1;
}

template <class real>
hipError_t launch_batch_dt_partials_impl(const BatchInst* insts, const DtList& list, int cnt, int gmax,
                                         hipStream_t s) {
    batch_dt_partials_kernel<real><<<dim3(gmax, cnt), 256, 0, s>>>(insts, list);
    return hipGetLastError();
}

template <class real>
hipError_t launch_ingest_impl(const double* ug, real* u, long plane, int pitch, int nx, int ny, hipStream_t s) {
    ingest_kernel<real><<<grid_for((long)nx * ny, 256), 256, 0, s>>>(ug, u, plane, pitch, nx, ny);
    return hipGetLastError();
}

template <class real>
hipError_t launch_materialize_impl(const real* u, long plane, int pitch, int nx, int ny, int bcx, int bcy,
                                   double* ug, hipStream_t s) {
    materialize_kernel<real><<<grid_for((long)(nx + 4) * (ny + 4), 256), 256, 0, s>>>(u, plane, pitch, nx, ny, bcx,
                                                                                        bcy, ug);
    return hipGetLastError();
}
#endif

}  // namespace amhd
