// amhd_f32.hip -- FP32 instantiation of the fused update (FastMath policy).
// Built with -ffp-contract=on (a*b+c fused per source expression, so the
// same body gives the same bits in every kernel it is inlined into).  FP32
// results are not bit-matched to the
// reference; their error is measured against the FP64 oracle.
#define AMHD_LAUNCH_IMPL
#include "amhd_launch.hpp"

namespace amhd {

hipError_t launch_stage_f32(int stage, bool powell, bool glm, const StageParams<float>& a, hipStream_t s,
                            hipEvent_t ev0, hipEvent_t ev1) {
    return launch_stage_impl<float, FastMath<float>>(stage, powell, glm, a, s, ev0, ev1);
}

hipError_t launch_batch_stage_f32(int stage, const BatchInst* insts, int n_inst, int n_tasks, int resident,
                                  hipStream_t s) {
    return launch_batch_stage_impl<float, FastMath<float>>(stage, insts, n_inst, n_tasks, resident, s);
}

int batch_waves_per_cu_f32() { return batch_waves_per_cu_impl<float, FastMath<float>>(); }

int stage_waves_per_cu_f32(bool powell) { return stage_waves_per_cu_impl<float, FastMath<float>>(powell); }

int launch_dt_partials_f32(const float* u, long plane, int pitch, int nx, int ny, double gamma, double* part,
                           int max_parts, hipStream_t s) {
    return launch_dt_partials_impl<float>(u, plane, pitch, nx, ny, gamma, part, max_parts, s);
}

hipError_t launch_batch_dt_partials_f32(const BatchInst* insts, const DtList& list, int cnt, int gmax,
                                        hipStream_t s) {
    return launch_batch_dt_partials_impl<float>(insts, list, cnt, gmax, s);
}

hipError_t launch_ingest_f32(const double* ug, float* u, long plane, int pitch, int nx, int ny, hipStream_t s) {
    return launch_ingest_impl<float>(ug, u, plane, pitch, nx, ny, s);
}

hipError_t launch_materialize_f32(const float* u, long plane, int pitch, int nx, int ny, int bcx, int bcy,
                                  double* ug, hipStream_t s) {
    return launch_materialize_impl<float>(u, plane, pitch, nx, ny, bcx, bcy, ug, s);
}

hipError_t launch_hlld_f32(const float* wl, const float* wr, float* f, float* smax_arr, long n, int dir, double gamma,
                           double ch, double floor_rho, double floor_pr, unsigned long long* smax_bits, hipStream_t s) {
    const FluxConst<float> kc = make_flux_const<float>(gamma, ch, floor_rho, floor_pr);
    hlld_kernel<float, FastMath<float>><<<grid_for(n, 256), 256, 0, s>>>(wl, wr, f, smax_arr, n, dir, kc, smax_bits);
    return hipGetLastError();
}

}  // namespace amhd

#ifdef AMHD_WAVE_TIMES
// Diagnostic builds only: copies the FP32 stage kernels' wave times
// (4 x 8192 x 5 unsigned 64-bit values) to host memory.  Not part of the
// ABI of include/aeromhd.h.
extern "C" int amhd_debug_wave_times(unsigned long long* out) {
    return hipMemcpyFromSymbol(out, HIP_SYMBOL(amhd::amhd_wave_times), sizeof(amhd::amhd_wave_times)) == hipSuccess
               ? 0 : -2;
}
#endif
