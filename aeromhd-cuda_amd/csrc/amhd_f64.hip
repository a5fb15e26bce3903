// amhd_f64.hip -- FP64 instantiation (ExactMath policy).  Built with
// -ffp-contract=off so every a*b+c is two IEEE operations, as in the
// reference's CuPy ufuncs; division and sqrt are the correctly rounded
// expansions.
#define AMHD_LAUNCH_IMPL
#include "amhd_launch.hpp"

namespace amhd {

hipError_t launch_stage_f64(int stage, bool powell, bool glm, const StageParams<double>& a, hipStream_t s,
                            hipEvent_t ev0, hipEvent_t ev1) {
    return launch_stage_impl<double, ExactMath<double>>(stage, powell, glm, a, s, ev0, ev1);
}

hipError_t launch_batch_stage_f64(int stage, const BatchInst* insts, int n_inst, int n_tasks, int resident,
                                  hipStream_t s) {
    return launch_batch_stage_impl<double, ExactMath<double>>(stage, insts, n_inst, n_tasks, resident, s);
}

int batch_waves_per_cu_f64() { return batch_waves_per_cu_impl<double, ExactMath<double>>(); }

int stage_waves_per_cu_f64(bool powell) { return stage_waves_per_cu_impl<double, ExactMath<double>>(powell); }

int launch_dt_partials_f64(const double* u, long plane, int pitch, int nx, int ny, double gamma, double* part,
                           int max_parts, hipStream_t s) {
    return launch_dt_partials_impl<double>(u, plane, pitch, nx, ny, gamma, part, max_parts, s);
}

hipError_t launch_batch_dt_partials_f64(const BatchInst* insts, const DtList& list, int cnt, int gmax,
                                        hipStream_t s) {
    return launch_batch_dt_partials_impl<double>(insts, list, cnt, gmax, s);
}

hipError_t launch_ingest_f64(const double* ug, double* u, long plane, int pitch, int nx, int ny, hipStream_t s) {
    return launch_ingest_impl<double>(ug, u, plane, pitch, nx, ny, s);
}

hipError_t launch_materialize_f64(const double* u, long plane, int pitch, int nx, int ny, int bcx, int bcy,
                                  double* ug, hipStream_t s) {
    return launch_materialize_impl<double>(u, plane, pitch, nx, ny, bcx, bcy, ug, s);
}

}  // namespace amhd
