// ref_launcher.hip -- TEST INFRASTRUCTURE ONLY.
//
// Host harness around the reference's HLLD kernel text (physics.py:43-275),
// which oracle/build_ref.py extracts into _ref/hlld_ref_kernel.inc.  It
// launches the kernel exactly as physics.py:359-367 does: one thread per
// face, block 256, SoA planes [v * n_faces + idx].  Host buffers in and out.
#include <hip/hip_runtime.h>

#include "hlld_ref_kernel.inc"

extern "C" int ref_hlld_flux_x(const double* wl, const double* wr, double* f, double* smax_arr, int n_faces,
                               double gamma, double ch, double floor_rho, double floor_pr) {
    if (n_faces <= 0) return 0;
    const size_t bytes = sizeof(double) * 10 * (size_t)n_faces;
    double *dl = nullptr, *dr = nullptr, *df = nullptr, *ds = nullptr;
    if (hipMalloc(&dl, bytes) != hipSuccess || hipMalloc(&dr, bytes) != hipSuccess ||
        hipMalloc(&df, bytes) != hipSuccess || hipMalloc(&ds, sizeof(double) * (size_t)n_faces) != hipSuccess)
        return -3;
    int rc = 0;
    if (hipMemcpy(dl, wl, bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dr, wr, bytes, hipMemcpyHostToDevice) != hipSuccess)
        rc = -2;
    if (!rc) {
        const int block = 256, grid = (n_faces + block - 1) / block;
        hlld_flux_x_kernel<<<grid, block>>>(dl, dr, df, ds, n_faces, gamma, ch, floor_rho, floor_pr);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = -2;
    }
    if (!rc && (hipMemcpy(f, df, bytes, hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(smax_arr, ds, sizeof(double) * (size_t)n_faces, hipMemcpyDeviceToHost) != hipSuccess))
        rc = -2;
    hipFree(dl); hipFree(dr); hipFree(df); hipFree(ds);
    return rc;
}
