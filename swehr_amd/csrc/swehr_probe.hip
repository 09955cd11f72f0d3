// swehr_probe.hip - the HBM rate this GPU reaches for the step kernel's access pattern
// (a diagnostic beside the step; no counterpart in the reference).
//
// A grid of 4 x 64-cell tiles (a wave reads 64 consecutive j of one row, rows at the
// step's pitch of a multiple of 64 cells) reads `in` planes and writes `out` planes of
// nx x ny cells with nontemporal loads and stores, as the step kernel streams its state.
// The rate counts the bytes the kernel must move (every plane read and written once);
// the best of `reps` timed launches is reported.  With the step's own plane counts it is
// the practical HBM ceiling of the step kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>

#include "swehr.h"
#include "swehr_lib_internal.h"

namespace {

template <typename T>
__global__ __launch_bounds__(256) void stream_kernel(const T *src, T *dst, int nx, int ny, int py, long long plane,
                                                     int nin, int nout, int tiles_j)
{
    const int tile = blockIdx.x;
    const int ti = tile / tiles_j, tj = tile - ti * tiles_j;
    const int i = ti * 4 + (int)(threadIdx.x >> 6), j = tj * 64 + (int)(threadIdx.x & 63);
    if (i >= nx || j >= ny) return;
    const long long c = (long long)i * py + j;
    T acc = T(0);
#pragma unroll 4
    for (int k = 0; k < nin; k++) acc += __builtin_nontemporal_load(src + k * plane + c);
#pragma unroll 4
    for (int k = 0; k < nout; k++) __builtin_nontemporal_store(acc + T(k), dst + k * plane + c);
}

template <typename T>
int stream_run(int nx, int ny, int nin, int nout, int reps, double *gbs)
{
    const int py = (ny + 63) / 64 * 64;
    const long long plane = (long long)nx * py;
    T *src = nullptr, *dst = nullptr;
    const size_t bin = sizeof(T) * (size_t)plane * std::max(nin, 1), bout = sizeof(T) * (size_t)plane * std::max(nout, 1);
    if (hipMalloc(&src, bin) != hipSuccess) {
        (void)hipGetLastError();
        swehr_set_error("hbm stream: no device memory%s%s", "", "");
        return SWEHR_ERR_OOM;
    }
    if (hipMalloc(&dst, bout) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(src);
        swehr_set_error("hbm stream: no device memory%s%s", "", "");
        return SWEHR_ERR_OOM;
    }
    int rc = SWEHR_OK;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    do {
        if (hipMemset(src, 0, bin) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
            hipEventCreate(&e1) != hipSuccess) {
            rc = SWEHR_ERR_HIP;
            break;
        }
        const int tiles_j = (ny + 63) / 64, tiles = ((nx + 3) / 4) * tiles_j;
        double best = 0;
        for (int r = 0; r <= reps && rc == SWEHR_OK; r++) {  // r = 0 warms up
            (void)hipEventRecord(e0, nullptr);
            hipLaunchKernelGGL(stream_kernel<T>, dim3(tiles), dim3(256), 0, nullptr, src, dst, nx, ny, py, plane, nin,
                               nout, tiles_j);
            (void)hipEventRecord(e1, nullptr);
            if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
                rc = SWEHR_ERR_HIP;
                break;
            }
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            const double bytes = (double)sizeof(T) * nx * (double)ny * (nin + nout);
            if (r > 0 && ms > 0) best = std::max(best, bytes / (ms * 1e-3) / 1e9);
        }
        *gbs = best;
    } while (0);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(src);
    (void)hipFree(dst);
    if (rc == SWEHR_ERR_HIP) swehr_set_error("hbm stream: %s%s", hipGetErrorString(hipGetLastError()), "");
    return rc;
}

}  // namespace

extern "C" int swehr_hbm_stream(int nx, int ny, int elem_bytes, int planes_in, int planes_out, int reps,
                                double *gbs)
{
    if (!gbs || nx < 1 || ny < 1 || planes_in < 0 || planes_out < 0 || planes_in + planes_out < 1 || reps < 1 ||
        (elem_bytes != 4 && elem_bytes != 8))
        return SWEHR_ERR_BAD_INPUT;
    *gbs = 0;
    return elem_bytes == 8 ? stream_run<double>(nx, ny, planes_in, planes_out, reps, gbs)
                           : stream_run<float>(nx, ny, planes_in, planes_out, reps, gbs);
}
