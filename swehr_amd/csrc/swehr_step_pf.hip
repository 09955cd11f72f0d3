// swehr_step_pf.hip - launcher of the prefetching FP64 step kernel (step_kernel<double,
// TI, TJ, false, true>: classes spanning two halo batches, the next batch's first
// classes loaded before the current batch is converted).
//
// It has a file of its own so that this kernel alone is built with the backend's
// max-ilp scheduling strategy (the Makefile's rule for this object), which orders its
// independent loads and arithmetic for latency within the same 168 registers:
// 4096^2 cells, S = 10, FP64: step kernel 1.993-1.996 -> 1.978-1.980 ms, results
// unchanged (scheduling does not change what is computed).  The same strategy spills
// in the plain variants and slows the wave-speed kernel, which keep the default.
#include <hip/hip_runtime.h>

#include "swehr_device.h"
#include "swehr_lib_internal.h"

namespace swehr {

// Tile ids 1-3 of the autotuner (8x32, 4x64, 16x16; 256 threads); the 128-thread tile
// has no prefetching variant.
void launch_step_prefetch(int id, int blocks, size_t lds, hipStream_t s, const StepArgs<double> &A)
{
    switch (id) {
    case 1: hipLaunchKernelGGL((step_kernel<double, 8, 32, false, true>), dim3(blocks), dim3(256), lds, s, A); break;
    case 2: hipLaunchKernelGGL((step_kernel<double, 4, 64, false, true>), dim3(blocks), dim3(256), lds, s, A); break;
    case 3: hipLaunchKernelGGL((step_kernel<double, 16, 16, false, true>), dim3(blocks), dim3(256), lds, s, A); break;
    default: break;
    }
}

}  // namespace swehr
