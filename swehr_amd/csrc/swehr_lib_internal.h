// swehr_lib_internal.h - definitions shared by the host-side translation units
// (swehr_lib.hip: single instance; swehr_step_pf.hip: the prefetching FP64 step
// kernel's launcher; swehr_batch.hip: ensembles and serving).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

#include "swehr.h"
#include "swehr_device.h"

// Record the calling thread's last error (returned by swehr_last_error()).
void swehr_set_error(const char *fmt, const char *a, const char *b);

namespace swehr {
// Launch the prefetching FP64 step kernel on tile id 1-3 (swehr_step_pf.hip).
void launch_step_prefetch(int id, int blocks, size_t lds, hipStream_t s, const StepArgs<double> &A);
}  // namespace swehr

#define SWEHR_HIPCHK(x)                                                                        \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) {                                                                \
            swehr_set_error("%s: %s", #x, hipGetErrorString(e_));                               \
            return SWEHR_ERR_HIP;                                                              \
        }                                                                                      \
    } while (0)

// One solver instance: its block of device memory and the kernel arguments that
// point into it.
struct swehr_handle {
    // device block and kernel arguments
    swehr_params p;
    int precision = SWEHR_FP64;
    int device = 0;
    int nx = 0, ny = 0, S = 0, py = 0;
    long long plane = 0;
    void *mem = nullptr;
    size_t bytes = 0;
    swehr::DevCtrl *ctrl = nullptr;   // device
    swehr::DevCtrl *hctrl = nullptr;  // pinned host mirror
    long long *hlimit = nullptr;
    int *hflag = nullptr;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_steps = 16;
    int tile = 0;
    int tiles_i = 0, tiles_j = 0, nblocks = 0, blocks_per_xcd = 0, bud_slots = 0;
    size_t lds = 0;
    int k1_blocks = 0;
    int stage_cap = 0;
    double z0sum = 0;
    double *scratch = nullptr;
    double *partials = nullptr;
    int budget_blocks = 0;
    swehr::StepArgs<double> a64{};
    swehr::StepArgs<float> a32{};
    double *prast[3] = {nullptr, nullptr, nullptr};  // parameter rasters (device), separate allocations
    bool member = false;                              // owned by an ensemble or a server
    bool own_stream = true;                           // false: the stream is the ensemble's or server's
};

namespace swehr {
// Tile shapes the step kernels are instantiated for (1-based ids).
int tile_count();
void tile_shape(int id, int *ti, int *tj);
// Dynamic LDS bytes of the step kernel of tile id for S classes at precision prec.
size_t tile_lds_bytes(int id, int prec, int S);
// Largest LDS a step-kernel workgroup may use.
constexpr size_t kLdsLimit = 64 * 1024;
// Ensemble launch of K2 for tile shape id (one block per map entry).
// Ensemble K1 on strip tiles: one block per entry of the wave-speed map (instance,
// strip tile), blocks a multiple of 8.
void launch_batch_wave_strip(int prec, int blocks, hipStream_t s, const void *args, const TileRef *map, int n,
                             int per_xcd);
// max_classes: the largest S of the members (selects the FP64 two-chunk variant above
// kLdsClasses<double>)
void launch_batch_step(int id, int prec, int blocks, size_t lds, hipStream_t s, const void *args, const TileRef *map,
                       int n, int per_xcd, int max_classes);
// Whether tile shape id at precision prec runs the SPLIT step variant for max_classes
// classes and lds bytes of tile LDS (the choice launch_batch_step makes).
bool batch_uses_split(int prec, int id, int max_classes, size_t lds);
}  // namespace swehr
