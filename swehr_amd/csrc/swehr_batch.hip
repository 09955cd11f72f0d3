// swehr_batch.hip - ensembles and the serving loop on one GPU.
//
// An ensemble steps many independent instances together: every stage of a time step
// (wave speeds, fused step, control) is ONE launch over the tiles of all unfinished
// instances, addressed through a work map of (instance, tile) entries; each instance
// keeps its own dt, time, rain table and outlet series in its own control block, so
// grids of different sizes and parameters share the launches without synchronising
// their clocks.  Steps run as captured HIP graphs; the graph ends with a gather of all
// control blocks into one array, so the host checks progress with a single copy.
//
// The server adds continuous batching: requests wait in a FIFO and are admitted while
// an instance slot is free and the device-memory budget holds their state; the work
// map and the graph are rebuilt whenever the running set changes.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <deque>
#include <map>
#include <new>
#include <vector>

#include "swehr.h"
#include "swehr_device.h"
#include "swehr_lib_internal.h"

using namespace swehr;

#define HIPCHK SWEHR_HIPCHK

namespace swehr {

__global__ void gather_ctrl_kernel(DevCtrl *const *ctrls, int n, DevCtrl *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = *ctrls[i];
}

__global__ void set_limits_kernel(DevCtrl *const *ctrls, int n, long long add)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        DevCtrl *c = ctrls[i];
        c->step_limit = add < 0 ? LLONG_MAX : c->steps + add;
    }
}

// The ensemble's steps in ONE cooperative launch (the persistent mode of an ensemble):
// every block loops over `steps` steps, and in each runs the wave-speed pass over the
// (instance, tile) entries, a grid barrier, the fused step over the same entries, a grid
// barrier, the control of the instances, a grid barrier.  The device functions are the
// per-stage kernels' own, so the results are bit-identical to the three launches per
// step.  Blocks take map entries in the XCD bands the per-stage kernels use (the
// dispatcher deals block b to XCD b mod 8): block b serves band b mod 8 at positions
// b / 8, b / 8 + grid / 8, ...  A block barrier separates consecutive tiles of a block,
// whose shared memory the next tile reuses.  The grid holds at most as many blocks as
// are resident at once (cooperative launch), so every grid barrier is reached.
// Register cost: the step body inside a loop over tiles does not fit the 3-wave (FP64)
// or 5-wave (FP32) budget that one tile per block fits without a spill.  Built as it
// is (this file without machine LICM, the thread index laundered in the loop body:
// step_tile's LOOP form) the compiler spills 44-72 bytes per lane, against ~420 with
// the step's constants and index arithmetic hoisted out of the loop.
template <typename R, int TI, int TJ, bool SPLIT>
__global__ __launch_bounds__(TI *TJ, (kStepWaves<R, SPLIT>)) void batch_persistent_kernel(
    const StepArgs<R> *__restrict__ args, const TileRef *__restrict__ map, int n, int per_xcd,
    const int *__restrict__ insts, int ninst, int steps)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cooperative_groups::grid_group grid = cooperative_groups::this_grid();
    const int b = blockIdx.x, xcd = b & 7, bx = b >> 3, per_b = (int)(gridDim.x >> 3);
    for (int s = 0; s < steps; s++) {
        for (int q = bx; q < per_xcd; q += per_b) {
            const int k = xcd * per_xcd + q;
            if (k < n) {
                const TileRef r = map[k];
                wave_speed_tile<R, TI, TJ>(args[r.inst], r.tile);
            }
            __syncthreads();
        }
        grid.sync();
        for (int q = bx; q < per_xcd; q += per_b) {
            const int k = xcd * per_xcd + q;
            if (k < n) {
                const TileRef r = map[k];
                step_tile<R, TI, TJ, SPLIT, fp64_own<TI, TJ, true>(), true>(args[r.inst], r.tile, smem_raw);
            }
            __syncthreads();
        }
        grid.sync();
        for (int k = b; k < ninst; k += (int)gridDim.x) {
            control_body(args[insts[k]]);
            __syncthreads();
        }
        grid.sync();
    }
}

template <typename R, bool SP>
const void *persistent_kernel_t(int id)
{
    switch (id) {
    case 1: return (const void *)&batch_persistent_kernel<R, 8, 32, SP>;
    case 2: return (const void *)&batch_persistent_kernel<R, 4, 64, SP>;
    case 3: return (const void *)&batch_persistent_kernel<R, 16, 16, SP>;
    default: return nullptr;  // 128-thread tiles: not built in this mode
    }
}

}  // namespace swehr

namespace {

// Launch engine shared by swehr_batch and swehr_server.
struct Engine {
    int precision = SWEHR_FP64;
    int tile = 0;
    hipStream_t stream = nullptr;
    std::vector<swehr_handle *> members;  // running instances, in map order
    // device tables
    void *d_args = nullptr;
    TileRef *d_map = nullptr;
    int *d_insts = nullptr;
    DevCtrl **d_ctrls = nullptr;
    DevCtrl *d_gather = nullptr;
    DevCtrl *h_gather = nullptr;  // pinned
    size_t cap_members = 0, cap_map = 0;
    int n_map = 0, per_xcd = 0;
    // the wave-speed map (instance, strip tile) follows the step map in d_map
    int n_map1 = 0, per_xcd1 = 0;
    size_t lds = 0;
    int max_classes = 0;  // largest S of the members
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_steps = 16;
    bool dirty = true;

    size_t arg_size() const { return precision == SWEHR_FP64 ? sizeof(StepArgs<double>) : sizeof(StepArgs<float>); }

    void drop_graph()
    {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
    }

    void release()
    {
        drop_graph();
        if (d_args) (void)hipFree(d_args);
        if (d_map) (void)hipFree(d_map);
        if (d_insts) (void)hipFree(d_insts);
        if (d_ctrls) (void)hipFree(d_ctrls);
        if (d_gather) (void)hipFree(d_gather);
        if (h_gather) (void)hipHostFree(h_gather);
        if (stream) (void)hipStreamDestroy(stream);
        d_args = nullptr;
        d_map = nullptr;
        d_insts = nullptr;
        d_ctrls = nullptr;
        d_gather = nullptr;
        h_gather = nullptr;
        stream = nullptr;
    }

    int init()
    {
        HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        return SWEHR_OK;
    }

    // Grow the member tables and the work map.  The capacities are zeroed before the old
    // buffers are freed, so an allocation that fails leaves capacity 0 and the next
    // rebuild allocates again instead of using buffers that no longer exist.
    int reserve(size_t members_needed, size_t map_needed)
    {
        if (members_needed > cap_members) {
            size_t c = std::max<size_t>(members_needed, 2 * cap_members);
            cap_members = 0;
            if (d_args) (void)hipFree(d_args);
            if (d_insts) (void)hipFree(d_insts);
            if (d_ctrls) (void)hipFree(d_ctrls);
            if (d_gather) (void)hipFree(d_gather);
            if (h_gather) (void)hipHostFree(h_gather);
            d_args = nullptr;
            d_insts = nullptr;
            d_ctrls = nullptr;
            d_gather = nullptr;
            h_gather = nullptr;
            HIPCHK(hipMalloc(&d_args, c * arg_size()));
            HIPCHK(hipMalloc((void **)&d_insts, c * sizeof(int)));
            HIPCHK(hipMalloc((void **)&d_ctrls, c * sizeof(DevCtrl *)));
            HIPCHK(hipMalloc((void **)&d_gather, c * sizeof(DevCtrl)));
            HIPCHK(hipHostMalloc((void **)&h_gather, c * sizeof(DevCtrl), 0));
            cap_members = c;
        }
        if (map_needed > cap_map) {
            size_t c = std::max<size_t>(map_needed, 2 * cap_map);
            cap_map = 0;
            if (d_map) (void)hipFree(d_map);
            d_map = nullptr;
            HIPCHK(hipMalloc((void **)&d_map, c * sizeof(TileRef)));
            cap_map = c;
        }
        return SWEHR_OK;
    }

    // Upload arguments, work map and control pointers of the current members.
    int rebuild()
    {
        drop_graph();
        const size_t n = members.size();
        size_t nmap = 0, nmap1 = 0;
        for (swehr_handle *h : members) {
            nmap += (size_t)h->nblocks;
            nmap1 += (size_t)k1_tiles(h->nx, h->ny);
        }
        int rc = reserve(std::max<size_t>(n, 1), std::max<size_t>(nmap + nmap1, 1));
        if (rc) return rc;
        std::vector<unsigned char> args(n * arg_size());
        std::vector<TileRef> map;
        map.reserve(nmap + nmap1);
        std::vector<int> insts(n);
        std::vector<DevCtrl *> ctrls(n);
        lds = 0;
        max_classes = 0;
        for (size_t i = 0; i < n; i++) {
            swehr_handle *h = members[i];
            if (precision == SWEHR_FP64) memcpy(&args[i * arg_size()], &h->a64, arg_size());
            else memcpy(&args[i * arg_size()], &h->a32, arg_size());
            for (int t = 0; t < h->nblocks; t++) map.push_back(TileRef{(int)i, t});
            insts[i] = (int)i;
            ctrls[i] = h->ctrl;
            lds = std::max(lds, tile_lds_bytes(tile, precision, h->S));
            max_classes = std::max(max_classes, h->S);
        }
        for (size_t i = 0; i < n; i++)
            for (int t = 0; t < k1_tiles(members[i]->nx, members[i]->ny); t++) map.push_back(TileRef{(int)i, t});
        n_map = (int)nmap;
        per_xcd = (n_map + 7) / 8;
        n_map1 = (int)nmap1;
        per_xcd1 = (n_map1 + 7) / 8;
        if (n) {
            HIPCHK(hipMemcpyAsync(d_args, args.data(), args.size(), hipMemcpyHostToDevice, stream));
            HIPCHK(hipMemcpyAsync(d_map, map.data(), map.size() * sizeof(TileRef), hipMemcpyHostToDevice, stream));
            HIPCHK(hipMemcpyAsync(d_insts, insts.data(), n * sizeof(int), hipMemcpyHostToDevice, stream));
            HIPCHK(hipMemcpyAsync(d_ctrls, ctrls.data(), n * sizeof(DevCtrl *), hipMemcpyHostToDevice, stream));
            HIPCHK(hipStreamSynchronize(stream));
        }
        dirty = false;
        return SWEHR_OK;
    }

    // Eager launches of one ensemble step each, with events between them; adds the
    // three launch durations (ms) to sum[0..2].
    int profile_steps(int steps, double *sum)
    {
        const int n = (int)members.size();
        if (!n) return SWEHR_OK;
        hipEvent_t ev[4];
        for (int k = 0; k < 4; k++) HIPCHK(hipEventCreate(&ev[k]));
        const int blocks = 8 * per_xcd;
        int rc = SWEHR_OK;
        for (int s = 0; s < steps && rc == SWEHR_OK; s++) {
            (void)hipEventRecord(ev[0], stream);
            launch_batch_wave_strip(precision, 8 * per_xcd1, stream, d_args, d_map + n_map, n_map1, per_xcd1);
            (void)hipEventRecord(ev[1], stream);
            launch_batch_step(tile, precision, blocks, lds, stream, d_args, d_map, n_map, per_xcd, max_classes);
            (void)hipEventRecord(ev[2], stream);
            if (precision == SWEHR_FP64)
                hipLaunchKernelGGL((batch_control_kernel<double>), dim3(n), dim3(256), 0, stream,
                                   (const StepArgs<double> *)d_args, (const int *)d_insts, n);
            else
                hipLaunchKernelGGL((batch_control_kernel<float>), dim3(n), dim3(256), 0, stream,
                                   (const StepArgs<float> *)d_args, (const int *)d_insts, n);
            (void)hipEventRecord(ev[3], stream);
            if (hipEventSynchronize(ev[3]) != hipSuccess) {
                rc = SWEHR_ERR_HIP;
                break;
            }
            for (int k = 0; k < 3; k++) {
                float ms = 0;
                (void)hipEventElapsedTime(&ms, ev[k], ev[k + 1]);
                sum[k] += ms;
            }
        }
        for (int k = 0; k < 4; k++) (void)hipEventDestroy(ev[k]);
        return rc;
    }

    int enqueue_steps(int steps)
    {
        const int n = (int)members.size();
        if (!n) return SWEHR_OK;
        int ti, tj;
        tile_shape(tile, &ti, &tj);
        const int blocks = 8 * per_xcd;
        for (int s = 0; s < steps; s++) {
            launch_batch_wave_strip(precision, 8 * per_xcd1, stream, d_args, d_map + n_map, n_map1, per_xcd1);
            launch_batch_step(tile, precision, blocks, lds, stream, d_args, d_map, n_map, per_xcd, max_classes);
            if (precision == SWEHR_FP64)
                hipLaunchKernelGGL((batch_control_kernel<double>), dim3(n), dim3(256), 0, stream,
                                   (const StepArgs<double> *)d_args, (const int *)d_insts, n);
            else
                hipLaunchKernelGGL((batch_control_kernel<float>), dim3(n), dim3(256), 0, stream,
                                   (const StepArgs<float> *)d_args, (const int *)d_insts, n);
        }
        hipLaunchKernelGGL(gather_ctrl_kernel, dim3((n + 127) / 128), dim3(128), 0, stream, (DevCtrl *const *)d_ctrls,
                           n, d_gather);
        HIPCHK(hipGetLastError());
        return SWEHR_OK;
    }

    int capture()
    {
        drop_graph();
        HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_steps(graph_steps);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(stream, &g);
        if (rc) return rc;
        if (e != hipSuccess) {
            swehr_set_error("hipStreamEndCapture: %s%s", hipGetErrorString(e), "");
            return SWEHR_ERR_HIP;
        }
        graph = g;
        HIPCHK(hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0));
        return SWEHR_OK;
    }

    // Set every member's step limit to its current step + add (add < 0: no limit).
    int set_limits(long long add)
    {
        const int n = (int)members.size();
        if (!n) return SWEHR_OK;
        hipLaunchKernelGGL(set_limits_kernel, dim3((n + 127) / 128), dim3(128), 0, stream, (DevCtrl *const *)d_ctrls,
                           n, add);
        HIPCHK(hipGetLastError());
        return SWEHR_OK;
    }

    // One graph launch (graph_steps steps) and the gathered control blocks.
    int run_graph()
    {
        if (dirty) {
            const int rc = rebuild();
            if (rc) return rc;
        }
        if (members.empty()) return SWEHR_OK;
        if (!gexec) {
            const int rc = capture();
            if (rc) return rc;
        }
        HIPCHK(hipGraphLaunch(gexec, stream));
        HIPCHK(hipMemcpyAsync(h_gather, d_gather, members.size() * sizeof(DevCtrl), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        return SWEHR_OK;
    }

    // Persistent mode: graph_steps steps in one cooperative launch instead of a graph
    // of three launches per step, then the same gather of the control blocks.
    bool persistent = false;

    int run_persistent()
    {
        if (dirty) {
            const int rc = rebuild();
            if (rc) return rc;
        }
        const int n = (int)members.size();
        if (!n) return SWEHR_OK;
        const bool sp = batch_uses_split(precision, tile, max_classes, lds);
        const void *k = precision == SWEHR_FP64
                            ? (sp ? persistent_kernel_t<double, true>(tile) : persistent_kernel_t<double, false>(tile))
                            : (sp ? persistent_kernel_t<float, true>(tile) : persistent_kernel_t<float, false>(tile));
        if (!k) {
            swehr_set_error("the persistent mode is built for the 256-thread tile shapes (1-3)%s%s", "", "");
            return SWEHR_ERR_BAD_INPUT;
        }
        int dev = 0, coop = 0, cus = 0, per_cu = 0;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
        HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 256, lds));
        const int grid = std::min(per_cu * cus, 8 * per_xcd) / 8 * 8;
        if (!coop || grid < 8) {
            swehr_set_error("no cooperative launch of the ensemble on this device%s%s", "", "");
            return SWEHR_ERR_HIP;
        }
        const void *a = d_args;
        const TileRef *m = d_map;
        const int *ins = d_insts;
        int nm = n_map, px = per_xcd, ni = n, st = graph_steps;
        void *kargs[] = {(void *)&a, (void *)&m, (void *)&nm, (void *)&px, (void *)&ins, (void *)&ni, (void *)&st};
        HIPCHK(hipLaunchCooperativeKernel(k, dim3(grid), dim3(256), kargs, (unsigned)lds, stream));
        hipLaunchKernelGGL(gather_ctrl_kernel, dim3((n + 127) / 128), dim3(128), 0, stream, (DevCtrl *const *)d_ctrls,
                           n, d_gather);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(h_gather, d_gather, members.size() * sizeof(DevCtrl), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        return SWEHR_OK;
    }

    int run() { return persistent ? run_persistent() : run_graph(); }

    bool member_idle(size_t i) const
    {
        const DevCtrl &c = h_gather[i];
        return c.done || c.print_due || c.steps >= c.step_limit;
    }
};

// A member instance: created like a single one, then moved onto the engine's stream
// (its own would only hold the create uploads; every later operation on a member, the
// ensemble launches, downloads, snapshots, runs in order on one stream, and an ensemble
// of 256 members keeps 1 stream instead of 257).
int make_instance(const swehr_params *p, const double *topo, const double *ks, const double *ds, const double *pf,
                  const double *rain, int precision, int tile, hipStream_t engine_stream, swehr_handle **out)
{
    swehr_handle *h = nullptr;
    int rc = swehr_create(p, topo, ks, ds, pf, rain, precision, &h);
    if (rc) return rc;
    rc = swehr_configure(h, 0, 0, tile);
    if (rc) {
        swehr_destroy(h);
        return rc;
    }
    if (hipStreamSynchronize(h->stream) != hipSuccess || hipStreamDestroy(h->stream) != hipSuccess) {
        swehr_set_error("member stream: %s%s", hipGetErrorString(hipGetLastError()), "");
        h->stream = nullptr;
        swehr_destroy(h);
        return SWEHR_ERR_HIP;
    }
    h->stream = engine_stream;
    h->own_stream = false;
    h->member = true;
    *out = h;
    return SWEHR_OK;
}

int pick_common_tile(int precision, int maxS, int requested)
{
    if (requested) return tile_lds_bytes(requested, precision, maxS) <= kLdsLimit ? requested : 0;
    const int order[] = {1, 3, 2, 4};
    for (int id : order)
        if (tile_lds_bytes(id, precision, maxS) <= kLdsLimit) return id;
    return 0;
}

}  // namespace

// ------------------------------------------------------------------ ensembles

struct swehr_batch {
    Engine eng;
    std::vector<swehr_handle *> all;
};

extern "C" int swehr_batch_create(int n, const swehr_params *params, const double *const *topo,
                                  const double *const *ks, const double *const *ds, const double *const *pfrac,
                                  const double *const *rain, int precision, int tile, swehr_batch **out)
{
    if (!out || n < 1 || !params || !topo || !ds || !pfrac || !rain) return SWEHR_ERR_BAD_INPUT;
    *out = nullptr;
    int maxS = 1;
    for (int i = 0; i < n; i++) maxS = std::max(maxS, params[i].nclass);
    const int t = pick_common_tile(precision, maxS, tile);
    if (!t) {
        swehr_set_error("no tile shape fits the ensemble%s%s", "", "");
        return SWEHR_ERR_BAD_INPUT;
    }
    swehr_batch *b = new (std::nothrow) swehr_batch();
    if (!b) return SWEHR_ERR_OOM;
    b->eng.precision = precision;
    b->eng.tile = t;
    int rc = b->eng.init();
    for (int i = 0; i < n && rc == SWEHR_OK; i++) {
        swehr_handle *h = nullptr;
        rc = make_instance(&params[i], topo[i], ks ? ks[i] : nullptr, ds[i], pfrac[i], rain[i], precision, t,
                           b->eng.stream, &h);
        if (rc == SWEHR_OK) b->all.push_back(h);
    }
    if (rc) {
        swehr_batch_destroy(b);
        return rc;
    }
    b->eng.members = b->all;
    b->eng.dirty = true;
    *out = b;
    return SWEHR_OK;
}

extern "C" int swehr_batch_advance(swehr_batch *b, int64_t max_steps, int *remaining)
{
    if (!b || max_steps < 0) return SWEHR_ERR_BAD_INPUT;
    Engine &E = b->eng;
    // drop finished members (their tiles would only launch empty blocks)
    if (E.dirty) {
        const int rc = E.rebuild();
        if (rc) return rc;
    }
    int rc = E.set_limits((long long)max_steps);
    if (rc) return rc;
    while (!E.members.empty() && max_steps > 0) {
        rc = E.run();
        if (rc) return rc;
        std::vector<swehr_handle *> keep;
        bool any_running = false;
        for (size_t i = 0; i < E.members.size(); i++) {
            const DevCtrl &c = E.h_gather[i];
            if (!c.done) keep.push_back(E.members[i]);
            if (!E.member_idle(i)) any_running = true;
        }
        if (keep.size() != E.members.size()) {
            E.members = keep;
            E.dirty = true;
        }
        if (!any_running) break;
    }
    int left = 0;
    for (swehr_handle *h : b->all) {
        swehr_step_info info;
        rc = swehr_advance(h, 0, &info);
        if (rc != SWEHR_OK && rc != SWEHR_ERR_CFL) return rc;
        if (!info.done) left++;
    }
    if (remaining) *remaining = left;
    return SWEHR_OK;
}

extern "C" int swehr_batch_handle(swehr_batch *b, int i, swehr_handle **h)
{
    if (!b || !h || i < 0 || i >= (int)b->all.size()) return SWEHR_ERR_BAD_INPUT;
    *h = b->all[i];
    return SWEHR_OK;
}

extern "C" int swehr_batch_autotune(swehr_batch *b, int nsteps, int *best, double *ms_per_tile)
{
    if (!b || nsteps < 1) return SWEHR_ERR_BAD_INPUT;
    Engine &E = b->eng;
    int maxS = 1;
    size_t total = 0;
    for (swehr_handle *h : b->all) {
        maxS = std::max(maxS, h->S);
        total += h->bytes;
    }
    // save every instance's state in one device buffer
    char *save = nullptr;
    if (hipMalloc((void **)&save, total) != hipSuccess) {
        (void)hipGetLastError();
        swehr_set_error("no device memory for the autotuner's state copy%s%s", "", "");
        return SWEHR_ERR_OOM;
    }
    auto copy_all = [&](bool to_save) -> int {
        size_t off = 0;
        for (swehr_handle *h : b->all) {
            HIPCHK(hipMemcpyAsync(to_save ? save + off : h->mem, to_save ? h->mem : save + off, h->bytes,
                                  hipMemcpyDeviceToDevice, E.stream));
            off += h->bytes;
        }
        HIPCHK(hipStreamSynchronize(E.stream));
        return SWEHR_OK;
    };
    int rc = copy_all(true);
    const int orig = E.tile;
    int best_id = orig;
    double best_ms = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = SWEHR_ERR_HIP;
    for (int id = 1; id <= tile_count() && rc == SWEHR_OK; id++) {
        if (ms_per_tile) ms_per_tile[id - 1] = 0;
        if (tile_lds_bytes(id, E.precision, maxS) > kLdsLimit) continue;
        for (swehr_handle *h : b->all)
            if ((rc = swehr_configure(h, -1, 0, id)) != SWEHR_OK) break;
        if (rc) break;
        E.tile = id;
        E.members.clear();
        for (swehr_handle *h : b->all) E.members.push_back(h);
        if ((rc = E.rebuild()) != SWEHR_OK) break;
        if ((rc = E.set_limits(-1)) != SWEHR_OK) break;
        if ((rc = E.enqueue_steps(1)) != SWEHR_OK) break;  // warm-up (code load)
        if (hipEventRecord(e0, E.stream) != hipSuccess) { rc = SWEHR_ERR_HIP; break; }
        if ((rc = E.enqueue_steps(nsteps)) != SWEHR_OK) break;
        if (hipEventRecord(e1, E.stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
            rc = SWEHR_ERR_HIP;
            break;
        }
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        const double per = ms / nsteps;
        if (ms_per_tile) ms_per_tile[id - 1] = per;
        if (best_ms == 0 || per < best_ms) {
            best_ms = per;
            best_id = id;
        }
        if ((rc = copy_all(false)) != SWEHR_OK) break;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(save);
    const int use = rc ? orig : best_id;
    for (swehr_handle *h : b->all) swehr_configure(h, -1, 0, use);
    E.tile = use;
    E.members.clear();
    for (swehr_handle *h : b->all) {
        swehr_step_info info;
        if (swehr_advance(h, 0, &info) == SWEHR_OK && !info.done) E.members.push_back(h);
    }
    E.dirty = true;
    if (best) *best = use;
    return rc;
}

extern "C" int swehr_batch_profile(swehr_batch *b, int nsteps, double *avg_ms)
{
    if (!b || nsteps < 1 || !avg_ms) return SWEHR_ERR_BAD_INPUT;
    Engine &E = b->eng;
    avg_ms[0] = avg_ms[1] = avg_ms[2] = 0;
    if (E.members.empty()) return SWEHR_OK;
    int rc = SWEHR_OK;
    if (E.dirty && (rc = E.rebuild()) != SWEHR_OK) return rc;
    if ((rc = E.set_limits(-1)) != SWEHR_OK) return rc;
    double sum[3] = {0, 0, 0};
    if ((rc = E.profile_steps(nsteps, sum)) != SWEHR_OK) return rc;
    for (int k = 0; k < 3; k++) avg_ms[k] = sum[k] / nsteps;
    return SWEHR_OK;
}

extern "C" int swehr_batch_set_persistent(swehr_batch *b, int on)
{
    if (!b) return SWEHR_ERR_BAD_INPUT;
    if (on && (b->eng.tile < 1 || b->eng.tile > 3)) {
        swehr_set_error("the persistent mode is built for the 256-thread tile shapes (1-3)%s%s", "", "");
        return SWEHR_ERR_BAD_INPUT;
    }
    b->eng.persistent = on != 0;
    return SWEHR_OK;
}

extern "C" int swehr_batch_destroy(swehr_batch *b)
{
    if (!b) return SWEHR_OK;
    if (b->eng.stream) (void)hipStreamSynchronize(b->eng.stream);
    for (swehr_handle *h : b->all) swehr_destroy(h);
    b->eng.release();
    delete b;
    return SWEHR_OK;
}

// ------------------------------------------------------------------ serving

namespace {

struct Request {
    int64_t id;
    swehr_params p;
    std::vector<double> topo, ks, ds, pf, rain;
    int64_t bytes;
    void *snap = nullptr;  // pinned host checkpoint of an evicted request's device state
    int64_t snap_bytes = 0;
};

void free_snapshot(Request &r)
{
    if (r.snap) (void)hipHostFree(r.snap);
    r.snap = nullptr;
}

}  // namespace

struct swehr_server {
    Engine eng;
    int64_t mem_budget = 0;
    int64_t mem_used = 0;
    int max_slots = 0;
    int precision = SWEHR_FP64;
    int tile_req = 0;
    int64_t next_id = 1;
    std::deque<Request> queue;
    std::map<int64_t, swehr_handle *> running, finished;
    std::map<swehr_handle *, int64_t> ids;
    std::deque<std::pair<int64_t, swehr_step_info>> done_fifo;
    std::map<int64_t, Request> live;       // inputs of the running requests (for re-admission)
    std::map<int64_t, int64_t> slice;      // steps taken since (re)admission
    int64_t quantum = 0, evictions = 0;
    // cell-class-steps stepped so far over every request, and each running request's
    // step count when it was last added in (kept across an eviction: the restored
    // instance resumes at the count it was checkpointed with)
    int64_t work = 0;
    std::map<int64_t, long long> seen;
};

namespace {

// Checkpoint running request id into pinned host memory, free its instance and queue
// it at the back.
int evict_request(swehr_server *s, int64_t id)
{
    auto it = s->running.find(id);
    if (it == s->running.end()) return SWEHR_ERR_BAD_INPUT;
    swehr_handle *h = it->second;
    Engine &E = s->eng;
    if (E.stream) HIPCHK(hipStreamSynchronize(E.stream));
    int64_t bytes = 0;
    int rc = swehr_snapshot_bytes(h, &bytes);
    if (rc) return rc;
    void *snap = nullptr;
    if (hipHostMalloc(&snap, (size_t)bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        swehr_set_error("no pinned host memory for the evicted state%s%s", "", "");
        return SWEHR_ERR_OOM;
    }
    rc = swehr_snapshot(h, snap);
    if (rc) {
        (void)hipHostFree(snap);
        return rc;
    }
    for (size_t i = 0; i < E.members.size(); i++)
        if (E.members[i] == h) {
            E.members.erase(E.members.begin() + (long)i);
            break;
        }
    E.dirty = true;
    s->running.erase(it);
    s->ids.erase(h);
    s->slice.erase(id);
    s->mem_used -= (int64_t)h->bytes;
    swehr_destroy(h);
    Request r = std::move(s->live[id]);
    s->live.erase(id);
    r.snap = snap;
    r.snap_bytes = bytes;
    s->queue.push_back(std::move(r));
    s->evictions++;
    return SWEHR_OK;
}

bool head_fits(const swehr_server *s)
{
    return !s->queue.empty() && (int)s->running.size() < s->max_slots &&
           s->mem_used + s->queue.front().bytes <= s->mem_budget;
}

}  // namespace

extern "C" int swehr_server_create(int64_t mem_bytes, int max_slots, int precision, int tile, swehr_server **out)
{
    if (!out || mem_bytes < 0 || max_slots < 1 || (precision != SWEHR_FP64 && precision != SWEHR_FP32))
        return SWEHR_ERR_BAD_INPUT;
    if (mem_bytes == 0) {  // admit by the device's free memory: 15/16 of what is free now
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
            swehr_set_error("no device to size the server by%s%s", "", "");
            return SWEHR_ERR_NO_DEVICE;
        }
        mem_bytes = (int64_t)(free_b / 16 * 15);
    }
    swehr_server *s = new (std::nothrow) swehr_server();
    if (!s) return SWEHR_ERR_OOM;
    s->mem_budget = mem_bytes;
    s->max_slots = max_slots;
    s->precision = precision;
    s->tile_req = tile;
    s->eng.precision = precision;
    // plan the shared tile shape for up to 10 size classes; requests with more classes
    // are accepted when the chosen shape still fits them
    s->eng.tile = pick_common_tile(precision, 10, tile);
    if (!s->eng.tile) s->eng.tile = pick_common_tile(precision, 1, tile);
    if (!s->eng.tile) {
        delete s;
        swehr_set_error("tile shape does not fit%s%s", "", "");
        return SWEHR_ERR_BAD_INPUT;
    }
    int rc = s->eng.init();
    if (rc) {
        delete s;
        return rc;
    }
    *out = s;
    return SWEHR_OK;
}

extern "C" int swehr_server_submit(swehr_server *s, const swehr_params *p, const double *topo, const double *ks,
                                   const double *ds, const double *pfrac, const double *rain, int64_t *id)
{
    if (!s || !p || !topo || !ds || !pfrac || !rain || !id) return SWEHR_ERR_BAD_INPUT;
    if (p->nx < 1 || p->ny < 1 || p->nclass < 1 || p->nclass > kMaxClasses || p->rnum < 1) return SWEHR_ERR_BAD_INPUT;
    if (tile_lds_bytes(s->eng.tile, s->precision, p->nclass) > kLdsLimit) {
        swehr_set_error("request's class count does not fit the server's tile shape%s%s", "", "");
        return SWEHR_ERR_BAD_INPUT;
    }
    const int64_t need = swehr_instance_bytes_for(p, s->precision, ks != nullptr);
    if (need < 0 || need > s->mem_budget) {
        swehr_set_error("request needs more device memory than the server's budget%s%s", "", "");
        return SWEHR_ERR_BUSY;
    }
    Request r;
    r.id = s->next_id++;
    r.p = *p;
    const size_t n = (size_t)p->nx * p->ny;
    r.topo.assign(topo, topo + n);
    if (ks) r.ks.assign(ks, ks + n);
    r.ds.assign(ds, ds + p->nclass);
    r.pf.assign(pfrac, pfrac + p->nclass);
    r.rain.assign(rain, rain + p->rnum);
    r.bytes = need;
    *id = r.id;
    s->queue.push_back(std::move(r));
    return SWEHR_OK;
}

extern "C" int swehr_server_pump(swehr_server *s, int steps, int *running, int *queued, int *finished)
{
    if (!s || steps < 0) return SWEHR_ERR_BAD_INPUT;
    Engine &E = s->eng;
    // time slicing: make room for the queue head by evicting requests that have had
    // their quantum, the longest-running first
    while (s->quantum > 0 && !s->queue.empty() && !head_fits(s)) {
        int64_t victim = -1, most = -1;
        for (const auto &kv : s->running) {
            const int64_t taken = s->slice[kv.first];
            if (taken >= s->quantum && taken > most) {
                most = taken;
                victim = kv.first;
            }
        }
        if (victim < 0) break;
        const int rc = evict_request(s, victim);
        if (rc) return rc;
    }
    // admission: FIFO, while a slot is free and the memory budget holds the request;
    // an evicted request is restored from its checkpoint
    while (head_fits(s)) {
        Request &r = s->queue.front();
        swehr_handle *h = nullptr;
        int rc = make_instance(&r.p, r.topo.data(), r.ks.empty() ? nullptr : r.ks.data(), r.ds.data(), r.pf.data(),
                               r.rain.data(), s->precision, E.tile, E.stream, &h);
        if (rc == SWEHR_ERR_OOM) break;  // the device is fuller than the budget says: retry later
        if (rc) return rc;
        if (r.snap) {
            int64_t bytes = 0;
            rc = swehr_snapshot_bytes(h, &bytes);
            if (!rc && bytes != r.snap_bytes) {
                swehr_set_error("re-admitted request's state size differs from its checkpoint%s%s", "", "");
                rc = SWEHR_ERR_BAD_INPUT;
            }
            if (!rc) rc = swehr_restore(h, r.snap);
            if (rc) {
                swehr_destroy(h);
                return rc;
            }
            free_snapshot(r);
        }
        s->mem_used += (int64_t)h->bytes;
        s->running[r.id] = h;
        s->ids[h] = r.id;
        s->slice[r.id] = 0;
        E.members.push_back(h);
        E.dirty = true;
        const int64_t id = r.id;
        s->live[id] = std::move(r);
        s->queue.pop_front();
    }
    if (steps > 0 && !E.members.empty()) {
        if (E.dirty) {
            int rc = E.rebuild();
            if (rc) return rc;
            rc = E.set_limits(-1);
            if (rc) return rc;
        }
        int taken = 0;
        while (taken < steps && !E.members.empty()) {
            int rc = E.run_graph();
            if (rc) return rc;
            taken += E.graph_steps;
            for (swehr_handle *h : E.members) s->slice[s->ids[h]] += E.graph_steps;
            std::vector<swehr_handle *> keep;
            for (size_t i = 0; i < E.members.size(); i++) {
                swehr_handle *h = E.members[i];
                const DevCtrl &c = E.h_gather[i];
                long long &seen = s->seen[s->ids[h]];
                s->work += (int64_t)h->nx * h->ny * h->S * (c.steps - seen);
                seen = c.steps;
                if (c.done) {
                    const int64_t id = s->ids[h];
                    swehr_step_info info;
                    info.t = c.t;
                    info.dt = c.dt;
                    info.courant = c.courant;
                    info.rain = c.rain;
                    info.t_prev = c.t_prev;
                    info.steps = c.steps;
                    info.done = c.done;
                    info.print_due = c.print_due;
                    info.stage_count = c.stagecnt;
                    info.status = c.status;
                    s->running.erase(id);
                    s->live.erase(id);
                    s->slice.erase(id);
                    s->seen.erase(id);
                    s->finished[id] = h;
                    s->done_fifo.push_back({id, info});
                } else {
                    keep.push_back(h);
                }
            }
            if (keep.size() != E.members.size()) {
                E.members = keep;
                E.dirty = true;
                break;  // let the caller admit new requests into the freed slots
            }
        }
    }
    if (running) *running = (int)s->running.size();
    if (queued) *queued = (int)s->queue.size();
    if (finished) *finished = (int)s->finished.size();
    return SWEHR_OK;
}

extern "C" int swehr_server_work(swehr_server *s, int64_t *cell_class_steps)
{
    if (!s || !cell_class_steps) return SWEHR_ERR_BAD_INPUT;
    *cell_class_steps = s->work;
    return SWEHR_OK;
}

extern "C" int swehr_server_next_done(swehr_server *s, int64_t *id, swehr_step_info *info)
{
    if (!s || !id) return SWEHR_ERR_BAD_INPUT;
    if (s->done_fifo.empty()) return SWEHR_ERR_BUSY;
    *id = s->done_fifo.front().first;
    if (info) *info = s->done_fifo.front().second;
    s->done_fifo.pop_front();
    return SWEHR_OK;
}

extern "C" int swehr_server_request(swehr_server *s, int64_t id, swehr_handle **h)
{
    if (!s || !h) return SWEHR_ERR_BAD_INPUT;
    auto it = s->finished.find(id);
    if (it == s->finished.end()) return SWEHR_ERR_BAD_INPUT;
    *h = it->second;
    return SWEHR_OK;
}

extern "C" int swehr_server_release(swehr_server *s, int64_t id)
{
    if (!s) return SWEHR_ERR_BAD_INPUT;
    auto it = s->finished.find(id);
    if (it == s->finished.end()) return SWEHR_ERR_BAD_INPUT;
    s->mem_used -= (int64_t)it->second->bytes;
    s->ids.erase(it->second);
    swehr_destroy(it->second);
    s->finished.erase(it);
    return SWEHR_OK;
}

extern "C" int swehr_server_stats(swehr_server *s, int64_t *mem_used, int64_t *mem_budget)
{
    if (!s) return SWEHR_ERR_BAD_INPUT;
    if (mem_used) *mem_used = s->mem_used;
    if (mem_budget) *mem_budget = s->mem_budget;
    return SWEHR_OK;
}

extern "C" int swehr_server_evict(swehr_server *s, int64_t id)
{
    if (!s) return SWEHR_ERR_BAD_INPUT;
    return evict_request(s, id);
}

extern "C" int swehr_server_set_quantum(swehr_server *s, int64_t quantum)
{
    if (!s || quantum < 0) return SWEHR_ERR_BAD_INPUT;
    s->quantum = quantum;
    return SWEHR_OK;
}

extern "C" int swehr_server_evictions(swehr_server *s, int64_t *evictions)
{
    if (!s || !evictions) return SWEHR_ERR_BAD_INPUT;
    *evictions = s->evictions;
    return SWEHR_OK;
}

extern "C" int swehr_server_destroy(swehr_server *s)
{
    if (!s) return SWEHR_OK;
    if (s->eng.stream) (void)hipStreamSynchronize(s->eng.stream);
    for (Request &r : s->queue) free_snapshot(r);
    for (auto &kv : s->running) swehr_destroy(kv.second);
    for (auto &kv : s->finished) swehr_destroy(kv.second);
    s->eng.release();
    delete s;
    return SWEHR_OK;
}
