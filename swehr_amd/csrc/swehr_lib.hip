// swehr_lib.hip - the C ABI of include/swehr.h for one solver instance on one GPU.
//
// Host side of the drop-in seam: allocation of the instance's planes in one block of
// HBM, upload of the initial condition (SWEHR.c:211-325, 656-803), graph-captured
// time stepping with all control state on the device, downloads of fields and of
// the outlet series, and the global water/sediment budgets.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "swehr.h"
#include "swehr_device.h"
#include "swehr_lib_internal.h"

using namespace swehr;

static thread_local char g_err[1024];

extern "C" const char *swehr_last_error(void) { return g_err; }

void swehr_set_error(const char *fmt, const char *a, const char *b)
{
    snprintf(g_err, sizeof g_err, fmt, a, b);
}

#define HIPCHK(x)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) {                                                                \
            swehr_set_error("%s: %s", #x, hipGetErrorString(e_));                               \
            return SWEHR_ERR_HIP;                                                              \
        }                                                                                      \
    } while (0)

// ------------------------------------------------------------------ tile shapes

namespace {

struct TileShape {
    int ti, tj;
};
constexpr TileShape kTiles[] = {{8, 32}, {4, 64}, {16, 16}, {4, 32}};
constexpr int kNumTiles = sizeof(kTiles) / sizeof(kTiles[0]);

// Waves per SIMD that the LDS of a block of `threads` admits: 160 KB per CU shared by
// its 4 SIMDs, with 1 KB of static LDS per block (the step kernel's tables) beside the
// dynamic size.
inline int lds_waves_per_simd(size_t lds, int threads)
{
    return (int)((163840 / (lds + 1024)) * (size_t)(threads / 64) / 4);
}

template <typename R>
size_t tile_lds(int id, int S)
{
    switch (id) {
    case 1: return TileGeom<R, 8, 32>::lds_bytes(S);
    case 2: return TileGeom<R, 4, 64>::lds_bytes(S);
    case 3: return TileGeom<R, 16, 16>::lds_bytes(S);
    default: return TileGeom<R, 4, 32>::lds_bytes(S);
    }
}

template <typename R, bool SPLIT>
void launch_step_v(int id, int blocks, size_t lds, hipStream_t s, const StepArgs<R> &A)
{
    switch (id) {
    case 1: hipLaunchKernelGGL((step_kernel<R, 8, 32, SPLIT>), dim3(blocks), dim3(256), lds, s, A); break;
    case 2: hipLaunchKernelGGL((step_kernel<R, 4, 64, SPLIT>), dim3(blocks), dim3(256), lds, s, A); break;
    case 3: hipLaunchKernelGGL((step_kernel<R, 16, 16, SPLIT>), dim3(blocks), dim3(256), lds, s, A); break;
    default: hipLaunchKernelGGL((step_kernel<R, 4, 32, SPLIT>), dim3(blocks), dim3(128), lds, s, A); break;
    }
}

// The variant for many classes (SPLIT) where the default one does not serve them
// well: in FP64 where there are more classes than LDS planes for them (two chunks),
// in FP32 where the tile's LDS admits fewer waves per SIMD than the 5 the default
// variant is compiled for (the SPLIT variant's 4 then cost nothing).
template <typename R>
bool use_split(int id, int S, size_t lds)
{
    if constexpr (kLdsClasses<R> < kMaxClasses) return S > kLdsClasses<R>;
    else return lds_waves_per_simd(lds, kTiles[id - 1].ti * kTiles[id - 1].tj) < kStepWaves<R, false>;
}

// The prefetching variant (PF) where FP64 classes span two halo batches in one LDS chunk
// (4096^2, S = 10: step kernel 2.028 -> 1.995 ms; FP32 and fewer classes gain nothing).
// The 128-thread tile keeps the plain variant (PF spills there).  Its launcher is in
// swehr_step_pf.hip, built with a scheduling strategy of its own.
template <typename R>
bool use_prefetch(int id, int S, bool split)
{
    return sizeof(R) == sizeof(double) && !split && S > kHaloBatch<R> && id != 4;
}

template <typename R>
void launch_step(int id, int blocks, size_t lds, hipStream_t s, const StepArgs<R> &A)
{
    const bool split = use_split<R>(id, A.S, lds);
    if (split) {
        launch_step_v<R, true>(id, blocks, lds, s, A);
        return;
    }
    if constexpr (sizeof(R) == sizeof(double)) {
        if (use_prefetch<R>(id, A.S, split)) {
            launch_step_prefetch(id, blocks, lds, s, A);
            return;
        }
    }
    launch_step_v<R, false>(id, blocks, lds, s, A);
}

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

}  // namespace

namespace swehr {

int tile_count() { return kNumTiles; }

void tile_shape(int id, int *ti, int *tj)
{
    *ti = kTiles[id - 1].ti;
    *tj = kTiles[id - 1].tj;
}

size_t tile_lds_bytes(int id, int prec, int S)
{
    return prec == SWEHR_FP64 ? tile_lds<double>(id, S) : tile_lds<float>(id, S);
}

template <typename R, bool SP>
static void batch_step_t(int id, int blocks, size_t lds, hipStream_t s, const StepArgs<R> *a, const TileRef *map,
                         int n, int per)
{
    switch (id) {
    case 1: hipLaunchKernelGGL((batch_step_kernel<R, 8, 32, SP>), dim3(blocks), dim3(256), lds, s, a, map, n, per); break;
    case 2: hipLaunchKernelGGL((batch_step_kernel<R, 4, 64, SP>), dim3(blocks), dim3(256), lds, s, a, map, n, per); break;
    case 3: hipLaunchKernelGGL((batch_step_kernel<R, 16, 16, SP>), dim3(blocks), dim3(256), lds, s, a, map, n, per); break;
    default: hipLaunchKernelGGL((batch_step_kernel<R, 4, 32, SP>), dim3(blocks), dim3(128), lds, s, a, map, n, per); break;
    }
}

void launch_batch_wave_strip(int prec, int blocks, hipStream_t s, const void *args, const TileRef *map, int n,
                             int per_xcd)
{
    if (prec == SWEHR_FP64)
        hipLaunchKernelGGL((batch_wave_strip_kernel<double>), dim3(blocks), dim3(256), 0, s,
                           (const StepArgs<double> *)args, map, n, per_xcd);
    else
        hipLaunchKernelGGL((batch_wave_strip_kernel<float>), dim3(blocks), dim3(256), 0, s,
                           (const StepArgs<float> *)args, map, n, per_xcd);
}

bool batch_uses_split(int prec, int id, int max_classes, size_t lds)
{
    return prec == SWEHR_FP64 ? use_split<double>(id, max_classes, lds) : use_split<float>(id, max_classes, lds);
}

void launch_batch_step(int id, int prec, int blocks, size_t lds, hipStream_t s, const void *args, const TileRef *map,
                       int n, int per_xcd, int max_classes)
{
    if (prec == SWEHR_FP64) {
        const StepArgs<double> *a = (const StepArgs<double> *)args;
        if (use_split<double>(id, max_classes, lds)) batch_step_t<double, true>(id, blocks, lds, s, a, map, n, per_xcd);
        else batch_step_t<double, false>(id, blocks, lds, s, a, map, n, per_xcd);
    } else {
        const StepArgs<float> *a = (const StepArgs<float> *)args;
        if (use_split<float>(id, max_classes, lds)) batch_step_t<float, true>(id, blocks, lds, s, a, map, n, per_xcd);
        else batch_step_t<float, false>(id, blocks, lds, s, a, map, n, per_xcd);
    }
}

}  // namespace swehr

extern "C" int swehr_tile_count(void) { return kNumTiles; }

extern "C" int swehr_tile_shape(int id, int *ti, int *tj)
{
    if (id < 1 || id > kNumTiles) return SWEHR_ERR_BAD_INPUT;
    if (ti) *ti = kTiles[id - 1].ti;
    if (tj) *tj = kTiles[id - 1].tj;
    return SWEHR_OK;
}

// ------------------------------------------------------------------ kernels used by the host API

namespace swehr {

__global__ void fill_f64_kernel(double *p, double v, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

// One field as a dense i-major FP64 raster (SWEHR.c:1718-1747 for the output files).
template <typename R>
__global__ void derive_kernel(StepArgs<R> A, int field, int k, double *out)
{
    const long long steps = A.ctrl->steps;
    const int p = (int)(steps & 1), p3 = (int)(steps % 3);
    const long long n = (long long)A.nx * A.ny;
    for (long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x; cell < n;
         cell += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(cell / A.ny), j = (int)(cell % A.ny);
        const long long c = (long long)i * A.py + j;
        double v = 0, h, v1, v2, ts;
        switch (field) {
        case SWEHR_F_DEPTH: v = (double)A.h[p][c]; break;
        case SWEHR_F_UH: v = (double)A.hu[p][c]; break;
        case SWEHR_F_VH: v = (double)A.hv[p][c]; break;
        case SWEHR_F_CH: v = (double)A.ch[p][(long long)(k - 1) * A.plane + c]; break;
        case SWEHR_F_M: v = (double)A.M[(long long)(k - 1) * A.plane + c]; break;
        case SWEHR_F_TOPO: v = A.z[p3][c]; break;
        case SWEHR_F_VINF: v = A.vinf[c]; break;
        case SWEHR_F_TOTM: v = (double)A.totm[c]; break;
        case SWEHR_F_DEP: v = (double)A.dep[p][c]; break;
        case SWEHR_F_TOTSED: derived(A, p, c, h, v1, v2, ts); v = ts; break;
        case SWEHR_F_SPEED: derived(A, p, c, h, v1, v2, ts); v = sqrt(v1 * v1 + v2 * v2); break;
        case SWEHR_F_MAXVEL: v = (double)A.maxvel[c]; break;
        case SWEHR_F_MAXDEPTH: v = (double)A.maxdepth[c]; break;
        case SWEHR_F_SED: {
            const R hh = A.h[p][c];
            if (hh > R(kHfilm)) {
                R sd = A.ch[p][(long long)(k - 1) * A.plane + c] / hh;
                if (sd < R(-1e-5)) sd = R(0);
                v = (double)sd;
            }
            break;
        }
        case SWEHR_F_VELX: derived(A, p, c, h, v1, v2, ts); v = v1; break;
        case SWEHR_F_VELY: derived(A, p, c, h, v1, v2, ts); v = v2; break;
        default: break;
        }
        out[cell] = v;
    }
}

// Compensated partial sums per block: state totals and the per-tile budget slots.
// out[block * 10 + n]: n = 0 h, 1 vinf - vinf0, 2 ch, 3 M, 4 z, 5..9 slot terms.
template <typename R>
__global__ __launch_bounds__(256) void budget_kernel(StepArgs<R> A, double *out)
{
    const long long steps = A.ctrl->steps;
    const int p = (int)(steps & 1), p3 = (int)(steps % 3);
    double s[10], c[10];
    for (int n = 0; n < 10; n++) s[n] = c[n] = 0;
    const long long ncell = (long long)A.nx * A.ny;
    for (long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x; cell < ncell;
         cell += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(cell / A.ny), j = (int)(cell % A.ny);
        const long long q = (long long)i * A.py + j;
        neumaier(s[0], c[0], (double)A.h[p][q]);
        neumaier(s[1], c[1], A.vinf[q] - A.vinf0);
        for (int k = 0; k < A.S; k++) {
            neumaier(s[2], c[2], (double)A.ch[p][k * A.plane + q]);
            neumaier(s[3], c[3], (double)A.M[k * A.plane + q]);
        }
        neumaier(s[4], c[4], A.z[p3][q]);
    }
    for (long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x; b < A.bud_slots;
         b += (long long)gridDim.x * blockDim.x)
        for (int k = 0; k < NB_TERMS; k++) {
            // running sum, compensation, pending partial
            const long long g = b * NB_TERMS + k;
            neumaier(s[5 + k], c[5 + k], A.bud[2 * g]);
            neumaier(s[5 + k], c[5 + k], A.bud[2 * g + 1]);
            neumaier(s[5 + k], c[5 + k], A.bud[2 * (long long)A.bud_slots * NB_TERMS + g]);
        }
    __shared__ double red[10][256];
    for (int n = 0; n < 10; n++) red[n][threadIdx.x] = s[n] + c[n];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int n = 0; n < 10; n++) red[n][threadIdx.x] += red[n][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int n = 0; n < 10; n++) out[blockIdx.x * 10 + n] = red[n][0];
}

}  // namespace swehr

// ------------------------------------------------------------------ the handle

namespace {

struct Layout {
    size_t total = 0;
    size_t add(size_t bytes)
    {
        const size_t off = total;
        total += align_up(bytes, 256);
        return off;
    }
};

int pick_pitch(int ny) { return (int)align_up((size_t)ny, 64); }

int choose_tile(int precision, int S, int requested)
{
    auto fits = [&](int id) {
        return (precision == SWEHR_FP64 ? tile_lds<double>(id, S) : tile_lds<float>(id, S)) <= kLdsLimit;
    };
    if (requested >= 1 && requested <= kNumTiles) return fits(requested) ? requested : 0;
    const int order[] = {1, 3, 2, 4};
    for (int id : order)
        if (fits(id)) return id;
    return 0;
}

int stage_capacity(double tend)
{
    double c = std::ceil(tend) + 3;
    if (!(c < kStageMax)) c = kStageMax;
    if (c < 4) c = 4;
    return (int)c;
}

// Plan the instance's device block; with base != null also fill the pointers.
template <typename R>
size_t plan(swehr_handle *h, char *base, bool ks_raster)
{
    Layout L;
    const size_t pl = sizeof(R) * (size_t)h->plane;
    const size_t pd = sizeof(double) * (size_t)h->plane;
    const int S = h->S;
    size_t o_h[2], o_hu[2], o_hv[2], o_ch[2], o_dep[2], o_z[3];
    for (int b = 0; b < 2; b++) {
        o_h[b] = L.add(pl);
        o_hu[b] = L.add(pl);
        o_hv[b] = L.add(pl);
        o_ch[b] = L.add(pl * S);
        o_dep[b] = L.add(pl);
    }
    for (int b = 0; b < 3; b++) o_z[b] = L.add(pd);
    const size_t o_M = L.add(pl * S), o_totm = L.add(pl), o_mv = L.add(pl), o_md = L.add(pl);
    const size_t o_vinf = L.add(pd);
    const size_t o_ks = ks_raster ? L.add(pd) : 0;
    const size_t o_scr = L.add(sizeof(double) * (size_t)h->nx * h->ny);
    const size_t o_rain = L.add(sizeof(double) * (size_t)h->p.rnum);
    const size_t o_bud = L.add(sizeof(double) * (size_t)h->bud_slots * NB_TERMS * 3);
    const size_t o_stage = L.add(sizeof(double) * (size_t)kStageCols * h->stage_cap);
    const size_t o_flow = L.add(sizeof(double) * (size_t)kFlowCols * h->stage_cap);
    const size_t o_outq = L.add(sizeof(double) * (size_t)h->nx), o_outs = L.add(sizeof(double) * (size_t)h->nx);
    const size_t o_disc = L.add(sizeof(double) * 2 * (size_t)h->stage_cap);
    const size_t o_part = L.add(sizeof(double) * 10 * (size_t)h->budget_blocks);
    const size_t o_ctrl = L.add(sizeof(DevCtrl));
    if (base) {
        StepArgs<R> &A = *reinterpret_cast<StepArgs<R> *>(h->precision == SWEHR_FP64 ? (void *)&h->a64 : (void *)&h->a32);
        for (int b = 0; b < 2; b++) {
            A.h[b] = (R *)(base + o_h[b]);
            A.hu[b] = (R *)(base + o_hu[b]);
            A.hv[b] = (R *)(base + o_hv[b]);
            A.ch[b] = (R *)(base + o_ch[b]);
            A.dep[b] = (R *)(base + o_dep[b]);
        }
        for (int b = 0; b < 3; b++) A.z[b] = (double *)(base + o_z[b]);
        A.M = (R *)(base + o_M);
        A.totm = (R *)(base + o_totm);
        A.maxvel = (R *)(base + o_mv);
        A.maxdepth = (R *)(base + o_md);
        A.vinf = (double *)(base + o_vinf);
        A.ks_raster = ks_raster ? (const double *)(base + o_ks) : nullptr;
        A.rain = (const double *)(base + o_rain);
        A.bud = (double *)(base + o_bud);
        A.stage = (double *)(base + o_stage);
        A.saveflow = (double *)(base + o_flow);
        A.outq = (double *)(base + o_outq);
        A.outs = (double *)(base + o_outs);
        A.discharge = (double *)(base + o_disc);
        A.ctrl = (DevCtrl *)(base + o_ctrl);
        h->scratch = (double *)(base + o_scr);
        h->partials = (double *)(base + o_part);
        h->ctrl = A.ctrl;
    }
    return L.total;
}

void geometry(swehr_handle *h, int tile)
{
    const TileShape ts = kTiles[tile - 1];
    h->tile = tile;
    h->tiles_i = (h->nx + ts.ti - 1) / ts.ti;
    h->tiles_j = (h->ny + ts.tj - 1) / ts.tj;
    h->nblocks = h->tiles_i * h->tiles_j;
    h->blocks_per_xcd = (h->nblocks + 7) / 8;
    const long long n = (long long)h->nx * h->ny;
    h->k1_blocks = 8 * ((k1_tiles(h->nx, h->ny) + 7) / 8);  // strips of the wave-speed kernel
    h->budget_blocks = (int)std::min<long long>((n + 255) / 256, 1024);
    h->lds = h->precision == SWEHR_FP64 ? tile_lds<double>(tile, h->S) : tile_lds<float>(tile, h->S);
    int slots = 0;
    for (int id = 1; id <= kNumTiles; id++) {
        const TileShape &t = kTiles[id - 1];
        const int n = ((h->nx + t.ti - 1) / t.ti) * ((h->ny + t.tj - 1) / t.tj) * (t.ti * t.tj / 64);
        slots = std::max(slots, n);  // one per wave
    }
    h->bud_slots = slots;
}

template <typename R>
void fill_args(swehr_handle *h, StepArgs<R> &A, const double *ds, const double *pf)
{
    const swehr_params &p = h->p;
    A.nx = h->nx;
    A.ny = h->ny;
    A.S = h->S;
    A.py = h->py;
    A.plane = h->plane;
    A.ks = p.ks;
    A.dx = p.dx;
    A.h0 = p.h0;
    A.mtstar0 = p.mtstar0;
    A.a = p.a;
    A.ad = p.ad;
    A.F = p.F;
    A.J = p.J;
    A.roughness = p.roughness;
    A.theta0 = p.theta0;
    A.thetas = p.thetas;
    A.hf = p.hf;
    A.cstable = p.cstable;
    A.inv_dx = 1.0 / p.dx;
    A.inv_6h0 = (R)1 / ((R)6 * (R)p.h0);
    A.maxstep = p.maxstep;
    A.tend = p.tend;
    A.rint = p.rint;
    A.vinf0 = p.vinf0;
    A.rnum = p.rnum;
    A.inflow = p.inflow;
    for (int k = 0; k < 4; k++) A.bndry[k] = p.bndry[k];
    A.tiles_i = h->tiles_i;
    A.tiles_j = h->tiles_j;
    A.nblocks = h->nblocks;
    A.blocks_per_xcd = h->blocks_per_xcd;
    A.stage_cap = h->stage_cap;
    A.bud_slots = h->bud_slots;
    A.coarse1 = A.coarse2 = 0;
    for (int k = 0; k < kMaxClasses; k++) {
        const double d = k < h->S ? ds[k] : 1.0;
        const double ak = 13.95 * kNu / d;
        A.ds[k] = (R)d;
        A.pf[k] = (R)(k < h->S ? pf[k] : 0.0);
        A.ak[k] = (R)ak;
        A.ak2[k] = (R)(ak * ak);
        if (d > 0.0001) A.coarse1 |= 1u << k;
        if (d > 0.002) A.coarse2 |= 1u << k;
        A.tds_gt[k] = thr_gt<R>(d);
    }
    A.t9h0_lt = thr_lt<R>(9 * p.h0);
    A.t9h0_gt = thr_gt<R>(9 * p.h0);
    A.th0_gt = thr_gt<R>(p.h0);
}

template <typename R>
int upload_raster(swehr_handle *h, R *dst, const double *src)
{
    if (sizeof(R) == sizeof(double)) {
        HIPCHK(hipMemcpy2DAsync(dst, sizeof(R) * h->py, src, sizeof(double) * h->ny, sizeof(double) * h->ny, h->nx,
                                hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    } else {
        std::vector<R> tmp((size_t)h->nx * h->ny);
        for (size_t n = 0; n < tmp.size(); n++) tmp[n] = (R)src[n];
        HIPCHK(hipMemcpy2DAsync(dst, sizeof(R) * h->py, tmp.data(), sizeof(R) * h->ny, sizeof(R) * h->ny, h->nx,
                                hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    return SWEHR_OK;
}

// The steps of one graph: the wave-speed pass and the step kernel per step; the
// wave-speed pass of a step also commits the previous step (its block 0), and the graph
// ends with the control kernel, which commits the last step, so that the control block
// the host reads after the graph is complete.
template <typename R>
int enqueue_steps(swehr_handle *h, const StepArgs<R> &A, int steps)
{
    for (int s = 0; s < steps; s++) {
        hipLaunchKernelGGL((wave_speed_kernel<R>), dim3(h->k1_blocks), dim3(256), 0, h->stream, A);
        launch_step<R>(h->tile, 8 * h->blocks_per_xcd, h->lds, h->stream, A);
    }
    hipLaunchKernelGGL((control_kernel<R>), dim3(1), dim3(256), 0, h->stream, A);
    HIPCHK(hipGetLastError());
    return SWEHR_OK;
}

int build_graph(swehr_handle *h)
{
    if (h->gexec) {
        (void)hipGraphExecDestroy(h->gexec);
        h->gexec = nullptr;
    }
    if (h->graph) {
        (void)hipGraphDestroy(h->graph);
        h->graph = nullptr;
    }
    HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    const int rc =
        h->precision == SWEHR_FP64 ? enqueue_steps(h, h->a64, h->graph_steps) : enqueue_steps(h, h->a32, h->graph_steps);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(h->stream, &g);
    if (rc != SWEHR_OK) return rc;
    if (e != hipSuccess) {
        swehr_set_error("hipStreamEndCapture: %s%s", hipGetErrorString(e), "");
        return SWEHR_ERR_HIP;
    }
    h->graph = g;
    HIPCHK(hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0));
    // move the executable graph to the device now, not at its first launch
    HIPCHK(hipGraphUpload(h->gexec, h->stream));
    return SWEHR_OK;
}

int check_params(const swehr_params *p)
{
    if (!p || p->nx < 1 || p->ny < 1 || p->nclass < 1 || p->nclass > kMaxClasses || p->rnum < 1 || !(p->dx > 0) ||
        !(p->rint > 0) || !std::isfinite(p->tend) || !(p->cstable > 0) || !(p->thetas != p->theta0)) {
        swehr_set_error("invalid parameters%s%s", "", "");
        return SWEHR_ERR_BAD_INPUT;
    }
    // the kernels address a plane with 32-bit byte offsets: at most 2^29 FP64 values
    if ((double)p->nx * (double)pick_pitch(p->ny) * 8.0 >= 4294967296.0) {
        swehr_set_error("grid %s too large for one instance (at most 2^29 cells per plane)%s", "", "");
        return SWEHR_ERR_BAD_INPUT;
    }
    return SWEHR_OK;
}

}  // namespace

extern "C" int64_t swehr_instance_bytes_for(const swehr_params *p, int precision, int ks_raster)
{
    if (!p || p->nx < 1 || p->ny < 1 || p->nclass < 1 || p->nclass > kMaxClasses || p->rnum < 1) return -1;
    swehr_handle h;
    h.p = *p;
    h.precision = precision;
    h.nx = p->nx;
    h.ny = p->ny;
    h.S = p->nclass;
    h.py = pick_pitch(p->ny);
    h.plane = (long long)h.nx * h.py;
    h.stage_cap = stage_capacity(p->tend);
    const int tile = choose_tile(precision, h.S, 0);
    if (tile == 0) return -1;
    geometry(&h, tile);
    return (int64_t)(precision == SWEHR_FP64 ? plan<double>(&h, nullptr, ks_raster != 0)
                                             : plan<float>(&h, nullptr, ks_raster != 0));
}

extern "C" int64_t swehr_instance_bytes(int nx, int ny, int nclass, int precision)
{
    swehr_handle h;
    swehr_params_init(&h.p);
    h.p.rnum = 4096;
    h.precision = precision;
    h.nx = nx;
    h.ny = ny;
    h.S = nclass;
    h.py = pick_pitch(ny);
    h.plane = (long long)nx * h.py;
    h.stage_cap = kStageMax;
    const int tile = choose_tile(precision, nclass, 0);
    if (tile == 0 || nx < 1 || ny < 1) return -1;
    geometry(&h, tile);
    return (int64_t)(precision == SWEHR_FP64 ? plan<double>(&h, nullptr, true) : plan<float>(&h, nullptr, true));
}

extern "C" int swehr_create(const swehr_params *p, const double *topo, const double *ks, const double *ds,
                            const double *pfrac, const double *rain, int precision, swehr_handle **out)
{
    if (!out) return SWEHR_ERR_BAD_INPUT;
    *out = nullptr;
    int rc = check_params(p);
    if (rc) return rc;
    if (!topo || !ds || !pfrac || !rain || (precision != SWEHR_FP64 && precision != SWEHR_FP32)) {
        swehr_set_error("missing input array or bad precision%s%s", "", "");
        return SWEHR_ERR_BAD_INPUT;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        swehr_set_error("no HIP device%s%s", "", "");
        return SWEHR_ERR_NO_DEVICE;
    }
    swehr_handle *h = new (std::nothrow) swehr_handle();
    if (!h) return SWEHR_ERR_OOM;
    h->p = *p;
    h->precision = precision;
    HIPCHK(hipGetDevice(&h->device));
    h->nx = p->nx;
    h->ny = p->ny;
    h->S = p->nclass;
    h->py = pick_pitch(p->ny);
    h->plane = (long long)h->nx * h->py;
    h->stage_cap = stage_capacity(p->tend);
    const int tile = choose_tile(precision, h->S, 0);
    if (!tile) {
        delete h;
        swehr_set_error("no tile shape fits the LDS for this class count%s%s", "", "");
        return SWEHR_ERR_BAD_INPUT;
    }
    geometry(h, tile);
    const bool ksr = ks != nullptr;
    h->bytes = precision == SWEHR_FP64 ? plan<double>(h, nullptr, ksr) : plan<float>(h, nullptr, ksr);
    if (hipMalloc(&h->mem, h->bytes) != hipSuccess) {
        (void)hipGetLastError();
        delete h;
        swehr_set_error("hipMalloc of the instance failed%s%s", "", "");
        return SWEHR_ERR_OOM;
    }
    if (precision == SWEHR_FP64) {
        plan<double>(h, (char *)h->mem, ksr);
        fill_args(h, h->a64, ds, pfrac);
    } else {
        plan<float>(h, (char *)h->mem, ksr);
        fill_args(h, h->a32, ds, pfrac);
    }
    rc = SWEHR_OK;
    do {
        if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) { rc = SWEHR_ERR_HIP; break; }
        if (hipHostMalloc((void **)&h->hctrl, sizeof(DevCtrl) + 64, 0) != hipSuccess) { rc = SWEHR_ERR_OOM; break; }
        h->hlimit = reinterpret_cast<long long *>(reinterpret_cast<char *>(h->hctrl) + sizeof(DevCtrl));
        h->hflag = reinterpret_cast<int *>(reinterpret_cast<char *>(h->hctrl) + sizeof(DevCtrl) + 16);
        if (hipMemsetAsync(h->mem, 0, h->bytes, h->stream) != hipSuccess) { rc = SWEHR_ERR_HIP; break; }
        // bed, with the lagged bed of step 1 equal to it (SWEHR.c:721-803)
        if (precision == SWEHR_FP64) {
            rc = upload_raster<double>(h, h->a64.z[0], topo);
            if (!rc) rc = upload_raster<double>(h, h->a64.z[2], topo);
            if (!rc && ksr) rc = upload_raster<double>(h, const_cast<double *>(h->a64.ks_raster), ks);
        } else {
            rc = upload_raster<double>(h, h->a32.z[0], topo);
            if (!rc) rc = upload_raster<double>(h, h->a32.z[2], topo);
            if (!rc && ksr) rc = upload_raster<double>(h, const_cast<double *>(h->a32.ks_raster), ks);
        }
        if (rc) break;
        double *vinf = precision == SWEHR_FP64 ? h->a64.vinf : h->a32.vinf;
        const double *rain_d = precision == SWEHR_FP64 ? h->a64.rain : h->a32.rain;
        hipLaunchKernelGGL(fill_f64_kernel, dim3(1024), dim3(256), 0, h->stream, vinf, p->vinf0, (long long)h->plane);
        if (hipMemcpyAsync((void *)rain_d, rain, sizeof(double) * p->rnum, hipMemcpyHostToDevice, h->stream) !=
            hipSuccess) { rc = SWEHR_ERR_HIP; break; }
        DevCtrl c;
        memset(&c, 0, sizeof c);
        c.t = 0;
        c.dt = 0.0001;
        c.stageinterval = 0;
        c.printinterval = std::fmax(1, p->tend / 1000);
        c.printint = std::fmax(1, p->tend / 1000);
        c.pause_on_print = 1;
        c.done = !(c.t < p->tend);
        *h->hctrl = c;
        if (hipMemcpyAsync(h->ctrl, h->hctrl, sizeof(DevCtrl), hipMemcpyHostToDevice, h->stream) != hipSuccess) {
            rc = SWEHR_ERR_HIP;
            break;
        }
        if (precision == SWEHR_FP64)
            hipLaunchKernelGGL(time_functions_kernel<double>, dim3(1), dim3(64), 0, h->stream, h->a64);
        else
            hipLaunchKernelGGL(time_functions_kernel<float>, dim3(1), dim3(64), 0, h->stream, h->a32);
        if (hipStreamSynchronize(h->stream) != hipSuccess) { rc = SWEHR_ERR_HIP; break; }
    } while (0);
    if (rc) {
        if (rc == SWEHR_ERR_HIP && !g_err[0]) swehr_set_error("HIP error during create: %s%s", hipGetErrorString(hipGetLastError()), "");
        swehr_destroy(h);
        return rc;
    }
    double s = 0, cc = 0;
    for (long long n = 0; n < (long long)h->nx * h->ny; n++) {
        const double v = topo[n];
        const double t = s + v;
        if (std::fabs(s) >= std::fabs(v)) cc += (s - t) + v;
        else cc += (v - t) + s;
        s = t;
    }
    h->z0sum = s + cc;
    *out = h;
    return SWEHR_OK;
}

extern "C" int swehr_configure(swehr_handle *h, int pause_on_print, int steps_per_graph, int tile)
{
    if (!h) return SWEHR_ERR_BAD_INPUT;
    HIPCHK(hipSetDevice(h->device));
    if (steps_per_graph > 0 && steps_per_graph != h->graph_steps) {
        h->graph_steps = steps_per_graph;
        if (h->gexec) {
            (void)hipGraphExecDestroy(h->gexec);
            h->gexec = nullptr;
        }
    }
    if (tile != 0 && tile != h->tile) {
        const int t = choose_tile(h->precision, h->S, tile);
        if (!t) {
            swehr_set_error("tile shape does not fit this instance%s%s", "", "");
            return SWEHR_ERR_BAD_INPUT;
        }
        // budget slots are allocated for the largest tile count of any shape; totals
        // are sums over all slots, so switching shapes keeps them
        HIPCHK(hipStreamSynchronize(h->stream));
        geometry(h, t);
        if (h->precision == SWEHR_FP64) {
            h->a64.tiles_i = h->tiles_i; h->a64.tiles_j = h->tiles_j;
            h->a64.nblocks = h->nblocks; h->a64.blocks_per_xcd = h->blocks_per_xcd;
        } else {
            h->a32.tiles_i = h->tiles_i; h->a32.tiles_j = h->tiles_j;
            h->a32.nblocks = h->nblocks; h->a32.blocks_per_xcd = h->blocks_per_xcd;
        }
        if (h->gexec) {
            (void)hipGraphExecDestroy(h->gexec);
            h->gexec = nullptr;
        }
    }
    if (pause_on_print == 0 || pause_on_print == 1) {
        *h->hflag = pause_on_print;
        HIPCHK(hipMemcpyAsync(&h->ctrl->pause_on_print, h->hflag, sizeof(int), hipMemcpyHostToDevice, h->stream));
    }
    // a changed graph length or tile shape is captured and instantiated here, so that
    // the next swehr_advance launches at once instead of first building the graph
    if (!h->gexec) {
        const int rc = build_graph(h);
        if (rc) return rc;
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    return SWEHR_OK;
}

static void fill_info(const DevCtrl &c, swehr_step_info *info)
{
    if (!info) return;
    info->t = c.t;
    info->dt = c.dt;
    info->courant = c.courant;
    info->rain = c.rain;
    info->t_prev = c.t_prev;
    info->steps = c.steps;
    info->done = c.done;
    info->print_due = c.print_due;
    info->stage_count = c.stagecnt;
    info->status = c.status;
}

extern "C" int swehr_advance(swehr_handle *h, int64_t max_steps, swehr_step_info *info)
{
    if (!h || max_steps < 0) return SWEHR_ERR_BAD_INPUT;
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipMemcpyAsync(h->hctrl, h->ctrl, sizeof(DevCtrl), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (h->hctrl->done || h->hctrl->print_due || max_steps == 0) {
        fill_info(*h->hctrl, info);
        return h->hctrl->status;
    }
    if (!h->gexec) {
        const int rc = build_graph(h);
        if (rc) return rc;
    }
    const long long target = h->hctrl->steps + (long long)max_steps;
    *h->hlimit = target;
    HIPCHK(hipMemcpyAsync(&h->ctrl->step_limit, h->hlimit, sizeof(long long), hipMemcpyHostToDevice, h->stream));
    for (;;) {
        const long long before = h->hctrl->steps;
        HIPCHK(hipGraphLaunch(h->gexec, h->stream));
        HIPCHK(hipMemcpyAsync(h->hctrl, h->ctrl, sizeof(DevCtrl), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        const DevCtrl &c = *h->hctrl;
        if (c.done || c.print_due || c.steps >= target) break;
        if (c.steps == before) {
            swehr_set_error("time stepping made no progress%s%s", "", "");
            return SWEHR_ERR_HIP;
        }
    }
    fill_info(*h->hctrl, info);
    return h->hctrl->status;
}

template <typename R>
static int profile_impl(swehr_handle *h, const StepArgs<R> &A, int nsteps, double *avg_ms, int *taken)
{
    hipEvent_t ev[4];
    for (int k = 0; k < 4; k++) HIPCHK(hipEventCreate(&ev[k]));
    double sum[3] = {0, 0, 0};
    int n = 0;
    int rc = SWEHR_OK;
    for (; n < nsteps; n++) {
        HIPCHK(hipMemcpyAsync(h->hctrl, h->ctrl, sizeof(DevCtrl), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        if (h->hctrl->done || h->hctrl->print_due) break;
        *h->hlimit = h->hctrl->steps + 1;
        HIPCHK(hipMemcpyAsync(&h->ctrl->step_limit, h->hlimit, sizeof(long long), hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipEventRecord(ev[0], h->stream));
        hipLaunchKernelGGL((wave_speed_kernel<R>), dim3(h->k1_blocks), dim3(256), 0, h->stream, A);
        HIPCHK(hipEventRecord(ev[1], h->stream));
        launch_step<R>(h->tile, 8 * h->blocks_per_xcd, h->lds, h->stream, A);
        HIPCHK(hipEventRecord(ev[2], h->stream));
        hipLaunchKernelGGL((control_kernel<R>), dim3(1), dim3(256), 0, h->stream, A);
        HIPCHK(hipEventRecord(ev[3], h->stream));
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventSynchronize(ev[3]));
        for (int k = 0; k < 3; k++) {
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
            sum[k] += ms;
        }
    }
    for (int k = 0; k < 4; k++) (void)hipEventDestroy(ev[k]);
    for (int k = 0; k < 3; k++) avg_ms[k] = n ? sum[k] / n : 0.0;
    if (taken) *taken = n;
    return rc;
}

extern "C" int swehr_profile(swehr_handle *h, int nsteps, double *avg_ms, int *taken)
{
    if (!h || !avg_ms || nsteps < 0) return SWEHR_ERR_BAD_INPUT;
    HIPCHK(hipSetDevice(h->device));
    return h->precision == SWEHR_FP64 ? profile_impl(h, h->a64, nsteps, avg_ms, taken)
                                      : profile_impl(h, h->a32, nsteps, avg_ms, taken);
}

// A snapshot is a header that records the instance's layout, then its whole device
// block (which holds no pointers, so it restores into any instance of that layout).
namespace {
// version 2: budget slots as running sums followed by the pending partials
constexpr uint64_t kSnapMagic = 0x3250414e53524857ull;  // "WHRSNAP2"
constexpr size_t kSnapHeader = 256;
struct SnapHeader {
    uint64_t magic;
    int64_t bytes;  // device block bytes that follow the header
    int32_t nx, ny, S, precision, ks_raster, rasters, stage_cap, bud_slots;
};
static_assert(sizeof(SnapHeader) <= kSnapHeader, "snapshot header");

SnapHeader snap_header(const swehr_handle *h)
{
    SnapHeader s;
    memset(&s, 0, sizeof s);
    s.magic = kSnapMagic;
    s.bytes = (int64_t)h->bytes;
    s.nx = h->nx;
    s.ny = h->ny;
    s.S = h->S;
    s.precision = h->precision;
    s.ks_raster = (h->precision == SWEHR_FP64 ? h->a64.ks_raster : h->a32.ks_raster) != nullptr;
    for (int k = 0; k < 3; k++)
        if (h->prast[k]) s.rasters |= 1 << k;
    s.stage_cap = h->stage_cap;
    s.bud_slots = h->bud_slots;
    return s;
}
}  // namespace

extern "C" int swehr_profile_window(swehr_handle *h, int nsteps, double *avg_ms, int *taken)
{
    if (!h || !avg_ms || nsteps < 0) return SWEHR_ERR_BAD_INPUT;
    HIPCHK(hipSetDevice(h->device));
    void *save = nullptr;
    if (hipMalloc(&save, h->bytes) != hipSuccess) {
        (void)hipGetLastError();
        swehr_set_error("no device memory for the profiler's state copy%s%s", "", "");
        return SWEHR_ERR_OOM;
    }
    int rc = SWEHR_OK;
    if (hipMemcpyAsync(save, h->mem, h->bytes, hipMemcpyDeviceToDevice, h->stream) != hipSuccess) rc = SWEHR_ERR_HIP;
    if (!rc) rc = swehr_profile(h, nsteps, avg_ms, taken);
    if (hipMemcpyAsync(h->mem, save, h->bytes, hipMemcpyDeviceToDevice, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
        rc = rc ? rc : SWEHR_ERR_HIP;
    (void)hipFree(save);
    if (rc == SWEHR_ERR_HIP && !g_err[0]) swehr_set_error("profile window: %s%s", hipGetErrorString(hipGetLastError()), "");
    return rc;
}

extern "C" int swehr_snapshot_bytes(swehr_handle *h, int64_t *bytes)
{
    if (!h || !bytes) return SWEHR_ERR_BAD_INPUT;
    *bytes = (int64_t)(kSnapHeader + h->bytes);
    return SWEHR_OK;
}

extern "C" int swehr_snapshot(swehr_handle *h, void *host)
{
    if (!h || !host) return SWEHR_ERR_BAD_INPUT;
    HIPCHK(hipSetDevice(h->device));
    const SnapHeader hd = snap_header(h);
    memset(host, 0, kSnapHeader);
    memcpy(host, &hd, sizeof hd);
    HIPCHK(hipMemcpyAsync((char *)host + kSnapHeader, h->mem, h->bytes, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return SWEHR_OK;
}

extern "C" int swehr_restore(swehr_handle *h, const void *host)
{
    if (!h || !host) return SWEHR_ERR_BAD_INPUT;
    SnapHeader got;
    memcpy(&got, host, sizeof got);
    const SnapHeader want = snap_header(h);
    if (got.magic != kSnapMagic) {
        swehr_set_error("restore: not a snapshot of this library%s%s", "", "");
        return SWEHR_ERR_BAD_INPUT;
    }
    if (got.bytes != want.bytes || got.nx != want.nx || got.ny != want.ny || got.S != want.S ||
        got.precision != want.precision || got.ks_raster != want.ks_raster || got.stage_cap != want.stage_cap ||
        got.bud_slots != want.bud_slots) {
        swehr_set_error("restore: the snapshot was taken from an instance of another layout "
                        "(grid, classes, precision, Ks raster or run length)%s%s", "", "");
        return SWEHR_ERR_BAD_INPUT;
    }
    if (got.rasters != want.rasters) {  // the constant parameter rasters are not in a snapshot
        swehr_set_error("restore: set the same parameter rasters (hf, dtheta, roughness) on the instance "
                        "before restoring%s%s", "", "");
        return SWEHR_ERR_BAD_INPUT;
    }
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipMemcpyAsync(h->mem, (const char *)host + kSnapHeader, h->bytes, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return SWEHR_OK;
}

extern "C" int swehr_autotune(swehr_handle *h, int nsteps, int *best, double *ms_per_tile)
{
    if (!h || nsteps < 1) return SWEHR_ERR_BAD_INPUT;
    HIPCHK(hipSetDevice(h->device));
    void *save = nullptr;
    if (hipMalloc(&save, h->bytes) != hipSuccess) {
        (void)hipGetLastError();
        swehr_set_error("no device memory for the autotuner's state copy%s%s", "", "");
        return SWEHR_ERR_OOM;
    }
    int rc = SWEHR_OK;
    int best_id = h->tile;
    double best_ms = 0;
    const int orig = h->tile;
    if (hipMemcpyAsync(save, h->mem, h->bytes, hipMemcpyDeviceToDevice, h->stream) != hipSuccess) rc = SWEHR_ERR_HIP;
    for (int id = 1; id <= kNumTiles && rc == SWEHR_OK; id++) {
        if (ms_per_tile) ms_per_tile[id - 1] = 0;
        if (!choose_tile(h->precision, h->S, id)) continue;
        rc = swehr_configure(h, -1, 0, id);
        if (rc) break;
        double ms[3];
        int taken = 0;
        rc = swehr_profile(h, nsteps, ms, &taken);
        if (rc) break;
        if (hipMemcpyAsync(h->mem, save, h->bytes, hipMemcpyDeviceToDevice, h->stream) != hipSuccess ||
            hipStreamSynchronize(h->stream) != hipSuccess) {
            rc = SWEHR_ERR_HIP;
            break;
        }
        if (taken == 0) continue;
        if (ms_per_tile) ms_per_tile[id - 1] = ms[1];
        if (best_ms == 0 || ms[1] < best_ms) {
            best_ms = ms[1];
            best_id = id;
        }
    }
    (void)hipFree(save);
    if (rc) {
        swehr_configure(h, -1, 0, orig);
        return rc;
    }
    rc = swehr_configure(h, -1, 0, best_id);
    if (best) *best = best_id;
    return rc;
}

extern "C" int swehr_ack_print(swehr_handle *h)
{
    if (!h) return SWEHR_ERR_BAD_INPUT;
    HIPCHK(hipSetDevice(h->device));
    *h->hflag = 0;
    HIPCHK(hipMemcpyAsync(&h->ctrl->print_due, h->hflag, sizeof(int), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return SWEHR_OK;
}

extern "C" int swehr_download(swehr_handle *h, int field, int k, double *host)
{
    if (!h || !host || field < 0 || field >= SWEHR_F_COUNT) return SWEHR_ERR_BAD_INPUT;
    if ((field == SWEHR_F_CH || field == SWEHR_F_M || field == SWEHR_F_SED) && (k < 1 || k > h->S))
        return SWEHR_ERR_BAD_INPUT;
    HIPCHK(hipSetDevice(h->device));
    const long long n = (long long)h->nx * h->ny;
    const int blocks = (int)std::min<long long>((n + 255) / 256, 4096);
    if (h->precision == SWEHR_FP64)
        hipLaunchKernelGGL(derive_kernel<double>, dim3(blocks), dim3(256), 0, h->stream, h->a64, field, k, h->scratch);
    else
        hipLaunchKernelGGL(derive_kernel<float>, dim3(blocks), dim3(256), 0, h->stream, h->a32, field, k, h->scratch);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(host, h->scratch, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return SWEHR_OK;
}

extern "C" int swehr_stage(swehr_handle *h, double *stage, double *saveflow, int capacity, int *count)
{
    if (!h || capacity < 0) return SWEHR_ERR_BAD_INPUT;
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipMemcpyAsync(h->hctrl, h->ctrl, sizeof(DevCtrl), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    const int cnt = h->hctrl->stagecnt;
    const int n = std::min(cnt, capacity);
    const double *ds = h->precision == SWEHR_FP64 ? h->a64.stage : h->a32.stage;
    const double *df = h->precision == SWEHR_FP64 ? h->a64.saveflow : h->a32.saveflow;
    if (n > 0) {
        if (stage)
            HIPCHK(hipMemcpy2DAsync(stage, sizeof(double) * capacity, ds, sizeof(double) * h->stage_cap,
                                    sizeof(double) * n, kStageCols, hipMemcpyDeviceToHost, h->stream));
        if (saveflow)
            HIPCHK(hipMemcpy2DAsync(saveflow, sizeof(double) * capacity, df, sizeof(double) * h->stage_cap,
                                    sizeof(double) * n, kFlowCols, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    if (count) *count = cnt;
    return SWEHR_OK;
}

extern "C" int swehr_discharge(swehr_handle *h, double *q, double *qs, int capacity, int *count)
{
    if (!h || capacity < 0) return SWEHR_ERR_BAD_INPUT;
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipMemcpyAsync(h->hctrl, h->ctrl, sizeof(DevCtrl), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    const int cnt = h->hctrl->stagecnt;
    const int n = std::min(cnt, capacity);
    const double *dd = h->precision == SWEHR_FP64 ? h->a64.discharge : h->a32.discharge;
    if (n > 0) {
        if (q) HIPCHK(hipMemcpyAsync(q, dd, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
        if (qs) HIPCHK(hipMemcpyAsync(qs, dd + h->stage_cap, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    if (count) *count = cnt;
    return SWEHR_OK;
}

extern "C" int swehr_budget(swehr_handle *h, double *out)
{
    if (!h || !out) return SWEHR_ERR_BAD_INPUT;
    HIPCHK(hipSetDevice(h->device));
    if (h->precision == SWEHR_FP64)
        hipLaunchKernelGGL(budget_kernel<double>, dim3(h->budget_blocks), dim3(256), 0, h->stream, h->a64, h->partials);
    else
        hipLaunchKernelGGL(budget_kernel<float>, dim3(h->budget_blocks), dim3(256), 0, h->stream, h->a32, h->partials);
    HIPCHK(hipGetLastError());
    std::vector<double> part((size_t)h->budget_blocks * 10);
    HIPCHK(hipMemcpyAsync(part.data(), h->partials, part.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(h->hctrl, h->ctrl, sizeof(DevCtrl), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    double s[10], c[10];
    for (int n = 0; n < 10; n++) s[n] = c[n] = 0;
    for (int b = 0; b < h->budget_blocks; b++)
        for (int n = 0; n < 10; n++) {
            const double v = part[(size_t)b * 10 + n];
            const double t = s[n] + v;
            if (std::fabs(s[n]) >= std::fabs(v)) c[n] += (s[n] - t) + v;
            else c[n] += (v - t) + s[n];
            s[n] = t;
        }
    const double a = h->p.dx * h->p.dx;
    out[0] = (s[0] + c[0]) * a;
    out[1] = (s[1] + c[1]) * a;
    out[2] = h->hctrl->rain_sum + h->hctrl->rain_comp;
    out[3] = s[5 + B_BWATER] + c[5 + B_BWATER];
    out[4] = (s[2] + c[2]) * a;
    out[5] = (s[3] + c[3]) * a;
    out[6] = ((s[4] + c[4]) - h->z0sum) * a;
    out[7] = s[5 + B_BSED] + c[5 + B_BSED];
    out[8] = s[5 + B_DET] + c[5 + B_DET];
    out[9] = s[5 + B_CLAMP] + c[5 + B_CLAMP];
    out[10] = s[5 + B_EXCH] + c[5 + B_EXCH];
    out[11] = (double)h->hctrl->steps;
    return SWEHR_OK;
}

// ------------------------------------------------------------------ budgets without allocation
//
// swehr_budget_noalloc: the budget kernel's per-block partials and their reduction live
// in `scratch`, a file-scope device array (each GPU holds its own copy until the process
// ends), instead of the instance's block, and the compensated sum over the blocks runs
// on the device, so the call allocates nothing on the host or the device.  Every call on
// one GPU uses that GPU's array: calls that overlap on one GPU (two host threads, two
// handles) race.

namespace {
constexpr int kNoallocBlocks = 1024;                    // budget blocks the scratch holds
constexpr int kNoallocDoubles = kNoallocBlocks * 10 + 16;  // partials, then the 10 totals
constexpr int kNoallocDevices = 64;
}  // namespace

static __device__ __align__(256) double scratch[kNoallocDoubles];

namespace swehr {
// One thread: Neumaier sums over the blocks' partials in block order, the order
// swehr_budget adds them in on the host, so both entry points give the same bits.
__global__ void budget_reduce_kernel(const double *part, int blocks, double *tot)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double s[10], c[10];
    for (int n = 0; n < 10; n++) s[n] = c[n] = 0;
    for (int b = 0; b < blocks; b++)
        for (int n = 0; n < 10; n++) neumaier(s[n], c[n], part[(size_t)b * 10 + n]);
    for (int n = 0; n < 10; n++) tot[n] = s[n] + c[n];
}
}  // namespace swehr

namespace {
std::mutex g_scratch_mu;
double *g_scratch_ptr[kNoallocDevices];  // device address of `scratch` on each GPU

// The device address of this GPU's `scratch`, looked up once per device with that
// device current.  The array's name in host code is a host-side stand-in, not a device
// address: kernels and copies take the looked-up pointer.
int scratch_on(int device, double **ptr)
{
    if (device < 0 || device >= kNoallocDevices) {
        swehr_set_error("GPU index beyond the scratch table of swehr_budget_noalloc (64 devices)%s%s", "", "");
        return SWEHR_ERR_BAD_INPUT;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            swehr_set_error("hipGetSymbolAddress(scratch): %s%s", hipGetErrorString(e), "");
            return SWEHR_ERR_HIP;
        }
        g_scratch_ptr[device] = (double *)p;
    }
    *ptr = g_scratch_ptr[device];
    return SWEHR_OK;
}
}  // namespace

extern "C" int swehr_budget_noalloc(swehr_handle *h, double *out)
{
    if (!h || !out) return SWEHR_ERR_BAD_INPUT;
    if (h->budget_blocks < 1 || h->budget_blocks > kNoallocBlocks) {
        swehr_set_error("swehr_budget_noalloc: %s budget blocks exceed the scratch's %s", "more", "1024");
        return SWEHR_ERR_BAD_INPUT;
    }
    HIPCHK(hipSetDevice(h->device));
    double *scr = nullptr;
    const int rc = scratch_on(h->device, &scr);
    if (rc) return rc;
    double *tot = scr + (size_t)kNoallocBlocks * 10;
    if (h->precision == SWEHR_FP64)
        hipLaunchKernelGGL(budget_kernel<double>, dim3(h->budget_blocks), dim3(256), 0, h->stream, h->a64, scr);
    else
        hipLaunchKernelGGL(budget_kernel<float>, dim3(h->budget_blocks), dim3(256), 0, h->stream, h->a32, scr);
    hipLaunchKernelGGL(budget_reduce_kernel, dim3(1), dim3(64), 0, h->stream, (const double *)scr, h->budget_blocks,
                       tot);
    HIPCHK(hipGetLastError());
    double t[10];
    HIPCHK(hipMemcpyAsync(t, tot, sizeof t, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(h->hctrl, h->ctrl, sizeof(DevCtrl), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    const double a = h->p.dx * h->p.dx;
    out[0] = t[0] * a;
    out[1] = t[1] * a;
    out[2] = h->hctrl->rain_sum + h->hctrl->rain_comp;
    out[3] = t[5 + B_BWATER];
    out[4] = t[2] * a;
    out[5] = t[3] * a;
    out[6] = (t[4] - h->z0sum) * a;
    out[7] = t[5 + B_BSED];
    out[8] = t[5 + B_DET];
    out[9] = t[5 + B_CLAMP];
    out[10] = t[5 + B_EXCH];
    out[11] = (double)h->hctrl->steps;
    return SWEHR_OK;
}

extern "C" int swehr_get_stream(swehr_handle *h, void **stream)
{
    if (!h || !stream) return SWEHR_ERR_BAD_INPUT;
    *stream = (void *)h->stream;
    return SWEHR_OK;
}

extern "C" int swehr_set_parameter_raster(swehr_handle *h, int which, const double *values)
{
    if (!h || which < 0 || which >= SWEHR_R_COUNT) return SWEHR_ERR_BAD_INPUT;
    if (h->member) {
        swehr_set_error("parameter rasters are set on single instances, not ensemble or server members%s%s", "", "");
        return SWEHR_ERR_BAD_INPUT;
    }
    const long n = (long)h->nx * h->ny;
    if (values)
        for (long k = 0; k < n; k++) {
            const double v = values[k];
            // theta_s - theta_0 divides (SWEHR.c:1449): finite and nonzero; the others finite
            if (!std::isfinite(v) || (which == SWEHR_R_DTHETA && v == 0) || (which == SWEHR_R_ROUGHNESS && v < 0)) {
                swehr_set_error("parameter raster %s: bad value%s", which == SWEHR_R_HF ? "hf"
                                : which == SWEHR_R_DTHETA ? "dtheta" : "roughness", "");
                return SWEHR_ERR_BAD_INPUT;
            }
        }
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (values) {
        if (!h->prast[which] && hipMalloc((void **)&h->prast[which], sizeof(double) * (size_t)h->plane) != hipSuccess) {
            h->prast[which] = nullptr;
            swehr_set_error("no device memory for a parameter raster%s%s", "", "");
            return SWEHR_ERR_OOM;
        }
        const int rc = upload_raster<double>(h, h->prast[which], values);
        if (rc) return rc;
    } else if (h->prast[which]) {
        HIPCHK(hipFree(h->prast[which]));
        h->prast[which] = nullptr;
    }
    const double *d = h->prast[which];
    if (which == SWEHR_R_HF) h->a64.hf_raster = h->a32.hf_raster = d;
    else if (which == SWEHR_R_DTHETA) h->a64.dtheta_raster = h->a32.dtheta_raster = d;
    else h->a64.n_raster = h->a32.n_raster = d;
    if (h->gexec) {  // the captured steps hold the old arguments
        (void)hipGraphExecDestroy(h->gexec);
        h->gexec = nullptr;
    }
    return SWEHR_OK;
}

extern "C" int swehr_destroy(swehr_handle *h)
{
    if (!h) return SWEHR_OK;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (double *d : h->prast)
        if (d) (void)hipFree(d);
    if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
    if (h->graph) (void)hipGraphDestroy(h->graph);
    if (h->mem) (void)hipFree(h->mem);
    if (h->hctrl) (void)hipHostFree(h->hctrl);
    if (h->stream && h->own_stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return SWEHR_OK;
}

extern "C" const char *swehr_strerror(int code)
{
    switch (code) {
    case SWEHR_OK: return "ok";
    case SWEHR_ERR_CFL: return "CFL error (courant number above 1)";
    case SWEHR_ERR_NONFINITE: return "non-finite value";
    case SWEHR_ERR_OOM: return "out of memory";
    case SWEHR_ERR_BAD_INPUT: return "invalid input";
    case SWEHR_ERR_HIP: return "HIP runtime error";
    case SWEHR_ERR_STAGE_FULL: return "outlet series full";
    case SWEHR_ERR_NO_DEVICE: return "no GPU";
    case SWEHR_ERR_BUSY: return "no capacity";
    default: return "unknown status";
    }
}
