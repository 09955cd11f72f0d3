// swehr_device.h - device-side data structures and the per-step kernels of the SWEHR
// coupled time step on gfx950.  Shared by the single-instance library
// (swehr_lib.hip) and the ensemble / serving runner (swehr_batch.hip).
//
// Memory layout in HBM (one instance): structure of arrays, one plane per variable,
// value (i, j) at [i * py + j] with j (downslope) unit-stride and the row pitch py
// rounded up to 64 elements, so every row starts on a 256 B (FP64) / 128 B (FP32)
// boundary and a wavefront reads a row segment in whole cache lines.  Fields read at
// neighbours (h, hu, hv, ch_k, DEP) are double-buffered by step parity; the bed z has
// three buffers (previous, current, next) because the reference reads geometry one
// step late (SWEHR.c:1135, 1243, 1293, 1304); fields read only at their own cell
// (M_k, TOTM, VINF, MAXVEL, MAXDEPTH) are updated in place.  SED, TOTSED, RHO and VEL
// are not stored: they are exact functions of (h, ch_k, hu, hv) (SWEHR.c:1516-1531,
// 1563-1564, 1578-1579) and are recomputed where needed.
//
// A step's work, guarded by the per-instance device control block so that a captured
// graph of many steps stops by itself at t >= tend, at a print pause or at a step limit:
//   K1 wave speeds  max |wave speed| over every face (SWEHR.c:1161, 1228) -> atomicMax
//                   on the bit pattern of a non-negative double
//   K2 step         boundary ghosts, fluxes, sources, dt, update (SWEHR.c:816-1589) on
//                   LDS-staged tiles; every face flux is computed once per tile and the
//                   per-class work streams through registers; one thread also decides
//                   the end of the step (t += dt, print pause, stop: SWEHR.c:1408-1413,
//                   1591-1592, 1679-1682) into the control block's `next`
//   commit          outlet and probe records (SWEHR.c:1593-1676) and the decided end
//                   into the control block: in a single instance's step graph by block
//                   0 of the next step's K1, and by K3 (the control kernel) at the end
//                   of the graph; in the ensemble modes and for steps launched one by
//                   one by K3 after every step
// The bodies are device functions; the single-instance kernels and the ensemble
// kernels (one launch over the tiles of many instances) call the same code.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "swehr_physics.h"

namespace swehr {

constexpr int kMaxClasses = 16;
constexpr int kStageCols = 24;
constexpr int kFlowCols = 18;
constexpr int kStageMax = 30000;

// per-tile budget terms accumulated across steps (sum and Neumaier compensation)
enum BudgetTerm { B_EXCH = 0, B_CLAMP = 1, B_DET = 2, B_BWATER = 3, B_BSED = 4, NB_TERMS = 5 };

// The end of a step as the control block will hold it (its scalar part): formed by the
// step kernel as soon as the step's dt is known, committed to the control block once
// every cell of the step is stored, together with the stage records that read them.
struct StepNext {
    double t, dt, courant, rain, t_prev;
    double stageinterval, printinterval, rain_sum, rain_comp;
    double in_h, in_hv, in_c, rain_now;
    long long steps;
    int stagecnt, status, done, print_due;
    int stage_due, stage_r, active, pad;
};

struct alignas(64) DevCtrl {
    double t, dt, courant, rain;
    double t_prev;  // time at the start of the last step (the CFL message prints it)
    double stageinterval, printinterval, printint, rain_sum;
    double rain_comp;
    // functions of t alone, set with t: the sediment-laden inflow state at j = ny
    // (SWEHR.c:1019-1033) and the rain rate (SWEHR.c:813-814)
    double in_h, in_hv, in_c, rain_now;
    long long steps, step_limit;
    int stagecnt, status, done, print_due;
    // pending: `next` holds the decision of a step whose kernel ran since the last
    // standalone commit (the control kernel clears it; a graph always ends with one)
    int pause_on_print, pending, pad[2];
    StepNext next;
    // max |wave speed| of a step as bit patterns of non-negative doubles, spread over 64
    // slots so that the per-block atomicMax of K1 does not serialise on one address;
    // one set per step parity, so that K1 of a step can fill its set while the commit
    // of the previous step clears the other
    unsigned long long smax_slots[2][64];
};

template <typename R>
struct StepArgs {
    int nx, ny, S, py;
    long long plane;
    R *h[2], *hu[2], *hv[2], *ch[2], *dep[2];
    double *z[3];  // bed elevation, FP64 in both precisions (slopes are differences of it)
    R *M, *totm, *maxvel, *maxdepth;
    double *vinf;
    const double *ks_raster;  // per-cell Ks (pitched like the state) or null
    // per-cell suction head, theta_s - theta_0 and Manning n0 (null: the scalars)
    const double *hf_raster, *dtheta_raster, *n_raster;
    const double *rain;
    double ks, dx, h0, mtstar0, a, ad, F, J, roughness, theta0, thetas, hf, cstable, maxstep, tend, rint, vinf0;
    int rnum, inflow, bndry[4];
    int tiles_i, tiles_j, nblocks, blocks_per_xcd;
    int bud_slots;  // budget slots allocated: one per wave of every tile, any tile shape
    int stage_cap;  // rows of the stage / saveflow series
    // per-class diameter and fraction, and the settling constants a_k = 13.95 nu / d_k
    // and a_k^2 (SWEHR.c:435), formed in FP64 and stored in the step's precision (the
    // class loop reads them so, with no conversion per class)
    R ds[kMaxClasses], pf[kMaxClasses];
    R ak[kMaxClasses], ak2[kMaxClasses];
    // bit k: d_k > 0.0001, resp. d_k > 0.002, compared in FP64 as the reference
    // compares them (the shallow-flow settling rules, SWEHR.c:1385-1388)
    unsigned coarse1, coarse2;
    // thresholds formed from parameters, rounded as thr_lt / thr_gt round: 9 h0 for
    // h < 9 h0 and h > 9 h0 (SWEHR.c:1384-1391, 1414), h0 for the depth normal to the
    // bed (SWEHR.c:1331, 1347), d_k for the entrainment condition h > d_k
    R t9h0_lt, t9h0_gt, th0_gt;
    R tds_gt[kMaxClasses];
    // quotients of parameters, divided on the host exactly as the step would (IEEE
    // division on both sides): 1/dx, and 1/(6 h0) in this precision
    double inv_dx;
    R inv_6h0;
    DevCtrl *ctrl;
    // budget slots: [bud_slots][NB_TERMS][2] running sum and compensation, followed by
    // [bud_slots][NB_TERMS] the last step's partials not yet folded into them (0 = none)
    double *bud;
    double *stage;     // [kStageCols][stage_cap]
    double *saveflow;  // [kFlowCols][stage_cap]
    // outlet discharge: per column i the outflow through the j = 1/2 face in the last
    // step (water m^3/s, sediment kg/s), and its sum at each stage record [2][stage_cap]
    double *outq, *outs, *discharge;
};

template <typename T>
__device__ __forceinline__ T wave_max(T v)
{
    for (int o = 32; o > 0; o >>= 1) {
        const T w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

// The pointers in an argument block are device (global) memory.  Typing them so lets
// the compiler use global (and scalar) instead of flat instructions where the block was
// read from memory rather than passed as a kernel argument (the ensemble kernels).
template <typename T>
using glob = T __attribute__((address_space(1)));
template <typename T>
__device__ __forceinline__ glob<T> *gmem(T *p)
{
    return (glob<T> *)p;
}

// Element at byte offset `off` of a plane.  Offsets are 32-bit (a plane is < 4 GiB,
// checked at create): the plane base stays in scalar registers and the cell's one
// offset register serves every plane of the same element size.
template <typename T>
__device__ __forceinline__ glob<T> &at(glob<T> *base, unsigned off)
{
    return *(glob<T> *)((glob<char> *)base + off);
}
template <typename T>
__device__ __forceinline__ const glob<T> &at(const glob<T> *base, unsigned off)
{
    return *(const glob<T> *)((const glob<char> *)base + off);
}
// The same with the nontemporal hint (`nt` loads and stores): for data a step touches
// once, so that it does not push out of L2 the halo rows that neighbouring tiles
// read again.  The step kernel streams its own-cell-only inputs (TOTM, MAXVEL,
// MAXDEPTH, VINF, Ks, M_k) and every output this way; on the 2048^2, S = 7 FP64 step
// that took the step kernel 0.457 -> 0.432 ms and the next wave-speed pass
// 0.049 -> 0.045 ms (streaming h, hu, hv, ch_k too, or the wave-speed loads, was slower).
// Bitwise equality: the in-place fields (M_k, TOTM, VINF) are stored only where the
// step changed them, which in dry cells it mostly does not (a wetting-phase step of the
// 4096^2, S = 10 case: 2.09 -> 1.82 ms).
__device__ __forceinline__ bool same_bits(double a, double b) { return __double_as_longlong(a) == __double_as_longlong(b); }
__device__ __forceinline__ bool same_bits(float a, float b) { return __float_as_int(a) == __float_as_int(b); }
template <typename T>
__device__ __forceinline__ T ld_stream(const glob<T> *base, unsigned off)
{
    return __builtin_nontemporal_load((const glob<T> *)((const glob<char> *)base + off));
}
template <typename T>
__device__ __forceinline__ void st_stream(glob<T> *base, unsigned off, T v)
{
    __builtin_nontemporal_store(v, (glob<T> *)((glob<char> *)base + off));
}

__device__ __forceinline__ bool step_active(const glob<DevCtrl> *C)
{
    return !C->done && !C->print_due && C->steps < C->step_limit;
}

// Max over the 64 wave-speed slots of the step with count `steps`; every lane of the
// calling wave gets it (all 64 lanes must be active).
__device__ __forceinline__ double ctrl_smax(const glob<DevCtrl> *C, long long steps)
{
    return wave_max(__longlong_as_double((long long)C->smax_slots[steps & 1][threadIdx.x & 63]));
}

__device__ __forceinline__ InflowState ctrl_inflow(const glob<DevCtrl> *C)
{
    InflowState in;
    in.h = C->in_h;
    in.hv = C->in_hv;
    in.c = C->in_c;
    return in;
}

// What the wave-speed pass of the next step needs of the control block: whether that
// step runs, its count (the parity of the buffers) and the inflow state at its time.
// While a decision is pending (the previous step's commit runs in the same launch, in
// block 0), they come from the decision, which no block of that launch writes.
struct K1View {
    bool active;
    long long steps;
    InflowState in;
};
__device__ __forceinline__ K1View k1_view(const glob<DevCtrl> *C)
{
    // both sets are read at once and selected (one round trip)
    const glob<StepNext> *N = &C->next;
    const bool pend = C->pending != 0;
    const bool act_c = step_active(C), act_n = N->active != 0;
    const long long steps_c = C->steps, steps_n = N->steps;
    const InflowState in_c = ctrl_inflow(C);
    const double nh = N->in_h, nhv = N->in_hv, nc = N->in_c;
    K1View v;
    v.active = pend ? act_n : act_c;
    v.steps = pend ? steps_n : steps_c;
    v.in.h = pend ? nh : in_c.h;
    v.in.hv = pend ? nhv : in_c.hv;
    v.in.c = pend ? nc : in_c.c;
    return v;
}

// Sets the functions of t kept in the control block (one thread).  The inflow state is
// read only where the inflow at j = ny is on; without it the pow of inflow_state is
// not evaluated on the step's critical path.
template <typename R>
__device__ __forceinline__ void set_time_functions(glob<DevCtrl> *C, const StepArgs<R> &A, double t)
{
    InflowState in = {0.0, 0.0, 0.0};
    if (A.inflow) in = inflow_state(t);
    C->in_h = in.h;
    C->in_hv = in.hv;
    C->in_c = in.c;
    C->rain_now = rain_rate(A.rain, A.rnum, A.rint, t);
}

template <typename R>
__global__ void time_functions_kernel(StepArgs<R> A)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) set_time_functions(gmem(A.ctrl), A, A.ctrl->t);
}

__device__ __forceinline__ double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ void neumaier(double &s, double &c, double x)
{
    const double t = s + x;
    if (fabs(s) >= fabs(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
}

// One-sided at the domain edges exactly as SWEHR.c:1240-1277; P is an LDS plane with
// row stride ZJ holding z (or the previous step's z) around the cell at (zi, zj).
// Computed in FP64 (the bed is FP64 in both precisions).
template <int ZJ>
__device__ __forceinline__ double slope_x(const double *P, int zi, int zj, int gi, int nx, double odx)
{
    if (gi != 0 && gi != nx - 1) return 0.5 * odx * (P[(zi + 1) * ZJ + zj] - P[(zi - 1) * ZJ + zj]);
    if (gi == nx - 1) return odx * (P[zi * ZJ + zj] - P[(zi - 1) * ZJ + zj]);
    return odx * (P[(zi + 1) * ZJ + zj] - P[zi * ZJ + zj]);
}

template <int ZJ>
__device__ __forceinline__ double slope_y(const double *P, int zi, int zj, int gj, int ny, double odx)
{
    if (gj != 0 && gj != ny - 1) return 0.5 * odx * (P[zi * ZJ + zj + 1] - P[zi * ZJ + zj - 1]);
    if (gj == ny - 1) return odx * (P[zi * ZJ + zj] - P[zi * ZJ + zj - 1]);
    return odx * (P[zi * ZJ + zj + 1] - P[zi * ZJ + zj]);
}

// Budget term g of an argument block's slots: folds its pending partial into the
// running sum (Neumaier) and clears it.  Most partials are zero (the step kernel stores
// only nonzero ones): then only the partial is read.
__device__ __forceinline__ void fold_budget(glob<double> *bud, long long nterms, long long g)
{
    glob<double> *pend = bud + 2 * nterms + g;
    const double x = *pend;
    if (x == 0.0) return;
    glob<double> *p = bud + 2 * g;
    double s = p[0], c = p[1];
    neumaier(s, c, x);
    p[0] = s;
    p[1] = c;
    *pend = 0.0;
}


// The scalar end of the current step (SWEHR.c:1408-1413 dt and the CFL stop,
// 1591-1592 the stage clock, 1679-1682 the print interval, the rain budget, and the
// functions of the new time), from the control block and the step's max wave speed smax,
// written to C->next.  One thread, while the step's kernels run; nothing of that launch
// reads C->next.
template <typename R>
__device__ __forceinline__ void decide_step(glob<DevCtrl> *C, const StepArgs<R> &A, double smax)
{
    const long long steps = C->steps;
    const double t = C->t;
    const double dt = step_dt(smax, A.cstable, A.dx, A.maxstep, t, A.tend);
    const double courant = (smax)*dt * (1.0 / A.dx);
    const double Rr = C->rain_now;
    const bool cfl = courant > 1;
    const double tnew = (cfl ? 2 * A.tend : t) + dt;
    const double stageinterval = C->stageinterval;
    const int r = C->stagecnt, status0 = C->status;
    const bool stage_due = tnew > stageinterval, stage_full = stage_due && r >= A.stage_cap;
    double printinterval = C->printinterval;
    int print_due = 0;
    if (tnew >= printinterval) {
        printinterval = printinterval + C->printint;
        if (C->pause_on_print) print_due = 1;
    }
    int status = status0;
    if (stage_full) status = 6;  // SWEHR_ERR_STAGE_FULL
    if (cfl) status = 1;         // SWEHR_ERR_CFL
    // dt = 0 (infinite wave speed) would never reach tend; the reference loops forever
    if (!(dt > 0) && tnew < A.tend) status = 2;  // SWEHR_ERR_NONFINITE
    double rs = C->rain_sum, rc = C->rain_comp;
    neumaier(rs, rc, dt * Rr * ((double)A.nx * A.ny) * A.dx * A.dx);
    const int done = (!(tnew < A.tend) || status != 0) ? 1 : 0;
    // the inflow state is read only where the inflow at j = ny is on; without it the pow
    // of inflow_state is not evaluated
    InflowState in = {0.0, 0.0, 0.0};
    if (A.inflow) in = inflow_state(tnew);
    glob<StepNext> *N = &C->next;
    N->t = tnew;
    N->dt = dt;
    N->courant = courant;
    N->rain = Rr;
    N->t_prev = t;
    N->stageinterval = stage_due ? stageinterval + 1 : stageinterval;
    N->printinterval = printinterval;
    N->rain_sum = rs;
    N->rain_comp = rc;
    N->in_h = in.h;
    N->in_hv = in.hv;
    N->in_c = in.c;
    N->rain_now = rain_rate(A.rain, A.rnum, A.rint, tnew);
    N->steps = steps + 1;
    N->stagecnt = stage_due && !stage_full ? r + 1 : r;
    N->status = status;
    N->done = done;
    N->print_due = print_due;
    N->stage_due = stage_due ? 1 : 0;
    N->stage_r = r;
    N->active = (!done && !print_due && steps + 1 < C->step_limit) ? 1 : 0;
    C->pending = 1;
}

// Derived quantities of a stored cell (SWEHR.c:1516-1531, 1563-1564).
template <typename R>
__device__ __forceinline__ void derived(const StepArgs<R> &A, int p, long long c, double &h, double &v1, double &v2,
                                        double &ts)
{
    const R hh = gmem(A.h[p])[c];
    h = (double)hh;
    v1 = v2 = ts = 0;
    if (hh > R(kHfilm)) {
        v1 = (double)(gmem(A.hu[p])[c] / hh);
        v2 = (double)(gmem(A.hv[p])[c] / hh);
        R s = R(0);
        const glob<R> *ch = gmem(A.ch[p]);
        for (int k = 0; k < A.S; k++) s += ch[k * A.plane + c] / hh / R(kRhos);
        ts = (double)s;
    }
}

// Commits the decision in C->next of the step just completed: its stage records
// (SWEHR.c:1593-1676, which read the step's new state), the control block's fields, and
// clears that step's wave-speed slots.  Run by all threads of one block (blocks of up to
// 1024 threads).
template <typename R>
__device__ __forceinline__ void commit_step(const StepArgs<R> &A, glob<DevCtrl> *C)
{
    __shared__ double s_best[16], s_q2[16];  // one per wave
    __shared__ int s_idx[16];
    StepNext N;
    {
        const glob<StepNext> *G = &C->next;
        N.t = G->t;
        N.steps = G->steps;
        N.stage_due = G->stage_due;
        N.stage_r = G->stage_r;
    }
    const double tnew = N.t;
    const int q = (int)(N.steps & 1);
    const int cap = A.stage_cap;
    const int r = N.stage_r;
    if (N.stage_due) {
        // outlet: deepest cell of row j = 1, first one on ties (SWEHR.c:1596-1604)
        double best = 0;
        int idx = 0x7fffffff;
#pragma unroll 8
        for (int i = threadIdx.x; i < A.nx; i += blockDim.x) {
            const double hv = (double)gmem(A.h[q])[(long long)i * A.py];
            if (hv > best) { best = hv; idx = i; }
        }
        // block arg-max: larger depth wins, the lower index on equal depths
        for (int o = 32; o > 0; o >>= 1) {
            const double ob = __shfl_xor(best, o, 64);
            const int oi = __shfl_xor(idx, o, 64);
            if (ob > best || (ob == best && oi < idx)) { best = ob; idx = oi; }
        }
        const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
        if (lane == 0) {
            s_best[wid] = best;
            s_idx[wid] = idx;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 1; w < (int)(blockDim.x >> 6); w++)
                if (s_best[w] > best || (s_best[w] == best && s_idx[w] < idx)) { best = s_best[w]; idx = s_idx[w]; }
            s_idx[0] = best > 0 ? idx : 0;
        }
        __syncthreads();
        const int outlet = s_idx[0];
        // outlet discharge of this step: the columns' outflows summed in a fixed order
        {
            double q = 0, qs = 0;
            const glob<double> *oq = gmem(A.outq), *os = gmem(A.outs);
            if (threadIdx.x < 256)  // stride 256 whatever the block size: the same order everywhere
#pragma unroll 8
                for (int i = threadIdx.x; i < A.nx; i += 256) {
                    q += oq[i];
                    qs += os[i];
                }
            q = wave_sum(q);
            qs = wave_sum(qs);
            __syncthreads();  // s_best is reused
            if (lane == 0) {
                s_best[wid] = q;
                s_q2[wid] = qs;
            }
            __syncthreads();
            if (threadIdx.x == 0 && r < cap) {
                double tq = 0, tqs = 0;
                for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
                    tq += s_best[w];
                    tqs += s_q2[w];
                }
                gmem(A.discharge)[r] = tq;
                gmem(A.discharge)[cap + r] = tqs;
            }
        }
        // one thread per stage column group: outlet state, its classes, the six probes
        // at (i = 2, j = 1, 10, ..., 50) (SWEHR.c:1605-1644)
        const int tid = threadIdx.x;
        if (r < cap) {
            glob<double> *st = gmem(A.stage);
            const long long co = (long long)outlet * A.py;
            if (tid == 0) {
                double h, v1, v2, ts;
                derived(A, q, co, h, v1, v2, ts);
                st[0 * cap + r] = tnew;
                st[1 * cap + r] = h;
                st[2 * cap + r] = sqrt(v1 * v1 + v2 * v2);
                st[3 * cap + r] = ts;
            } else if (tid >= 1 && tid <= 10) {
                const int k = tid - 1;
                st[(4 + k) * cap + r] = k < A.S ? (double)gmem(A.M)[k * A.plane + co] : 0.0;
                st[(14 + k) * cap + r] = k < A.S ? (double)gmem(A.ch[q])[k * A.plane + co] : 0.0;
            } else if (tid >= 11 && tid <= 16) {
                const int n = tid - 11;
                const int yp = n == 0 ? 1 : 10 * n;
                double d = 0, sp = 0, tt = 0;
                if (A.nx >= 2 && yp <= A.ny) {
                    double hh, w1, w2;
                    derived(A, q, (long long)1 * A.py + (yp - 1), hh, w1, w2, tt);
                    d = hh;
                    sp = sqrt(w1 * w1 + w2 * w2);
                }
                glob<double> *sf = gmem(A.saveflow);
                sf[(3 * n + 0) * cap + r] = d;
                sf[(3 * n + 1) * cap + r] = sp;
                sf[(3 * n + 2) * cap + r] = tt;
            }
        }
    }
    __syncthreads();  // every thread has read the decision
    if (threadIdx.x == 0) {
        const glob<StepNext> *G = &C->next;
        C->t = G->t;
        C->dt = G->dt;
        C->courant = G->courant;
        C->rain = G->rain;
        C->t_prev = G->t_prev;
        C->stageinterval = G->stageinterval;
        C->printinterval = G->printinterval;
        C->rain_sum = G->rain_sum;
        C->rain_comp = G->rain_comp;
        C->in_h = G->in_h;
        C->in_hv = G->in_hv;
        C->in_c = G->in_c;
        C->rain_now = G->rain_now;
        C->stagecnt = G->stagecnt;
        if (G->status != C->status) C->status = G->status;
        if (G->print_due) C->print_due = 1;
        if (G->done) C->done = 1;
        C->steps = G->steps;
    }
    if (threadIdx.x < 64) C->smax_slots[(N.steps - 1) & 1][threadIdx.x] = 0ull;
}

}  // namespace swehr

#include "swehr_cell.h"

namespace swehr {

// --------------------------------------------------------------------------- K1

// Value held by the lane below (L - 1) by a DPP wave shift; lane 0 keeps its own.
__device__ __forceinline__ int lane_below_i(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x138, 0xf, 0xf, false); }
__device__ __forceinline__ float lane_below(float v) { return __int_as_float(lane_below_i(__float_as_int(v))); }
__device__ __forceinline__ double lane_below(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = lane_below_i((int)(unsigned)b), hi = lane_below_i((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
// Value held by the lane above (L + 1) by a DPP wave shift; lane 63 keeps its own.
__device__ __forceinline__ int lane_above_i(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x130, 0xf, 0xf, false); }
__device__ __forceinline__ float lane_above(float v) { return __int_as_float(lane_above_i(__float_as_int(v))); }
__device__ __forceinline__ double lane_above(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = lane_above_i((int)(unsigned)b), hi = lane_above_i((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}


// Wave-speed bound of the faces a cell owns: i+1/2 and j+1/2, plus 1/2 at i = 0 or
// j = 0.  Ghost states follow the boundary rules of SWEHR.c:816-1035; GZ is that of
// the previous step's flux pass, from the lagged bed zp (read from global memory).
template <typename R>
__device__ __forceinline__ R wave_speed_cell(const StepArgs<R> &A, const glob<R> *h, const glob<R> *hu,
                                             const glob<R> *hv, const glob<double> *zp, int i, int j,
                                             const InflowState &in)
{
    const int nx = A.nx, ny = A.ny;
    const unsigned py = (unsigned)A.py, c = (unsigned)i * py + (unsigned)j;
    // neighbours clamped into the grid, so every load is unconditional; the one-sided
    // differences at the edges (SWEHR.c:1240-1277) then read the cell itself
    const unsigned im = i > 0 ? c - py : c, ip = i < nx - 1 ? c + py : c;
    const unsigned jm = j > 0 ? c - 1 : c, jp = j < ny - 1 ? c + 1 : c;
    const double odx = A.inv_dx;
    const double fx = (i != 0 && i != nx - 1) ? 0.5 * odx : odx;
    const double fy = (j != 0 && j != ny - 1) ? 0.5 * odx : odx;
    const double sxd = fx * (at(zp, ip * 8u) - at(zp, im * 8u));
    const double syd = fy * (at(zp, jp * 8u) - at(zp, jm * 8u));
    const unsigned w = (unsigned)sizeof(R);
    const R hc = at(h, c * w), huc = at(hu, c * w), hvc = at(hv, c * w);
    R hx = at(h, ip * w), qx = at(hu, ip * w);
    R hy = at(h, jp * w), qy = at(hv, jp * w);
    const R sx = R(sxd), sy = R(syd);
    const R g = R(kG) * mrsqrt(sx * sx + sy * sy + R(1));
    if (i == nx - 1) qx = A.bndry[2] == 1 ? -huc : huc;  // hx = hc already
    if (j == ny - 1) {
        if (A.inflow) { hy = R(in.h); qy = R(in.hv); }
        else qy = A.bndry[3] == 1 ? -hvc : hvc;
    }
    R smax = mfmax(face_speed(hc, hx, huc, qx, g), face_speed(hc, hy, hvc, qy, g));
    if (i == 0)  // face 1/2: ghost on the left, GZ of cell 1 (SWEHR.c:1054-1059)
        smax = mfmax(smax, face_speed(hc, hc, A.bndry[0] == 1 ? -huc : huc, huc, g));
    if (j == 0) smax = mfmax(smax, face_speed(hc, hc, A.bndry[1] == 1 ? -hvc : hvc, hvc, g));
    return smax;
}

// Block-wide max of v, then one atomicMax on slot `slot` (0-63) of the instance.
__device__ __forceinline__ void block_max_to(unsigned long long *slots, int slot, double v)
{
    unsigned long long *dst = slots + (slot & 63);
    __shared__ double red[16];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    v = wave_max(v);
    if (lane == 0) red[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)((blockDim.x + 63) >> 6); w++) v = red[w] > v ? red[w] : v;
        atomicMax(dst, (unsigned long long)__double_as_longlong(v));
    }
}

// K1 on strips of rows: wave w of a block takes rows ib .. ib + K1_RPW - 1 (ib = first
// row of the block + w K1_RPW) of 64 columns, lanes 0 and 63 holding the columns left
// and right of the 62 whose faces it computes.  A cell's velocities (hu/h, hv/h) are
// computed once and passed to the cells whose faces need them: the row above through
// registers, the column to the left through a lane shift; each row of each field is
// loaded once per wave.  Every face speed is the one wave_speed_cell computes (the same
// velocities, celerities and GZ, the same boundary rules).  Tiles are dealt to XCDs in
// contiguous runs; the grid is the tile count rounded up to a multiple of 8.
constexpr int kK1Rows = 4;               // rows per wave
constexpr int kK1TI = 4 * kK1Rows, kK1TJ = 62;  // rows and computed columns per block
__host__ __device__ inline int k1_tiles(int nx, int ny) { return ((nx + kK1TI - 1) / kK1TI) * ((ny + kK1TJ - 1) / kK1TJ); }

// The wave speeds of strip tile `tile` of an instance (the strip geometry above; the
// block's 256 threads are 4 waves of kK1Rows rows each): the largest face speed of this
// thread's cells.  `view` gives the step count and the inflow state.
template <typename R>
__device__ __forceinline__ R wave_speed_strip(const StepArgs<R> &A, const K1View &view, int tile)
{
    constexpr int RPW = kK1Rows;
    const long long steps = view.steps;
    const int p = (int)(steps & 1), p3 = (int)(steps % 3);
    const InflowState in = view.in;
    const int nx = A.nx, ny = A.ny;
    const int tiles_j = (ny + kK1TJ - 1) / kK1TJ;
    R s = R(0);
    {
        const int ti = tile / tiles_j, tj = tile - ti * tiles_j;
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        const int j = tj * kK1TJ - 1 + lane;  // this lane's column (-1 or >= ny at the edges)
        const int jc = j < 0 ? 0 : (j >= ny ? ny - 1 : j);
        const bool comp = lane >= 1 && lane <= kK1TJ && j < ny;
        const int ib = ti * kK1TI + w * RPW;
        const unsigned py = (unsigned)A.py, w8 = (unsigned)sizeof(R);
        const glob<R> *h = gmem(A.h[p]), *hu = gmem(A.hu[p]), *hv = gmem(A.hv[p]);
        const glob<double> *zp = gmem(A.z[(p3 + 2) % 3]);
        auto cell = [&](int row) {
            row = row < 0 ? 0 : (row >= nx ? nx - 1 : row);
            return (unsigned)row * py + (unsigned)jc;
        };
        // the lagged bed of rows ib-1 .. ib+RPW, depth and momenta of rows ib .. ib+RPW
        // (hv to ib+RPW-1), rows clamped into the grid
        double zr[RPW + 2];
        R hr[RPW + 1], hur[RPW + 1], hvr[RPW];
#pragma unroll
        for (int r = 0; r < RPW + 1; r++) {
            const unsigned c = cell(ib + r);
            hr[r] = at(h, c * w8);
            hur[r] = at(hu, c * w8);
            if (r < RPW) hvr[r] = at(hv, c * w8);
        }
#pragma unroll
        for (int r = 0; r < RPW + 2; r++) zr[r] = at(zp, cell(ib - 1 + r) * 8u);
        const R hfilm = R(kHfilm);
        const double odx = A.inv_dx;
        // velocities of each row (0 below the film depth), one reciprocal per cell
        R ur[RPW + 1], vr[RPW];
        bool wet[RPW + 1];
#pragma unroll
        for (int r = 0; r < RPW + 1; r++) {
            wet[r] = hr[r] > hfilm;
            const R ih = wet[r] ? mrcp(hr[r]) : R(0);
            ur[r] = wet[r] ? hur[r] * ih : R(0);
            if (r < RPW) vr[r] = wet[r] ? hvr[r] * ih : R(0);
        }
        // the inflow ghost at j = ny (SWEHR.c:1019-1035)
        const R hin = R(in.h), vin = hin > hfilm ? R(in.hv) * mrcp(hin) : R(0);
        const double fy = (j != 0 && j != ny - 1) ? 0.5 * odx : odx;
#pragma unroll
        for (int r = 0; r < RPW; r++) {
            const int i = ib + r;
            // neighbours across the lane shifts (the loads are clamped, so at the domain's
            // edges they hold the cell itself, as the one-sided slopes need)
            const double zjp = lane_above(zr[r + 1]), zjm = lane_below(zr[r + 1]);
            const R hjp = lane_above(hr[r]), vjp = lane_above(vr[r]);
            if (comp && i < nx) {
                const double fx = (i != 0 && i != nx - 1) ? 0.5 * odx : odx;
                const R sx = R(fx * (zr[r + 2] - zr[r])), sy = R(fy * (zjp - zjm));
                const R g = R(kG) * mrsqrt(sx * sx + sy * sy + R(1));
                const R hc = hr[r], cl = msqrt(g * hc);
                const R uc = ur[r], vc = vr[r];
                // x face i+1/2; at i = nx-1 the ghost mirrors (or copies) the cell
                R hx = hr[r + 1], ux = ur[r + 1];
                if (i == nx - 1) {
                    hx = hc;
                    ux = wet[r] ? (A.bndry[2] == 1 ? -uc : uc) : R(0);
                }
                R sm = face_speed_uc(hc, hx, uc, ux, cl, msqrt(g * hx));
                // y face j+1/2; at j = ny-1 the inflow ghost or the mirrored cell
                R hy = hjp, vy = vjp;
                if (j == ny - 1) {
                    if (A.inflow) {
                        hy = hin;
                        vy = vin;
                    } else {
                        hy = hc;
                        vy = wet[r] ? (A.bndry[3] == 1 ? -vc : vc) : R(0);
                    }
                }
                sm = mfmax(sm, face_speed_uc(hc, hy, vc, vy, cl, msqrt(g * hy)));
                // faces 1/2: the ghost on the low side, GZ of the first cell (SWEHR.c:1054-1059)
                if (i == 0) sm = mfmax(sm, face_speed_uc(hc, hc, wet[r] ? (A.bndry[0] == 1 ? -uc : uc) : R(0), uc, cl, cl));
                if (j == 0) sm = mfmax(sm, face_speed_uc(hc, hc, wet[r] ? (A.bndry[1] == 1 ? -vc : vc) : R(0), vc, cl, cl));
                s = mfmax(s, sm);
            }
        }
    }
    return s;
}

template <typename R>
__global__ __launch_bounds__(256) void wave_speed_kernel(StepArgs<R> A)
{
    glob<DevCtrl> *C = gmem(A.ctrl);
    const int b = blockIdx.x;
    const K1View view = k1_view(C);
    // block 0 also commits the previous step (its stage records read that step's stored
    // cells, which the launch boundary has made visible)
    if (!view.active && b != 0) return;
    const int ntiles = k1_tiles(A.nx, A.ny);
    const int per = (int)(gridDim.x >> 3);
    const int tile = (b & 7) * per + (b >> 3);
    R s = R(0);
    if (view.active && tile < ntiles) s = wave_speed_strip(A, view, tile);
    if (view.active) {
        // the step kernel's budget partials of the previous step
        const long long nterms = (long long)A.bud_slots * NB_TERMS;
        for (long long g = (long long)b * blockDim.x + threadIdx.x; g < nterms;
             g += (long long)gridDim.x * blockDim.x)
            fold_budget(gmem(A.bud), nterms, g);
        block_max_to((unsigned long long *)C->smax_slots[view.steps & 1], b, (double)s);
    }
    if (b == 0) {
        __shared__ int s_commit;
        if (threadIdx.x == 0) s_commit = C->pending && C->steps + 1 == C->next.steps;
        __syncthreads();
        if (s_commit) commit_step(A, C);
    }
}

// Wave speeds of one TI x TJ tile (the persistent ensemble mode, one tile at a time).
template <typename R, int TI, int TJ>
__device__ __forceinline__ void wave_speed_tile(const StepArgs<R> &A, int tile)
{
    glob<DevCtrl> *C = gmem(A.ctrl);
    const K1View view = k1_view(C);
    if (!view.active) return;
    const long long steps = view.steps;
    const int p = (int)(steps & 1), p3 = (int)(steps % 3);
    const InflowState in = view.in;
    const int ti = tile / A.tiles_j, tj = tile - (tile / A.tiles_j) * A.tiles_j;
    // (the thread index through an empty asm statement, as in step_tile's LOOP form:
    // this function runs in the persistent kernel's loop over tiles)
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int il = tid / TJ, jl = tid - (tid / TJ) * TJ;
    const int i = ti * TI + il, j = tj * TJ + jl;
    R s = R(0);
    if (i < A.nx && j < A.ny)
        s = wave_speed_cell(A, gmem(A.h[p]), gmem(A.hu[p]), gmem(A.hv[p]), gmem(A.z[(p3 + 2) % 3]), i, j, in);
    // the step kernel's budget partials of this tile's slots (previous step)
    if (tid < (TI * TJ / 64) * NB_TERMS)
        fold_budget(gmem(A.bud), (long long)A.bud_slots * NB_TERMS, (long long)tile * (TI * TJ / 64) * NB_TERMS + tid);
    block_max_to((unsigned long long *)C->smax_slots[steps & 1], tile, (double)s);
}

// --------------------------------------------------------------------------- K2

// Waves per SIMD the step kernels are compiled for (VGPR budget 512 / waves): the
// LDS of a tile admits 3 blocks per CU in FP64 and 5 in FP32 at the default tile.
template <typename R, bool SPLIT = false>
constexpr int kStepWaves = sizeof(R) == 4 ? (SPLIT ? 4 : 5) : 3;

// Size classes whose halo the step kernel loads per batch in phase 1 (FP32 keeps fewer
// in flight to stay at 5 waves per SIMD).
template <typename R>
constexpr int kHaloBatch = sizeof(R) == 4 ? 4 : 8;

// FP32: classes whose M_k the step kernel loads with the own-cell values (the rest
// one class ahead in the class loop); the 96-VGPR budget of 5 waves/SIMD spills from
// 2 on.
constexpr int kFp32MOwn = 1;
// FP64 two-chunk variant: classes whose M_k is loaded up front, and whose ch_k stays in
// registers (7 and 7: with 8 M_k up front the 168-VGPR budget spills 12 bytes per lane)
constexpr int kSplitMOwn = 7;
constexpr int kSplitOwn = 7;
// FP64: classes whose ch_k and M_k of the own cell stay in registers from the halo
// staging on (the rest are read one class ahead in the class loop, where the first of
// them waits for memory): as many as the 168 VGPRs of 3 waves per SIMD hold without a
// spill for the tile shape, one fewer in the ensemble kernel, which also holds its
// work-map entry (4096^2, S = 10: 2.17 ms per step with 7, 2.10 ms with 10)
template <int TI, int TJ, bool BATCH>
constexpr int fp64_own()
{
    return (TI == 8 && TJ == 32) || (TI == 16 && TJ == 16) ? (BATCH ? 9 : 10) : (BATCH && TI == 4 && TJ == 32 ? 8 : 9);
}
// FP32 variant for many classes: classes whose ch_k and M_k of the own cell stay in
// registers (within the 128-VGPR budget of 4 waves per SIMD)
constexpr int kFp32WideOwn = 10;

// Size classes whose upwind concentrations have LDS planes of their own (the first
// chunk); in FP64 the classes above share the planes of the velocities, TOTSED, DEP,
// the exchanged faces and the first chunk, which are dead by the time the class loop
// reaches them.  An FP64 tile of 8 x 32 with 10 class planes fits 3 blocks per CU.
// FP32 keeps every class in its own planes: its second-chunk variant needs more
// registers than 5 waves per SIMD allow and measured slower than the single chunk.
template <typename R>
constexpr int kLdsClasses = sizeof(R) == sizeof(double) ? 10 : kMaxClasses;

template <typename R, int TI, int TJ>
struct TileGeom {
    static constexpr int NT = TI * TJ;
    static constexpr int EI = TI + 2, EJ = TJ + 2, EXT = EI * EJ;  // 1-cell halo
    static constexpr int ZI = TI + 4, ZJ = TJ + 4, ZEXT = ZI * ZJ;  // 2-cell halo (bed)
    // face fluxes crossing between waves: the rows of a wave (RW = 64 / TJ) pass their
    // x faces within the wave; LDS holds the x faces at the first row of each wave (NXW
    // rows of TJ faces: the tile's low edge, then each wave's last row) and the y faces
    // at the tile's low edge (TI)
    static constexpr int RW = 64 / TJ, NXW = TI / RW;
    static constexpr int NFXW = NXW * TJ, NFYE = TI;
    static constexpr int KS = kLdsClasses<R>;
    // the second chunk (classes KS .. S-1) overlays everything from the u plane on
    static_assert((4 + KS) * EXT + 3 * NFXW + 3 * NFYE >= (kMaxClasses - KS) * EXT, "second class chunk");
    static_assert(64 % TJ == 0 && TI % RW == 0, "rows of a tile fill whole waves");
    static size_t lds_bytes(int S)
    {
        const int s0 = S < KS ? S : KS;
        return sizeof(double) * 2 * ZEXT + sizeof(R) * ((size_t)(5 + s0) * EXT + 3 * NFXW + 3 * NFYE);
    }
};


// The fused step for one tile: ghosts, face fluxes, sources, update.
//
// Arithmetic follows SWEHR.c operation for operation except where a division or a
// transcendental call is replaced by an identity exact in real arithmetic (a
// reciprocal computed once and multiplied, 1/sqrt via rsqrt, c* = |d|/4 in the
// Riemann solver, the class-invariant source factors hoisted out of the class loop);
// those change results at the rounding level only.
// SPLIT: the variant for many classes.  In FP64 it serves any S up to kMaxClasses by
// staging the classes beyond kLdsClasses in a second LDS chunk (without it S <=
// kLdsClasses).  In FP32, where a tile's LDS for that many classes admits 4 blocks
// per CU instead of 5, it is compiled for 4 waves per SIMD and spends the registers
// this frees on the own cell's ch_k and M_k of the first classes, which are then read
// once (with the halo staging) instead of again in the class loop.
// LOOP: the body of a loop over tiles (the persistent ensemble kernel).  The thread
// index is then passed through an empty asm statement, so that the compiler cannot
// hoist the per-thread index arithmetic out of the loop, where it would stay live
// through every tile's step and crowd its registers.
// PF: more classes than one batch of halo loads (kHaloBatch): the next batch's first
// classes are loaded before the current batch is converted (a variant of its own, so
// that the kernel for fewer classes keeps its code).
template <typename R, int TI, int TJ, bool SPLIT, int KOWN, bool LOOP = false, bool PF = false>
__device__ __forceinline__ void step_tile(const StepArgs<R> &A, int tile, unsigned char *smem_raw)
{
    using T = TileGeom<R, TI, TJ>;
    constexpr int NT = T::NT, EJ = T::EJ, EXT = T::EXT, ZJ = T::ZJ, ZEXT = T::ZEXT;
    constexpr int RW = T::RW, NFXW = T::NFXW, NFYE = T::NFYE, KS = T::KS;
    glob<DevCtrl> *C = gmem(A.ctrl);
    if (!step_active(C)) return;

    const int ti = tile / A.tiles_j, tj = tile - (tile / A.tiles_j) * A.tiles_j;
    const int i0 = ti * TI, j0 = tj * TJ;

    const long long steps = C->steps;
    const double t = C->t;
    const double smax = ctrl_smax(C, steps);
    const double dtd = step_dt(smax, A.cstable, A.dx, A.maxstep, t, A.tend);
    const double Rd = C->rain_now;
    // the end of this step, decided by one thread of the instance's first tile while the
    // step runs; it is committed once every tile has stored its cells
    if (tile == 0 && threadIdx.x == 0) decide_step(C, A, smax);
    const int p = (int)(steps & 1), q = p ^ 1, p3 = (int)(steps % 3);
    const int nx = A.nx, ny = A.ny, py = A.py, S = A.S;
    const long long plane = A.plane;
    const glob<R> *hin = gmem(A.h[p]), *huin = gmem(A.hu[p]), *hvin = gmem(A.hv[p]), *chin = gmem(A.ch[p]);
    const glob<R> *depin = gmem(A.dep[p]);
    const glob<double> *zc = gmem(A.z[p3]), *zp = gmem(A.z[(p3 + 2) % 3]);
    glob<R> *const hout = gmem(A.h[q]), *const huout = gmem(A.hu[q]), *const hvout = gmem(A.hv[q]);
    glob<R> *const chout = gmem(A.ch[q]), *const depout = gmem(A.dep[q]);
    glob<double> *const zout = gmem(A.z[(p3 + 1) % 3]);
    glob<R> *const Mp = gmem(A.M), *const totmp = gmem(A.totm), *const maxvelp = gmem(A.maxvel);
    glob<R> *const maxdepthp = gmem(A.maxdepth);
    glob<double> *const vinfp = gmem(A.vinf);

    double *s_z = reinterpret_cast<double *>(smem_raw), *s_zp = s_z + ZEXT;
    R *s_h = reinterpret_cast<R *>(s_zp + ZEXT);
    // 2^(j/32) for the FP64 exponentials of phase 3, then kLogTab for its logarithms
    __shared__ double s_tab[32 + 69];
    double *const s_exp2 = s_tab, *const s_log = s_tab + 32;
    R *s_u = s_h + EXT, *s_v = s_u + EXT, *s_ts = s_v + EXT, *s_dep = s_ts + EXT;
    R *s_fxw = s_dep + EXT, *s_fye = s_fxw + 3 * NFXW;
    R *s_sed = s_fye + 3 * NFYE;

    const R hfilm = R(kHfilm);
    const R inv_rhos = R(1.0 / kRhos);
    int tid = threadIdx.x;
    if constexpr (LOOP) asm volatile("" : "+v"(tid));
    // The tables go to LDS by direct (LDS-DMA) loads of the first wave, one dword per
    // lane (lane L to byte 4L after the instruction's base): no registers and no wait
    // until the barrier of phase 1 (a copy through registers waits on each load before
    // the wave issues any of the tile's loads).
    // (The prefetching variant copies them through registers: with the direct loads its
    // register allocation spills.)
    if constexpr (sizeof(R) == sizeof(double) && PF) {
        if (tid < 32) s_exp2[tid] = kExp2Frac[tid];
        if (tid < 69) s_log[tid] = (&kLogTab[0][0])[tid];
    } else if constexpr (sizeof(R) == sizeof(double)) {
        if (tid < 64) {
            typedef __attribute__((address_space(3))) void lds_void;
            const unsigned *e2 = reinterpret_cast<const unsigned *>(kExp2Frac);
            const unsigned *lg = reinterpret_cast<const unsigned *>(&kLogTab[0][0]);
            __builtin_amdgcn_global_load_lds(e2 + tid, (lds_void *)s_tab, 4, 0, 0);  // 32 doubles
#pragma unroll
            for (int c = 0; c < 3; c++)  // 69 doubles = 138 dwords
                if (c * 64 + tid < 138)
                    __builtin_amdgcn_global_load_lds(lg + c * 64 + tid, (lds_void *)((char *)s_log + c * 256), 4, 0, 0);
        }
    }

    // Own-cell values of the update, loaded first so that their latency overlaps the
    // halo staging below (every load of a phase is issued before the first wait).
    const int il = tid / TJ, jl = tid - (tid / TJ) * TJ;
    const int gi = i0 + il, gj = j0 + jl;
    const bool own = gi < nx && gj < ny;
    // cell offsets within a plane are 32-bit unsigned (a plane holds < 2^32 values on any
    // GPU memory): with the plane base in scalar registers every access is base + one
    // shared offset register, with no per-access 64-bit address arithmetic
    const unsigned c = (unsigned)gi * (unsigned)py + (unsigned)gj;
    const unsigned cR = c * (unsigned)sizeof(R), c8 = c * 8u;  // byte offsets
    // the same for loads of phase 3, which threads outside the domain take part in when
    // the classes come in two chunks (they read cell 0 and store nothing)
    const unsigned cRL = own ? cR : 0u, c8L = own ? c8 : 0u;
    // own-cell values: threads outside the domain read cell 0 (and never store)
    R hu_o, hv_o, totm_o, mv_o, md_o;
    double vinf_o, ks_o = A.ks;
    const int lane = tid & 63, wid = tid >> 6;
    // ch_k of the own cell for the first KO classes, loaded with the halo staging (the
    // same lines, so the requests merge in L2) and kept in registers for phase 3: read
    // again there they would come from memory, the halo having left L2 by then
    // (in FP32 only in the SPLIT variant: the register budget of 5 waves per SIMD has
    // no room)
    constexpr int KO = sizeof(R) == sizeof(double) ? (SPLIT ? kSplitOwn : KOWN) : (SPLIT ? kFp32WideOwn : 0);
    R ch_own[KO > 0 ? KO : 1];
#pragma unroll
    for (int k = 0; k < KO; k++) ch_own[k] = k < S ? at(chin + k * plane, cRL) : R(0);
    hu_o = at(huin, cRL);
    hv_o = at(hvin, cRL);
    totm_o = ld_stream(totmp, cRL);
    mv_o = ld_stream(maxvelp, cRL);
    md_o = ld_stream(maxdepthp, cRL);
    vinf_o = ld_stream(vinfp, c8L);
    if (A.ks_raster) ks_o = ld_stream(gmem(A.ks_raster), c8L);
    // M_k of the own cell for the first MO classes, loaded here with the other own-cell
    // values so that their latency hides behind phases 1-2 (one class ahead inside the
    // class loop it did not: 0.428 -> 0.403 ms on the 2048^2, S = 7 FP64 step); the
    // classes above MO load M_k one class ahead
    constexpr int MO = sizeof(R) == sizeof(double) ? (SPLIT ? kSplitMOwn : KO) : (SPLIT ? kFp32WideOwn : kFp32MOwn);
    R m_own[MO > 0 ? MO : 1];
#pragma unroll
    for (int k = 0; k < MO; k++) m_own[k] = k < S ? ld_stream(Mp + k * plane, cRL) : R(0);
    R m_a = MO < S ? ld_stream(Mp + MO * plane, cRL) : R(0);

    // ---- phase 1: ghosts and the 1-cell halo of the conserved state (SWEHR.c:816-1035)
    constexpr int EIT = (EXT + NT - 1) / NT, ZIT = (ZEXT + NT - 1) / NT;
    double zc_r[ZIT], zp_r[ZIT];
#pragma unroll
    for (int r = 0; r < ZIT; r++) {
        const int e = tid + r * NT;
        if (e < ZEXT) {
            const int zi = e / ZJ, zj = e - (e / ZJ) * ZJ;
            int ggi = i0 - 2 + zi, ggj = j0 - 2 + zj;
            ggi = ggi < 0 ? 0 : (ggi >= nx ? nx - 1 : ggi);
            ggj = ggj < 0 ? 0 : (ggj >= ny ? ny - 1 : ggj);
            const unsigned cz = (unsigned)ggi * (unsigned)py + (unsigned)ggj;
            zc_r[r] = at(zc, cz * 8u);
            zp_r[r] = at(zp, cz * 8u);
        }
    }
    // ext cell r of this thread: clamped source cell, and its state
    unsigned c_r[EIT];
    R h_r[EIT], hu_r[EIT], hv_r[EIT], dep_r[EIT];
#pragma unroll
    for (int r = 0; r < EIT; r++) {
        const int e = tid + r * NT;
        c_r[r] = 0;
        if (e < EXT) {
            const int ei = e / EJ, ej = e - (e / EJ) * EJ;
            const int ggi = i0 - 1 + ei, ggj = j0 - 1 + ej;
            const int ci = ggi < 0 ? 0 : (ggi >= nx ? nx - 1 : ggi);
            const int cj = ggj < 0 ? 0 : (ggj >= ny ? ny - 1 : ggj);
            c_r[r] = ((unsigned)ci * (unsigned)py + (unsigned)cj) * (unsigned)sizeof(R);  // bytes
            h_r[r] = at(hin, c_r[r]);
            hu_r[r] = at(huin, c_r[r]);
            hv_r[r] = at(hvin, c_r[r]);
            dep_r[r] = at(depin, c_r[r]);
        }
    }
    // suspended mass of the ext cells, up to KC classes per batch of loads
    constexpr int KC = kHaloBatch<R>;
    R ch_r[KC][EIT];
#pragma unroll
    for (int kk = 0; kk < KC; kk++)
#pragma unroll
        for (int r = 0; r < EIT; r++)
            if (kk < S && tid + r * NT < EXT) ch_r[kk][r] = at(chin + kk * plane, c_r[r]);

    // boundary rules and velocities of the ext cells
    const InflowState in = ctrl_inflow(C);
    R ih_r[EIT], ts_r[EIT], sdin_r[EIT];
    bool wet_r[EIT], ghost_r[EIT];
#pragma unroll
    for (int r = 0; r < EIT; r++) {
        const int e = tid + r * NT;
        ih_r[r] = R(0);
        ts_r[r] = R(0);
        sdin_r[r] = R(0);
        wet_r[r] = false;
        ghost_r[r] = false;
        if (e < EXT) {
            const int ei = e / EJ, ej = e - (e / EJ) * EJ;
            const int ggi = i0 - 1 + ei, ggj = j0 - 1 + ej;
            R h = h_r[r], hu = hu_r[r], hv = hv_r[r];
            if (ggi < 0 && A.bndry[0] == 1) hu = -hu;
            if (ggi >= nx && A.bndry[2] == 1) hu = -hu;
            if (ggj < 0 && A.bndry[1] == 1) hv = -hv;
            const bool inflow_ghost = ggj >= ny && A.inflow;
            if (ggj >= ny && !A.inflow && A.bndry[3] == 1) hv = -hv;
            if (inflow_ghost) {
                h = R(in.h);
                hu = R(0);
                hv = R(in.hv);
                sdin_r[r] = R(in.c) / h;
            }
            R u = R(0), v = R(0);
            if (h > hfilm) {
                const R ih = mrcp(h);
                ih_r[r] = ih;
                u = hu * ih;
                v = hv * ih;
            }
            wet_r[r] = h > hfilm;
            ghost_r[r] = inflow_ghost;
            s_h[e] = h;
            s_u[e] = u;
            s_v[e] = v;
            s_dep[e] = dep_r[r];
        }
    }
    // PF: the next batch's first KP classes are loaded before this batch is converted,
    // so that their latency overlaps the conversion (the registers of the depth and
    // momenta above are free by then)
    constexpr int KP = PF ? 2 : 0;
    R ch_p[KP > 0 ? KP : 1][EIT];
    for (int k0 = 0; k0 < S; k0 += KC) {
        if (k0 > 0) {
#pragma unroll
            for (int kk = 0; kk < KC; kk++)
#pragma unroll
                for (int r = 0; r < EIT; r++)
                    if (k0 + kk < S && tid + r * NT < EXT) {
                        if (kk < KP) ch_r[kk][r] = ch_p[kk < KP ? kk : 0][r];
                        else ch_r[kk][r] = at(chin + (k0 + kk) * plane, c_r[r]);
                    }
        }
        if (KP > 0 && k0 + KC < S) {
#pragma unroll
            for (int kk = 0; kk < KP; kk++)
#pragma unroll
                for (int r = 0; r < EIT; r++)
                    if (k0 + KC + kk < S && tid + r * NT < EXT) ch_p[kk][r] = at(chin + (k0 + KC + kk) * plane, c_r[r]);
        }
#pragma unroll
        for (int kk = 0; kk < KC; kk++) {
            if (k0 + kk >= S) break;
#pragma unroll
            for (int r = 0; r < EIT; r++) {
                const int e = tid + r * NT;
                if (e < EXT) {
                    R sd = R(0);
                    if (ghost_r[r]) {
                        sd = sdin_r[r];
                    } else if (wet_r[r]) {
                        sd = ch_r[kk][r] * ih_r[r];
                        ts_r[r] += sd * inv_rhos;
                        if (sd < R(-1e-5)) sd = R(0);
                    }
                    if (k0 + kk < KS) s_sed[(k0 + kk) * EXT + e] = sd;
                }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < EIT; r++)
        if (tid + r * NT < EXT) s_ts[tid + r * NT] = ts_r[r];
#pragma unroll
    for (int r = 0; r < ZIT; r++) {
        const int e = tid + r * NT;
        if (e < ZEXT) {
            s_z[e] = zc_r[r];
            s_zp[e] = zp_r[r];
        }
    }
    if constexpr (sizeof(R) == sizeof(double) && !PF) __builtin_amdgcn_s_waitcnt(0);  // the tables' LDS-DMA loads
    __syncthreads();

    const double odxd = A.inv_dx;
    const R g = R(kG);
    // GZ of the previous step's flux pass at global cell (gi, gj) (SWEHR.c:1279, read
    // lagged at 1135 and 1203); gi, gj are inside the domain.
    auto gz_lag = [&](int gi, int gj) -> R {
        const int zi = gi - (i0 - 2), zj = gj - (j0 - 2);
        const R a = R(slope_x<ZJ>(s_zp, zi, zj, gi, nx, odxd));
        const R bb = R(slope_y<ZJ>(s_zp, zi, zj, gj, ny, odxd));
        return g * mrsqrt(a * a + bb * bb + R(1));
    };

    // ---- phase 2: face fluxes of the tile, each computed once.  A cell's own faces
    // i+1/2 and j+1/2 (sharing its lagged GZ and celerity) stay in its thread's
    // registers.  Its j-1/2 face is the j+1/2 face of the lane below, in the same row;
    // its i-1/2 face the i+1/2 face of the row above, in the same wave except at a
    // wave's first row.  LDS holds what crosses waves: the i+1/2 faces of each wave's
    // last row and the faces of the tile's low edges.
    const int rw = il % RW;
    R fxr[3] = {R(0), R(0), R(0)}, fyr[3] = {R(0), R(0), R(0)};
    if (own) {
        const R gz = gz_lag(gi, gj);
        const int e = (il + 1) * EJ + jl + 1;
        const R cl = msqrt(gz * s_h[e]);
        face_flux_v(s_h[e], s_h[e + EJ], s_u[e], s_u[e + EJ], s_v[e], s_v[e + EJ], cl, msqrt(gz * s_h[e + EJ]), gz,
                    fxr[0], fxr[1], fxr[2]);
        // y face: normal velocity v (SWEHR.c:1203), fluxes stored swapped (SWEHR.c:1216-1218)
        R f0, f1, f2;
        face_flux_v(s_h[e], s_h[e + 1], s_v[e], s_v[e + 1], s_u[e], s_u[e + 1], cl, msqrt(gz * s_h[e + 1]), gz, f0,
                    f1, f2);
        fyr[0] = f0;
        fyr[1] = f2;  // hu flux (tangential)
        fyr[2] = f1;  // hv flux (normal)
        if (rw == RW - 1 && il < TI - 1)
#pragma unroll
            for (int m = 0; m < 3; m++) s_fxw[(m * (NFXW / TJ) + (il + 1) / RW) * TJ + jl] = fxr[m];
    }
    {
        // faces on the tile's low edges belong to the cell outside (or to the ghost, which
        // takes the GZ of the first cell, SWEHR.c:1054-1088): TJ x faces then TI y faces,
        // one per thread from thread 0, in one pass of the first wave(s)
        const int ef = tid;
        if (ef < TJ + TI) {
            const bool isx = ef < TJ;
            const int fj = isx ? ef : 0, fi = isx ? 0 : ef - TJ;
            const bool ok = isx ? (j0 + fj < ny) : (i0 + fi < nx);
            if (ok) {
                const int el = isx ? fj + 1 : (fi + 1) * EJ, er = isx ? el + EJ : el + 1;
                const R *un = isx ? s_u : s_v, *ut = isx ? s_v : s_u;
                const R gz = isx ? gz_lag(i0 > 0 ? i0 - 1 : 0, j0 + fj) : gz_lag(i0 + fi, j0 > 0 ? j0 - 1 : 0);
                R f0, f1, f2;
                face_flux_v(s_h[el], s_h[er], un[el], un[er], ut[el], ut[er], msqrt(gz * s_h[el]), msqrt(gz * s_h[er]),
                            gz, f0, f1, f2);
                if (isx) {
                    s_fxw[fj] = f0;
                    s_fxw[NFXW + fj] = f1;
                    s_fxw[2 * NFXW + fj] = f2;
                } else {
                    s_fye[fi] = f0;
                    s_fye[NFYE + fi] = f2;      // hu flux (tangential)
                    s_fye[2 * NFYE + fi] = f1;  // hv flux (normal), SWEHR.c:1216-1218
                }
            }
        }
    }
    __syncthreads();
    // the faces i-1/2 and j-1/2 (lane exchanges: every thread takes part)
    R fxl[3], fyl[3];
#pragma unroll
    for (int m = 0; m < 3; m++) {
        const R up = __shfl_up(fxr[m], TJ, 64), dn = lane_below(fyr[m]);
        fxl[m] = rw > 0 ? up : s_fxw[(m * (NFXW / TJ) + il / RW) * TJ + jl];
        fyl[m] = jl > 0 ? dn : s_fye[m * NFYE + il];
    }

    // ---- phase 3: sources and update of the tile's own cells.  With more than KS
    // classes every thread of the block takes part (the second chunk of classes is
    // staged between two barriers); threads outside the domain store nothing.
    double bud[NB_TERMS] = {0, 0, 0, 0, 0};
    constexpr bool TWO_CHUNK = SPLIT && KS < kMaxClasses;
    if (own || (TWO_CHUNK && S > KS)) {
        const int e = (il + 1) * EJ + (jl + 1);
        const int zi = il + 2, zj = jl + 2;
        // the cell's neighbourhood from the tile's LDS planes; neighbour slopes and DEP
        // at the domain's edges are the cell's own (SWEHR.c:1292-1312)
        const int du = gi < nx - 1 ? 1 : 0, dd = gi > 0 ? 1 : 0;
        const int ju = gj < ny - 1 ? 1 : 0, jd = gj > 0 ? 1 : 0;
        CellIn<R> ci;
        ci.h = s_h[e];
        ci.hu = hu_o;
        ci.hv = hv_o;
        ci.ts = s_ts[e];
        ci.ts_xm = s_ts[e - EJ];
        ci.ts_xp = s_ts[e + EJ];
        ci.ts_ym = s_ts[e - 1];
        ci.ts_yp = s_ts[e + 1];
        ci.dep = s_dep[e];
        ci.dep_xu = s_dep[e + du * EJ];
        ci.dep_xd = s_dep[e - dd * EJ];
        ci.dep_yu = s_dep[e + ju];
        ci.dep_yd = s_dep[e - jd];
        ci.sxn = R(slope_x<ZJ>(s_z, zi, zj, gi, nx, odxd));
        ci.syo = R(slope_y<ZJ>(s_zp, zi, zj, gj, ny, odxd));
        ci.syn = R(slope_y<ZJ>(s_z, zi, zj, gj, ny, odxd));
        ci.sx_up = du ? R(slope_x<ZJ>(s_zp, zi + 1, zj, gi + 1, nx, odxd)) : ci.sxn;
        ci.sx_dn = dd ? R(slope_x<ZJ>(s_z, zi - 1, zj, gi - 1, nx, odxd)) : ci.sxn;
#pragma unroll
        for (int m = 0; m < 3; m++) {
            ci.fxl[m] = fxl[m];
            ci.fxr[m] = fxr[m];
            ci.fyl[m] = fyl[m];
            ci.fyr[m] = fyr[m];
        }
        ci.totm_o = totm_o;
        ci.mv_o = mv_o;
        ci.md_o = md_o;
        ci.vinf_o = vinf_o;
        ci.ks_o = ks_o;
        ci.zn = s_z[zi * ZJ + zj];
        Cell<R> X;
        X.gi = gi;
        X.gj = gj;
        X.own = own;
        const R syn = ci.syn;
        cell_begin(X, A, ci, dtd, Rd, steps, c, s_exp2, s_log, c8L, [&](R &up, R &dn) {
            up = ju ? R(slope_y<ZJ>(s_zp, zi, zj + 1, gj + 1, ny, odxd)) : syn;
            dn = jd ? R(slope_y<ZJ>(s_z, zi, zj - 1, gj - 1, ny, odxd)) : syn;
        });

        // upwind cell of each face (SWEHR.c:1152-1157, 1219-1224): class-invariant
        const int uxl = X.fx0l > R(0) ? e - EJ : e, uxr = X.fx0r > R(0) ? e : e + EJ;
        const int uyl = X.fy0l > R(0) ? e - 1 : e, uyr = X.fy0r > R(0) ? e : e + 1;
        // class k from its suspended mass ch_k, deposited mass m_k and the LDS plane sk
        // of its concentrations
        auto update_class = [&](int k, R ch_k, R m_k, const R *sk) {
            cell_class(X, A, k, ch_k, m_k, X.fx0l * sk[uxl], X.fx0r * sk[uxr], X.fy0l * sk[uyl], X.fy0r * sk[uyr],
                       s_exp2, chout, Mp, plane, cR);
        };
        // first chunk, classes 0 .. S0-1.  Below KO: unrolled, ch_k from registers; from
        // KO on ch_k comes from memory, one class ahead.  M_k from registers below MO.
        const int S0 = S < KS ? S : KS;
#pragma unroll
        for (int k = 0; k < KO; k++)
            if (k < S0) update_class(k, ch_own[k], m_own[k], s_sed + k * EXT);
        const int kf = KO < S0 ? KO : KS;  // first class whose ch_k is read below
        R ch_a = kf < S ? at(chin + kf * plane, cRL) : R(0);
#pragma unroll
        for (int k = KO; k < MO; k++) {
            if (k < S0) {
                const R ch_k = ch_a;
                if (k + 1 < S) ch_a = at(chin + (k + 1) * plane, cRL);
                update_class(k, ch_k, m_own[k], s_sed + k * EXT);
            }
        }
        for (int k = (KO > MO ? KO : MO); k < S0; k++) {
            const R ch_k = ch_a, m_k = m_a;
            if (k + 1 < S) {
                ch_a = at(chin + (k + 1) * plane, cRL);
                m_a = ld_stream(Mp + (k + 1) * plane, cRL);
            }
            update_class(k, ch_k, m_k, s_sed + k * EXT);
        }
        // second chunk, classes KS .. S-1: their upwind concentrations are staged as in
        // phase 1 over the planes from s_u on, which no thread reads after this point
        if (TWO_CHUNK && S > KS) {
            __syncthreads();
            R *const s_sed1 = s_u;
            constexpr int KB = 2;  // classes per batch of halo loads
            for (int k0 = KS; k0 < S; k0 += KB) {
                R chb[KB][EIT];
                unsigned cb[EIT];
#pragma unroll
                for (int r = 0; r < EIT; r++) {
                    const int e = tid + r * NT;
                    cb[r] = 0;
                    if (e < EXT) {
                        const int ei = e / EJ, ej = e - (e / EJ) * EJ;
                        const int ggi = i0 - 1 + ei, ggj = j0 - 1 + ej;
                        const int ci = ggi < 0 ? 0 : (ggi >= nx ? nx - 1 : ggi);
                        const int cj = ggj < 0 ? 0 : (ggj >= ny ? ny - 1 : ggj);
                        cb[r] = ((unsigned)ci * (unsigned)py + (unsigned)cj) * (unsigned)sizeof(R);
                    }
                }
#pragma unroll
                for (int kk = 0; kk < KB; kk++)
#pragma unroll
                    for (int r = 0; r < EIT; r++)
                        if (k0 + kk < S && tid + r * NT < EXT) chb[kk][r] = at(chin + (k0 + kk) * plane, cb[r]);
#pragma unroll
                for (int r = 0; r < EIT; r++) {
                    const int e = tid + r * NT;
                    if (e < EXT) {
                        const int ggj = j0 - 1 + (e - (e / EJ) * EJ);
                        const R hx = s_h[e];
                        const bool ghost = ggj >= ny && A.inflow, wet = hx > hfilm;
                        const R ih = wet ? mrcp(hx) : R(0);
#pragma unroll
                        for (int kk = 0; kk < KB; kk++) {
                            if (k0 + kk < S) {
                                R sd = R(0);
                                if (ghost) {
                                    sd = R(in.c) / hx;
                                } else if (wet) {
                                    sd = chb[kk][r] * ih;
                                    if (sd < R(-1e-5)) sd = R(0);
                                }
                                s_sed1[(k0 + kk - KS) * EXT + e] = sd;
                            }
                        }
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (int k = KS; k < MO; k++) {
                if (k < S) {
                    const R ch_k = ch_a;
                    if (k + 1 < S) ch_a = at(chin + (k + 1) * plane, cRL);
                    update_class(k, ch_k, m_own[k], s_sed1 + (k - KS) * EXT);
                }
            }
            for (int k = (MO > KS ? MO : KS); k < S; k++) {
                const R ch_k = ch_a, m_k = m_a;
                if (k + 1 < S) {
                    ch_a = at(chin + (k + 1) * plane, cRL);
                    m_a = ld_stream(Mp + (k + 1) * plane, cRL);
                }
                update_class(k, ch_k, m_k, s_sed1 + (k - KS) * EXT);
            }
        }

        cell_end(X, A, dtd, s_exp2, s_log, hout, huout, hvout, chout, depout, zout, Mp, totmp, maxvelp, maxdepthp,
                 plane, cR, c8);
#pragma unroll
        for (int k = 0; k < NB_TERMS; k++) bud[k] = X.bud[k];
    }

    // ---- budget partials: each wave stores its nonzero sums of this step as the
    // pending partials of its own slot (tile, wave); the wave-speed kernel of the next
    // step folds them into the slot's compensated running sums (fold_budget) and clears
    // them.  No barrier and no read-modify-write at the end of the tile.
    {
        double mine = 0;
#pragma unroll
        for (int k = 0; k < NB_TERMS; k++) {
            double v = own ? bud[k] : 0.0;
            if (__ballot(v != 0.0) != 0ull) v = wave_sum(v);  // most terms are zero in most tiles
            if (lane == k) mine = v;
        }
        if (lane < NB_TERMS && mine != 0.0)
            gmem(A.bud)[(size_t)A.bud_slots * NB_TERMS * 2 + ((size_t)tile * (NT / 64) + wid) * NB_TERMS + lane] = mine;
    }
}

template <typename R, int TI, int TJ, bool SPLIT, bool PF = false>
__global__ __launch_bounds__(TI *TJ, (kStepWaves<R, SPLIT>)) void step_kernel(StepArgs<R> A)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    // XCD-aware tile order: consecutive block ids land on different XCDs, so give
    // each XCD a contiguous band of tiles (their halos then share that XCD's L2).
    const int b = blockIdx.x;
    const int tile = (b & 7) * A.blocks_per_xcd + (b >> 3);
    if (tile >= A.nblocks) return;
    step_tile<R, TI, TJ, SPLIT, fp64_own<TI, TJ, false>(), false, PF>(A, tile, smem_raw);
}

// --------------------------------------------------------------------------- K3

// The standalone commit: the control kernel after a step's kernels (in the step graph
// after the last step; every step in the ensemble modes and where steps are launched
// one by one).  Commits the pending decision unless the wave-speed pass of a following
// step already did, and clears the pending mark.
template <typename R>
__device__ __forceinline__ void control_body(const StepArgs<R> &A)
{
    glob<DevCtrl> *C = gmem(A.ctrl);
    __shared__ int s_todo;
    if (threadIdx.x == 0) s_todo = C->pending ? (C->steps + 1 == C->next.steps ? 2 : 1) : 0;
    __syncthreads();
    const int todo = s_todo;
    if (todo == 0) return;
    if (todo == 2) commit_step(A, C);
    if (threadIdx.x == 0) C->pending = 0;
}

template <typename R>
__global__ __launch_bounds__(256) void control_kernel(StepArgs<R> A)
{
    control_body(A);
}

// --------------------------------------------------------------------------- ensembles

// Work map of an ensemble launch: entry k -> (instance, tile of that instance).  Block b
// takes entry (b % 8) * per_xcd + b / 8, so each XCD works through a contiguous run
// of entries (neighbouring tiles of one instance share that XCD's L2).
struct TileRef {
    int inst, tile;
};

// K1 of the ensemble on strip tiles: one block per (instance, strip tile) entry of the
// wave-speed map.  Each block also folds its share of its instance's budget partials
// (the strip tiles of an instance split the slots between them).
template <typename R>
__global__ __launch_bounds__(256) void batch_wave_strip_kernel(const StepArgs<R> *args, const TileRef *map, int n,
                                                                int per_xcd)
{
    const int b = blockIdx.x;
    const int k = (b & 7) * per_xcd + (b >> 3);
    if (k >= n) return;
    const TileRef r = map[k];
    const StepArgs<R> &A = args[r.inst];
    glob<DevCtrl> *C = gmem(A.ctrl);
    const K1View view = k1_view(C);
    if (!view.active) return;
    const R s = wave_speed_strip(A, view, r.tile);
    const long long nterms = (long long)A.bud_slots * NB_TERMS;
    const long long share = (nterms + k1_tiles(A.nx, A.ny) - 1) / k1_tiles(A.nx, A.ny);
    const long long g0 = (long long)r.tile * share, g1 = g0 + share < nterms ? g0 + share : nterms;
    for (long long g = g0 + threadIdx.x; g < g1; g += blockDim.x) fold_budget(gmem(A.bud), nterms, g);
    block_max_to((unsigned long long *)C->smax_slots[view.steps & 1], r.tile, (double)s);
}

template <typename R, int TI, int TJ, bool SPLIT>
__global__ __launch_bounds__(TI *TJ, (kStepWaves<R, SPLIT>)) void batch_step_kernel(const StepArgs<R> *__restrict__ args,
                                                                           const TileRef *__restrict__ map, int n,
                                                                           int per_xcd)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int b = blockIdx.x;
    const int k = (b & 7) * per_xcd + (b >> 3);
    if (k >= n) return;
    const TileRef r = map[k];
    step_tile<R, TI, TJ, SPLIT, fp64_own<TI, TJ, true>()>(args[r.inst], r.tile, smem_raw);
}

template <typename R>
__global__ __launch_bounds__(256) void batch_control_kernel(const StepArgs<R> *args, const int *insts, int n)
{
    if ((int)blockIdx.x >= n) return;
    control_body(args[insts[blockIdx.x]]);
}

}  // namespace swehr
