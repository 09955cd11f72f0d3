/*
 * swehr.h - C ABI of the MI355X-native SWEHR coupled time step.
 *
 * The reference (SWEHR.c) has no library seam: the time step is written inline in
 * main() (SWEHR.c:808-1699) and every parameter either comes from the five input
 * files (SWEHR.c:468-522) or is hard-coded (SWEHR.c:236-246, 617-650).  This header
 * is the seam: a host driver keeps main()'s control flow (parse, initial condition,
 * outputs) and calls these entry points for the loop body (boundary fill, flux and
 * source pass, CFL dt, update pass; SWEHR.c:810-1591) plus the outlet/probe sampling
 * (SWEHR.c:1593-1676) that the loop does after each step.
 *
 * Conventions
 *  - Plain C types only.  Host arrays are borrowed for the duration of a call.
 *  - Rasters are "i-major": value (i,j), i=1..nx across, j=1..ny downslope, sits at
 *    [(i-1)*ny + (j-1)], the order topoin.txt is read in (SWEHR.c:223-227).
 *  - Every entry point returns an int status (enum swehr_status); the reference's
 *    exit(1)/print-and-stop paths (SWEHR.c:29-36, 1413) become codes.
 *  - One host thread per handle; different handles may be used from different threads.
 */
#ifndef SWEHR_H
#define SWEHR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SWEHR_MAX_CLASSES 16
#define SWEHR_STAGE_COLS 24   /* STAGE[1..24] columns, SWEHR.c:600, 1608-1651 */
#define SWEHR_FLOW_COLS 18    /* SAVEFLOW[1..18] columns, SWEHR.c:601, 1653-1675 */
#define SWEHR_STAGE_MAX 30000 /* records the reference can hold, SWEHR.c:600-601 */

enum swehr_status {
    SWEHR_OK = 0,
    SWEHR_ERR_CFL = 1,       /* courant > 1 (SWEHR.c:1413): the step is taken, then the run stops */
    SWEHR_ERR_NONFINITE = 2, /* non-finite value where a finite one is required */
    SWEHR_ERR_OOM = 3,       /* device or host allocation failed (SWEHR.c:116-169 exit(1)) */
    SWEHR_ERR_BAD_INPUT = 4, /* invalid sizes or parameters */
    SWEHR_ERR_HIP = 5,       /* HIP runtime error */
    SWEHR_ERR_STAGE_FULL = 6,/* more than SWEHR_STAGE_MAX outlet records */
    SWEHR_ERR_NO_DEVICE = 7, /* no usable GPU */
    SWEHR_ERR_BUSY = 8       /* capacity exhausted (serving admission) */
};

enum swehr_precision {
    SWEHR_FP64 = 0,          /* plain: all state and arithmetic in FP64, like the reference */
    SWEHR_FP32 = 1           /* state and fluxes FP32; dt, bed elevation, cumulative infiltration
                                and budgets FP64.  Increments below the float spacing of the
                                value they change are added with unbiased rounding, whose random
                                bits hash the cell and the step: runs are reproducible, and an
                                instance gives the same bits alone and in an ensemble or server */
};

/* Field ids for swehr_download().  k is the 1-based size class for per-class fields. */
enum swehr_field {
    SWEHR_F_DEPTH = 0,   /* U[1]  h (m)                         depth.txt   */
    SWEHR_F_UH = 1,      /* U[2]  h*u (m^2/s)                   uh.txt      */
    SWEHR_F_VH = 2,      /* U[3]  h*v (m^2/s)                   vh.txt      */
    SWEHR_F_CH = 3,      /* U[3+k] suspended mass (kg/m^2)      ch.txt (k=1)*/
    SWEHR_F_M = 4,       /* M[k]  deposited mass (kg/m^2)                   */
    SWEHR_F_TOPO = 5,    /* TOPO  bed elevation (m)             topoout.txt */
    SWEHR_F_VINF = 6,    /* VINF  cumulative infiltration (m)               */
    SWEHR_F_TOTM = 7,    /* TOTM  deposited mass, all classes   m.txt       */
    SWEHR_F_DEP = 8,     /* DEP   deposited thickness (m)                   */
    SWEHR_F_TOTSED = 9,  /* TOTSED volumetric concentration     c.txt       */
    SWEHR_F_SPEED = 10,  /* |VEL| (m/s)                         vel.txt     */
    SWEHR_F_MAXVEL = 11, /* running max |VEL|                   maxvel.txt  */
    SWEHR_F_MAXDEPTH = 12,/* running max h                      maxdepth.txt*/
    SWEHR_F_SED = 13,    /* SED[k] concentration (kg/m^3)                   */
    SWEHR_F_VELX = 14,   /* VEL[1] (m/s)                                    */
    SWEHR_F_VELY = 15,   /* VEL[2] (m/s)                                    */
    SWEHR_F_COUNT = 16
};

/*
 * Run parameters.  The first block is input.txt (SWEHR.c:474-504, field number in
 * brackets); the second block holds values the reference hard-codes, exposed so the
 * optional sidecar file can set them.  swehr_params_init() fills the reference values.
 */
typedef struct swehr_params {
    int nx;            /* [1] cells across slope (i)                               */
    int ny;            /* [2] cells downslope (j); j=1 is the outlet row            */
    double dx;         /* [3] grid spacing (m)                                      */
    double F;          /* [4] effective fraction of stream power in entrainment     */
    double manning0;   /* [5] read and ignored, as in the reference (SWEHR.c:480)  */
    double tend;       /* [6] simulated time (s)                                    */
    double a;          /* [7] rainsplash detachability of original soil            */
    double ad;         /* [8] rainsplash detachability of deposited sediment       */
    double h0;         /* [9] critical depth for rainsplash                         */
    double mtstar0;    /* [10] deposited mass that shields the soil                 */
    int rnum;          /* [11] rain table entries                                   */
    double rint;       /* [12] seconds between rain entries                         */
    int nclass;        /* [13] size classes S (1..SWEHR_MAX_CLASSES)                */
    double J;          /* [14] energy to entrain a unit mass                        */
    /* hard-coded in the reference */
    double cstable;    /* 0.05  target courant number   (SWEHR.c:632) */
    double maxstep;    /* 0.25  largest dt (s)          (SWEHR.c:633) */
    int bndry[4];      /* {1,0,1,1} i=1, j=1, i=nx, j=ny: 1 solid, 0 transmissive (SWEHR.c:617-621) */
    int inflow;        /* 1     water+sediment inflow at j=ny (SWEHR.c:1010-1035) */
    double roughness;  /* 0.1   Manning n               (SWEHR.c:237) */
    double theta0;     /* 0.05  initial soil moisture   (SWEHR.c:240) */
    double thetas;     /* 0.4   saturated soil moisture (SWEHR.c:241) */
    double hf;         /* 0.001 Green-Ampt suction head (SWEHR.c:242) */
    double ks;         /* 0     saturated conductivity (m/s) when no raster (SWEHR.c:239) */
    double vinf0;      /* 1e-4  initial cumulative infiltration (SWEHR.c:243) */
} swehr_params;

/* State after the last step taken by swehr_advance(). */
typedef struct swehr_step_info {
    double t;          /* simulated time (s)                                  */
    double dt;         /* dt of the last step                                  */
    double courant;    /* Smax*dt/dx of the last step (SWEHR.c:1412)           */
    double rain;       /* rain rate R of the last step (m/s)                   */
    double t_prev;     /* time at the start of the last step                   */
    int64_t steps;     /* steps taken since create                             */
    int32_t done;      /* 1 once t >= tend or the run stopped                  */
    int32_t print_due; /* 1: the last step crossed a print interval (SWEHR.c:1679); */
                       /*    the run pauses until swehr_ack_print()            */
    int32_t stage_count; /* outlet/probe records so far                        */
    int32_t status;    /* enum swehr_status of the run                          */
} swehr_step_info;

/* Mass-budget terms (FP64, compensated sums over all cells), in m^3 of water and kg of
 * sediment.  index: see swehr_budget(). */
#define SWEHR_BUDGET_N 12

typedef struct swehr_handle swehr_handle;

/* Fill *p with the reference's defaults (input fields zero). */
int swehr_params_init(swehr_params *p);

/*
 * Create a solver instance on the current GPU and upload the initial condition the
 * reference sets up (SWEHR.c:211-325, 537-803): bed from topo, zero depth, momentum
 * and sediment, VINF=vinf0, lagged geometry seeded from topo.
 *   topo   nx*ny bed elevations, i-major (topoin.txt order)
 *   ks     nx*ny saturated conductivity raster (m/s), or NULL for the scalar p->ks
 *   ds     nclass particle diameters (m)           (particlesizein.txt)
 *   pfrac  nclass fractions in the original soil   (particlepercentin.txt)
 *   rain   rnum rain rates (m/s)                    (rain.txt)
 *   precision  enum swehr_precision
 */
int swehr_create(const swehr_params *p, const double *topo, const double *ks,
                 const double *ds, const double *pfrac, const double *rain,
                 int precision, swehr_handle **out);

/*
 * Take up to max_steps time steps (each: boundary fill, fluxes and sources, CFL dt,
 * update, outlet sampling).  Stops early when t >= tend, when a print interval is
 * crossed (info->print_due; call swehr_ack_print() to continue) or on error.
 * Steps run as captured HIP graphs; the host reads back only *info.
 */
int swehr_advance(swehr_handle *h, int64_t max_steps, swehr_step_info *info);

/* Clear the print pause raised by swehr_advance(). */
int swehr_ack_print(swehr_handle *h);

/* Copy one field (enum swehr_field; k = 1-based class for per-class fields) to a host
 * raster of nx*ny doubles, i-major. */
int swehr_download(swehr_handle *h, int field, int k, double *host);

/* Copy the outlet series: stage[c*capacity + r] for column c < SWEHR_STAGE_COLS and
 * saveflow[c*capacity + r] for c < SWEHR_FLOW_COLS, r < *count (SWEHR.c:1608-1675). */
int swehr_stage(swehr_handle *h, double *stage, double *saveflow, int capacity, int *count);

/* Per-cell parameters in place of scalars the reference hard-codes (it has no such
 * rasters; SWEHR.c:237-242): suction head HF (m), theta_s - theta_0, and Manning's n0.
 * values: nx*ny, i-major like topoin.txt, or NULL to return to the scalar.  Set on a
 * single instance before stepping (refused for ensemble and server members); the
 * rasters are constant, so snapshots do not hold them. */
enum swehr_param_raster { SWEHR_R_HF = 0, SWEHR_R_DTHETA = 1, SWEHR_R_ROUGHNESS = 2, SWEHR_R_COUNT = 3 };
int swehr_set_parameter_raster(swehr_handle *h, int which, const double *values);

/* Outlet discharge at each stage record (the times of swehr_stage): the water (m^3/s)
 * and sediment (kg/s) leaving through the j = 1/2 face - the open side with the
 * reference's boundary types - in the step that ended at that time.  No counterpart in
 * the reference, which records the outlet cell's depth and velocity only (SWEHR.c:
 * 1593-1644); this is its outlet hydrograph and sedigraph.  q, qs: capacity values each
 * (either may be NULL); count: records so far. */
int swehr_discharge(swehr_handle *h, double *q, double *qs, int capacity, int *count);

/* Global budgets, FP64 compensated sums:
 *  [0] water on the surface  sum h*dx^2            [1] water infiltrated sum (VINF-vinf0)*dx^2
 *  [2] rain added            sum dt*R*dx^2 (cells)  [3] net boundary water outflow (m^3)
 *  [4] suspended sediment    sum ch*dx^2 (kg)      [5] deposited sediment sum M*dx^2 (kg)
 *  [6] bed change            sum (z-z0)*dx^2 (m^3) [7] net boundary sediment outflow (kg)
 *  [8] sediment detached from the original soil (kg) [9] water added by film clamps (m^3)
 *  [10] water from sediment-volume exchange (m^3)  [11] steps accumulated */
int swehr_budget(swehr_handle *h, double *out);

/* swehr_budget without any allocation: the same 12 terms, bit for bit, with the
 * per-block partial sums and their compensated reduction in a fixed device array of the
 * library (80 KiB per GPU, held until the process ends; at most 1024 partial blocks,
 * which every instance's budget needs at most) instead of the instance's memory and a
 * host vector.  No counterpart in the reference; it is for callers that cannot
 * allocate (tight serving loops, harnesses that forbid allocation).  All calls on one
 * GPU share that array: calls that overlap on one GPU race. */
int swehr_budget_noalloc(swehr_handle *h, double *out);

/* Run-loop options.  pause_on_print: 1 (default) pause swehr_advance() at each print
 * interval so the caller can write the reference's progress line and movie frame
 * (SWEHR.c:1679-1698); 0 never pause.  steps_per_graph: time steps captured per HIP
 * graph launch (default 16; <= 0 keeps the current value).  tile: kernel tile shape
 * id (0 = automatic choice; see swehr_tile_count()).  The step graph is captured and
 * instantiated here, so the next swehr_advance() starts stepping at once. */
int swehr_configure(swehr_handle *h, int pause_on_print, int steps_per_graph, int tile);

/* Number of tile shapes the step kernel is built for, and the shape of id (1-based):
 * cells across (i) and downslope (j) per workgroup. */
int swehr_tile_count(void);
int swehr_tile_shape(int id, int *ti, int *tj);

/* The HIP stream (hipStream_t) the instance's kernels run on, for event timing. */
int swehr_get_stream(swehr_handle *h, void **stream);

/* Diagnostic, no counterpart in the reference: take up to nsteps steps launched one by
 * one (not as a graph) with HIP events around every kernel on the instance's stream,
 * and return the average duration in milliseconds of the wave-speed, step and control
 * kernels in avg_ms[0..2]; *taken receives the number of steps run. */
int swehr_profile(swehr_handle *h, int nsteps, double *avg_ms, int *taken);

/* As swehr_profile, then the instance's state is put back (from a device copy taken
 * before the first profiled step), so the steps profiled are exactly the ones the next
 * swehr_advance takes.  Diagnostic; no counterpart in the reference. */
int swehr_profile_window(swehr_handle *h, int nsteps, double *avg_ms, int *taken);

/* Diagnostic, no counterpart in the reference: the HBM rate (GB/s of bytes that must
 * move) of a kernel that reads planes_in and writes planes_out planes of nx x ny cells
 * of elem_bytes (4 or 8) each, laid out and traversed as the step kernel's state planes
 * (rows at a pitch of a multiple of 64 cells, one row segment of 64 cells per wave,
 * nontemporal loads and stores); best of reps timed launches on the current device.
 * With the step's plane counts this is the step kernel's practical HBM ceiling. */
int swehr_hbm_stream(int nx, int ny, int elem_bytes, int planes_in, int planes_out, int reps, double *gbs);

/* Checkpoint of the instance's whole device state (snapshot/restore): *bytes is the
 * snapshot size; swehr_snapshot copies the state into host memory of that size, and
 * swehr_restore loads a snapshot taken from an instance created with the same
 * parameters.  The snapshot starts with a header of the instance's layout (grid,
 * classes, precision, Ks raster, run length, which parameter rasters are set); a
 * restore into an instance of another layout is refused with SWEHR_ERR_BAD_INPUT.
 * The constant parameter rasters (swehr_set_parameter_raster) are not part of a
 * snapshot: set them on the target instance before restoring.  The host buffer must
 * hold swehr_snapshot_bytes() bytes.  No counterpart in the reference (its final text
 * rasters are lossy). */
int swehr_snapshot_bytes(swehr_handle *h, int64_t *bytes);
int swehr_snapshot(swehr_handle *h, void *host);
int swehr_restore(swehr_handle *h, const void *host);

/* Tile-shape autotuner: time the step kernel for nsteps steps with every tile shape
 * that fits (state is saved on the device first and restored after each trial),
 * select the fastest and return it in *best; ms_per_tile (swehr_tile_count()
 * entries, may be NULL) receives the average kernel time of each shape (0 = unfit). */
int swehr_autotune(swehr_handle *h, int nsteps, int *best, double *ms_per_tile);

/* ---- ensembles: many instances (different sizes and parameters, one precision)
 * stepped together, each with its own dt; every stage of a step is one launch over
 * the tiles of all unfinished instances, and steps run as captured HIP graphs. */
typedef struct swehr_batch swehr_batch;

/* n instances; per-instance arguments as for swehr_create (ks[i] may be NULL).
 * tile: tile shape id for all instances (0 = automatic). */
int swehr_batch_create(int n, const swehr_params *params, const double *const *topo, const double *const *ks,
                       const double *const *ds, const double *const *pfrac, const double *const *rain,
                       int precision, int tile, swehr_batch **out);
/* Up to max_steps steps of every unfinished instance; *remaining receives the number
 * of instances not yet done. */
int swehr_batch_advance(swehr_batch *b, int64_t max_steps, int *remaining);
/* Instance i as a handle (swehr_download, swehr_stage, swehr_budget, and
 * swehr_advance with max_steps = 0 for its state). Owned by the batch. */
int swehr_batch_handle(swehr_batch *b, int i, swehr_handle **h);
/* Autotune the shared tile shape over the whole ensemble (as swehr_autotune). */
int swehr_batch_autotune(swehr_batch *b, int nsteps, int *best, double *ms_per_tile);
/* Diagnostic, no counterpart in the reference: nsteps ensemble steps launched one by
 * one with HIP events around each launch; avg_ms[0..2] receives the average duration
 * of the wave-speed, step and control launches (the steps advance the ensemble). */
int swehr_batch_profile(swehr_batch *b, int nsteps, double *avg_ms);
/* Persistent mode (on = 1), no counterpart in the reference: swehr_batch_advance runs
 * the ensemble's steps as one cooperative launch per 16 steps (wave-speed pass, step and
 * control of every instance separated by grid barriers) instead of a graph of three
 * launches per step; results are bit-identical.  Needs a 256-thread tile shape (1-3):
 * SWEHR_ERR_BAD_INPUT otherwise. */
int swehr_batch_set_persistent(swehr_batch *b, int on);
int swehr_batch_destroy(swehr_batch *b);

/* ---- serving: continuous batching of solve requests.  Requests wait in a queue
 * and are admitted while an instance slot is free and the server's device-memory
 * budget holds their state; admitted requests step together (ensemble launches) and
 * stay resident in HBM until released. */
typedef struct swehr_server swehr_server;

/* mem_bytes: the device-memory budget for resident states; 0 sizes it by the device's
 * free memory at creation (15/16 of it).  An admission whose allocation still fails
 * (the device is fuller than the budget says) leaves the request queued. */
int swehr_server_create(int64_t mem_bytes, int max_slots, int precision, int tile, swehr_server **out);
/* Queue a request (inputs are copied); *id identifies it.  SWEHR_ERR_BUSY if its
 * state can never fit the memory budget. */
int swehr_server_submit(swehr_server *s, const swehr_params *p, const double *topo, const double *ks,
                        const double *ds, const double *pfrac, const double *rain, int64_t *id);
/* Admit what fits, take up to steps time steps of every running request, retire the
 * finished ones.  Counts after the call go to *running, *queued, *finished. */
int swehr_server_pump(swehr_server *s, int steps, int *running, int *queued, int *finished);
/* Pop one finished request not yet reported: its id and final step info;
 * SWEHR_ERR_BUSY when there is none. */
int swehr_server_next_done(swehr_server *s, int64_t *id, swehr_step_info *info);
/* The finished request's instance (fields, stage series, budgets); owned by the server. */
int swehr_server_request(swehr_server *s, int64_t id, swehr_handle **h);
/* Free a finished request's device memory. */
int swehr_server_release(swehr_server *s, int64_t id);
/* Device memory held by admitted requests and the budget (bytes). */
int swehr_server_stats(swehr_server *s, int64_t *mem_used, int64_t *mem_budget);
/* Work stepped so far over every request the server has run, finished or running:
 * the sum of cells x size classes x steps (cell-class-steps), updated after each graph
 * of a pump.  Sampled between pumps it gives the server's throughput over any window,
 * unfinished requests included.  No counterpart in the reference (it has no serving
 * loop). */
int swehr_server_work(swehr_server *s, int64_t *cell_class_steps);
/* Evict a running request: its whole device state is checkpointed into pinned host
 * memory (as swehr_snapshot), its device memory freed, and it joins the back of the
 * queue; when admitted again it is restored and continues exactly where it stopped.
 * No counterpart in the reference (it has no serving loop). */
int swehr_server_evict(swehr_server *s, int64_t id);
/* Time slicing: with quantum > 0 steps, a pump whose queue head cannot be admitted
 * first evicts running requests that have taken at least quantum steps since their
 * admission (the longest-running first) until the head fits; 0 (the default) never
 * evicts by itself.  *evictions counts every eviction so far. */
int swehr_server_set_quantum(swehr_server *s, int64_t quantum);
int swehr_server_evictions(swehr_server *s, int64_t *evictions);
int swehr_server_destroy(swehr_server *s);

/* Last error message of the calling thread (empty when none). */
const char *swehr_last_error(void);

/* Upper bound of the device bytes of one instance of these sizes (a per-cell Ks
 * raster, 4096 rain entries and a full 30000-record outlet series included). */
int64_t swehr_instance_bytes(int nx, int ny, int nclass, int precision);
/* Exact device bytes of an instance with these parameters (what swehr_create
 * allocates; ks_raster: 1 if a per-cell Ks raster is given). */
int64_t swehr_instance_bytes_for(const swehr_params *p, int precision, int ks_raster);

/* Release the instance and its device memory. */
int swehr_destroy(swehr_handle *h);

/* Text for a status code. */
const char *swehr_strerror(int code);

#ifdef __cplusplus
}
#endif
#endif /* SWEHR_H */
