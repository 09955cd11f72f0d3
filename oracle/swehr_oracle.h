/*
 * swehr_oracle.h - TEST INFRASTRUCTURE.  CPU FP64 restatement of the SWEHR coupled
 * time step, used only as the checker by tests/, __graft_entry__.smoke() and as the
 * timed CPU baseline in bench.py.  The product (swehr_amd/) never links it.
 *
 * It follows SWEHR.c operation for operation (same libm calls, same evaluation order,
 * built without FMA contraction) so that, run on the same inputs, it reproduces the
 * reference's results bit for bit; the golden tests check exactly that against
 * outputs of the compiled reference.  Unlike the reference it is deterministic under
 * OpenMP: the geometry the reference reads one step late is recomputed from a copy of
 * the bed elevation of the previous step instead of being read while it is rewritten.
 */
#ifndef SWEHR_ORACLE_H
#define SWEHR_ORACLE_H
#include "swehr.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle oracle_t;

/* Same arguments as swehr_create() (FP64 only).  Returns NULL on bad input. */
oracle_t *oracle_create(const swehr_params *p, const double *topo, const double *ks,
                        const double *ds, const double *pfrac, const double *rain);

/* One time step (SWEHR.c:810-1676).  *print_due is set when the step crossed a print
 * interval (SWEHR.c:1679).  Returns enum swehr_status. */
int oracle_step(oracle_t *o, int *print_due);

/* Steps until t >= tend or n steps; returns status.  Print intervals are ignored. */
int oracle_run(oracle_t *o, long n);

void oracle_info(const oracle_t *o, swehr_step_info *info);
int oracle_download(const oracle_t *o, int field, int k, double *host);
int oracle_stage(const oracle_t *o, double *stage, double *saveflow, int capacity, int *count);
/* outlet discharge at each stage record: water (m^3/s) and sediment (kg/s) through the
 * j = 1/2 face in the step just taken */
/* per-cell suction head (m), theta_s - theta_0 and Manning n0 in place of the scalars
 * (each may be NULL: keep the scalar); nx*ny values, i-major; set before stepping */
int oracle_set_rasters(oracle_t *o, const double *hf, const double *dtheta, const double *roughness);
int oracle_discharge(const oracle_t *o, double *q, double *qs, int capacity, int *count);
int oracle_budget(const oracle_t *o, double *out);
long oracle_fail_events(const oracle_t *o);
void oracle_set_threads(int n);
int oracle_get_threads(void);
void oracle_destroy(oracle_t *o);

#ifdef __cplusplus
}
#endif
#endif
