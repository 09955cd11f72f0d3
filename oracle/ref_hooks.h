/*
 * ref_hooks.h - TEST INFRASTRUCTURE.  Parameter hooks for the instrumented build of
 * the reference (oracle/_ref/swehr_ref_hooked).  oracle/Makefile streams
 * /root/reference/SWEHR.c through a handful of sed substitutions straight into gcc
 * (no patched copy is written to disk) and force-includes this header first.
 *
 * Every hook defaults to the reference's own hard-coded value, so with no REFH_*
 * variable set the hooked binary computes what the unmodified one does.  Hooks:
 *   REFH_KS / REFH_KS_FILE   saturated conductivity, scalar or i-major raster file (SWEHR.c:239)
 *   REFH_THETA0 / REFH_THETA0_FILE, REFH_THETAS / REFH_THETAS_FILE, REFH_HF / REFH_HF_FILE,
 *   REFH_ROUGHNESS / REFH_ROUGHNESS_FILE
 *                            soil moisture, suction head and Manning n0: scalar or i-major
 *                            raster file, as for Ks (SWEHR.c:237-242)
 *   REFH_CSTABLE             target courant number (SWEHR.c:632)
 *   REFH_BNDRY="a,b,c,d"     boundary types (SWEHR.c:618-621)
 *   REFH_INFLOW=0            switch off the inflow at j=ny (SWEHR.c:1010-1035)
 *   REFH_MOVIES=0            skip the progress line and movie frames (SWEHR.c:1679-1698)
 *   REFH_MAXSTEPS=n          stop after n steps
 *   REFH_DUMP=path           write the final state in full precision (format below)
 *
 * Heap memory is forced to zero (malloc -> calloc): the reference reads arrays it
 * never initialises (depth, deposited mass, ...), which only works because a fresh
 * heap is zero.
 *
 * Dump format (little endian): "SWEHRDMP", int32 nx, ny, nn, stagecnt, int64 steps,
 * double t, then nx*ny doubles (i-major) for each of h, hu, hv, ch_1..S, M_1..S,
 * TOPO, VINF, TOTM, DEP, TOTSED, MAXVEL, MAXDEPTH, VEL1, VEL2, then STAGE (24 columns
 * of stagecnt) and SAVEFLOW (18 columns of stagecnt).
 */
#ifndef REF_HOOKS_H
#define REF_HOOKS_H
#include <stdlib.h>
#include <stdio.h>
#include <string.h>
#include <stdint.h>

#define malloc(n) calloc(1, (size_t)(n))

static double refh_env(const char *name, double dflt)
{
    const char *s = getenv(name);
    return (s && *s) ? atof(s) : dflt;
}

/* Value at cell (i, j) (1-based) of a per-cell parameter: from the raster file named by
 * the variable file_var (nx*ny numbers, i-major: row i holds the ny values of j) when it
 * is set, else the scalar variable scalar_var, else the reference's value dflt.  slot
 * caches one raster per parameter (0 Ks, 1 THETA0, 2 THETAS, 3 HF, 4 ROUGHNESS). */
static double refh_raster(int slot, const char *file_var, const char *scalar_var, double dflt, int i, int j,
                          int nx, int ny)
{
    static double *ras[5] = {NULL, NULL, NULL, NULL, NULL};
    static int loaded[5] = {0, 0, 0, 0, 0};
    if (!loaded[slot]) {
        const char *path = getenv(file_var);
        loaded[slot] = 1;
        if (path && *path) {
            FILE *f = fopen(path, "r");
            if (!f) { fprintf(stderr, "refh: cannot open %s\n", path); exit(2); }
            ras[slot] = (double *)calloc((size_t)nx * ny, sizeof(double));
            for (long n = 0; n < (long)nx * ny; n++)
                if (fscanf(f, "%lf", &ras[slot][n]) != 1) { fprintf(stderr, "refh: short raster %s\n", path); exit(2); }
            fclose(f);
        }
    }
    if (ras[slot]) return ras[slot][(long)(i - 1) * ny + (j - 1)];
    return refh_env(scalar_var, dflt);
}

static double refh_ks(int i, int j, int nx, int ny) { return refh_raster(0, "REFH_KS_FILE", "REFH_KS", 0.0, i, j, nx, ny); }

static int refh_bndry(int side, int dflt)
{
    const char *s = getenv("REFH_BNDRY");
    int v[4];
    if (s && sscanf(s, "%d,%d,%d,%d", &v[0], &v[1], &v[2], &v[3]) == 4) return v[side - 1];
    return dflt;
}

static long refh_steps = 0;

static int refh_continue(void)
{
    static long maxsteps = -2;
    if (maxsteps == -2) maxsteps = (long)refh_env("REFH_MAXSTEPS", -1);
    if (maxsteps >= 0 && refh_steps >= maxsteps) return 0;
    refh_steps++;
    return 1;
}

static void refh_put(FILE *f, double **a, int nx, int ny)
{
    for (int i = 1; i <= nx; i++) fwrite(&a[i][1], sizeof(double), (size_t)ny, f);
}

static void refh_dump(int nx, int ny, int nn, double t, int stagecnt, double ***U, double **TOPO,
                      double ***M, double **TOTM, double **DEP, double **VINF, double **TOTSED,
                      double **MAXVEL, double **MAXDEPTH, double ***VEL, double **STAGE,
                      double **SAVEFLOW)
{
    const char *path = getenv("REFH_DUMP");
    if (!path || !*path) return;
    FILE *f = fopen(path, "wb");
    if (!f) { fprintf(stderr, "refh: cannot write %s\n", path); return; }
    int32_t hdr[4] = {nx, ny, nn, stagecnt};
    int64_t st = refh_steps;
    fwrite("SWEHRDMP", 1, 8, f);
    fwrite(hdr, sizeof(int32_t), 4, f);
    fwrite(&st, sizeof(int64_t), 1, f);
    fwrite(&t, sizeof(double), 1, f);
    for (int k = 1; k <= nn; k++) refh_put(f, U[k], nx, ny);
    for (int k = 1; k <= nn - 3; k++) refh_put(f, M[k], nx, ny);
    refh_put(f, TOPO, nx, ny);
    refh_put(f, VINF, nx, ny);
    refh_put(f, TOTM, nx, ny);
    refh_put(f, DEP, nx, ny);
    refh_put(f, TOTSED, nx, ny);
    refh_put(f, MAXVEL, nx, ny);
    refh_put(f, MAXDEPTH, nx, ny);
    refh_put(f, VEL[1], nx, ny);
    refh_put(f, VEL[2], nx, ny);
    for (int c = 1; c <= 24; c++) fwrite(&STAGE[c][1], sizeof(double), (size_t)stagecnt, f);
    for (int c = 1; c <= 18; c++) fwrite(&SAVEFLOW[c][1], sizeof(double), (size_t)stagecnt, f);
    fclose(f);
}

#define REFH_DUMP_STATE() \
    refh_dump(nx, ny, nn, t, stagecnt, U, TOPO, M, TOTM, DEP, VINF, TOTSED, MAXVEL, MAXDEPTH, VEL, STAGE, SAVEFLOW)

#endif
