/*
 * oracle_main.c - TEST INFRASTRUCTURE.  Runs the CPU oracle with the reference's file
 * contract (reads the five input files of a directory, writes the same outputs and
 * stdout), so the golden tests can compare whole runs with the compiled reference
 * byte for byte.  Usage: swehr_oracle_run [dir [maxsteps]] (dir defaults to ".";
 * maxsteps > 0 stops after that many steps).
 */
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include "swehr_io.h"
#include "swehr_oracle.h"

int main(int argc, char **argv)
{
    const char *dir = argc > 1 ? argv[1] : ".";
    const long maxsteps = argc > 2 ? atol(argv[2]) : 0;
    char err[512] = {0};
    swehr_inputs in;
    int rc = swehr_io_read(dir, &in, err, sizeof err);
    if (rc != SWEHR_OK) {
        fprintf(stderr, "swehr_oracle_run: %s\n", err);
        return 1;
    }
    swehr_movies mv;
    if (swehr_io_open_movies(dir, &mv) != 0) {
        fprintf(stderr, "swehr_oracle_run: cannot open movie files in %s\n", dir);
        return 1;
    }
    swehr_io_print_setup(stdout, &in);
    oracle_t *o = oracle_create(&in.p, in.topo, in.ks, in.ds, in.pfrac, in.rain);
    if (o && (in.hf || in.dtheta || in.roughness)) oracle_set_rasters(o, in.hf, in.dtheta, in.roughness);
    if (!o) {
        fprintf(stderr, "swehr_oracle_run: bad parameters\n");
        return 1;
    }
    const int nx = in.p.nx, ny = in.p.ny;
    const long n = (long)nx * ny;
    double *fld[SWEHR_F_COUNT];
    for (int f = 0; f < SWEHR_F_COUNT; f++) fld[f] = (double *)malloc(sizeof(double) * (size_t)n);
    time_t t0 = time(NULL);
    swehr_step_info info;
    oracle_info(o, &info);
    while (!info.done && (maxsteps <= 0 || info.steps < maxsteps)) {
        const double t_before = info.t;
        int print_due = 0;
        const int st = oracle_step(o, &print_due);
        oracle_info(o, &info);
        if (st == SWEHR_ERR_CFL) printf("!!! CFL Error at t=%lf\n", t_before);
        if (print_due && in.movies) {
            swehr_io_print_progress(stdout, info.t, info.courant, info.dt, info.rain);
            oracle_download(o, SWEHR_F_TOPO, 0, fld[SWEHR_F_TOPO]);
            oracle_download(o, SWEHR_F_DEPTH, 0, fld[SWEHR_F_DEPTH]);
            oracle_download(o, SWEHR_F_VELX, 0, fld[SWEHR_F_VELX]);
            oracle_download(o, SWEHR_F_VELY, 0, fld[SWEHR_F_VELY]);
            oracle_download(o, SWEHR_F_TOTSED, 0, fld[SWEHR_F_TOTSED]);
            oracle_download(o, SWEHR_F_TOTM, 0, fld[SWEHR_F_TOTM]);
            swehr_io_movie_frame(&mv, fld[SWEHR_F_TOPO], fld[SWEHR_F_DEPTH], fld[SWEHR_F_VELX], fld[SWEHR_F_VELY],
                                 fld[SWEHR_F_TOTSED], fld[SWEHR_F_TOTM], nx, ny);
        }
        if (st != SWEHR_OK && st != SWEHR_ERR_CFL) {
            fprintf(stderr, "swehr_oracle_run: %d\n", st);
            break;
        }
    }
    time_t t1 = time(NULL);
    printf("Elapsed wall clock time: %ld\n", (long)(t1 - t0));
    for (int f = 0; f < SWEHR_F_COUNT; f++) {
        const int k = (f == SWEHR_F_CH || f == SWEHR_F_M || f == SWEHR_F_SED) ? 1 : 0;
        oracle_download(o, f, k, fld[f]);
    }
    double *stage = (double *)malloc(sizeof(double) * SWEHR_STAGE_COLS * SWEHR_STAGE_MAX);
    double *flow = (double *)malloc(sizeof(double) * SWEHR_FLOW_COLS * SWEHR_STAGE_MAX);
    int count = 0;
    oracle_stage(o, stage, flow, SWEHR_STAGE_MAX, &count);
    rc = swehr_io_write_final(dir, nx, ny, (const double *const *)fld, stage, flow, SWEHR_STAGE_MAX, count);
    swehr_io_close_movies(&mv);
    oracle_destroy(o);
    swehr_io_free(&in);
    return rc ? 1 : 0;
}
