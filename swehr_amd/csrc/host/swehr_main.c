/*
 * swehr_main.c - drop-in replacement for the reference program (SWEHR.c main(),
 * SWEHR.c:446-1784) running the time step on the GPU through include/swehr.h.
 *
 * Usage: swehr [dir]   (dir defaults to the working directory, as in the reference)
 *
 * Reads input.txt, rain.txt, particlesizein.txt, particlepercentin.txt, topoin.txt
 * and the optional sidecar swehr_ext.txt from dir; writes the twelve output files,
 * the five movie files and the reference's stdout lines.  Exit status 0 on success
 * (a CFL stop also ends normally, as in the reference), 1 on errors.
 */
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include "swehr.h"
#include "swehr_io.h"

/* Movie frames are written by a background thread from one of two host copies of the
 * six frame fields, so the GPU resumes stepping as soon as a frame is downloaded. */
enum { FRAME_FIELDS = 6 };
static const int frame_ids[FRAME_FIELDS] = {SWEHR_F_TOPO, SWEHR_F_DEPTH, SWEHR_F_VELX,
                                            SWEHR_F_VELY, SWEHR_F_TOTSED, SWEHR_F_TOTM};

typedef struct frame_job {
    swehr_movies *mv;
    double *f[FRAME_FIELDS];
    int nx, ny, rc;
} frame_job;

static void *write_frame(void *arg)
{
    frame_job *j = (frame_job *)arg;
    j->rc = swehr_io_movie_frame(j->mv, j->f[0], j->f[1], j->f[2], j->f[3], j->f[4], j->f[5], j->nx, j->ny);
    return NULL;
}

static int fail(const char *what, int rc)
{
    const char *msg = swehr_last_error();
    fprintf(stderr, "swehr: %s: %s%s%s\n", what, swehr_strerror(rc), msg && *msg ? ": " : "", msg ? msg : "");
    return 1;
}

int main(int argc, char **argv)
{
    const char *dir = argc > 1 ? argv[1] : ".";
    char err[512] = {0};
    swehr_inputs in;
    int rc = swehr_io_read(dir, &in, err, sizeof err);
    if (rc != SWEHR_OK) {
        fprintf(stderr, "swehr: %s\n", err);
        return 1;
    }
    if (in.short_reads) fprintf(stderr, "swehr: %d input values missing, read as 0\n", in.short_reads);
    swehr_movies mv;
    if (swehr_io_open_movies(dir, &mv) != 0) {
        fprintf(stderr, "swehr: cannot open movie files in %s\n", dir);
        return 1;
    }
    swehr_io_print_setup(stdout, &in);
    swehr_handle *h = NULL;
    rc = swehr_create(&in.p, in.topo, in.ks, in.ds, in.pfrac, in.rain, in.precision, &h);
    if (rc != SWEHR_OK) return fail("create", rc);
    rc = swehr_configure(h, in.movies ? 1 : 0, 0, 0);
    if (rc != SWEHR_OK) return fail("configure", rc);
    const double *rasters[SWEHR_R_COUNT] = {in.hf, in.dtheta, in.roughness};
    for (int k = 0; k < SWEHR_R_COUNT; k++)
        if (rasters[k] && (rc = swehr_set_parameter_raster(h, k, rasters[k])) != SWEHR_OK)
            return fail("parameter raster", rc);
    const int nx = in.p.nx, ny = in.p.ny;
    const long n = (long)nx * ny;
    double *fld[SWEHR_F_COUNT];
    for (int f = 0; f < SWEHR_F_COUNT; f++) {
        fld[f] = (double *)malloc(sizeof(double) * (size_t)n);
        if (!fld[f]) return fail("host allocation", SWEHR_ERR_OOM);
    }
    frame_job jobs[2];
    for (int b = 0; b < 2; b++) {
        jobs[b].mv = &mv;
        jobs[b].nx = nx;
        jobs[b].ny = ny;
        jobs[b].rc = 0;
        for (int q = 0; q < FRAME_FIELDS; q++) {
            jobs[b].f[q] = (double *)malloc(sizeof(double) * (size_t)n);
            if (!jobs[b].f[q]) return fail("host allocation", SWEHR_ERR_OOM);
        }
    }
    pthread_t writer;
    int writing = 0, cur = 0;
    time_t t0 = time(NULL);
    swehr_step_info info;
    double t_before = 0;
    for (;;) {
        rc = swehr_advance(h, (int64_t)1 << 40, &info);
        if (rc == SWEHR_ERR_CFL) {
            /* the reference prints this before its last step (SWEHR.c:1413) */
            t_before = info.t_prev;
            printf("!!! CFL Error at t=%lf\n", t_before);
        } else if (rc != SWEHR_OK) {
            return fail("advance", rc);
        }
        if (info.print_due) {
            swehr_io_print_progress(stdout, info.t, info.courant, info.dt, info.rain);
            /* the set written two frames ago is free: its writer was joined last frame */
            for (int q = 0; q < FRAME_FIELDS; q++)
                if ((rc = swehr_download(h, frame_ids[q], 0, jobs[cur].f[q])) != SWEHR_OK) return fail("download", rc);
            if ((rc = swehr_ack_print(h)) != SWEHR_OK) return fail("ack", rc);
            if (writing) { /* frames are appended in order: one writer at a time */
                pthread_join(writer, NULL);
                writing = 0;
            }
            if (pthread_create(&writer, NULL, write_frame, &jobs[cur]) == 0) writing = 1;
            else write_frame(&jobs[cur]);
            cur ^= 1;
        }
        if (info.done) break;
    }
    if (writing) pthread_join(writer, NULL);
    if (jobs[0].rc || jobs[1].rc) fprintf(stderr, "swehr: writing a movie frame failed\n");
    time_t t1 = time(NULL);
    printf("Elapsed wall clock time: %ld\n", (long)(t1 - t0));
    for (int f = 0; f < SWEHR_F_COUNT; f++) {
        const int k = (f == SWEHR_F_CH || f == SWEHR_F_M || f == SWEHR_F_SED) ? 1 : 0;
        if ((rc = swehr_download(h, f, k, fld[f])) != SWEHR_OK) return fail("download", rc);
    }
    int count = 0;
    rc = swehr_stage(h, NULL, NULL, 0, &count);
    if (rc != SWEHR_OK) return fail("stage", rc);
    const int cap = count > 0 ? count : 1;
    double *stage = (double *)malloc(sizeof(double) * SWEHR_STAGE_COLS * (size_t)cap);
    double *flow = (double *)malloc(sizeof(double) * SWEHR_FLOW_COLS * (size_t)cap);
    if (!stage || !flow) return fail("host allocation", SWEHR_ERR_OOM);
    if ((rc = swehr_stage(h, stage, flow, cap, &count)) != SWEHR_OK) return fail("stage", rc);
    if (swehr_io_write_final(dir, nx, ny, (const double *const *)fld, stage, flow, cap, count) != 0) {
        fprintf(stderr, "swehr: cannot write outputs in %s\n", dir);
        return 1;
    }
    if (in.discharge) { /* outlet hydrograph and sedigraph at the stage times */
        double *q = (double *)malloc(sizeof(double) * (size_t)cap), *qs = (double *)malloc(sizeof(double) * (size_t)cap);
        if (!q || !qs) return fail("host allocation", SWEHR_ERR_OOM);
        if ((rc = swehr_discharge(h, q, qs, cap, &count)) != SWEHR_OK) return fail("discharge", rc);
        if (swehr_io_write_discharge(dir, stage, q, qs, count < cap ? count : cap) != 0) {
            fprintf(stderr, "swehr: cannot write discharge.txt in %s\n", dir);
            return 1;
        }
        free(q);
        free(qs);
    }
    swehr_io_close_movies(&mv);
    for (int b = 0; b < 2; b++)
        for (int q = 0; q < FRAME_FIELDS; q++) free(jobs[b].f[q]);
    swehr_destroy(h);
    swehr_io_free(&in);
    return 0;
}
