/*
 * swehr_io.h - the reference's file contract, for the host driver.
 *
 * Inputs (SWEHR.c:468-522, 218-227): input.txt (14 numbers), rain.txt (rnum rates,
 * m/s), particlesizein.txt and particlepercentin.txt (S values each), topoin.txt
 * (nx*ny elevations, i-major).  Optional sidecar swehr_ext.txt (key = value lines)
 * sets values the reference hard-codes; without it the run is the reference's.
 *
 * Outputs (SWEHR.c:1683-1697, 1718-1761): rasters written j-outer/i-inner with
 * "%10.10lf" and tab/newline separators, the 24-column stage.txt and 18-column
 * saveflow.txt with "%lf", five movie files, and the stdout lines of SWEHR.c:524-535,
 * 695-718, 1681, 1702.
 */
#ifndef SWEHR_IO_H
#define SWEHR_IO_H
#include <stdio.h>
#include "swehr.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct swehr_inputs {
    swehr_params p;
    double *topo;   /* nx*ny, i-major */
    double *ks;     /* nx*ny raster (sidecar ks_file) or NULL for p.ks */
    double *hf, *dtheta, *roughness; /* sidecar hf_file, dtheta_file, roughness_file or NULL */
    double *ds;     /* nclass diameters (m) */
    double *pfrac;  /* nclass fractions */
    double *rain;   /* rnum rates (m/s) */
    int movies;     /* 1: write movie frames at print intervals (reference behaviour) */
    int discharge;  /* 1: also write discharge.txt (sidecar "discharge = 1"; not in the reference) */
    int precision;  /* enum swehr_precision requested by the sidecar */
    int short_reads;/* values missing from the input files (read as 0, like the reference) */
} swehr_inputs;

/* Read the five input files and the optional sidecar from dir.  On failure returns a
 * status and writes a message to err. */
int swehr_io_read(const char *dir, swehr_inputs *in, char *err, size_t errlen);
void swehr_io_free(swehr_inputs *in);

/* stdout lines printed before the time loop, SWEHR.c:524-535 and 695-718 */
void swehr_io_print_setup(FILE *out, const swehr_inputs *in);
/* progress line at a print interval, SWEHR.c:1681 */
void swehr_io_print_progress(FILE *out, double t, double courant, double dt, double rain);

/* One raster, j outer and i inner, "%10.10lf" + tab / newline (SWEHR.c:1718-1747).
 * v is i-major (nx*ny).  Returns 0 or -1 on a write error. */
int swehr_io_write_raster(FILE *f, const double *v, int nx, int ny);
/* speed raster pow(vx^2+vy^2, 0.5) written like swehr_io_write_raster (vel.txt) */
int swehr_io_write_speed(FILE *f, const double *vx, const double *vy, int nx, int ny);
/* Column series (stage.txt / saveflow.txt): rows r < count of ncols columns stored
 * col-major with stride capacity, "%lf" + tab / newline (SWEHR.c:1749-1761). */
int swehr_io_write_series(FILE *f, const double *cols, int ncols, int capacity, int count);

/* printf("%.<prec>f", v) into out (room for SWEHR_FIXED_MAX bytes), from the exact
 * binary value with the same ties-to-even rounding; returns the length.  prec 1..18. */
#define SWEHR_FIXED_MAX 400
int swehr_io_format_fixed(char *out, double v, int prec);
/* n doubles from text as successive fscanf("%lf") calls read them: returns the count of
 * values missing at the end (set to 0), or -1 if a token is not entirely a number. */
long swehr_io_parse_doubles(const char *text, long len, double *out, long n);

/* Movie files, opened (and truncated) at start as the reference does (SWEHR.c:462-466). */
typedef struct swehr_movies { FILE *f[5]; } swehr_movies;
int swehr_io_open_movies(const char *dir, swehr_movies *m);
/* Append one frame: topo, depth, speed (from vx, vy), TOTSED, TOTM; all i-major. */
int swehr_io_movie_frame(swehr_movies *m, const double *topo, const double *depth, const double *vx,
                         const double *vy, const double *totsed, const double *totm, int nx, int ny);
void swehr_io_close_movies(swehr_movies *m);

/* discharge.txt: one row per stage record, "%lf" time, outlet water (m^3/s) and
 * sediment (kg/s) discharge, tab separated (an output the reference does not have). */
int swehr_io_write_discharge(const char *dir, const double *t, const double *q, const double *qs, int count);

/* Write the twelve final output files of SWEHR.c:1705-1761 into dir.  fields holds
 * i-major rasters indexed by enum swehr_field (CH = class 1); stage/saveflow as in
 * swehr_io_write_series. */
int swehr_io_write_final(const char *dir, int nx, int ny, const double *const *fields,
                         const double *stage, const double *saveflow, int capacity, int count);

#ifdef __cplusplus
}
#endif
#endif
