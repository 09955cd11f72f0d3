/*
 * swehr_io.c - reading the reference's input files and writing its outputs
 * (SWEHR.c:462-535, 218-227, 695-718, 1681-1761).  Plain C, no GPU code.
 */
#include "swehr_io.h"

#include <ctype.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

int swehr_params_init(swehr_params *p)
{
    if (!p) return SWEHR_ERR_BAD_INPUT;
    memset(p, 0, sizeof(*p));
    p->cstable = 0.05;
    p->maxstep = 0.25;
    p->bndry[0] = 1;
    p->bndry[1] = 0;
    p->bndry[2] = 1;
    p->bndry[3] = 1;
    p->inflow = 1;
    p->roughness = 0.1;
    p->theta0 = 0.05;
    p->thetas = 0.4;
    p->hf = 0.001;
    p->ks = 0;
    p->vinf0 = 1e-4;
    return SWEHR_OK;
}

static FILE *open_in(const char *dir, const char *name, char *path, size_t plen)
{
    snprintf(path, plen, "%s/%s", dir, name);
    return fopen(path, "r");
}

/* fscanf-style read of n doubles; missing values stay 0 and are counted */
static int read_doubles(FILE *f, double *v, long n)
{
    int missing = 0;
    for (long k = 0; k < n; k++) {
        if (!f || fscanf(f, "%lf", &v[k]) != 1) {
            v[k] = 0;
            missing++;
        }
    }
    return missing;
}

static char *trim(char *s)
{
    while (isspace((unsigned char)*s)) s++;
    char *e = s + strlen(s);
    while (e > s && isspace((unsigned char)e[-1])) *--e = 0;
    return s;
}

static int read_sidecar(const char *dir, swehr_inputs *in, char *err, size_t errlen)
{
    char path[4096], line[4096];
    FILE *f = open_in(dir, "swehr_ext.txt", path, sizeof path);
    if (!f) return SWEHR_OK;
    swehr_params *p = &in->p;
    int lineno = 0;
    while (fgets(line, sizeof line, f)) {
        lineno++;
        char *hash = strchr(line, '#');
        if (hash) *hash = 0;
        char *s = trim(line);
        if (!*s) continue;
        char *eq = strchr(s, '=');
        if (!eq) {
            snprintf(err, errlen, "swehr_ext.txt:%d: expected key = value", lineno);
            fclose(f);
            return SWEHR_ERR_BAD_INPUT;
        }
        *eq = 0;
        char *key = trim(s), *val = trim(eq + 1);
        int ok = 1;
        if (!strcmp(key, "cstable")) p->cstable = atof(val);
        else if (!strcmp(key, "maxstep")) p->maxstep = atof(val);
        else if (!strcmp(key, "inflow")) p->inflow = atoi(val);
        else if (!strcmp(key, "bndry"))
            ok = sscanf(val, "%d %d %d %d", &p->bndry[0], &p->bndry[1], &p->bndry[2], &p->bndry[3]) == 4 ||
                 sscanf(val, "%d,%d,%d,%d", &p->bndry[0], &p->bndry[1], &p->bndry[2], &p->bndry[3]) == 4;
        else if (!strcmp(key, "ks")) p->ks = atof(val);
        else if (!strcmp(key, "hf")) p->hf = atof(val);
        else if (!strcmp(key, "theta0")) p->theta0 = atof(val);
        else if (!strcmp(key, "thetas")) p->thetas = atof(val);
        else if (!strcmp(key, "roughness")) p->roughness = atof(val);
        else if (!strcmp(key, "vinf0")) p->vinf0 = atof(val);
        else if (!strcmp(key, "movies")) in->movies = atoi(val);
        else if (!strcmp(key, "precision")) {
            if (!strcmp(val, "fp64")) in->precision = SWEHR_FP64;
            else if (!strcmp(val, "fp32")) in->precision = SWEHR_FP32;
            else ok = 0;
        } else if (!strcmp(key, "ks_file")) {
            char kp[4096];
            FILE *kf = open_in(dir, val, kp, sizeof kp);
            if (!kf) kf = fopen(val, "r");
            if (!kf) {
                snprintf(err, errlen, "swehr_ext.txt:%d: cannot open ks_file %s", lineno, val);
                fclose(f);
                return SWEHR_ERR_BAD_INPUT;
            }
            const long n = (long)p->nx * p->ny;
            free(in->ks);
            in->ks = (double *)malloc(sizeof(double) * (size_t)n);
            const int miss = read_doubles(kf, in->ks, n);
            fclose(kf);
            if (miss) {
                snprintf(err, errlen, "ks_file %s: %d of %ld values missing", val, miss, n);
                fclose(f);
                return SWEHR_ERR_BAD_INPUT;
            }
        } else {
            snprintf(err, errlen, "swehr_ext.txt:%d: unknown key '%s'", lineno, key);
            fclose(f);
            return SWEHR_ERR_BAD_INPUT;
        }
        if (!ok) {
            snprintf(err, errlen, "swehr_ext.txt:%d: bad value for %s", lineno, key);
            fclose(f);
            return SWEHR_ERR_BAD_INPUT;
        }
    }
    fclose(f);
    return SWEHR_OK;
}

int swehr_io_read(const char *dir, swehr_inputs *in, char *err, size_t errlen)
{
    char path[4096];
    memset(in, 0, sizeof(*in));
    swehr_params_init(&in->p);
    in->movies = 1;
    in->precision = SWEHR_FP64;
    if (!dir || !*dir) dir = ".";

    double v[14] = {0};
    FILE *f = open_in(dir, "input.txt", path, sizeof path);
    if (!f) {
        snprintf(err, errlen, "cannot open %s", path);
        return SWEHR_ERR_BAD_INPUT;
    }
    in->short_reads += read_doubles(f, v, 14);
    fclose(f);
    swehr_params *p = &in->p;
    /* field mapping and int conversions of SWEHR.c:477-504 */
    p->nx = (int)v[0];
    p->ny = (int)v[1];
    p->dx = v[2];
    p->F = v[3];
    p->manning0 = v[4];
    p->tend = v[5];
    p->a = v[6];
    p->ad = v[7];
    p->h0 = v[8];
    p->mtstar0 = v[9];
    p->rnum = (int)v[10];
    p->rint = v[11];
    p->nclass = (int)v[12];
    p->J = v[13];
    if (p->nx < 1 || p->ny < 1 || p->nclass < 1 || p->nclass > SWEHR_MAX_CLASSES || p->rnum < 1 || !(p->dx > 0)) {
        snprintf(err, errlen, "input.txt: unsupported sizes nx=%d ny=%d S=%d rnum=%d dx=%g", p->nx, p->ny,
                 p->nclass, p->rnum, p->dx);
        return SWEHR_ERR_BAD_INPUT;
    }
    const long n = (long)p->nx * p->ny;
    in->topo = (double *)malloc(sizeof(double) * (size_t)n);
    in->rain = (double *)malloc(sizeof(double) * (size_t)p->rnum);
    in->ds = (double *)malloc(sizeof(double) * (size_t)p->nclass);
    in->pfrac = (double *)malloc(sizeof(double) * (size_t)p->nclass);
    if (!in->topo || !in->rain || !in->ds || !in->pfrac) {
        snprintf(err, errlen, "out of host memory");
        return SWEHR_ERR_OOM;
    }
    f = open_in(dir, "rain.txt", path, sizeof path);
    in->short_reads += read_doubles(f, in->rain, p->rnum);
    if (f) fclose(f);
    FILE *fs = open_in(dir, "particlesizein.txt", path, sizeof path);
    FILE *fp = open_in(dir, "particlepercentin.txt", path, sizeof path);
    in->short_reads += read_doubles(fs, in->ds, p->nclass);
    in->short_reads += read_doubles(fp, in->pfrac, p->nclass);
    if (fs) fclose(fs);
    if (fp) fclose(fp);
    f = open_in(dir, "topoin.txt", path, sizeof path);
    in->short_reads += read_doubles(f, in->topo, n);
    if (f) fclose(f);
    return read_sidecar(dir, in, err, errlen);
}

void swehr_io_free(swehr_inputs *in)
{
    free(in->topo);
    free(in->ks);
    free(in->ds);
    free(in->pfrac);
    free(in->rain);
    memset(in, 0, sizeof(*in));
}

void swehr_io_print_setup(FILE *out, const swehr_inputs *in)
{
    const swehr_params *p = &in->p;
    const int nx = p->nx, ny = p->ny;
    fprintf(out, "Grid Size (Spacing): %d x %d (%lf)\n", nx, ny, p->dx);
    fprintf(out, "Number of conserved variables: %d\n", 3 + p->nclass);
    fprintf(out, "Number of particle size classes: %d\n", p->nclass);
    fprintf(out, "H-R Parameters: \n");
    fprintf(out, "a0: %lf\n", p->a);
    fprintf(out, "ad0: %lf\n", p->ad);
    fprintf(out, "F: %lf\n", p->F);
    fprintf(out, "J: %lf\n", p->J);
    fprintf(out, "h0: %lf\n", p->h0);
    fprintf(out, "mtstar0: %lf\n", p->mtstar0);
    fprintf(out, "b: %lf\n", 1.0);
    /* lowest bed cell of the outlet row, SWEHR.c:685-696 */
    int outlet = 1;
    double elev = 99999;
    for (int i = 1; i <= nx; i++) {
        const double z = in->topo[(long)(i - 1) * ny];
        if (z < elev) { elev = z; outlet = i; }
    }
    fprintf(out, "Basin Outlet Pixel: %d\n", outlet);
    fprintf(out, "Basin Outlet Elevation (m): %lf\n", elev);
    /* probe points (2,1), (2,10) ... (2,50), SWEHR.c:710-718; points outside the grid
     * (the reference reads out of bounds there) are reported with elevation 0 */
    static const int yp[6] = {1, 10, 20, 30, 40, 50};
    for (int k = 0; k < 6; k++) {
        const double z = (2 <= nx && yp[k] <= ny) ? in->topo[(long)(2 - 1) * ny + (yp[k] - 1)] : 0.0;
        fprintf(out, "Point %d at (%d,%d) has elevation %lf\n", k + 1, 2, yp[k], z);
    }
}

void swehr_io_print_progress(FILE *out, double t, double courant, double dt, double rain)
{
    fprintf(out, "%lf\t %lf\t %lf\t %lf\n", t, courant, dt, rain);
}

int swehr_io_write_raster(FILE *f, const double *v, int nx, int ny)
{
    for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++)
            if (fprintf(f, i != nx ? "%10.10lf\t" : "%10.10lf\n", v[(long)(i - 1) * ny + (j - 1)]) < 0) return -1;
    return 0;
}

int swehr_io_write_speed(FILE *f, const double *vx, const double *vy, int nx, int ny)
{
    for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
            const long c = (long)(i - 1) * ny + (j - 1);
            if (fprintf(f, i != nx ? "%10.10lf\t" : "%10.10lf\n", pow(vx[c] * vx[c] + vy[c] * vy[c], 0.5)) < 0) return -1;
        }
    return 0;
}

int swehr_io_write_series(FILE *f, const double *cols, int ncols, int capacity, int count)
{
    for (int r = 0; r < count; r++)
        for (int c = 0; c < ncols; c++)
            if (fprintf(f, c != ncols - 1 ? "%lf\t" : "%lf\n", cols[(long)c * capacity + r]) < 0) return -1;
    return 0;
}

static const char *const movie_names[5] = {"topomovie.txt", "depthmovie.txt", "velocitymovie.txt", "cmovie.txt",
                                           "Mmovie.txt"};

int swehr_io_open_movies(const char *dir, swehr_movies *m)
{
    char path[4096];
    for (int k = 0; k < 5; k++) {
        snprintf(path, sizeof path, "%s/%s", dir, movie_names[k]);
        m->f[k] = fopen(path, "w");
        if (!m->f[k]) return -1;
    }
    return 0;
}

int swehr_io_movie_frame(swehr_movies *m, const double *topo, const double *depth, const double *vx,
                         const double *vy, const double *totsed, const double *totm, int nx, int ny)
{
    int rc = 0;
    rc |= swehr_io_write_raster(m->f[0], topo, nx, ny);
    rc |= swehr_io_write_raster(m->f[1], depth, nx, ny);
    rc |= swehr_io_write_speed(m->f[2], vx, vy, nx, ny);
    rc |= swehr_io_write_raster(m->f[3], totsed, nx, ny);
    rc |= swehr_io_write_raster(m->f[4], totm, nx, ny);
    return rc;
}

void swehr_io_close_movies(swehr_movies *m)
{
    for (int k = 0; k < 5; k++)
        if (m->f[k]) fclose(m->f[k]);
}

int swehr_io_write_final(const char *dir, int nx, int ny, const double *const *fields, const double *stage,
                         const double *saveflow, int capacity, int count)
{
    static const struct { const char *name; int field; } rasters[] = {
        {"topoout.txt", SWEHR_F_TOPO}, {"depth.txt", SWEHR_F_DEPTH}, {"uh.txt", SWEHR_F_UH},
        {"vh.txt", SWEHR_F_VH},        {"ch.txt", SWEHR_F_CH},       {"m.txt", SWEHR_F_TOTM},
        {"c.txt", SWEHR_F_TOTSED},     {"vel.txt", -1},              {"maxvel.txt", SWEHR_F_MAXVEL},
        {"maxdepth.txt", SWEHR_F_MAXDEPTH}};
    char path[4096];
    int rc = 0;
    for (size_t k = 0; k < sizeof rasters / sizeof rasters[0]; k++) {
        snprintf(path, sizeof path, "%s/%s", dir, rasters[k].name);
        FILE *f = fopen(path, "w");
        if (!f) return -1;
        if (rasters[k].field < 0) rc |= swehr_io_write_speed(f, fields[SWEHR_F_VELX], fields[SWEHR_F_VELY], nx, ny);
        else rc |= swehr_io_write_raster(f, fields[rasters[k].field], nx, ny);
        rc |= fclose(f);
    }
    snprintf(path, sizeof path, "%s/stage.txt", dir);
    FILE *f = fopen(path, "w");
    if (!f) return -1;
    rc |= swehr_io_write_series(f, stage, SWEHR_STAGE_COLS, capacity, count);
    rc |= fclose(f);
    snprintf(path, sizeof path, "%s/saveflow.txt", dir);
    f = fopen(path, "w");
    if (!f) return -1;
    rc |= swehr_io_write_series(f, saveflow, SWEHR_FLOW_COLS, capacity, count);
    rc |= fclose(f);
    return rc ? -1 : 0;
}
