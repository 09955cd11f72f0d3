/*
 * swehr_io.c - reading the reference's input files and writing its outputs
 * (SWEHR.c:462-535, 218-227, 695-718, 1681-1761).  Plain C, no GPU code.
 */
#include "swehr_io.h"

#include <ctype.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

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

/* ------------------------------------------------------------------ threads */

/* Worker count for the text I/O: SWEHR_IO_THREADS, else the online CPUs, at most 16. */
static int io_threads(void)
{
    const char *e = getenv("SWEHR_IO_THREADS");
    long n = e ? strtol(e, NULL, 10) : sysconf(_SC_NPROCESSORS_ONLN);
    if (n < 1) n = 1;
    if (n > 16) n = 16;
    return (int)n;
}

typedef struct par_job {
    void (*fn)(long lo, long hi, void *ctx);
    void *ctx;
    long n, per;
    int k;
} par_job;

static void *par_run(void *arg)
{
    par_job *j = (par_job *)arg;
    long lo = (long)j->k * j->per, hi = lo + j->per;
    if (hi > j->n) hi = j->n;
    if (lo < hi) j->fn(lo, hi, j->ctx);
    return NULL;
}

/* fn(lo, hi, ctx) over [0, n) split into contiguous ranges, one per thread */
static void par_for(long n, void (*fn)(long, long, void *), void *ctx)
{
    int t = io_threads();
    if (t > n) t = (int)(n > 0 ? n : 1);
    if (t <= 1) {
        if (n > 0) fn(0, n, ctx);
        return;
    }
    pthread_t th[16];
    par_job jobs[16];
    const long per = (n + t - 1) / t;
    int started = 0;
    for (int k = 0; k < t; k++) {
        jobs[k] = (par_job){fn, ctx, n, per, k};
        if (k > 0 && pthread_create(&th[k], NULL, par_run, &jobs[k]) == 0) started |= 1 << k;
    }
    par_run(&jobs[0]);
    for (int k = 1; k < t; k++) {
        if (started & (1 << k)) pthread_join(th[k], NULL);
        else par_run(&jobs[k]); /* thread not available: run it here */
    }
}

/* ------------------------------------------------------------------ numbers as text */

static const uint64_t kPow10[19] = {1ull, 10ull, 100ull, 1000ull, 10000ull, 100000ull, 1000000ull, 10000000ull,
                                    100000000ull, 1000000000ull, 10000000000ull, 100000000000ull,
                                    1000000000000ull, 10000000000000ull, 100000000000000ull,
                                    1000000000000000ull, 10000000000000000ull, 100000000000000000ull,
                                    1000000000000000000ull};

/* Writes v as printf("%.<prec>f") does (prec 1..18) and returns the length.  The
 * digits come from the exact binary value: v = m 2^e, N = round(|v| 10^prec) computed
 * in 128-bit integers with ties to even, the rounding glibc applies; values whose N
 * does not fit 64 bits, and NaN / inf, go through snprintf. */
int swehr_io_format_fixed(char *out, double v, int prec)
{
    uint64_t bits;
    memcpy(&bits, &v, sizeof bits);
    const int neg = (int)(bits >> 63), ex = (int)((bits >> 52) & 0x7ff);
    const uint64_t frac = bits & ((1ull << 52) - 1);
    if (ex == 0x7ff || prec < 1 || prec > 18) return snprintf(out, SWEHR_FIXED_MAX, "%.*f", prec, v);
    const uint64_t m = ex ? frac | (1ull << 52) : frac;
    const int e = ex ? ex - 1075 : -1074; /* v = m 2^e */
    const uint64_t p10 = kPow10[prec];
    unsigned __int128 N;
    if (m == 0) {
        N = 0;
    } else if (e >= 0) {
        if (e > 60) return snprintf(out, SWEHR_FIXED_MAX, "%.*f", prec, v);
        const unsigned __int128 a = (unsigned __int128)m * p10;
        if (a >> (127 - e)) return snprintf(out, SWEHR_FIXED_MAX, "%.*f", prec, v);
        N = a << e;
    } else {
        const int sh = -e;
        const unsigned __int128 num = (unsigned __int128)m * p10; /* < 2^113 */
        if (sh >= 127) {
            N = 0; /* num < 2^113 < half */
        } else {
            unsigned __int128 q = num >> sh;
            const unsigned __int128 r = num - (q << sh), half = (unsigned __int128)1 << (sh - 1);
            if (r > half || (r == half && (q & 1))) q++;
            N = q;
        }
    }
    if (N >> 64) return snprintf(out, SWEHR_FIXED_MAX, "%.*f", prec, v);
    const uint64_t n = (uint64_t)N, ip = n / p10, fp = n % p10;
    char *o = out;
    if (neg) *o++ = '-';
    char digits[24];
    int nd = 0;
    uint64_t x = ip;
    do {
        digits[nd++] = (char)('0' + x % 10);
        x /= 10;
    } while (x);
    while (nd) *o++ = digits[--nd];
    *o++ = '.';
    x = fp;
    for (int k = prec - 1; k >= 0; k--) {
        o[k] = (char)('0' + x % 10);
        x /= 10;
    }
    o += prec;
    *o = 0;
    return (int)(o - out);
}

/* Parses n doubles from text as successive fscanf("%lf") calls would (strtod, the same
 * conversion): returns the number of values missing at the end (set to 0), or -1 when
 * a token is not entirely a number (the caller then reads the file with fscanf). */
typedef struct parse_ctx {
    const char *text;
    long len;
    const long *cut; /* chunk boundaries, at whitespace */
    long *count;     /* tokens per chunk */
    const long *first;
    double *out;
    long n;
    int bad;
} parse_ctx;

static void parse_count(long lo, long hi, void *arg)
{
    parse_ctx *c = (parse_ctx *)arg;
    for (long k = lo; k < hi; k++) {
        long cnt = 0;
        for (long i = c->cut[k]; i < c->cut[k + 1];) {
            while (i < c->cut[k + 1] && isspace((unsigned char)c->text[i])) i++;
            if (i >= c->cut[k + 1]) break;
            cnt++;
            while (i < c->cut[k + 1] && !isspace((unsigned char)c->text[i])) i++;
        }
        c->count[k] = cnt;
    }
}

static void parse_values(long lo, long hi, void *arg)
{
    parse_ctx *c = (parse_ctx *)arg;
    for (long k = lo; k < hi; k++) {
        long idx = c->first[k];
        for (long i = c->cut[k]; i < c->cut[k + 1] && idx < c->n;) {
            while (i < c->cut[k + 1] && isspace((unsigned char)c->text[i])) i++;
            if (i >= c->cut[k + 1]) break;
            long e = i;
            while (e < c->cut[k + 1] && !isspace((unsigned char)c->text[e])) e++;
            char tok[128];
            const long L = e - i;
            if (L >= (long)sizeof tok) {
                c->bad = 1;
                return;
            }
            memcpy(tok, c->text + i, (size_t)L);
            tok[L] = 0;
            char *end;
            c->out[idx++] = strtod(tok, &end);
            if (end != tok + L) {
                c->bad = 1;
                return;
            }
            i = e;
        }
    }
}

long swehr_io_parse_doubles(const char *text, long len, double *out, long n)
{
    const int T = 64;
    long cut[65], count[64], first[64];
    cut[0] = 0;
    for (int k = 1; k < T; k++) {
        long c = len * k / T;
        if (c < cut[k - 1]) c = cut[k - 1];
        while (c < len && !isspace((unsigned char)text[c])) c++;
        cut[k] = c;
    }
    cut[T] = len;
    parse_ctx ctx = {text, len, cut, count, first, out, n, 0};
    par_for(T, parse_count, &ctx);
    long tot = 0;
    for (int k = 0; k < T; k++) {
        first[k] = tot;
        tot += count[k];
    }
    par_for(T, parse_values, &ctx);
    if (ctx.bad) return -1;
    for (long i = tot; i < n; i++) out[i] = 0;
    return tot < n ? n - tot : 0;
}

/* Reads the rest of f into memory (NUL-terminated); NULL on failure. */
static char *slurp(FILE *f, long *len)
{
    size_t cap = 1 << 20, used = 0;
    char *b = (char *)malloc(cap + 1);
    if (!b) return NULL;
    for (;;) {
        const size_t got = fread(b + used, 1, cap - used, f);
        used += got;
        if (used < cap) break;
        cap *= 2;
        char *nb = (char *)realloc(b, cap + 1);
        if (!nb) {
            free(b);
            return NULL;
        }
        b = nb;
    }
    b[used] = 0;
    *len = (long)used;
    return b;
}

static FILE *open_in(const char *dir, const char *name, char *path, size_t plen)
{
    snprintf(path, plen, "%s/%s", dir, name);
    return fopen(path, "r");
}

/* fscanf-style read of n doubles; missing values stay 0 and are counted */
static int read_doubles_scanf(FILE *f, double *v, long n)
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

/* The same values as read_doubles_scanf; large inputs (rasters) are parsed in parallel
 * from memory, falling back to fscanf when a token is not a plain number. */
static int read_doubles(FILE *f, double *v, long n)
{
    if (!f || n < 4096) return read_doubles_scanf(f, v, n);
    const long pos = ftell(f);
    long len = 0;
    char *text = slurp(f, &len);
    if (text) {
        const long miss = swehr_io_parse_doubles(text, len, v, n);
        free(text);
        if (miss >= 0) return (int)miss;
    }
    if (pos < 0 || fseek(f, pos, SEEK_SET) != 0) {
        for (long k = 0; k < n; k++) v[k] = 0;
        return (int)n;
    }
    return read_doubles_scanf(f, v, n);
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
        else if (!strcmp(key, "discharge")) in->discharge = atoi(val);
        else if (!strcmp(key, "precision")) {
            if (!strcmp(val, "fp64")) in->precision = SWEHR_FP64;
            else if (!strcmp(val, "fp32")) in->precision = SWEHR_FP32;
            else ok = 0;
        } else if (!strcmp(key, "ks_file") || !strcmp(key, "hf_file") || !strcmp(key, "dtheta_file") ||
                   !strcmp(key, "roughness_file")) {
            /* a per-cell raster in topoin.txt's layout (nx*ny values, i-major) */
            double **dst = !strcmp(key, "ks_file")       ? &in->ks
                           : !strcmp(key, "hf_file")     ? &in->hf
                           : !strcmp(key, "dtheta_file") ? &in->dtheta
                                                         : &in->roughness;
            char kp[4096];
            FILE *kf = open_in(dir, val, kp, sizeof kp);
            if (!kf) kf = fopen(val, "r");
            if (!kf) {
                snprintf(err, errlen, "swehr_ext.txt:%d: cannot open %s %s", lineno, key, val);
                fclose(f);
                return SWEHR_ERR_BAD_INPUT;
            }
            const long n = (long)p->nx * p->ny;
            free(*dst);
            *dst = (double *)malloc(sizeof(double) * (size_t)n);
            const int miss = *dst ? read_doubles(kf, *dst, n) : (int)n;
            fclose(kf);
            if (miss) {
                snprintf(err, errlen, "%s %s: %d of %ld values missing", key, val, miss, n);
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
    free(in->hf);
    free(in->dtheta);
    free(in->roughness);
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

/* Raster rows j0 <= j < j0 + nb formatted into per-row buffers (the writers below). */
typedef struct raster_ctx {
    const double *v, *vy; /* vy != NULL: the speed raster pow(v^2 + vy^2, 0.5) */
    int nx, ny, j0;
    char **row;
    size_t *cap;
    long *rowlen;
    int fail;
} raster_ctx;

static void format_rows(long lo, long hi, void *arg)
{
    raster_ctx *c = (raster_ctx *)arg;
    for (long r = lo; r < hi; r++) {
        const int j = c->j0 + (int)r;
        size_t used = 0;
        for (int i = 1; i <= c->nx; i++) {
            if (c->cap[r] - used < SWEHR_FIXED_MAX + 16) { /* room for the longest value */
                const size_t ncap = 2 * c->cap[r] + SWEHR_FIXED_MAX;
                char *nr = (char *)realloc(c->row[r], ncap);
                if (!nr) {
                    c->fail = 1;
                    return;
                }
                c->row[r] = nr;
                c->cap[r] = ncap;
            }
            const long k = (long)(i - 1) * c->ny + (j - 1);
            const double x = c->vy ? pow(c->v[k] * c->v[k] + c->vy[k] * c->vy[k], 0.5) : c->v[k];
            char *o = c->row[r] + used;
            int L = swehr_io_format_fixed(o, x, 10);
            if (L < 10) { /* "%10.10lf": only nan / inf are shorter than the width */
                memmove(o + (10 - L), o, (size_t)L);
                memset(o, ' ', (size_t)(10 - L));
                L = 10;
            }
            o[L] = i != c->nx ? '\t' : '\n';
            used += (size_t)L + 1;
        }
        c->rowlen[r] = (long)used;
    }
}

static int write_rows(FILE *f, const double *v, const double *vy, int nx, int ny)
{
    enum { B = 256 }; /* rows per batch */
    char *row[B];
    size_t cap[B];
    long rowlen[B];
    int rc = 0;
    for (int r = 0; r < B; r++) {
        cap[r] = (size_t)nx * 24 + SWEHR_FIXED_MAX + 16;
        row[r] = (char *)malloc(cap[r]);
        if (!row[r]) rc = -1;
    }
    for (int j0 = 1; j0 <= ny && !rc; j0 += B) {
        const int nb = ny - j0 + 1 < B ? ny - j0 + 1 : B;
        raster_ctx c = {v, vy, nx, ny, j0, row, cap, rowlen, 0};
        par_for(nb, format_rows, &c);
        if (c.fail) rc = -1;
        for (int r = 0; r < nb && !rc; r++)
            if (fwrite(row[r], 1, (size_t)rowlen[r], f) != (size_t)rowlen[r]) rc = -1;
    }
    for (int r = 0; r < B; r++) free(row[r]);
    return rc;
}

int swehr_io_write_raster(FILE *f, const double *v, int nx, int ny) { return write_rows(f, v, NULL, nx, ny); }

int swehr_io_write_speed(FILE *f, const double *vx, const double *vy, int nx, int ny)
{
    return write_rows(f, vx, vy, nx, ny);
}

int swehr_io_write_series(FILE *f, const double *cols, int ncols, int capacity, int count)
{
    char line[(SWEHR_FIXED_MAX + 1) * 32];
    for (int r = 0; r < count; r++) {
        char *o = line;
        for (int c = 0; c < ncols; c++) {
            o += swehr_io_format_fixed(o, cols[(long)c * capacity + r], 6); /* "%lf" */
            *o++ = c != ncols - 1 ? '\t' : '\n';
        }
        if (fwrite(line, 1, (size_t)(o - line), f) != (size_t)(o - line)) return -1;
    }
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

int swehr_io_write_discharge(const char *dir, const double *t, const double *q, const double *qs, int count)
{
    char path[4096];
    snprintf(path, sizeof path, "%s/discharge.txt", dir);
    FILE *f = fopen(path, "w");
    if (!f) return -1;
    char line[3 * (SWEHR_FIXED_MAX + 1)];
    int rc = 0;
    for (int r = 0; r < count && !rc; r++) {
        char *o = line;
        o += swehr_io_format_fixed(o, t[r], 6);
        *o++ = '\t';
        o += swehr_io_format_fixed(o, q[r], 6);
        *o++ = '\t';
        o += swehr_io_format_fixed(o, qs[r], 6);
        *o++ = '\n';
        if (fwrite(line, 1, (size_t)(o - line), f) != (size_t)(o - line)) rc = -1;
    }
    return fclose(f) || rc ? -1 : 0;
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
