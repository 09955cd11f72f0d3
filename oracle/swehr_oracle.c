/*
 * swehr_oracle.c - TEST INFRASTRUCTURE.  CPU FP64 restatement of the SWEHR coupled
 * time step (SWEHR.c:808-1699).  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline use it: as the checker the HIP path is compared against, and as the
 * timed CPU baseline.  The product never links it.
 *
 * Fidelity: every floating-point expression keeps the reference's operands, libm
 * calls and left-to-right evaluation order, and the file is built with
 * -ffp-contract=off, so results equal the reference's bit for bit (checked against
 * the compiled reference by tests/test_oracle_golden.py).  Where a libm call is
 * repeated with the same argument in the reference it is evaluated once here; libm
 * functions are pure, so the value is the same.
 *
 * Determinism: the reference reads some geometry of the previous step while the same
 * pass overwrites it (GZ in the fluxes, SLOPEY in GX, the right/upper neighbour slopes
 * in the slope-failure test; SWEHR.c:1135, 1203, 1243, 1293, 1304).  Every such value
 * is a function of the bed elevation at the previous step's flux pass, so the oracle
 * keeps that elevation (zprev) and recomputes them; both passes are then free of
 * loop-order effects and run in parallel over i with OpenMP.
 *
 * Layout: one plane per variable, halo of two cells like the reference's arrays,
 * value (i,j) at IX(i,j) with j unit-stride.  Interior solid cells (SOLID) cannot be
 * set through any input of the reference and are not represented.
 */
#include "swehr_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* constants of the reference, SWEHR.c:489-494, 633-650 */
#define G_ACC 9.81
#define HFILM 1e-5
#define PHI 0.4
#define RHOS 2600.0
#define RHOW 1000.0
#define NU 1.2e-6
#define FRICTION_ANGLE 0.8693
#define COHESION 400.0
#define LAMBDA 0.65
#define CTHRESHOLD 0.4
#define BETABEDLOAD 0.5
#define HCFRICTION 0.003
#define DDEXP (-0.33)
#define B_EXP 1.0
#define AF 1.0
#define AFR 1.0

struct oracle {
    swehr_params p;
    int nx, ny, S, nn;
    long py, plane;
    double oneoverdx, rho0, printint;
    /* persistent state */
    double *h, *hu, *hv, *ch, *M, *z, *zprev, *vinf, *totm, *dep, *maxvel, *maxdepth;
    /* derived, stored after each update pass as the reference does */
    double *sed, *totsed, *rho, *vel1, *vel2;
    /* parameters */
    double *ks, *ds, *pf, *rain;
    /* per-cell suction head, theta_s - theta_0 and Manning n0 (the scalars of SWEHR.c:
     * 237-242 unless a parameter raster is set; no raster in the reference) */
    double *hfc, *dth, *n0c;
    /* step-transient */
    double *fx, *fy, *sx, *sy, *gmx, *gmy, *dfm, *btx, *bty, *gzn, *E, *vf;
    char *fail;
    /* driver state */
    double t, dt, courant, R, stageinterval, printinterval, t_prev;
    long steps;
    int stagecnt, status, done;
    long fail_events; /* cells flagged unstable, summed over steps (diagnostic) */
    double *stage, *saveflow;
    double *discharge;     /* [2][SWEHR_STAGE_MAX]: outlet water (m^3/s) and sediment (kg/s) */
    double qout, sout;     /* outflow through the j = 1/2 face in the last flux pass */
    double z0sum;
    /* budgets: Neumaier sum and compensation per term */
    double bud[SWEHR_BUDGET_N], bud_c[SWEHR_BUDGET_N];
};

#define IX(i, j) ((long)((i) + 1) * o->py + (j) + 1)

static int g_threads = 0;
void oracle_set_threads(int n) { g_threads = n; }
int oracle_get_threads(void)
{
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

static void nsum(double *s, double *c, double x)
{
    double t = *s + x;
    if (fabs(*s) >= fabs(x)) *c += (*s - t) + x;
    else *c += (x - t) + *s;
    *s = t;
}

static double *alloc0(long n) { return (double *)calloc((size_t)n, sizeof(double)); }

/* ---------------------------------------------------------------- point physics */

/* HLL flux across one face, SWEHR.c:327-409.  (qn, qt) are the normal and tangential
 * discharges.  f[0] mass, f[1] normal momentum, f[2] tangential momentum.  Returns the
 * largest |wave speed|.  If every speed is NaN no flux branch fires; the reference
 * then keeps whatever its shared TFLUX held, here the flux is NaN. */
static double face_flux(double hl, double hr, double qnl, double qnr, double qtl, double qtr,
                        double g, double f[3])
{
    double ul, vl, ur, vr;
    if (hl > HFILM) { ul = qnl / hl; vl = qtl / hl; }
    else { ul = 0; vl = 0; }
    if (hr > HFILM) { ur = qnr / hr; vr = qtr / hr; }
    else { ur = 0; vr = 0; }
    const double cl = pow(g * hl, .5), cr = pow(g * hr, .5);
    const double hstar = pow(ul + 2 * cl - ur - 2 * cr, 2) / (16 * g);
    const double ustar = .5 * (ul + ur) + cl - cr;
    const double cstar = pow(g * hstar, .5);
    double sl = ul - cl, sm = ustar, sr = ur + cr;
    if (ustar - cstar < sl) sl = ustar - cstar;
    if (ustar + cstar > sr) sr = ustar + cstar;
    if (hl <= HFILM) { sl = ur - 2 * cr; sm = sl; sr = ur + cr; }
    if (hr <= HFILM) { sl = ul - cl; sr = ul + 2 * cl; sm = sr; }
    f[0] = NAN; f[1] = NAN;
    if (sl >= 0) {
        f[0] = hl * ul;
        f[1] = hl * ul * ul + .5 * g * hl * hl;
    }
    if (sl <= 0 && sr >= 0) {
        f[0] = (sr * (hl * ul) - sl * (hr * ur) + sl * sr * (hr - hl)) / (sr - sl);
        f[1] = (sr * (hl * ul * ul + .5 * g * hl * hl) - sl * (hr * ur * ur + .5 * g * hr * hr)
                + sl * sr * (hr * ur - hl * ul)) / (sr - sl);
    }
    if (sr <= 0) {
        f[0] = hr * ur;
        f[1] = hr * ur * ur + .5 * g * hr * hr;
    }
    f[2] = (sm >= 0) ? f[0] * vl : f[0] * vr;
    double s = 0;
    if (fabs(sl) > s) s = fabs(sl);
    if (fabs(sr) > s) s = fabs(sr);
    if (fabs(sm) > s) s = fabs(sm);
    return s;
}

/* Stream-power entrainment, SWEHR.c:411-426 */
static double stream_power(double uh, double vh, double h, double beta, double rho, double g,
                           double sedconc, double sf, double minmorph, double uc)
{
    const double q = pow(uh * uh + vh * vh, 0.5);
    const double omega = rho * g * sf * q;
    double e = 0;
    if (omega > uc && sedconc < 0.6 && h > minmorph) e = beta * (omega - uc);
    return e;
}

/* Settling velocity in still water (Cao 2006), SWEHR.c:428-438 */
static double settling(double d, double g)
{
    const double s = RHOS / RHOW - 1;
    return pow(pow(13.95 * NU / d, 2) + 1.09 * s * g * d, 0.5) - 13.95 * NU / d;
}

/* bed slopes from an elevation plane, SWEHR.c:1240-1277 (one-sided at the edges) */
static double slope_x(const struct oracle *o, const double *zz, int i, int j)
{
    const int nx = o->nx;
    if (i != 1 && i != nx) return 0.5 * o->oneoverdx * (zz[IX(i + 1, j)] - zz[IX(i - 1, j)]);
    if (i == nx) return o->oneoverdx * (zz[IX(i, j)] - zz[IX(i - 1, j)]);
    return o->oneoverdx * (zz[IX(i + 1, j)] - zz[IX(i, j)]);
}

static double slope_y(const struct oracle *o, const double *zz, int i, int j)
{
    const int ny = o->ny;
    if (j != 1 && j != ny) return 0.5 * o->oneoverdx * (zz[IX(i, j + 1)] - zz[IX(i, j - 1)]);
    if (j == ny) return o->oneoverdx * (zz[IX(i, j)] - zz[IX(i, j - 1)]);
    return o->oneoverdx * (zz[IX(i, j + 1)] - zz[IX(i, j)]);
}

/* GZ of cell (i,j) as of the previous step's flux pass (halo cells take the edge value) */
static double gz_lagged(const struct oracle *o, int i, int j)
{
    if (i < 1) i = 1;
    if (j < 1) j = 1;
    const double a = slope_x(o, o->zprev, i, j), b = slope_y(o, o->zprev, i, j);
    return G_ACC / pow(a * a + b * b + 1, 0.5);
}

/* ------------------------------------------------------------------- boundaries */

/* Ghost cells i=0, nx+1, j=0, ny+1 of h, hu, hv, ch_k, sed_k and the bed halos,
 * SWEHR.c:816-1052.  The second ghost layer is never read by a first-order step. */
static void fill_halo(struct oracle *o)
{
    const int nx = o->nx, ny = o->ny, S = o->S;
    const long P = o->plane;
    const int *bn = o->p.bndry;
    for (int j = -1; j <= ny + 2; j++) {
        const int src[2] = {1, nx}, dst[2] = {0, nx + 1}, side[2] = {0, 2};
        for (int s = 0; s < 2; s++) {
            for (int g = 0; g < 2; g++) {
                const int d = s == 0 ? dst[0] - g : dst[1] + g;
                o->h[IX(d, j)] = o->h[IX(src[s], j)];
                o->hu[IX(d, j)] = bn[side[s]] == 1 ? -o->hu[IX(src[s], j)] : o->hu[IX(src[s], j)];
                o->hv[IX(d, j)] = o->hv[IX(src[s], j)];
                for (int k = 0; k < S; k++) {
                    o->ch[k * P + IX(d, j)] = o->ch[k * P + IX(src[s], j)];
                    o->sed[k * P + IX(d, j)] = o->sed[k * P + IX(src[s], j)];
                }
            }
        }
    }
    for (int i = -1; i <= nx + 2; i++) {
        const int src[2] = {1, ny}, side[2] = {1, 3};
        for (int s = 0; s < 2; s++) {
            for (int g = 0; g < 2; g++) {
                const int d = s == 0 ? 0 - g : ny + 1 + g;
                o->h[IX(i, d)] = o->h[IX(i, src[s])];
                o->hu[IX(i, d)] = o->hu[IX(i, src[s])];
                o->hv[IX(i, d)] = bn[side[s]] == 1 ? -o->hv[IX(i, src[s])] : o->hv[IX(i, src[s])];
                for (int k = 0; k < S; k++) {
                    o->ch[k * P + IX(i, d)] = o->ch[k * P + IX(i, src[s])];
                    o->sed[k * P + IX(i, d)] = o->sed[k * P + IX(i, src[s])];
                }
            }
        }
    }
    if (o->p.inflow) {
        /* water and sediment inflow at j=ny, SWEHR.c:1010-1035 */
        const double t = o->t;
        for (int i = -1; i <= nx + 2; i++) {
            for (int d = ny + 1; d <= ny + 2; d++) {
                o->hu[IX(i, d)] = 0;
                o->hv[IX(i, d)] = fmax(-0.035, t / 100 * -0.035);
                o->h[IX(i, d)] = pow(fabs(o->hv[IX(i, d)]) * 0.045 / pow(0.7, 0.5), 0.6);
                for (int k = 0; k < S; k++) o->ch[k * P + IX(i, d)] = fmin(500, t / 100 * 500) * o->h[IX(i, d)];
                for (int k = 0; k < S; k++) o->sed[k * P + IX(i, d)] = o->ch[k * P + IX(i, d)] / o->h[IX(i, d)];
            }
        }
    }
    /* bed halos (read only by the one-sided slopes of 1-cell-wide grids) */
    for (int j = -1; j <= ny + 2; j++) {
        for (int d = -1; d <= 0; d++) { o->z[IX(d, j)] = o->z[IX(1, j)]; o->zprev[IX(d, j)] = o->zprev[IX(1, j)]; }
        for (int d = nx + 1; d <= nx + 2; d++) { o->z[IX(d, j)] = o->z[IX(nx, j)]; o->zprev[IX(d, j)] = o->zprev[IX(nx, j)]; }
    }
    for (int i = -1; i <= nx + 2; i++) {
        for (int d = -1; d <= 0; d++) { o->z[IX(i, d)] = o->z[IX(i, 1)]; o->zprev[IX(i, d)] = o->zprev[IX(i, 1)]; }
        for (int d = ny + 1; d <= ny + 2; d++) { o->z[IX(i, d)] = o->z[IX(i, ny)]; o->zprev[IX(i, d)] = o->zprev[IX(i, ny)]; }
    }
}

/* ------------------------------------------------------------ flux + source pass */

/* sources of cell (i,j), SWEHR.c:1233-1401 */
static void cell_sources(struct oracle *o, int i, int j)
{
    const int nx = o->nx, ny = o->ny, S = o->S;
    const long P = o->plane, c = IX(i, j);
    const double odx = o->oneoverdx, g = G_ACC, h = o->h[c];
    const double *ps = &o->p.a;
    (void)ps;

    /* geometry: own slopes now, SLOPEY of the previous step in GX */
    const double sxn = slope_x(o, o->z, i, j);
    const double syo = slope_y(o, o->zprev, i, j);
    const double gx = g * sxn / pow(sxn * sxn + syo * syo + 1, 0.5);
    o->sx[c] = -gx * h;
    const double syn = slope_y(o, o->z, i, j);
    const double nrm = pow(sxn * sxn + syn * syn + 1, 0.5);
    const double gy = g * syn / nrm;
    o->sy[c] = -gy * h;
    const double gz = g / nrm;
    o->gzn[c] = gz;
    const double slope = pow(sxn * sxn + syn * syn, 0.5);

    /* variable-density forcing, SWEHR.c:1282-1290 */
    const double rho = o->rho[c];
    double gmx = 0, gmy = 0;
    if (i > 1 && i < nx) gmx -= (RHOS - rho) * gz * h * h / (2 * rho) * .5 * odx * (o->totsed[IX(i + 1, j)] - o->totsed[IX(i - 1, j)]);
    if (j > 1 && j < ny) gmy -= (RHOS - rho) * gz * h * h / (2 * rho) * .5 * odx * (o->totsed[IX(i, j + 1)] - o->totsed[IX(i, j - 1)]);
    o->gmx[c] = gmx;
    o->gmy[c] = gmy;

    /* slope stability of the deposited layer, SWEHR.c:1292-1312.  Right/upper neighbour
     * slopes are those of the previous step, left/lower ones (and the cell's own at the
     * edges) those of this step. */
    const double dep = o->dep[c], rho0 = o->rho0;
    const int iu = i < nx ? i + 1 : nx, id = i > 1 ? i - 1 : 1;
    const int ju = j < ny ? j + 1 : ny, jd = j > 1 ? j - 1 : 1;
    const double sx_up = (iu == i) ? sxn : slope_x(o, o->zprev, iu, j);
    const double sx_dn = (id == i) ? sxn : slope_x(o, o->z, id, j);
    const double sy_up = (ju == j) ? syn : slope_y(o, o->zprev, i, ju);
    const double sy_dn = (jd == j) ? syn : slope_y(o, o->z, i, jd);
    char fail = 0;
    const double atsy = atan(fabs(syn));
    const double load = rho0 * gz * dep + rho * gz * h;
    const double lhs = load * sin(atsy);
    const double base = load * cos(atsy) - RHOW * gz * (dep + h);
    double pgfx = dep * rho0 * gz * 0.5 * odx * (o->dep[IX(iu, j)] * cos(atan(fabs(sx_up))) - o->dep[IX(id, j)] * cos(atan(fabs(sx_dn))));
    if (pgfx * sxn > 0) pgfx = -1 * fabs(pgfx);
    if (pgfx * sxn < 0) pgfx = fabs(pgfx);
    if (fail == 0 && lhs > (COHESION + base * FRICTION_ANGLE + pgfx)) fail = 1;
    double pgfy = dep * rho0 * gz * 0.5 * odx * (o->dep[IX(i, ju)] * cos(atan(fabs(sy_up))) - o->dep[IX(i, jd)] * cos(atan(fabs(sy_dn))));
    if (pgfy * syn > 0) pgfy = -1 * fabs(pgfy);
    if (pgfy * syn < 0) pgfy = fabs(pgfy);
    if (fail == 0 && lhs > (COHESION + base * FRICTION_ANGLE + pgfy)) fail = 1;
    o->fail[c] = fail;

    /* debris-flow forcing, SWEHR.c:1314-1320 */
    const double ts = o->totsed[c];
    double dfm = fmin(1, (ts - 0.2) / 0.2);
    if (h < 9 * o->p.h0 || ts < 0.2) dfm = 0;
    double btx = -(1 - LAMBDA) * gz * h * FRICTION_ANGLE;
    double bty = -(1 - LAMBDA) * gz * h * FRICTION_ANGLE;
    if (o->hu[c] < 0) btx = -1 * btx;
    if (o->hv[c] < 0) bty = -1 * bty;
    o->dfm[c] = dfm;
    o->btx[c] = btx;
    o->bty[c] = bty;

    /* Hairsine-Rose sources per size class, SWEHR.c:1322-1401 */
    const double h0 = o->p.h0, totm = o->totm[c], R = o->R;
    const double hvert = h / cos(atan(fabs(slope)));
    double mtstar = o->p.mtstar0;
    if (hvert > h0) mtstar = o->p.mtstar0 * pow(h0 / hvert, B_EXP);
    const double H = fmin(1, totm / mtstar);
    const double n0 = o->n0c[c];
    const double manning = (h <= HCFRICTION) ? n0 * pow(h / HCFRICTION, DDEXP) : n0;
    const double q = pow(o->hu[c] * o->hu[c] + o->hv[c] * o->hv[c], 0.5);
    const double sf = manning * manning * q * q / pow(h, 3.333);
    const double hindered = pow(1 - fmin(ts, 0.99), 4);
    const int active = h > 10 * HFILM && ts < 0.6;
    for (int k = 0; k < S; k++) {
        const double mk = o->M[k * P + c], dk = o->ds[k];
        double e1 = 0, e2 = 0, e3 = 0, e4 = 0;
        if (active) {
            e1 = (1 - H) * o->pf[k] * o->p.a * R;
            if (totm > 1e-8 && mk > 1e-8) e2 = H * mk / totm * o->p.ad * R;
            if (hvert > h0) {
                e1 = e1 * pow(h0 / hvert, B_EXP);
                e2 = e2 * pow(h0 / hvert, B_EXP);
            }
        }
        if (active) {
            e3 = stream_power(o->hu[c], o->hv[c], h, AF, rho, gz, ts, sf, dk, 0.0);
            e3 = (1 - H) * o->pf[k] * o->p.F / o->p.J * e3;
            if (totm > 1e-6 && mk > 0) {
                e4 = stream_power(o->hu[c], o->hv[c], h, AFR, rho, gz, ts, sf, dk, 0.0);
                e4 = (1 - BETABEDLOAD) * H * mk / totm * o->p.F / ((RHOS - rho) / RHOS * gz * h) * e4;
            } else {
                e4 = 0;
            }
        }
        double vf = settling(dk, gz);
        vf = vf * hindered;
        if (dk > 0.0001 && h < 9 * h0) vf = 2e-3 + fmax(h - 3 * h0, 0) * (vf - 2e-3) / (6 * h0);
        if (dk > 0.002 && h < 9 * h0) vf = 2e-3 + fmax(h - 3 * h0, 0) * (vf - 2e-3) / (6 * h0);
        if (h > 9 * h0 && ts > CTHRESHOLD) { e1 = e2 = e3 = e4 = 0; vf = 0; }
        double *E = o->E + (long)4 * k * P;
        E[c] = e1; E[P + c] = e2; E[2 * P + c] = e3; E[3 * P + c] = e4;
        o->vf[k * P + c] = vf;
    }
}

static double flux_pass(struct oracle *o)
{
    const int nx = o->nx, ny = o->ny, S = o->S, nn = o->nn;
    const long P = o->plane;
    double smax = 0;
    long fails = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) reduction(max : smax) reduction(+ : fails) num_threads(oracle_get_threads())
#endif
    for (int i = 0; i <= nx; i++) {
        for (int j = 0; j <= ny; j++) {
            double f[3], s;
            if (j != 0) { /* face i+1/2, SWEHR.c:1097-1163 */
                const long a = IX(i, j), b = IX(i + 1, j), fo = IX(i + 1, j);
                s = face_flux(o->h[a], o->h[b], o->hu[a], o->hu[b], o->hv[a], o->hv[b], gz_lagged(o, i, j), f);
                o->fx[0 * P + fo] = f[0];
                o->fx[1 * P + fo] = f[1];
                o->fx[2 * P + fo] = f[2];
                for (int k = 0; k < S; k++)
                    o->fx[(3 + k) * P + fo] = f[0] > 0 ? f[0] * o->sed[k * P + a] : f[0] * o->sed[k * P + b];
                if (s > smax) smax = s;
            }
            if (i != 0) { /* face j+1/2, normal momentum is hv, SWEHR.c:1165-1230 */
                const long a = IX(i, j), b = IX(i, j + 1), fo = IX(i, j + 1);
                s = face_flux(o->h[a], o->h[b], o->hv[a], o->hv[b], o->hu[a], o->hu[b], gz_lagged(o, i, j), f);
                o->fy[0 * P + fo] = f[0];
                o->fy[2 * P + fo] = f[1];
                o->fy[1 * P + fo] = f[2];
                for (int k = 0; k < S; k++)
                    o->fy[(3 + k) * P + fo] = f[0] > 0 ? f[0] * o->sed[k * P + a] : f[0] * o->sed[k * P + b];
                if (s > smax) smax = s;
            }
            if (i != 0 && j != 0) {
                cell_sources(o, i, j);
                fails += o->fail[IX(i, j)];
            }
        }
    }
    o->fail_events += fails;
    (void)nn;
    return smax;
}

/* ------------------------------------------------------------------ update pass */

static void update_pass(struct oracle *o)
{
    const int nx = o->nx, ny = o->ny, S = o->S;
    const long P = o->plane;
    const double dt = o->dt, R = o->R, odx = o->oneoverdx, dx2 = o->p.dx * o->p.dx;
    const double kbed = 1 / (RHOS * (1 - PHI));
    double b_exch = 0, b_exch_c = 0, b_clamp = 0, b_clamp_c = 0, b_det = 0, b_det_c = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(oracle_get_threads())
#endif
    for (int i = 1; i <= nx; i++) {
        double le = 0, lec = 0, lc = 0, lcc = 0, ld = 0, ldc = 0;
        for (int j = 1; j <= ny; j++) {
            const long c = IX(i, j), xr = IX(i + 1, j), yr = IX(i, j + 1);
            double chtemp = 0, redetach;
            /* conservative update, SWEHR.c:1424-1441 */
            double h = o->h[c] - dt * odx * (o->fx[0 * P + xr] - o->fx[0 * P + c]) - dt * odx * (o->fy[0 * P + yr] - o->fy[0 * P + c]) + dt * R;
            double hu = o->hu[c] - dt * odx * (o->fx[1 * P + xr] - o->fx[1 * P + c]) - dt * odx * (o->fy[1 * P + yr] - o->fy[1 * P + c])
                        + dt * o->sx[c] + dt * o->gmx[c] + dt * o->dfm[c] * o->btx[c];
            double hv = o->hv[c] - dt * odx * (o->fx[2 * P + xr] - o->fx[2 * P + c]) - dt * odx * (o->fy[2 * P + yr] - o->fy[2 * P + c])
                        + dt * o->sy[c] + dt * o->gmy[c] + dt * o->dfm[c] * o->bty[c];
            /* Green-Ampt infiltration; negative when h < hfilm (film clamp), SWEHR.c:1448-1452 */
            const double ks = o->ks[c];
            const double zf = o->vinf[c] / o->dth[c];
            const double infl = fmin(h - HFILM, dt * ks * (zf + o->hfc[c] + h) / zf);
            o->vinf[c] += infl;
            h -= infl;
            double chk[SWEHR_MAX_CLASSES];
            for (int k = 0; k < S; k++)
                chk[k] = o->ch[k * P + c] - dt * odx * (o->fx[(3 + k) * P + xr] - o->fx[(3 + k) * P + c])
                         - dt * odx * (o->fy[(3 + k) * P + yr] - o->fy[(3 + k) * P + c]);
            /* Hairsine-Rose exchange per class, SWEHR.c:1461-1511 */
            double totm = 0, dep = 0, z = o->z[c];
            const int fail = o->fail[c];
            for (int k = 0; k < S; k++) {
                const double *E = o->E + (long)4 * k * P;
                const double e1 = E[c], e2 = E[P + c], e3 = E[2 * P + c], e4 = E[3 * P + c];
                const double vf = o->vf[k * P + c];
                double m = o->M[k * P + c], u = chk[k];
                if (h < 10 * HFILM) {
                    m += dt * e1 + u;
                    nsum(&ld, &ldc, dt * e1);
                    u = 0;
                } else {
                    chtemp = u;
                    redetach = fmin(m, dt * (e2 + e4));
                    const double decay = exp(-vf / h * (0.5 * dt));
                    u = decay * u;
                    u = u + dt * (e1 + e3) + redetach;
                    u = decay * u;
                    z = z - 1 / (RHOS * (1 - PHI)) * (u - chtemp);
                    m = m + dt * (e1 + e3) - (u - chtemp);
                    nsum(&ld, &ldc, dt * (e1 + e3));
                    if (fail == 1) {
                        z = z - 0.99 * m / (RHOS * (1 - PHI));
                        u = u + 0.99 * m;
                        totm = 0;
                        m = 0.01 * m;
                        dep = 0;
                    }
                }
                const double dh = (u - chtemp) / (RHOS * (1 - PHI));
                h = h + dh;
                nsum(&le, &lec, dh);
                totm += m;
                dep += m / (RHOS);
                chk[k] = u;
                o->M[k * P + c] = m;
            }
            /* concentrations and the dry-cell clamp, SWEHR.c:1513-1551 */
            double ts = 0, v1, v2;
            if (h > HFILM) {
                v1 = hu / h;
                v2 = hv / h;
                for (int k = 0; k < S; k++) {
                    double sd = chk[k] / h;
                    ts += sd / RHOS;
                    if (sd < -1e-5) sd = 0;
                    o->sed[k * P + c] = sd;
                }
            } else {
                nsum(&lc, &lcc, HFILM - h);
                h = HFILM;
                hu = 0;
                hv = 0;
                v1 = 0;
                v2 = 0;
                for (int k = 0; k < S; k++) {
                    z = z + 1 / (RHOS * (1 - PHI)) * chk[k];
                    o->M[k * P + c] = o->M[k * P + c] + chk[k];
                    chk[k] = 0;
                    o->sed[k * P + c] = 0;
                }
            }
            /* implicit friction (LeVeque), SWEHR.c:1553-1565 */
            if (h > HFILM) {
                const double n0 = o->n0c[c];
                const double manning = (h <= HCFRICTION) ? n0 * pow(h / HCFRICTION, DDEXP) : n0;
                const double sf = manning * manning * o->gzn[c] * pow(hu * hu + hv * hv, 0.5) * 1 / pow(h, 2.333);
                hu = hu / (1 + dt * sf);
                hv = hv / (1 + dt * sf);
                v1 = hu / h;
                v2 = hv / h;
            }
            /* density, maxima, SWEHR.c:1578-1588 */
            double rho = RHOW * (1 - ts) + RHOS * ts;
            if (rho > 0.8 * RHOS) rho = 0.8 * RHOS;
            const double speed = pow(v1 * v1 + v2 * v2, 0.5);
            if (speed > o->maxvel[c]) o->maxvel[c] = speed;
            if (h > o->maxdepth[c]) o->maxdepth[c] = h;
            o->h[c] = h;
            o->hu[c] = hu;
            o->hv[c] = hv;
            for (int k = 0; k < S; k++) o->ch[k * P + c] = chk[k];
            o->z[c] = z;
            o->totm[c] = totm;
            o->dep[c] = dep;
            o->totsed[c] = ts;
            o->rho[c] = rho;
            o->vel1[c] = v1;
            o->vel2[c] = v2;
        }
#ifdef _OPENMP
#pragma omp critical
#endif
        {
            nsum(&b_exch, &b_exch_c, le + lec);
            nsum(&b_clamp, &b_clamp_c, lc + lcc);
            nsum(&b_det, &b_det_c, ld + ldc);
        }
    }
    (void)kbed;
    nsum(&o->bud[10], &o->bud_c[10], (b_exch + b_exch_c) * dx2);
    nsum(&o->bud[9], &o->bud_c[9], (b_clamp + b_clamp_c) * dx2);
    nsum(&o->bud[8], &o->bud_c[8], (b_det + b_det_c) * dx2);
}

/* net outflow through the four edges during this step (water m^3, sediment kg) */
static void boundary_budget(struct oracle *o)
{
    const int nx = o->nx, ny = o->ny, S = o->S;
    const long P = o->plane;
    const double w = o->dt * o->p.dx;
    double wq = 0, wc = 0, sq = 0, scc = 0;
    for (int j = 1; j <= ny; j++) {
        nsum(&wq, &wc, -o->fx[IX(1, j)]);
        nsum(&wq, &wc, o->fx[IX(nx + 1, j)]);
        for (int k = 0; k < S; k++) {
            nsum(&sq, &scc, -o->fx[(3 + k) * P + IX(1, j)]);
            nsum(&sq, &scc, o->fx[(3 + k) * P + IX(nx + 1, j)]);
        }
    }
    for (int i = 1; i <= nx; i++) {
        nsum(&wq, &wc, -o->fy[IX(i, 1)]);
        nsum(&wq, &wc, o->fy[IX(i, ny + 1)]);
        for (int k = 0; k < S; k++) {
            nsum(&sq, &scc, -o->fy[(3 + k) * P + IX(i, 1)]);
            nsum(&sq, &scc, o->fy[(3 + k) * P + IX(i, ny + 1)]);
        }
    }
    nsum(&o->bud[3], &o->bud_c[3], (wq + wc) * w);
    nsum(&o->bud[7], &o->bud_c[7], (sq + scc) * w);
    /* outlet discharge: outflow through the j = 1/2 face, the one open side by default
     * (SWEHR.c:618-621); per unit width times dx, summed over i (not in the reference) */
    double q = 0, qs = 0;
    for (int i = 1; i <= nx; i++) {
        q += -o->fy[IX(i, 1)] * o->p.dx;
        double si = 0;
        for (int k = 0; k < S; k++) si -= o->fy[(3 + k) * P + IX(i, 1)];
        qs += si * o->p.dx;
    }
    o->qout = q;
    o->sout = qs;
}

/* outlet and probe records, SWEHR.c:1593-1676 */
static void sample_stage(struct oracle *o)
{
    const int nx = o->nx, ny = o->ny, S = o->S;
    const long P = o->plane;
    int outlet = 1;
    double best = 0;
    for (int i = 1; i <= nx; i++)
        if (o->h[IX(i, 1)] > best) { best = o->h[IX(i, 1)]; outlet = i; }
    o->stageinterval += 1;
    if (o->stagecnt >= SWEHR_STAGE_MAX) { o->status = SWEHR_ERR_STAGE_FULL; return; }
    const int r = o->stagecnt++;
    double *st = o->stage;
    const long c = IX(outlet, 1);
    st[0 * SWEHR_STAGE_MAX + r] = o->t;
    st[1 * SWEHR_STAGE_MAX + r] = o->h[c];
    st[2 * SWEHR_STAGE_MAX + r] = pow(o->vel1[c] * o->vel1[c] + o->vel2[c] * o->vel2[c], 0.5);
    st[3 * SWEHR_STAGE_MAX + r] = o->totsed[c];
    o->discharge[0 * SWEHR_STAGE_MAX + r] = o->qout;
    o->discharge[1 * SWEHR_STAGE_MAX + r] = o->sout;
    for (int k = 0; k < 10; k++) {
        st[(4 + k) * SWEHR_STAGE_MAX + r] = k < S ? o->M[k * P + c] : 0;
        st[(14 + k) * SWEHR_STAGE_MAX + r] = k < S ? o->ch[k * P + c] : 0;
    }
    static const int yp[6] = {1, 10, 20, 30, 40, 50};
    for (int n = 0; n < 6; n++) {
        double d = 0, sp = 0, ts = 0;
        if (2 <= nx && yp[n] <= ny) {
            const long q = IX(2, yp[n]);
            d = o->h[q];
            sp = pow(o->vel1[q] * o->vel1[q] + o->vel2[q] * o->vel2[q], 0.5);
            ts = o->totsed[q];
        }
        o->saveflow[(3 * n + 0) * SWEHR_STAGE_MAX + r] = d;
        o->saveflow[(3 * n + 1) * SWEHR_STAGE_MAX + r] = sp;
        o->saveflow[(3 * n + 2) * SWEHR_STAGE_MAX + r] = ts;
    }
}

int oracle_step(oracle_t *o, int *print_due)
{
    if (print_due) *print_due = 0;
    if (o->done) return o->status;
    const int nx = o->nx, ny = o->ny;
    /* rain rate, SWEHR.c:812-814 (int conversion as in the reference) */
    int rind = (int)fmin(ceil(o->t / o->p.rint + 0.01), o->p.rnum);
    if (rind < 1) rind = 1;
    o->R = o->rain[rind - 1];
    fill_halo(o);
    const double smax = flux_pass(o);
    /* adaptive time step, SWEHR.c:1409-1413 */
    double dt = o->p.cstable * o->p.dx / (smax);
    if (dt > o->p.maxstep) dt = o->p.maxstep;
    if (o->t + dt > o->p.tend) dt = o->p.tend - o->t;
    o->dt = dt;
    o->courant = (smax)*dt * o->oneoverdx;
    int cfl = 0;
    if (o->courant > 1) cfl = 1;
    /* the bed of this flux pass becomes the lagged bed of the next one */
    memcpy(o->zprev, o->z, sizeof(double) * (size_t)o->plane);
    boundary_budget(o);
    update_pass(o);
    nsum(&o->bud[2], &o->bud_c[2], dt * o->R * (double)nx * ny * o->p.dx * o->p.dx);
    o->bud[11] += 1;
    o->t_prev = o->t;
    if (cfl) { o->t = 2 * o->p.tend; o->status = SWEHR_ERR_CFL; }
    o->t += dt;
    o->steps++;
    if (o->t > o->stageinterval) sample_stage(o);
    if (o->t >= o->printinterval) {
        o->printinterval += o->printint;
        if (print_due) *print_due = 1;
    }
    if (!(o->t < o->p.tend) || o->status != SWEHR_OK) o->done = 1;
    return o->status;
}

int oracle_run(oracle_t *o, long n)
{
    for (long s = 0; s < n && !o->done; s++) oracle_step(o, NULL);
    return o->status;
}

void oracle_info(const oracle_t *o, swehr_step_info *info)
{
    info->t = o->t;
    info->dt = o->dt;
    info->courant = o->courant;
    info->rain = o->R;
    info->t_prev = o->t_prev;
    info->steps = o->steps;
    info->done = o->done;
    info->print_due = 0;
    info->stage_count = o->stagecnt;
    info->status = o->status;
}

oracle_t *oracle_create(const swehr_params *p, const double *topo, const double *ks,
                        const double *ds, const double *pfrac, const double *rain)
{
    if (!p || p->nx < 1 || p->ny < 1 || p->nclass < 1 || p->nclass > SWEHR_MAX_CLASSES || p->rnum < 1 || !(p->dx > 0))
        return NULL;
    struct oracle *o = (struct oracle *)calloc(1, sizeof(*o));
    o->p = *p;
    o->nx = p->nx;
    o->ny = p->ny;
    o->S = p->nclass;
    o->nn = 3 + o->S;
    o->py = o->ny + 4;
    o->plane = (long)(o->nx + 4) * o->py;
    o->oneoverdx = 1 / p->dx;
    o->rho0 = RHOW * PHI + RHOS * (1 - PHI);
    const long P = o->plane;
    const int S = o->S;
    o->h = alloc0(P); o->hu = alloc0(P); o->hv = alloc0(P);
    o->ch = alloc0(S * P); o->M = alloc0(S * P); o->sed = alloc0(S * P);
    o->z = alloc0(P); o->zprev = alloc0(P); o->vinf = alloc0(P); o->totm = alloc0(P); o->dep = alloc0(P);
    o->maxvel = alloc0(P); o->maxdepth = alloc0(P); o->totsed = alloc0(P); o->rho = alloc0(P);
    o->vel1 = alloc0(P); o->vel2 = alloc0(P); o->ks = alloc0(P);
    o->hfc = alloc0(P); o->dth = alloc0(P); o->n0c = alloc0(P);
    o->fx = alloc0((3 + S) * P); o->fy = alloc0((3 + S) * P);
    o->sx = alloc0(P); o->sy = alloc0(P); o->gmx = alloc0(P); o->gmy = alloc0(P);
    o->dfm = alloc0(P); o->btx = alloc0(P); o->bty = alloc0(P); o->gzn = alloc0(P);
    o->E = alloc0(4L * S * P); o->vf = alloc0(S * P);
    o->fail = (char *)calloc((size_t)P, 1);
    o->ds = alloc0(S); o->pf = alloc0(S); o->rain = alloc0(p->rnum);
    o->stage = alloc0((long)SWEHR_STAGE_COLS * SWEHR_STAGE_MAX);
    o->saveflow = alloc0((long)SWEHR_FLOW_COLS * SWEHR_STAGE_MAX);
    o->discharge = alloc0(2L * SWEHR_STAGE_MAX);
    memcpy(o->ds, ds, sizeof(double) * S);
    memcpy(o->pf, pfrac, sizeof(double) * S);
    memcpy(o->rain, rain, sizeof(double) * p->rnum);
    /* initial condition, SWEHR.c:223-286, 656-667: zero depth, momentum and sediment
     * (the reference never fills U[1]; its heap is zero), VINF = vinf0, RHO = rhow */
    double zs = 0, zc = 0;
    for (int i = 1; i <= o->nx; i++)
        for (int j = 1; j <= o->ny; j++) {
            const long c = IX(i, j);
            const long n = (long)(i - 1) * o->ny + (j - 1);
            o->z[c] = topo[n];
            o->zprev[c] = topo[n];
            o->vinf[c] = p->vinf0;
            o->rho[c] = RHOW;
            o->ks[c] = ks ? ks[n] : p->ks;
            o->hfc[c] = p->hf;
            o->dth[c] = p->thetas - p->theta0; /* the difference SWEHR.c:1449 forms */
            o->n0c[c] = p->roughness;
            nsum(&zs, &zc, topo[n]);
        }
    o->z0sum = zs + zc;
    o->t = 0;
    o->dt = 0.0001;
    o->stageinterval = 0;
    o->printinterval = fmax(1, p->tend / 1000);
    o->printint = fmax(1, p->tend / 1000);
    o->status = SWEHR_OK;
    o->done = !(o->t < p->tend);
    return o;
}

static double speed_of(const struct oracle *o, long c)
{
    return pow(o->vel1[c] * o->vel1[c] + o->vel2[c] * o->vel2[c], 0.5);
}

int oracle_download(const oracle_t *o, int field, int k, double *host)
{
    const long P = o->plane;
    if ((field == SWEHR_F_CH || field == SWEHR_F_M || field == SWEHR_F_SED) && (k < 1 || k > o->S))
        return SWEHR_ERR_BAD_INPUT;
    for (int i = 1; i <= o->nx; i++)
        for (int j = 1; j <= o->ny; j++) {
            const long c = IX(i, j);
            double v;
            switch (field) {
            case SWEHR_F_DEPTH: v = o->h[c]; break;
            case SWEHR_F_UH: v = o->hu[c]; break;
            case SWEHR_F_VH: v = o->hv[c]; break;
            case SWEHR_F_CH: v = o->ch[(k - 1) * P + c]; break;
            case SWEHR_F_M: v = o->M[(k - 1) * P + c]; break;
            case SWEHR_F_TOPO: v = o->z[c]; break;
            case SWEHR_F_VINF: v = o->vinf[c]; break;
            case SWEHR_F_TOTM: v = o->totm[c]; break;
            case SWEHR_F_DEP: v = o->dep[c]; break;
            case SWEHR_F_TOTSED: v = o->totsed[c]; break;
            case SWEHR_F_SPEED: v = speed_of(o, c); break;
            case SWEHR_F_MAXVEL: v = o->maxvel[c]; break;
            case SWEHR_F_MAXDEPTH: v = o->maxdepth[c]; break;
            case SWEHR_F_SED: v = o->sed[(k - 1) * P + c]; break;
            case SWEHR_F_VELX: v = o->vel1[c]; break;
            case SWEHR_F_VELY: v = o->vel2[c]; break;
            default: return SWEHR_ERR_BAD_INPUT;
            }
            host[(long)(i - 1) * o->ny + (j - 1)] = v;
        }
    return SWEHR_OK;
}

int oracle_set_rasters(oracle_t *o, const double *hf, const double *dtheta, const double *roughness)
{
    if (!o) return SWEHR_ERR_BAD_INPUT;
    for (int i = 1; i <= o->nx; i++)
        for (int j = 1; j <= o->ny; j++) {
            const long c = IX(i, j);
            const long n = (long)(i - 1) * o->ny + (j - 1);
            if (hf) o->hfc[c] = hf[n];
            if (dtheta) o->dth[c] = dtheta[n];
            if (roughness) o->n0c[c] = roughness[n];
        }
    return SWEHR_OK;
}

int oracle_discharge(const oracle_t *o, double *q, double *qs, int capacity, int *count)
{
    const int n = o->stagecnt < capacity ? o->stagecnt : capacity;
    if (q) memcpy(q, o->discharge, sizeof(double) * n);
    if (qs) memcpy(qs, o->discharge + SWEHR_STAGE_MAX, sizeof(double) * n);
    if (count) *count = o->stagecnt;
    return SWEHR_OK;
}

int oracle_stage(const oracle_t *o, double *stage, double *saveflow, int capacity, int *count)
{
    const int n = o->stagecnt < capacity ? o->stagecnt : capacity;
    for (int c = 0; c < SWEHR_STAGE_COLS; c++)
        if (stage) memcpy(stage + (long)c * capacity, o->stage + (long)c * SWEHR_STAGE_MAX, sizeof(double) * n);
    for (int c = 0; c < SWEHR_FLOW_COLS; c++)
        if (saveflow) memcpy(saveflow + (long)c * capacity, o->saveflow + (long)c * SWEHR_STAGE_MAX, sizeof(double) * n);
    if (count) *count = o->stagecnt;
    return SWEHR_OK;
}

int oracle_budget(const oracle_t *o, double *out)
{
    const long P = o->plane;
    const double a = o->p.dx * o->p.dx;
    double s[8] = {0}, c[8] = {0};
    for (int i = 1; i <= o->nx; i++)
        for (int j = 1; j <= o->ny; j++) {
            const long q = IX(i, j);
            nsum(&s[0], &c[0], o->h[q]);
            nsum(&s[1], &c[1], o->vinf[q] - o->p.vinf0);
            for (int k = 0; k < o->S; k++) {
                nsum(&s[4], &c[4], o->ch[k * P + q]);
                nsum(&s[5], &c[5], o->M[k * P + q]);
            }
            nsum(&s[6], &c[6], o->z[q]);
        }
    out[0] = (s[0] + c[0]) * a;
    out[1] = (s[1] + c[1]) * a;
    out[2] = o->bud[2] + o->bud_c[2];
    out[3] = o->bud[3] + o->bud_c[3];
    out[4] = (s[4] + c[4]) * a;
    out[5] = (s[5] + c[5]) * a;
    out[6] = ((s[6] + c[6]) - o->z0sum) * a;
    for (int n = 7; n < SWEHR_BUDGET_N; n++) out[n] = o->bud[n] + o->bud_c[n];
    return SWEHR_OK;
}

long oracle_fail_events(const oracle_t *o) { return o->fail_events; }

void oracle_destroy(oracle_t *o)
{
    if (!o) return;
    double *ptrs[] = {o->h, o->hu, o->hv, o->ch, o->M, o->sed, o->z, o->zprev, o->vinf, o->totm, o->dep,
                      o->maxvel, o->maxdepth, o->totsed, o->rho, o->vel1, o->vel2, o->ks, o->hfc, o->dth, o->n0c,
                      o->fx, o->fy,
                      o->sx, o->sy, o->gmx, o->gmy, o->dfm, o->btx, o->bty, o->gzn, o->E, o->vf,
                      o->ds, o->pf, o->rain, o->stage, o->saveflow, o->discharge};
    for (size_t n = 0; n < sizeof(ptrs) / sizeof(ptrs[0]); n++) free(ptrs[n]);
    free(o->fail);
    free(o);
}
