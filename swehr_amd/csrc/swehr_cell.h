// swehr_cell.h - the update of one cell in the SWEHR step (SWEHR.c:1239-1589), shared
// by the step kernels that gather a cell's neighbourhood in different ways (LDS tiles,
// register-marching strips).  A kernel supplies the cell's neighbourhood (TOTSED and
// DEP of the four neighbours, the slopes, the four faces' fluxes), then:
//   cell_begin   geometry, variable-density and debris forcing, slope failure, the
//                class-invariant factors of the Hairsine-Rose sources, the conservative
//                update of h, hu, hv and Green-Ampt infiltration (SWEHR.c:1239-1452)
//   cell_class   one size class: flux divergence from the four faces' class fluxes,
//                sources, settling and the exchange with the bed, slope-failure
//                remobilisation, the water carried by the exchanged sediment volume
//                (SWEHR.c:1322-1399, 1454-1511), for k = 0 .. S-1 in order
//   cell_end     dry-cell clamp, implicit friction, maxima, stores and the cell's
//                boundary-outflow budget terms (SWEHR.c:1513-1589)
// Every expression keeps the operation order the tile kernel had when it held this code
// inline, so both kernels give the same bits.
// Included by swehr_device.h after the argument block and the memory helpers it uses.
#pragma once

namespace swehr {

// What a kernel gathers for one cell before its update.  Neighbour values in a
// direction the cell has no neighbour in (the domain's edge) are the cell's own, as
// the one-sided rules of SWEHR.c:1240-1312 have it.
template <typename R>
struct CellIn {
    R h, hu, hv, ts;            // own depth (from the staged state), momenta, TOTSED
    R ts_xm, ts_xp, ts_ym, ts_yp;  // TOTSED at i-1, i+1, j-1, j+1 (read only inside the domain)
    R dep, dep_xu, dep_xd;      // DEP own, at i+1 and at i-1 (own at the edges)
    R dep_yu, dep_yd;           // DEP at j+1 and j-1 (own at the edges)
    R sxn, syo, syn;            // SLOPEX of this step, SLOPEY of the previous and of this step
    R sx_up, sx_dn;             // SLOPEX of the previous step at i+1, of this step at i-1
    R fxl[3], fxr[3];           // x faces i-1/2, i+1/2: mass, hu and hv fluxes
    R fyl[3], fyr[3];           // y faces j-1/2, j+1/2: mass, hu and hv fluxes
    R totm_o, mv_o, md_o;       // TOTM, MAXVEL, MAXDEPTH before the step
    double vinf_o, ks_o;        // VINF and Ks of the cell
    double zn;                  // bed elevation before the step
};

// The cell's state through the class loop.
template <typename R>
struct Cell {
    int gi, gj;
    bool own, edge, fail, film, e2ok, e4ok, shallow_v, left_suspended;
    R h, dt, hn, hun, hvn, dh_rest, gz, rough, hdt, dx_;
    R f1, f2, f3, f4, omega, vf_keep, hind, vshal, cg;
    R fx0l, fx0r, fy0l, fy0r;
    R totm, depn, chtemp, mv_o, md_o, totm_o;
    double zn_c, sb_out, sq_out;
    double bud[NB_TERMS];
    // FP32: the class loop's bed, deposition and exchange increments summed in FP32 and
    // added to zn_c and the budget terms once per cell (cell_end)
    R dz, bdet, bexch;
    unsigned rbits;
};

// ---------------------------------------------------------------------------------
// ysl() returns {SLOPEY of the previous step at j+1, SLOPEY of this step at j-1} (the
// cell's own slopes at the edges); it is called only when the x check of the slope
// failure test does not already fail (SWEHR.c:1296-1307).
template <typename R, typename YSlopes>
__device__ __forceinline__ void cell_begin(Cell<R> &X, const StepArgs<R> &A, const CellIn<R> &in, double dtd,
                                           double Rd, long long steps, unsigned c, const double *s_exp2,
                                           const double *s_log, unsigned c8L, YSlopes ysl)
{
    const int nx = A.nx, ny = A.ny, gi = X.gi, gj = X.gj;
    const R hfilm = R(kHfilm);
    const R rhos = R(kRhos), rhow = R(kRhow);
    const R g = R(kG);
    const R odx = R(A.inv_dx);
    const R h = in.h, hu = in.hu, hv = in.hv;
    const R ts = in.ts;
    const R dt = R(dtd), Rr = R(Rd);
    const R h0 = R(A.h0);
    X.h = h;
    X.dt = dt;
    for (int k = 0; k < NB_TERMS; k++) X.bud[k] = 0;

    // geometry of this step; GX takes SLOPEY of the previous step (SWEHR.c:1239-1280)
    const R sxn = in.sxn, syo = in.syo, syn = in.syn;
    const R gx = g * sxn * mrsqrt(sxn * sxn + syo * syo + R(1));
    const R sxs = -gx * h;
    const R n2 = sxn * sxn + syn * syn + R(1);
    const R rn = mrsqrt(n2);
    const R gy = g * syn * rn;
    const R sys = -gy * h;
    const R gz = g * rn;
    X.gz = gz;

    // mixture density from the previous update (SWEHR.c:1578-1579)
    R rho = rhow * (R(1) - ts) + rhos * ts;
    if (rho > thr_gt<R>(0.8 * kRhos)) rho = R(0.8) * rhos;

    // variable-density forcing (SWEHR.c:1282-1290); the same factor in x and y
    R gmx = R(0), gmy = R(0);
    const bool gx_in = gi > 0 && gi < nx - 1, gy_in = gj > 0 && gj < ny - 1;
    if (gx_in || gy_in) {
        const R fac = (rhos - rho) * gz * h * h * mrcp(R(2) * rho) * R(.5) * odx;
        if (gx_in) gmx -= fac * (in.ts_xp - in.ts_xm);
        if (gy_in) gmy -= fac * (in.ts_yp - in.ts_ym);
    }

    // stability of the deposited layer (SWEHR.c:1292-1312): right/upper neighbour
    // slopes of the previous step, left/lower ones of this step, own at the edges;
    // cos(atan|x|) = 1/msqrt(1+x^2), sin(atan|x|) = |x|/msqrt(1+x^2)
    const R rho0 = R(kRhow * kPhi + kRhos * (1 - kPhi));
    const R dep = in.dep;
    const R cy = mrsqrt(R(1) + syn * syn);
    const R load = rho0 * gz * dep + rho * gz * h;
    const R lhs = load * (mfabs(syn) * cy);
    const R base = load * cy - rhow * gz * (dep + h);
    const R kfa = R(kFrictionAngle), kco = R(kCohesion);
    const R pgf0 = dep * rho0 * gz * R(0.5) * odx;
    R pgfx = pgf0 * (in.dep_xu * mrsqrt(R(1) + in.sx_up * in.sx_up) - in.dep_xd * mrsqrt(R(1) + in.sx_dn * in.sx_dn));
    if (pgfx * sxn > R(0)) pgfx = R(-1) * mfabs(pgfx);
    if (pgfx * sxn < R(0)) pgfx = mfabs(pgfx);
    bool fail = lhs > (kco + base * kfa + pgfx);
    if (!fail) {
        R sy_up, sy_dn;
        ysl(sy_up, sy_dn);
        R pgfy = pgf0 * (in.dep_yu * mrsqrt(R(1) + sy_up * sy_up) - in.dep_yd * mrsqrt(R(1) + sy_dn * sy_dn));
        if (pgfy * syn > R(0)) pgfy = R(-1) * mfabs(pgfy);
        if (pgfy * syn < R(0)) pgfy = mfabs(pgfy);
        fail = lhs > (kco + base * kfa + pgfy);
    }
    X.fail = fail;

    // debris-flow forcing (SWEHR.c:1314-1320)
    R dfm = mfmin(R(1), (ts - R(0.2)) * R(5));
    if (h < A.t9h0_lt || ts < thr_lt<R>(0.2)) dfm = R(0);
    const R bt = -(R(1) - R(kLambda)) * gz * h * kfa;
    const R btx = hu < R(0) ? R(-1) * bt : bt;
    const R bty = hv < R(0) ? R(-1) * bt : bt;

    // class-invariant parts of the Hairsine-Rose sources (SWEHR.c:1329-1378);
    // depth normal to the bed h / cos(atan|slope|) = h * msqrt(1 + slope^2)
    const R totm_old = in.totm_o;
    const R hvert = h * (n2 * rn);
    const bool deep_v = hvert > A.th0_gt;
    const R ihv = deep_v ? mrcp(hvert) : R(0);
    const R mtstar = deep_v ? R(A.mtstar0) * (h0 * ihv) : R(A.mtstar0);
    const R H = mfmin(R(1), totm_old * mrcp(mtstar));  // mtstar = 0: NaN, which fmin drops as inf/NaN
    const R shield = deep_v ? h0 * ihv : R(1);
    const R rough = A.n_raster ? R(at(gmem(A.n_raster), c8L)) : R(A.roughness);
    X.rough = rough;
    const bool active = h > thr_gt<R>(10 * kHfilm) && ts < thr_lt<R>(0.6);
    const R oneh = R(1) - mfmin(ts, R(0.99));
    X.hind = (oneh * oneh) * (oneh * oneh);
    const bool shutoff = h > A.t9h0_gt && ts > thr_gt<R>(kCthreshold);
    X.shallow_v = h < A.t9h0_lt;
    X.vshal = mfmax(h - R(3) * h0, R(0)) * A.inv_6h0;
    X.cg = R(1.09) * R(kRhos / kRhow - 1) * gz;
    // rainsplash and stream-power factors; class k multiplies by P_k or M_k
    R f1 = R(0), f2 = R(0), f3 = R(0), f4 = R(0), omega = R(0);
    bool e2ok = false, e4ok = false;
    if (active) {
        const R qq = hu * hu + hv * hv;
        const R qm = msqrt(qq);
        const R sf = manning2_over_hpow<3>(h, rough, s_exp2, s_log) * qq;
        omega = rho * gz * sf * qm;
        const R oneH = R(1) - H;
        f1 = oneH * R(A.a) * Rr * shield;
        e2ok = totm_old > thr_gt<R>(1e-8);
        e4ok = totm_old > thr_gt<R>(1e-6);
        const R itm = e2ok ? mrcp(totm_old) : R(0);
        f2 = H * itm * R(A.ad) * Rr * shield;
        f3 = oneH * R(A.F) / R(A.J) * omega;
        if (e4ok) f4 = (R(1) - R(kBetaBedload)) * H * itm * R(A.F) * mrcp((rhos - rho) * R(1.0 / kRhos) * gz * h) * omega;
    }
    // debris shutoff (SWEHR.c:1391-1399) zeroes every source and the settling: with
    // the factors zeroed each class's E1..E4 are exactly 0 (E2, E4 are taken only
    // where M_k is finite and positive), and vf is scaled by 0
    X.vf_keep = shutoff ? R(0) : R(1);
    if (shutoff) f1 = f2 = f3 = f4 = R(0);
    X.f1 = f1;
    X.f2 = f2;
    X.f3 = f3;
    X.f4 = f4;
    X.omega = omega;
    X.e2ok = e2ok;
    X.e4ok = e4ok;

    // conservative update (SWEHR.c:1424-1441)
    const R dx_ = dt * odx;
    X.dx_ = dx_;
    // FP32: the depth's increments of the step (fluxes, rain, infiltration, exchange)
    // are summed apart in dh_rest and added to the depth once, without bias,
    // after the class loop.  In deep water they are a few float spacings at h each
    // step (rain 1.3e-6 m against 9.5e-7 m at 8 m), and rounded to nearest with the
    // same fraction every step they would gain or lose water systematically.
    R hn, dh_rest = R(0);
    bool film = false;  // FP32: the infiltration term clamped the cell to the film depth
    if constexpr (sizeof(R) == sizeof(double)) {
        hn = h - dx_ * (in.fxr[0] - in.fxl[0]) - dx_ * (in.fyr[0] - in.fyl[0]) + dt * Rr;
    } else {
        dh_rest = dt * Rr - dx_ * (in.fxr[0] - in.fxl[0]) - dx_ * (in.fyr[0] - in.fyl[0]);
        hn = h + dh_rest;
    }
    X.hun = hu - dx_ * (in.fxr[1] - in.fxl[1]) - dx_ * (in.fyr[1] - in.fyl[1]) + dt * sxs + dt * gmx + dt * dfm * btx;
    X.hvn = hv - dx_ * (in.fxr[2] - in.fxl[2]) - dx_ * (in.fyr[2] - in.fyl[2]) + dt * sys + dt * gmy + dt * dfm * bty;

    // Green-Ampt infiltration in FP64; negative below the film depth (SWEHR.c:1448-1452)
    {
        const double ks = in.ks_o;
        const double vin = in.vinf_o;
        const double dth = A.dtheta_raster ? at(gmem(A.dtheta_raster), c8L) : A.thetas - A.theta0;
        const double hf_c = A.hf_raster ? at(gmem(A.hf_raster), c8L) : A.hf;
        const double zf = vin * mrcp(dth);
        const double hd = sizeof(R) == sizeof(double) ? (double)hn : (double)h + (double)dh_rest;
        // the film depth as stored in this precision, so a dry cell stays exactly dry
        // zf = 0 gives NaN where division gives inf (or NaN): fmin returns h - hfilm for both
        const double infl = mfmin(hd - (double)hfilm, dtd * ks * (zf + hf_c + hd) * mrcp(zf));
        const double vin_new = vin + infl;
        if (X.own && !same_bits(vin_new, vin)) st_stream(gmem(A.vinf), c * 8u, vin_new);
        hn = R(hd - infl);
        if constexpr (sizeof(R) == sizeof(float)) {
            if (infl == hd - (double)hfilm) {  // the film clamp: exactly the film depth
                film = true;
                dh_rest = R(0);
            } else {
                dh_rest = dh_rest - R(infl);
            }
        }
    }
    X.hn = hn;
    X.dh_rest = dh_rest;
    X.film = film;

    // sediment: flux divergence, sources and exchange per class (SWEHR.c:1454-1511)
    X.fx0l = in.fxl[0];
    X.fx0r = in.fxr[0];
    X.fy0l = in.fyl[0];
    X.fy0r = in.fyr[0];
    X.hdt = R(0.5) * dt;
    X.zn_c = in.zn;
    X.dz = X.bdet = X.bexch = R(0);
    X.totm = R(0);
    X.depn = R(0);
    X.chtemp = R(0);
    X.left_suspended = false;  // some class ends the step with ch_k != 0
    X.sb_out = 0;              // sediment leaving through all edges
    X.sq_out = 0;              // ... and through j = 1/2
    X.mv_o = in.mv_o;
    X.md_o = in.md_o;
    X.totm_o = in.totm_o;
    // random bits of the unbiased roundings (FP32): a hash of the cell and the step,
    // then a Weyl sequence over the values the cell stores (cell_next_bits), so
    // results are reproducible and do not depend on the tile shape or on the ensemble
    X.rbits = sizeof(R) == sizeof(float) ? mix32(c ^ mix32((unsigned)steps)) : 0u;
    X.edge = gi == 0 || gi == nx - 1 || gj == 0 || gj == ny - 1;
}

// The next 32 random bits of the cell's unbiased roundings: a Weyl sequence (the
// golden-ratio increment 0x9e3779b9, odd, so it visits every 32-bit value) from the
// hash of the cell and the step; each draw is uniform because its start is, which is
// all an unbiased rounding needs, at one integer add per draw.
template <typename R>
__device__ __forceinline__ unsigned cell_next_bits(Cell<R> &X)
{
    X.rbits += 0x9e3779b9u;
    return X.rbits;
}

// Class k from its suspended mass ch_k and deposited mass m_k and the class fluxes of
// the four faces (mass flux times the upwind cell's concentration, SWEHR.c:1152-1157,
// 1219-1224).  Stores the new ch_k (chout) and M_k (in place, where it changed).
template <typename R>
__device__ __forceinline__ void cell_class(Cell<R> &X, const StepArgs<R> &A, int k, R ch_k, R m_k, R fkxl, R fkxr,
                                           R fkyl, R fkyr, const double *s_exp2, glob<R> *chout, glob<R> *Mp,
                                           long long plane, unsigned cR)
{
    const R hfilm = R(kHfilm);
    const R rsp = R(kRhos) * (R(1) - R(kPhi));
    const R kbed = R(1) / rsp;
    const int nx = A.nx, ny = A.ny, gi = X.gi, gj = X.gj;
    const R dt = X.dt, dx_ = X.dx_;
    if (X.edge) {  // a branch: most waves hold no cell on the domain's edge
        if (gi == 0) X.sb_out -= (double)fkxl;
        if (gi == nx - 1) X.sb_out += (double)fkxr;
        if (gj == 0) {
            X.sb_out -= (double)fkyl;
            X.sq_out -= (double)fkyl;
        }
        if (gj == ny - 1) X.sb_out += (double)fkyr;
    }
    R u = ch_k - dx_ * (fkxr - fkxl) - dx_ * (fkyr - fkyl);
    R m = m_k;
    // sources of class k from the state before this update (SWEHR.c:1322-1399)
    // (outside active cells f1..f4 and omega are 0 and e2ok, e4ok false, so every
    // E is 0 as the reference's guard makes it)
    const R dk = A.ds[k], pk = A.pf[k];
    const R e1 = pk * X.f1;
    const R e2 = (X.e2ok && m > thr_gt<R>(1e-8)) ? m * X.f2 : R(0);
    const bool ent = X.omega > R(0) && X.h > A.tds_gt[k];
    const R e3 = ent ? pk * X.f3 : R(0);
    const R e4 = (ent && X.e4ok && m > R(0)) ? m * X.f4 : R(0);
    // settling velocity sqrt(ak^2 + cg d) - ak (SWEHR.c:435).  For fine classes
    // cg d is ~1e-6 of ak^2 and the difference keeps only the last digits of the
    // sqrt: FP64 evaluates it as written, FP32 as cg d / (sqrt(ak^2 + cg d) + ak),
    // equal in real arithmetic and free of the cancellation
    const R ak = A.ak[k];
    R vf;
    if constexpr (sizeof(R) == sizeof(double)) {
        vf = msqrt_pos(A.ak2[k] + X.cg * dk) - ak;
    } else {
        const R b = X.cg * dk;
        vf = b * mrcp(msqrt_pos(A.ak2[k] + b) + ak);
    }
    vf = vf * X.hind;
    if (((A.coarse1 >> k) & 1u) && X.shallow_v) vf = R(2e-3) + X.vshal * (vf - R(2e-3));
    if (((A.coarse2 >> k) & 1u) && X.shallow_v) vf = R(2e-3) + X.vshal * (vf - R(2e-3));
    vf = vf * X.vf_keep;
    // exchange with the deposited layer
    // the running depth: in FP32 the depth before the step plus the increments so far
    const R hcur = sizeof(R) == sizeof(double) ? X.hn : (X.film ? hfilm : X.h) + X.dh_rest;
    if (hcur < thr_lt<R>(10 * kHfilm)) {
        if constexpr (sizeof(R) == sizeof(double)) {
            m += dt * e1 + u;
            X.bud[B_DET] += (double)(dt * e1);
        } else {
            m = add_unbiased(m, dt * e1 + u, cell_next_bits(X));
            X.bdet += dt * e1;
        }
        u = R(0);
    } else {
        X.chtemp = u;
        const R red = mfmin(m, dt * (e2 + e4));
        if constexpr (sizeof(R) == sizeof(double)) {
            const R decay = exp_tab(-vf * X.hdt * mrcp(hcur), s_exp2);
            u = decay * u;
            u = u + dt * (e1 + e3) + red;
            u = decay * u;
            X.zn_c = X.zn_c - (double)(kbed * (u - X.chtemp));
            m = m + dt * (e1 + e3) - (u - X.chtemp);
        } else {
            // FP32: the decay factor d = exp(-vf dt / 2h) of a fine class differs
            // from 1 by less than the spacing of floats there (2 um at h = 1 cm:
            // 1 - 4.5e-8), so d * u as written rounds to u or overstates the
            // settling.  The update d (d u + s) is taken as one increment of u,
            // u em1 (2 + em1) + s (1 + em1) with em1 = d - 1 (mexpm1_neg), added
            // with one unbiased rounding; M_k takes the difference of the stored
            // values and is updated the same way
            const R em1 = mexpm1_neg(-vf * X.hdt * mrcp(hcur));
            const R src = dt * (e1 + e3) + red;
            const R inc = u * (em1 * (R(2) + em1)) + src * (R(1) + em1);
            u = add_unbiased(u, inc, cell_next_bits(X));
            const R dch = u - X.chtemp;
            X.dz += kbed * dch;
            m = add_unbiased(m, dt * (e1 + e3) - dch, cell_next_bits(X));
        }
        if constexpr (sizeof(R) == sizeof(double)) X.bud[B_DET] += (double)(dt * (e1 + e3));
        else X.bdet += dt * (e1 + e3);
        if (X.fail) {
            if constexpr (sizeof(R) == sizeof(double)) {
                X.zn_c = X.zn_c - (double)(R(0.99) * m * kbed);
                u = u + R(0.99) * m;
                m = R(0.01) * m;
            } else {
                // FP32: the float 0.99 and 0.01 do not sum to 1, so the two
                // products would not conserve the class's mass; the mass kept is
                // rounded once and the rest moves into suspension
                const R keep = R(0.01) * m, moved = m - keep;
                X.dz += moved * kbed;
                u = add_unbiased(u, moved, cell_next_bits(X));
                m = keep;
            }
            X.totm = R(0);
            X.depn = R(0);
        }
    }
    const R dh = (u - X.chtemp) * kbed;
    if constexpr (sizeof(R) == sizeof(double)) X.hn = X.hn + dh;
    else X.dh_rest += dh;
    if constexpr (sizeof(R) == sizeof(double)) X.bud[B_EXCH] += (double)dh;
    else X.bexch += dh;
    X.totm += m;
    X.depn += m * R(1.0 / kRhos);
    X.left_suspended = X.left_suspended || u != R(0);
    if (X.own) {
        st_stream(chout + k * plane, cR, u);
        if (!same_bits(m, m_k)) st_stream(Mp + k * plane, cR, m);
    }
}

// After the class loop: the FP32 depth, the dry clamp, friction, maxima, the stores of
// the cell's state and its budget terms.  hout.. are the output planes of this step.
template <typename R>
__device__ __forceinline__ void cell_end(Cell<R> &X, const StepArgs<R> &A, double dtd, const double *s_exp2,
                                         const double *s_log, glob<R> *hout, glob<R> *huout, glob<R> *hvout,
                                         glob<R> *chout, glob<R> *depout, glob<double> *zout, glob<R> *Mp,
                                         glob<R> *totmp, glob<R> *maxvelp, glob<R> *maxdepthp, long long plane,
                                         unsigned cR, unsigned c8)
{
    const R hfilm = R(kHfilm);
    const R rsp = R(kRhos) * (R(1) - R(kPhi));
    const R kbed = R(1) / rsp;
    const int nx = A.nx, ny = A.ny, gi = X.gi, gj = X.gj, S = A.S;
    const R dt = X.dt;
    R hn = X.hn, hun = X.hun, hvn = X.hvn;
    if constexpr (sizeof(R) == sizeof(float)) {
        hn = add_unbiased(X.film ? hfilm : X.h, X.dh_rest, cell_next_bits(X));
        X.zn_c = X.zn_c - (double)X.dz;
        X.bud[B_DET] += (double)X.bdet;
        X.bud[B_EXCH] += (double)X.bexch;
    }

    // dry-cell clamp (SWEHR.c:1532-1551)
    if (!(hn > hfilm) && X.own) {
        X.bud[B_CLAMP] += (double)(hfilm - hn);
        hn = hfilm;
        hun = R(0);
        hvn = R(0);
        // the suspended mass goes to the deposited layer and the bed; where every
        // class ended with ch_k = 0 (a cell dry through the whole step, the common
        // case) that changes nothing but the sign of a zero, and the classes are not
        // read back and rewritten, which would double the cell's traffic
        glob<R> *chq = chout;
        R u_a = X.left_suspended ? at(chq, cR) : R(0), mm_a = X.left_suspended ? at(Mp, cR) : R(0);
        for (int k = 0; k < (X.left_suspended ? S : 0); k++) {
            const R u = u_a, mm = mm_a;
            if (k + 1 < S) {
                u_a = at(chq + (k + 1) * plane, cR);
                mm_a = at(Mp + (k + 1) * plane, cR);
            }
            X.zn_c = X.zn_c + (double)(kbed * u);
            // FP32: u is often far below the float spacing at mm
            if constexpr (sizeof(R) == sizeof(double)) at(Mp + k * plane, cR) = mm + u;
            else at(Mp + k * plane, cR) = add_unbiased(mm, u, cell_next_bits(X));
            at(chq + k * plane, cR) = R(0);
        }
    }
    // implicit friction (SWEHR.c:1553-1565) and maxima (SWEHR.c:1587-1588)
    R v1 = R(0), v2 = R(0);
    if (hn > hfilm) {
        const R sf2 = manning2_over_hpow<2>(hn, X.rough, s_exp2, s_log) * X.gz * msqrt(hun * hun + hvn * hvn);
        const R den = mrcp(R(1) + dt * sf2);
        hun = hun * den;
        hvn = hvn * den;
        const R ihn = mrcp(hn);
        v1 = hun * ihn;
        v2 = hvn * ihn;
    }
    const R speed = msqrt(v1 * v1 + v2 * v2);
    if (X.own) {
        if (speed > X.mv_o) st_stream(maxvelp, cR, speed);
        if (hn > X.md_o) st_stream(maxdepthp, cR, hn);
        st_stream(hout, cR, hn);
        st_stream(huout, cR, hun);
        st_stream(hvout, cR, hvn);
        st_stream(zout, c8, X.zn_c);
        if (!same_bits(X.totm, X.totm_o)) st_stream(totmp, cR, X.totm);
        st_stream(depout, cR, X.depn);
    }

    // water and sediment leaving through the domain edges this step
    double wb = 0;
    if (gi == 0) wb -= (double)X.fx0l;
    if (gi == nx - 1) wb += (double)X.fx0r;
    if (gj == 0) wb -= (double)X.fy0l;
    if (gj == ny - 1) wb += (double)X.fy0r;
    X.bud[B_BWATER] = wb * dtd * A.dx;
    X.bud[B_BSED] = X.sb_out * dtd * A.dx;
    if (gj == 0 && X.own) {  // outlet discharge of this column (per unit width times dx)
        gmem(A.outq)[gi] = -(double)X.fy0l * A.dx;
        gmem(A.outs)[gi] = X.sq_out * A.dx;
    }
    const double a2 = A.dx * A.dx;
    X.bud[B_EXCH] *= a2;
    X.bud[B_CLAMP] *= a2;
    X.bud[B_DET] *= a2;
}

}  // namespace swehr
