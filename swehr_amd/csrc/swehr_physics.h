// swehr_physics.h - point physics of the SWEHR step for gfx950, templated on the
// storage/arithmetic type R (double for the plain solver, float for the precision form).
//
// Each function restates one piece of SWEHR.c (cited per function).  Expressions keep
// the reference's operand order so the FP64 build stays within a few ulp of it; the
// transcendental calls and divisions are strength-reduced where an identity is exact in
// real arithmetic: pow(x,.5) -> sqrt(x), pow(x,2) -> x*x, a/b -> a*(1/b) with the
// reciprocal shared, sqrt(g*d^2/(16g)) -> |d|/4, and repeated calls with equal
// arguments are made once.
#pragma once
#include <hip/hip_runtime.h>

namespace swehr {

// constants of SWEHR.c:489-494, 632-650
constexpr double kG = 9.81;
constexpr double kHfilm = 1e-5;
constexpr double kPhi = 0.4;
constexpr double kRhos = 2600.0;
constexpr double kRhow = 1000.0;
constexpr double kNu = 1.2e-6;
constexpr double kFrictionAngle = 0.8693;
constexpr double kCohesion = 400.0;
constexpr double kLambda = 0.65;
constexpr double kCthreshold = 0.4;
constexpr double kBetaBedload = 0.5;
constexpr double kHcFriction = 0.003;
constexpr double kDepthExp = -0.33;

// HLL flux across one face, SWEHR.c:327-409.  (qn, qt): normal and tangential
// discharge.  Writes mass, normal- and tangential-momentum flux; returns the largest
// |wave speed|.  The flux branches are sequential ifs: the last true one wins (so
// hl = hr = 0 yields zero flux, not the 0/0 of the middle branch).
template <typename R>
__device__ __forceinline__ R face_flux(R hl, R hr, R qnl, R qnr, R qtl, R qtr, R g, R &f0, R &f1, R &f2)
{
    const R hfilm = R(kHfilm);
    R ul, vl, ur, vr;
    if (hl > hfilm) { const R r = R(1) / hl; ul = qnl * r; vl = qtl * r; }
    else { ul = R(0); vl = R(0); }
    if (hr > hfilm) { const R r = R(1) / hr; ur = qnr * r; vr = qtr * r; }
    else { ur = R(0); vr = R(0); }
    const R cl = sqrt(g * hl), cr = sqrt(g * hr);
    const R d = ul + R(2) * cl - ur - R(2) * cr;
    // h* = d^2 / (16 g), so c* = sqrt(g h*) = |d| / 4
    const R ustar = R(.5) * (ul + ur) + cl - cr;
    const R cstar = fabs(d) * R(0.25);
    R sl = ul - cl, sm = ustar, sr = ur + cr;
    if (ustar - cstar < sl) sl = ustar - cstar;
    if (ustar + cstar > sr) sr = ustar + cstar;
    if (hl <= hfilm) { sl = ur - R(2) * cr; sm = sl; sr = ur + cr; }
    if (hr <= hfilm) { sl = ul - cl; sr = ul + R(2) * cl; sm = sr; }
    const R hg = R(.5) * g;
    R a0 = R(NAN), a1 = R(NAN);
    if (sl >= R(0)) {
        a0 = hl * ul;
        a1 = hl * ul * ul + hg * hl * hl;
    }
    if (sl <= R(0) && sr >= R(0)) {
        const R iw = R(1) / (sr - sl);
        a0 = (sr * (hl * ul) - sl * (hr * ur) + sl * sr * (hr - hl)) * iw;
        a1 = (sr * (hl * ul * ul + hg * hl * hl) - sl * (hr * ur * ur + hg * hr * hr) + sl * sr * (hr * ur - hl * ul)) *
             iw;
    }
    if (sr <= R(0)) {
        a0 = hr * ur;
        a1 = hr * ur * ur + hg * hr * hr;
    }
    f0 = a0;
    f1 = a1;
    f2 = (sm >= R(0)) ? a0 * vl : a0 * vr;
    R s = R(0);
    if (fabs(sl) > s) s = fabs(sl);
    if (fabs(sr) > s) s = fabs(sr);
    if (fabs(sm) > s) s = fabs(sm);
    return s;
}

// Wave-speed bound of the same face without the fluxes (SWEHR.c:333-376, 404-408).
template <typename R>
__device__ __forceinline__ R face_speed(R hl, R hr, R qnl, R qnr, R g)
{
    const R hfilm = R(kHfilm);
    const R ul = (hl > hfilm) ? qnl * (R(1) / hl) : R(0);
    const R ur = (hr > hfilm) ? qnr * (R(1) / hr) : R(0);
    const R cl = sqrt(g * hl), cr = sqrt(g * hr);
    const R d = ul + R(2) * cl - ur - R(2) * cr;
    const R ustar = R(.5) * (ul + ur) + cl - cr;
    const R cstar = fabs(d) * R(0.25);
    R sl = ul - cl, sm = ustar, sr = ur + cr;
    if (ustar - cstar < sl) sl = ustar - cstar;
    if (ustar + cstar > sr) sr = ustar + cstar;
    if (hl <= hfilm) { sl = ur - R(2) * cr; sm = sl; sr = ur + cr; }
    if (hr <= hfilm) { sl = ul - cl; sr = ul + R(2) * cl; sm = sr; }
    R s = R(0);
    if (fabs(sl) > s) s = fabs(sl);
    if (fabs(sr) > s) s = fabs(sr);
    if (fabs(sm) > s) s = fabs(sm);
    return s;
}

// Sediment-laden inflow state at j = ny at time t, SWEHR.c:1019-1033 (FP64 always).
struct InflowState {
    double h, hv, c;  // depth, discharge, suspended mass per area per class
};
__device__ __forceinline__ InflowState inflow_state(double t)
{
    InflowState s;
    s.hv = fmax(-0.035, t / 100 * -0.035);
    s.h = pow(fabs(s.hv) * 0.045 / sqrt(0.7), 0.6);
    s.c = fmin(500.0, t / 100 * 500) * s.h;
    return s;
}

// Rain rate for time t, SWEHR.c:813-814: entry ceil(t/rint + 0.01), capped at rnum.
__device__ __forceinline__ double rain_rate(const double *rain, int rnum, double rint, double t)
{
    int rind = (int)fmin(ceil(t / rint + 0.01), (double)rnum);
    if (rind < 1) rind = 1;
    return rain[rind - 1];
}

// Adaptive dt, SWEHR.c:1409-1412.
__device__ __forceinline__ double step_dt(double smax, double cstable, double dx, double maxstep, double t,
                                          double tend)
{
    double dt = cstable * dx / smax;
    if (dt > maxstep) dt = maxstep;
    if (t + dt > tend) dt = tend - t;
    return dt;
}

}  // namespace swehr
