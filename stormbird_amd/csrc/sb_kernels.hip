// gfx950 kernels of the dynamic lifting-line step.
//
// Reference (stormbird 0.9.0):
//   singularity elements  lifting_line/singularity_elements/{vortex_line,panel,symmetry_condition}.rs
//   wake update           lifting_line/wake/dynamic_wake/{initialization,update_data}.rs
//   induced velocities    lifting_line/wake/dynamic_wake/induced_velocity_calc.rs
//   frozen wake           lifting_line/wake/frozen_wake.rs
//   solver                lifting_line/solvers/{simple_iterative,velocity_corrections}.rs
//   sectional model       line_force_model/force_calculations.rs, section_models/foil.rs
//   forces                common_utils/{forces_and_moments,flow_derivatives}.rs
//
// Compiled with -ffp-contract=off: the EXACT kernels evaluate every expression
// in the reference's order with IEEE division and square root, so results are
// bit-identical to the f64 CPU restatement.  The FAST pair kernels use fma and
// a refined reciprocal square root instead.
#include <hip/hip_runtime.h>

#include "sb_device.h"
#include "sb_wind.h"
#include "sb_mlp.h"

namespace sb {

constexpr int WAVE = 64;

// ---------------------------------------------------------------------------
// Singularity elements (exact: reference operation order)
// ---------------------------------------------------------------------------

// vortex_line.rs:59-84 + 90-106 ; rc.powi(4) = (rc*rc)*(rc*rc)
__device__ inline double viscous_core_term(V3 l0, V3 l1, V3 x, double rc, double r1sq, double r2sq) {
    V3 rel_line = l1 - l0;
    V3 rel_pt = x - l0;
    double line_len_sq = len2(rel_line);
    double line_len = sqrt(line_len_sq);
    V3 dir = rel_line / line_len;
    double par = dot(rel_pt, dir);
    double d2;
    if (par < 0.0) d2 = r1sq;
    else if (par > line_len) d2 = r2sq;
    else d2 = len2(rel_pt) - par * par;
    double rc2 = rc * rc;
    double den = sqrt(rc2 * rc2 + d2 * d2);
    return den > 0.0 ? d2 / den : 0.0;
}

// vortex_line.rs:18-55
__device__ inline V3 segment_unit(V3 l0, V3 l1, V3 x, double rc) {
    V3 r1 = x - l0;
    V3 r2 = x - l1;
    double r1sq = len2(r1), r2sq = len2(r2);
    double r1l = sqrt(r1sq), r2l = sqrt(r2sq);
    double r1r2 = r1l * r2l;
    double den = r1r2 * (r1r2 + dot(r1, r2));
    if (fabs(den) > CLOSENESS_ERROR) {
        double core = rc == 0.0 ? 1.0 : viscous_core_term(l0, l1, x, rc, r1sq, r2sq);
        double k = (r1l + r2l) / den;
        return (cross(r1, r2) * core) * (k * FOUR_PI_INVERSE);
    }
    return V3{0.0, 0.0, 0.0};
}

// Corners of panel k (dynamic_wake/mod.rs:82-103): rows s and s+1, columns
// i + wing(i) and +1; `cbase` = s*P + i + wing(i) is stored in the record.
__device__ inline void panel_corners(const Inst& d, const PanelRec& r, int k, V3 c[4]) {
    if (d.corners) {
        const double* q = d.corners + 12L * k;
        c[0] = load3(q);
        c[1] = load3(q + 3);
        c[2] = load3(q + 6);
        c[3] = load3(q + 9);
        return;
    }
    long a = (long)r.cbase;
    long b = a + d.P;
    c[0] = load3(d.pts + 3 * a);
    c[1] = load3(d.pts + 3 * (a + 1));
    c[2] = load3(d.pts + 3 * (b + 1));
    c[3] = load3(d.pts + 3 * b);
}

// segment_unit for the ring edge from corner i to corner j with the quantities
// that do not depend on the edge's direction shared: r = x - corner, |r|^2 and
// |r| of both corners, and the edge's unit direction and length (computed per
// panel by panels_body with the operations viscous_core_term uses).  Every
// value equals the one segment_unit computes, so the result is the same bits.
__device__ inline V3 ring_edge(V3 r1, V3 r2, double r1sq, double r2sq, double r1l, double r2l, const double* seg,
                               double rc) {
    double r1r2 = r1l * r2l;
    double den = r1r2 * (r1r2 + dot(r1, r2));
    if (fabs(den) > CLOSENESS_ERROR) {
        double core = 1.0;
        if (rc != 0.0) {  // vortex_line.rs:59-84 + 90-106 with rel_pt = r1, |rel_pt|^2 = r1sq
            V3 dir{seg[0], seg[1], seg[2]};
            double line_len = seg[3];
            double par = dot(r1, dir);
            double d2;
            if (par < 0.0) d2 = r1sq;
            else if (par > line_len) d2 = r2sq;
            else d2 = r1sq - par * par;
            double rc2 = rc * rc;
            double cden = sqrt(rc2 * rc2 + d2 * d2);
            core = cden > 0.0 ? d2 / cden : 0.0;
        }
        double k = (r1l + r2l) / den;
        return (cross(r1, r2) * core) * (k * FOUR_PI_INVERSE);
    }
    return V3{0.0, 0.0, 0.0};
}

// 1/sqrt(a) from the hardware estimate and one Newton step (~1 ulp); valid for
// positive normal a, which is and-ed into `ok`.
__device__ inline double rsqrt_fast(double a, bool& ok) {
    ok = ok & __builtin_amdgcn_class(a, 0x100);  // positive normal
    double y = __builtin_amdgcn_rsq(a);
    return fma(0.5 * y, fma(-a * y, y, 1.0), y);
}

// IEEE f64 square root and division without their operand scaling.  gfx950
// expands sqrt(a) into a scaling of a < 2^-767, v_rsq_f64, two Goldschmidt /
// Newton corrections, the unscaling and a zero/inf select; and n / d into
// v_div_scale of both operands, a reciprocal refined by two Newton steps, a
// quotient corrected by one fma residual step and v_div_fixup.  For a in
// [2^-767, 2^1000], and for |d| in [2^-300, 2^300] with |n| in [2^-600, 2^400],
// the scalings, the select and the fix-up return their input unchanged, so
// these sequences (the rest of the expansions, step for step) return the same
// bits as sqrt(a) and n / d.  Callers check the ranges.
__device__ inline double sqrt_unscaled(double a) {
    double y = __builtin_amdgcn_rsq(a);
    double h = a * y;
    double g = y * 0.5;
    double e = fma(-g, h, 0.5);
    h = fma(h, e, h);
    double rr = fma(-h, h, a);
    g = fma(g, e, g);
    h = fma(rr, g, h);
    rr = fma(-h, h, a);
    return fma(rr, g, h);
}

__device__ inline double div_unscaled(double n, double d) {
    double v = __builtin_amdgcn_rcp(d);
    double t = fma(-d, v, 1.0);
    v = fma(v, t, v);
    t = fma(-d, v, 1.0);
    v = fma(v, t, v);
    double q = n * v;
    double e = fma(-d, q, n);
    return fma(e, v, q);
}

// ring_edge with sqrt_unscaled / div_unscaled.  Equal to ring_edge bit for bit
// when |r1|^2, |r2|^2 lie in [2^-700, 2^140], rc is 0 or |rc| in [2^-150, 2^69]
// (then |den| <= 2^281, the core denominator lies in [2^-300, 2^141] and the
// numerators below 2^141) and, with a core, d2 >= 2^-600: the last condition
// is and-ed into `ok`, the others are the caller's.
__device__ inline V3 ring_edge_unscaled(V3 r1, V3 r2, double r1sq, double r2sq, double r1l, double r2l,
                                        const double* seg, double rc, bool& ok) {
    double r1r2 = r1l * r2l;
    double den = r1r2 * (r1r2 + dot(r1, r2));
    if (fabs(den) > CLOSENESS_ERROR) {
        double core = 1.0;
        if (rc != 0.0) {
            V3 dir{seg[0], seg[1], seg[2]};
            double line_len = seg[3];
            double par = dot(r1, dir);
            double d2;
            if (par < 0.0) d2 = r1sq;
            else if (par > line_len) d2 = r2sq;
            else d2 = r1sq - par * par;
            double rc2 = rc * rc;
            ok = ok & (d2 >= 0x1p-600);
            core = div_unscaled(d2, sqrt_unscaled(rc2 * rc2 + d2 * d2));
        }
        double k = div_unscaled(r1l + r2l, den);
        return (cross(r1, r2) * core) * (k * FOUR_PI_INVERSE);
    }
    return V3{0.0, 0.0, 0.0};
}

// |r|^2 in the range ring_edge_unscaled needs
__device__ inline bool ring_sq_ok(double sq) { return (sq >= 0x1p-700) & (sq <= 0x1p140); }

// ring_edge in FAST arithmetic: fused dot products, square roots and divisions
// from the hardware reciprocal (square root) with one Newton step (~1 ulp).
// Straight-line code like ring_edge_unscaled; operands outside the estimates'
// ranges clear `ok` (for an edge that contributes).
__device__ inline V3 ring_edge_fast(V3 r1, V3 r2, double r1sq, double r2sq, double r1l, double r2l,
                                    const double* seg, double rc, bool& ok) {
    double r1r2 = r1l * r2l;
    double den = r1r2 * (r1r2 + fma(r1.x, r2.x, fma(r1.y, r2.y, r1.z * r2.z)));
    const bool take = fabs(den) > CLOSENESS_ERROR;
    double core = 1.0;
    bool eok = fabs(den) <= 0x1p1000;
    if (rc != 0.0) {  // uniform
        double par = fma(r1.x, seg[0], fma(r1.y, seg[1], r1.z * seg[2]));
        double d2 = par < 0.0 ? r1sq : (par > seg[3] ? r2sq : fma(-par, par, r1sq));
        double rc2 = rc * rc;
        double a = fma(rc2, rc2, d2 * d2);
        core = d2 * rsqrt_fast(a, eok);
    }
    double v = __builtin_amdgcn_rcp(den);
    v = fma(v, fma(-den, v, 1.0), v);
    ok = ok & (!take | eok);
    double kk = (r1l + r2l) * v * FOUR_PI_INVERSE * core;
    V3 w{fma(r1.y, r2.z, -r1.z * r2.y) * kk, fma(r1.z, r2.x, -r1.x * r2.z) * kk, fma(r1.x, r2.y, -r1.y * r2.x) * kk};
    return take ? w : V3{0.0, 0.0, 0.0};
}


// One vortex ring from its corners c0..c3 (ring order) and its ring-edge
// geometry g (16 doubles: per edge e, from corner e to e+1, the unit direction
// and the length), for target x.  The corner pointers may point to global
// memory or to the shared tile.  Each corner's r, |r|^2 and |r| serve the two
// edges that meet there.  FAST: fused operations and refined hardware
// estimates.  EXACT: sqrt_unscaled / div_unscaled, and the IEEE operators for
// lanes outside their ranges (a target at a corner or on an edge), so the
// result is ring_edge's, the reference's operations, bit for bit.
template <int MODE>
__device__ inline V3 rings_at(const double* c0, const double* c1, const double* c2, const double* c3,
                              const double* g, double rc, V3 x) {
    const double* cp[4] = {c0, c1, c2, c3};
    V3 u{0.0, 0.0, 0.0};
    if (MODE == MODE_FAST) {
        // lanes outside the estimates' ranges (a target at a corner) take the
        // EXACT ring
        bool ok = true;
        const V3 r0 = x - load3(cp[0]);
        const double sq0 = fma(r0.x, r0.x, fma(r0.y, r0.y, r0.z * r0.z)), ln0 = sq0 * rsqrt_fast(sq0, ok);
        V3 ra = r0;
        double sqa = sq0, lna = ln0;
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            const V3 rb = x - load3(cp[e + 1]);
            const double sqb = fma(rb.x, rb.x, fma(rb.y, rb.y, rb.z * rb.z)), lnb = sqb * rsqrt_fast(sqb, ok);
            u = u + ring_edge_fast(ra, rb, sqa, sqb, lna, lnb, g + 4 * e, rc, ok);
            ra = rb;
            sqa = sqb;
            lna = lnb;
        }
        u = u + ring_edge_fast(ra, r0, sqa, sq0, lna, ln0, g + 12, rc, ok);
        if (ok) return u;
        return rings_at<MODE_EXACT>(c0, c1, c2, c3, g, rc, x);
    }
    {
        bool ok = (rc == 0.0) | ((fabs(rc) >= 0x1p-150) & (fabs(rc) <= 0x1p69));
        const V3 r0 = x - load3(cp[0]);
        const double sq0 = len2(r0), ln0 = sqrt_unscaled(sq0);
        ok = ok & ring_sq_ok(sq0);
        V3 ra = r0;
        double sqa = sq0, lna = ln0;
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            const V3 rb = x - load3(cp[e + 1]);
            const double sqb = len2(rb), lnb = sqrt_unscaled(sqb);
            ok = ok & ring_sq_ok(sqb);
            u = u + ring_edge_unscaled(ra, rb, sqa, sqb, lna, lnb, g + 4 * e, rc, ok);
            ra = rb;
            sqa = sqb;
            lna = lnb;
        }
        u = u + ring_edge_unscaled(ra, r0, sqa, sq0, lna, ln0, g + 12, rc, ok);
        if (ok) return u;
    }
    u = V3{0.0, 0.0, 0.0};
    const V3 r0 = x - load3(cp[0]);
    const double sq0 = len2(r0), ln0 = sqrt(sq0);
    V3 ra = r0;
    double sqa = sq0, lna = ln0;
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        const V3 rb = x - load3(cp[e + 1]);
        const double sqb = len2(rb), lnb = sqrt(sqb);
        u = u + ring_edge(ra, rb, sqa, sqb, lna, lnb, g + 4 * e, rc);
        ra = rb;
        sqa = sqb;
        lna = lnb;
    }
    return u + ring_edge(ra, r0, sqa, sq0, lna, ln0, g + 12, rc);
}

// panel.rs:99-118 (vortex ring) — used for near pairs only.  The panel's
// near-field record (corners and edge geometry, NEARD doubles) comes from
// Inst::segs, or with NB from its copy `nb` in the shared tile.
template <int MODE = MODE_EXACT, bool NB = false>
__device__ inline V3 panel_rings(const Inst& d, const PanelRec& r, int k, V3 x, const double* nb = nullptr,
                                 const double* nbc = nullptr) {
    double rc = NB ? *nbc : d.core[k];
    if (NB || (d.segs && !d.corners)) {
        const double* q = NB ? nb : d.segs + (long)NEARD * k;
        return rings_at<MODE>(q, q + 3, q + 6, q + 9, q + 12, rc, x);
    }
    V3 c[4];
    panel_corners(d, r, k, c);
    V3 u{0.0, 0.0, 0.0};
    u = u + segment_unit(c[0], c[1], x, rc);
    u = u + segment_unit(c[1], c[2], x, rc);
    u = u + segment_unit(c[2], c[3], x, rc);
    u = u + segment_unit(c[3], c[0], x, rc);
    return u;
}

// The point-doublet quotient num / (d2 * d2 * sqrt(d2)) of panel.rs:86-95 with
// IEEE square root and division, in fewer instructions.  gfx950 expands an
// IEEE f64 sqrt into operand scaling, v_rsq_f64 and two Goldschmidt/Newton
// corrections, then a zero/inf select; and each IEEE f64 division into operand
// scaling (v_div_scale), a reciprocal refined by two Newton steps, a quotient
// corrected by one fma residual step (Markstein), and v_div_fixup.  For
// d2 in [2^-120, 2^120] (so the denominator lies in [2^-300, 2^300]) and
// numerators with 2^-600 <= |n| <= 2^400, every scaling step, the select and
// the fix-up return their input unchanged, so the sequences below, which keep
// the rest of those expansions step for step and share the refined reciprocal
// among the three numerators, return the same bits.  Other lanes (including a
// zero numerator) take the plain IEEE operators.
__device__ inline V3 doublet_quotients(double d2, V3 num) {
    double mx = fmax(fmax(fabs(num.x), fabs(num.y)), fabs(num.z));
    double mn = fmin(fmin(fabs(num.x), fabs(num.y)), fabs(num.z));
    bool fast = d2 >= 0x1p-120 && d2 <= 0x1p120 && mx <= 0x1p400 && mn >= 0x1p-600;
    if (fast) {
        double y = __builtin_amdgcn_rsq(d2);
        double h = d2 * y;
        double g = y * 0.5;
        double e = fma(-g, h, 0.5);
        h = fma(h, e, h);
        double rr = fma(-h, h, d2);
        g = fma(g, e, g);
        h = fma(rr, g, h);
        rr = fma(-h, h, d2);
        h = fma(rr, g, h);  // = sqrt(d2), correctly rounded
        double den = d2 * d2 * h;
        double v = __builtin_amdgcn_rcp(den);
        double t = fma(-den, v, 1.0);
        v = fma(v, t, v);
        t = fma(-den, v, 1.0);
        v = fma(v, t, v);
        double qx = num.x * v, qy = num.y * v, qz = num.z * v;
        double ex = fma(-den, qx, num.x), ey = fma(-den, qy, num.y), ez = fma(-den, qz, num.z);
        return V3{fma(ex, v, qx), fma(ey, v, qy), fma(ez, v, qz)};
    }
    double d5 = d2 * d2 * sqrt(d2);
    return V3{num.x / d5, num.y / d5, num.z / d5};
}

// ---- single precision (MODE_FP32_KAHAN) ----
// Relative vectors are formed in f64 (where the cancellation happens) and
// rounded to f32; the element formulas then run in f32.  The near/far switch is
// decided in f64, as the reference decides it, so every pair takes the same
// branch as in the f64 path.
struct F3 {
    float x, y, z;
};
__device__ inline F3 f3_of(V3 a) { return F3{(float)a.x, (float)a.y, (float)a.z}; }
__device__ inline float f3dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline F3 f3cross(F3 a, F3 b) {
    return F3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// vortex_line.rs:18-106 in f32
// One vortex segment in fp32 (vortex_line.rs:18-106), in a form that does not
// cancel near the segment's line: with sg = l1 - l0 and r2 = r1 - sg,
// r1 x r2 = sg x r1 (no difference of nearly equal products), and
// |r1||r2| + r1.r2 = |r1 x r2|^2 / (|r1||r2| - r1.r2) where r1.r2 < 0; the
// distance to the line is |sg x r1| / |sg|.  The relative error of those terms
// grows as |r1| / h for a target at distance h from the line, so a lane with a
// target over the segment closer than |r1| / 256 evaluates it in f64
// (segment_unit); `rcd` is the core length in f64.
__device__ inline F3 segment_unit_f32(V3 l0, V3 l1, V3 x, float rc, double rcd) {
    F3 r1 = f3_of(x - l0), r2 = f3_of(x - l1), sg = f3_of(l1 - l0);
    float r1sq = f3dot(r1, r1), r2sq = f3dot(r2, r2);
    float r1l = sqrtf(r1sq), r2l = sqrtf(r2sq);
    float r1r2 = r1l * r2l;
    F3 c = f3cross(sg, r1);
    float c2 = f3dot(c, c);
    float dt = f3dot(r1, r2);
    float sum = dt >= 0.0f ? r1r2 + dt : c2 / (r1r2 - dt);
    float den = r1r2 * sum;
    float sg2 = f3dot(sg, sg);
    float ll = sqrtf(sg2);
    float par = f3dot(r1, sg) / ll;
    bool over = par >= 0.0f && par <= ll;
    bool cancel = over && c2 < (r1sq * sg2) * 0x1p-16f;
    F3 out{0.0f, 0.0f, 0.0f};
    if (fabsf(den) > (float)CLOSENESS_ERROR) {
        float core = 1.0f;
        if (rc != 0.0f) {
            float d2 = par < 0.0f ? r1sq : (par > ll ? r2sq : c2 / sg2);
            float rc2 = rc * rc;
            float cden = sqrtf(rc2 * rc2 + d2 * d2);
            core = cden > 0.0f ? d2 / cden : 0.0f;
        }
        float kk = (r1l + r2l) / den * (core * (float)FOUR_PI_INVERSE);
        out = F3{c.x * kk, c.y * kk, c.z * kk};
    }
    if (cancel) {
        V3 u = segment_unit(l0, l1, x, rcd);
        out = F3{(float)u.x, (float)u.y, (float)u.z};
    }
    return out;
}

template <bool NB = false>
__device__ inline F3 panel_unit_f32(const Inst& d, const PanelRec& r, int k, V3 x, const double* nb = nullptr,
                                    const double* nbc = nullptr) {
    double tx = x.x - r.cx, ty = x.y - r.cy, tz = x.z - r.cz;
    double d2 = tx * tx + ty * ty + tz * tz;
    if (d2 <= r.far2) {
        V3 c[4];
        if (NB) {
            c[0] = load3(nb);
            c[1] = load3(nb + 3);
            c[2] = load3(nb + 6);
            c[3] = load3(nb + 9);
        } else {
            panel_corners(d, r, k, c);
        }
        const double rcd = NB ? *nbc : d.core[k];
        const float rc = (float)rcd;
        F3 a = segment_unit_f32(c[0], c[1], x, rc, rcd), b = segment_unit_f32(c[1], c[2], x, rc, rcd);
        F3 e = segment_unit_f32(c[2], c[3], x, rc, rcd), f = segment_unit_f32(c[3], c[0], x, rc, rcd);
        return F3{((a.x + b.x) + e.x) + f.x, ((a.y + b.y) + e.y) + f.y, ((a.z + b.z) + e.z) + f.z};
    }
    F3 t{(float)tx, (float)ty, (float)tz};
    float q = (float)d2;
    float y = __builtin_amdgcn_rsqf(q);
    float y2 = y * y;
    float sc = (float)r.area_term * ((y2 * y2) * y);
    float h3 = 3.0f * (t.x * (float)r.nx + t.y * (float)r.ny + t.z * (float)r.nz);
    return F3{(t.x * h3 - (float)r.nx * q) * sc, (t.y * h3 - (float)r.ny * q) * sc, (t.z * h3 - (float)r.nz * q) * sc};
}

template <int MODE, bool NB = false>
__device__ inline V3 panel_unit(const Inst& d, const PanelRec& r, int k, V3 x, const double* nb = nullptr,
                                const double* nbc = nullptr) {
    if (MODE == MODE_FP32_KAHAN) {
        F3 u = panel_unit_f32<NB>(d, r, k, x, nb, nbc);
        return V3{(double)u.x, (double)u.y, (double)u.z};
    }
    // panel.rs:53-76: near/far switch on |x - c|^2 <= far^2
    double tx = x.x - r.cx, ty = x.y - r.cy, tz = x.z - r.cz;
    if (MODE == MODE_EXACT) {
        double d2 = tx * tx + ty * ty + tz * tz;
        if (d2 <= r.far2) return panel_rings<MODE_EXACT, NB>(d, r, k, x, nb, nbc);
        // panel.rs:82-96 point doublet
        double nh = tx * r.nx + ty * r.ny + tz * r.nz;
        double h3 = 3.0 * nh;
        V3 num{(tx * h3 - r.nx * d2) * r.area_term, (ty * h3 - r.ny * d2) * r.area_term,
               (tz * h3 - r.nz * d2) * r.area_term};
        return doublet_quotients(d2, num);
    } else {
        double d2 = fma(tx, tx, fma(ty, ty, tz * tz));
        if (d2 <= r.far2) return panel_rings<MODE_FAST, NB>(d, r, k, x, nb, nbc);
        double y = __builtin_amdgcn_rsq(d2);
        // one Newton step on 1/sqrt(d2): y <- y (3 - d2 y^2) / 2
        double e = fma(-d2 * y, y, 1.0);
        y = fma(0.5 * y, e, y);
        double y2 = y * y;
        double inv5 = (y2 * y2) * y;
        double nh = fma(tx, r.nx, fma(ty, r.ny, tz * r.nz));
        double h3 = 3.0 * nh;
        double s = r.area_term * inv5;
        return V3{fma(tx, h3, -r.nx * d2) * s, fma(ty, h3, -r.ny * d2) * s, fma(tz, h3, -r.nz * d2) * s};
    }
}

// The point-doublet branch of panel_unit without control flow, so that several
// panels' evaluations interleave: returns true when the lane takes that branch
// with operands in the range of the fast quotient sequence, and then `out`
// equals panel_unit's result bit for bit (same operations, same order).  Other
// lanes (near field, or operands outside the range) must call panel_unit.
template <int MODE>
__device__ inline bool far_fast(const PanelRec& r, V3 x, V3& out) {
    double tx = x.x - r.cx, ty = x.y - r.cy, tz = x.z - r.cz;
    if (MODE == MODE_EXACT) {
        double d2 = tx * tx + ty * ty + tz * tz;
        double nh = tx * r.nx + ty * r.ny + tz * r.nz;
        double h3 = 3.0 * nh;
        V3 num{(tx * h3 - r.nx * d2) * r.area_term, (ty * h3 - r.ny * d2) * r.area_term,
               (tz * h3 - r.nz * d2) * r.area_term};
        double mx = fmax(fmax(fabs(num.x), fabs(num.y)), fabs(num.z));
        double mn = fmin(fmin(fabs(num.x), fabs(num.y)), fabs(num.z));
        bool ok = d2 > r.far2 && d2 >= 0x1p-120 && d2 <= 0x1p120 && mx <= 0x1p400 && mn >= 0x1p-600;
        double y = __builtin_amdgcn_rsq(d2);
        double h = d2 * y;
        double g = y * 0.5;
        double e = fma(-g, h, 0.5);
        h = fma(h, e, h);
        double rr = fma(-h, h, d2);
        g = fma(g, e, g);
        h = fma(rr, g, h);
        rr = fma(-h, h, d2);
        h = fma(rr, g, h);
        double den = d2 * d2 * h;
        double v = __builtin_amdgcn_rcp(den);
        double t = fma(-den, v, 1.0);
        v = fma(v, t, v);
        t = fma(-den, v, 1.0);
        v = fma(v, t, v);
        double qx = num.x * v, qy = num.y * v, qz = num.z * v;
        double ex = fma(-den, qx, num.x), ey = fma(-den, qy, num.y), ez = fma(-den, qz, num.z);
        out = V3{fma(ex, v, qx), fma(ey, v, qy), fma(ez, v, qz)};
        return ok;
    } else {
        double d2 = fma(tx, tx, fma(ty, ty, tz * tz));
        double y = __builtin_amdgcn_rsq(d2);
        double e = fma(-d2 * y, y, 1.0);
        y = fma(0.5 * y, e, y);
        double y2 = y * y;
        double inv5 = (y2 * y2) * y;
        double nh = fma(tx, r.nx, fma(ty, r.ny, tz * r.nz));
        double h3 = 3.0 * nh;
        double s = r.area_term * inv5;
        out = V3{fma(tx, h3, -r.nx * d2) * s, fma(ty, h3, -r.ny * d2) * s, fma(tz, h3, -r.nz * d2) * s};
        return d2 > r.far2;
    }
}

// induced_velocity_calc.rs:182-198 with symmetry_condition.rs:27-43
// The mirror flips one coordinate (multiplying by -1 is exact) and the
// correction adds s*u_m component-wise, which equals the reference's
// u -/+ u_m bit for bit (a - b is a + (-b) in IEEE arithmetic).  Both points
// go through one evaluation site: with two separately inlined near-field
// evaluations the gfx950 compiler returned the first result for both.
template <int MODE, bool NB = false>
__device__ inline V3 unit_velocity_sym(const Inst& d, const PanelRec& r, int k, V3 x, const double* nb = nullptr,
                                       const double* nbc = nullptr) {
    int nq = d.symmetry == 0 ? 1 : 2;
    V3 s{d.symmetry == 1 ? -1.0 : 1.0, d.symmetry == 2 ? -1.0 : 1.0, d.symmetry == 3 ? -1.0 : 1.0};
    V3 out{0.0, 0.0, 0.0};
#pragma nounroll
    for (int q = 0; q < nq; ++q) {
        V3 p = q == 0 ? x : V3{x.x * s.x, x.y * s.y, x.z * s.z};
        V3 v = panel_unit<MODE, NB>(d, r, k, p, nb, nbc);
        out = q == 0 ? v : V3{out.x + s.x * v.x, out.y + s.y * v.y, out.z + s.z * v.z};
    }
    return out;
}

// ---------------------------------------------------------------------------
// Biot-Savart sums
// ---------------------------------------------------------------------------
constexpr int TILE = 256;  // panel records per LDS tile (20 KiB)
constexpr int RECD = (int)(sizeof(PanelRec) / sizeof(double));

template <int MODE, int SYM, bool NB = false>
__device__ inline V3 pair_unit(const Inst& d, const PanelRec& r, int k, V3 x, const double* nb = nullptr,
                               const double* nbc = nullptr) {
    if (SYM) return unit_velocity_sym<MODE, NB>(d, r, k, x, nb, nbc);
    return panel_unit<MODE, NB>(d, r, k, x, nb, nbc);
}

// Near-field staging (NB): the staged records, the block's bounding box of
// evaluation points and the per-wave scratch of the tile test.  A tile's
// near-field records are staged only when some panel of the tile may be within
// its near-field radius of the box; otherwise no lane of the block can take the
// vortex-ring branch for that tile and the buffer is not read.
struct NearShared {
    double red[6][16];  // per-wave partial bounds
    double box[6];      // the block's box: min x, y, z, then -max x, y, z
    int flag[16];       // per-wave: a panel of the tile may be near
};
struct NearCtx {
    double* nbuf = nullptr;       // [TL][NEARD] shared
    double* ncore = nullptr;      // [TL] shared: the panels' viscous core lengths
    NearShared* sh = nullptr;     // shared
};

// Bounding box of the block's evaluation points: the targets and, with a
// symmetry plane, their mirror images (unit_velocity_sym).  Every thread of
// the block calls it.
__device__ inline void block_box(const Inst& d, V3 x, NearCtx& nc) {
    V3 m = x;
    if (d.symmetry != 0)
        m = V3{d.symmetry == 1 ? -x.x : x.x, d.symmetry == 2 ? -x.y : x.y, d.symmetry == 3 ? -x.z : x.z};
    double v[6] = {fmin(x.x, m.x), fmin(x.y, m.y), fmin(x.z, m.z), -fmax(x.x, m.x), -fmax(x.y, m.y), -fmax(x.z, m.z)};
#pragma unroll
    for (int i = 0; i < 6; ++i)
        for (int off = WAVE / 2; off > 0; off >>= 1) v[i] = fmin(v[i], __shfl_xor(v[i], off, WAVE));
    const int nw = (int)blockDim.x / WAVE;
    if ((threadIdx.x & (WAVE - 1)) == 0)
        for (int i = 0; i < 6; ++i) nc.sh->red[i][threadIdx.x / WAVE] = v[i];
    __syncthreads();
    if (threadIdx.x < 6) {
        double a = nc.sh->red[threadIdx.x][0];
        for (int q = 1; q < nw; ++q) a = fmin(a, nc.sh->red[threadIdx.x][q]);
        nc.sh->box[threadIdx.x] = a;
    }
    __syncthreads();
}

// Publishes the record tile (barrier) and stages the tile's near-field records
// when a panel of it may be near the block's box.  The test keeps a relative
// margin of 2^-20 on the radius squared, far above the rounding of the lanes'
// own distance (a lane inside the box is no closer to the panel centre than the
// box is), and treats NaN as near.
__device__ inline void stage_near(const Inst& d, NearCtx& nc, long base, int cnt) {
    int want = 0;
    const double* bx = nc.sh->box;
    for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
        const PanelRec& r = d.rec[base + q];
        double dx = fmax(fmax(bx[0] - r.cx, r.cx + bx[3]), 0.0);
        double dy = fmax(fmax(bx[1] - r.cy, r.cy + bx[4]), 0.0);
        double dz = fmax(fmax(bx[2] - r.cz, r.cz + bx[5]), 0.0);
        double db2 = dx * dx + dy * dy + dz * dz;
        want |= (r.strength != 0.0) && !(db2 > r.far2 * (1.0 + 0x1p-20));
    }
    const unsigned long long bal = __ballot(want);
    if ((threadIdx.x & (WAVE - 1)) == 0) nc.sh->flag[threadIdx.x / WAVE] = bal != 0ull;
    __syncthreads();
    int need = 0;
    for (int q = 0; q < (int)blockDim.x / WAVE; ++q) need |= nc.sh->flag[q];
    if (need) {
        const double* nsrc = d.segs + (long)NEARD * base;
        for (int q = threadIdx.x; q < cnt * NEARD; q += blockDim.x) nc.nbuf[q] = nsrc[q];
        for (int q = threadIdx.x; q < cnt; q += blockDim.x) nc.ncore[q] = d.core[base + q];
        __syncthreads();
    }
}

// Folds panels [p0, p1) into `acc` for this lane's target x, in ascending
// panel order: acc = acc + strength * u  (induced_velocity_calc.rs:111-126;
// a zero-strength panel adds nothing, :202-210).  Every thread of the block
// calls it with the same range: records are staged through the shared tile,
// and the next record is read while the current one is evaluated.
// Compensated (Kahan) f32 summation of f32 terms over panels [p0, p1): the
// sum s and its running error c per component; s - c is added to acc.
__device__ inline void kahan_add(float& s, float& c, float v) {
    float y = v - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
}

template <int SYM, int TL = TILE, bool NB = false>
__device__ inline V3 sweep_kahan(const Inst& d, PanelRec* tile, V3 x, V3 acc, long p0, long p1, NearCtx& nc) {
    float sx = 0.0f, sy = 0.0f, sz = 0.0f, cx = 0.0f, cy = 0.0f, cz = 0.0f;
    for (long base = p0; base < p1; base += TL) {
        int cnt = (int)((p1 - base) < TL ? (p1 - base) : TL);
        __syncthreads();
        const double* src = reinterpret_cast<const double*>(d.rec + base);
        double* dst = reinterpret_cast<double*>(tile);
        for (int q = threadIdx.x; q < cnt * RECD; q += blockDim.x) dst[q] = src[q];
        if (NB) stage_near(d, nc, base, cnt);
        else __syncthreads();
        for (int p = 0; p < cnt; ++p) {
            const PanelRec& r = tile[p];
            if (r.strength == 0.0) continue;
            const double* nb = NB ? nc.nbuf + NEARD * p : nullptr;
            const double* nbc = NB ? nc.ncore + p : nullptr;
            F3 u;
            if constexpr (SYM == 0) {
                u = panel_unit_f32<NB>(d, r, (int)(base + p), x, nb, nbc);  // the fp32 terms, no f64 round trip
            } else {
                V3 w = pair_unit<MODE_FP32_KAHAN, SYM, NB>(d, r, (int)(base + p), x, nb, nbc);
                u = F3{(float)w.x, (float)w.y, (float)w.z};
            }
            float g = (float)r.strength;
            kahan_add(sx, cx, (float)u.x * g);
            kahan_add(sy, cy, (float)u.y * g);
            kahan_add(sz, cz, (float)u.z * g);
        }
    }
    return acc + V3{(double)sx - (double)cx, (double)sy - (double)cy, (double)sz - (double)cz};
}

// With NB (free wake with ring-edge geometry, no explicit corner array) the
// tile also stages each panel's near-field data (NEARD doubles in `nbuf`), so
// that the lanes taking the vortex-ring branch read it from the shared tile.
template <int MODE, int SYM, int TL = TILE, bool NB = false>
__device__ inline V3 sweep(const Inst& d, PanelRec* tile, V3 x, V3 acc, long p0, long p1, NearCtx& nc) {
    if constexpr (MODE == MODE_FP32_KAHAN) return sweep_kahan<SYM, TL, NB>(d, tile, x, acc, p0, p1, nc);
    for (long base = p0; base < p1; base += TL) {
        int cnt = (int)((p1 - base) < TL ? (p1 - base) : TL);
        __syncthreads();
        const double* src = reinterpret_cast<const double*>(d.rec + base);
        double* dst = reinterpret_cast<double*>(tile);
        for (int q = threadIdx.x; q < cnt * RECD; q += blockDim.x) dst[q] = src[q];
        if (NB) stage_near(d, nc, base, cnt);
        else __syncthreads();
        // Two records in flight, in alternating registers: b is read while a
        // is evaluated and the next a while b is evaluated.
        PanelRec a = tile[0];
        for (int p = 0; p < cnt; p += 2) {
            PanelRec b = tile[p + 1 < cnt ? p + 1 : p];
            if (a.strength != 0.0)
                acc = acc + pair_unit<MODE, SYM, NB>(d, a, (int)(base + p), x, NB ? nc.nbuf + NEARD * p : nullptr,
                                                     NB ? nc.ncore + p : nullptr) *
                                a.strength;
            if (p + 1 >= cnt) break;
            a = tile[p + 2 < cnt ? p + 2 : p + 1];
            if (b.strength != 0.0)
                acc = acc + pair_unit<MODE, SYM, NB>(d, b, (int)(base + p + 1), x,
                                                     NB ? nc.nbuf + NEARD * (p + 1) : nullptr,
                                                     NB ? nc.ncore + p + 1 : nullptr) *
                                b.strength;
        }
    }
    return acc;
}

// The target a lane of a pair-sum launch evaluates: with a target map (McBuffers::tmap)
// the map's entry for the slot, else the slot itself; a negative value -(i + 1) marks a
// padding slot that evaluates target i and stores nothing.
__device__ inline long target_of_slot(const int* tmap, long slot, long n_targets) {
    if (tmap) return (long)tmap[slot];
    return slot < n_targets ? slot : -n_targets;
}

typedef const __attribute__((address_space(4))) PanelRec* ScalarRecs;

// Record k read through the constant address space: a wave-uniform k gives
// scalar loads into SGPRs.
__device__ inline PanelRec scalar_rec(ScalarRecs recs, long k) {
    PanelRec o;
    o.cx = recs[k].cx;
    o.cy = recs[k].cy;
    o.cz = recs[k].cz;
    o.nx = recs[k].nx;
    o.ny = recs[k].ny;
    o.nz = recs[k].nz;
    o.far2 = recs[k].far2;
    o.area_term = recs[k].area_term;
    o.strength = recs[k].strength;
    o.cbase = recs[k].cbase;
    return o;
}

// sweep with the panel records read by scalar loads (the panel index is
// wave-uniform) instead of through a shared tile: the records take no LDS
// bandwidth, a tile barrier is needed only for the near-field staging, and the
// zero-strength skip is a scalar branch.  The same pairs in the same order, so
// the same bits as sweep.
template <int MODE, int SYM, int TL, bool NB>
__device__ inline V3 sweep_scalar(const Inst& d, V3 x, V3 acc, long p0l, long p1l, NearCtx& nc) {
    const ScalarRecs recs = (ScalarRecs)d.rec;
    const int p0 = __builtin_amdgcn_readfirstlane((int)p0l), p1 = __builtin_amdgcn_readfirstlane((int)p1l);
    for (int base = p0; base < p1; base += TL) {
        const int cnt = p1 - base < TL ? p1 - base : TL;
        if (NB) {
            __syncthreads();  // every wave is done with the previous tile's near-field records
            stage_near(d, nc, base, cnt);
        }
        // Scalar loads return out of order, so waiting for one waits for all: each record
        // is taken (its strength tested) before the next one's load is issued, and the
        // scheduler barrier keeps the load behind the test; the load then runs under the
        // current record's evaluation.
        PanelRec a = scalar_rec(recs, base);
        for (int p = 0; p < cnt; p += 2) {
            const bool live_a = a.strength != 0.0;
            __builtin_amdgcn_sched_barrier(0);
            PanelRec b = scalar_rec(recs, base + (p + 1 < cnt ? p + 1 : p));
            if (live_a)
                acc = acc + pair_unit<MODE, SYM, NB>(d, a, base + p, x, NB ? nc.nbuf + NEARD * p : nullptr,
                                                     NB ? nc.ncore + p : nullptr) *
                                a.strength;
            if (p + 1 >= cnt) break;
            const bool live_b = b.strength != 0.0;
            __builtin_amdgcn_sched_barrier(0);
            a = scalar_rec(recs, base + (p + 2 < cnt ? p + 2 : p + 1));
            if (live_b)
                acc = acc + pair_unit<MODE, SYM, NB>(d, b, base + p + 1, x, NB ? nc.nbuf + NEARD * (p + 1) : nullptr,
                                                     NB ? nc.ncore + p + 1 : nullptr) *
                                b.strength;
        }
    }
    return acc;
}

// The pair sum of the single-instance kernels: scalar-loaded records for the f64
// modes, the shared record tile for FP32-Kahan.
template <int MODE, int SYM, int TL, bool NB>
__device__ inline V3 sweep_targets(const Inst& d, PanelRec* tile, V3 x, V3 acc, long p0, long p1, NearCtx& nc) {
    if constexpr (MODE == MODE_FP32_KAHAN) return sweep_kahan<SYM, TL, NB>(d, tile, x, acc, p0, p1, nc);
    else return sweep_scalar<MODE, SYM, TL, NB>(d, x, acc, p0, p1, nc);
}

// NB: near-field data staged with the records (sweep), in tiles of NB_TILE
// panels so that four workgroups still fit a CU's shared memory.
constexpr int NB_TILE = 128;

// One target per lane, the whole panel range per block.  out[t] += acc when
// `add` is set (update_velocity_at_points, update_data.rs:125-141), else
// out[t] = acc.
template <int MODE, int SYM, bool NB = false>
__global__ __launch_bounds__(256) void k_targets(Inst d, const double* __restrict__ targets, long n_targets,
                                                 long n_panels, double* __restrict__ out, int add,
                                                 const int* __restrict__ tmap) {
    // staged tiles of 64 panels (19.5 KB): eight workgroups per CU, as without staging
    constexpr int TL = NB ? 64 : TILE;
    __shared__ PanelRec tile[MODE == MODE_FP32_KAHAN ? TL : 1];
    __shared__ double nbuf[NB ? TL * NEARD : 1];
    __shared__ double ncore[NB ? TL : 1];
    __shared__ NearShared nsh;
    const long slot = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long t = target_of_slot(tmap, slot, n_targets);
    V3 x = load3(targets + 3 * (t >= 0 ? t : -t - 1));
    NearCtx nc;
    nc.nbuf = nbuf;
    nc.ncore = ncore;
    nc.sh = &nsh;
    if (NB) block_box(d, x, nc);
    V3 acc = sweep_targets<MODE, SYM, TL, NB>(d, tile, x, V3{0.0, 0.0, 0.0}, 0, n_panels, nc);
    if (t >= 0) {
        if (add) store3(out + 3 * t, load3(out + 3 * t) + acc);
        else store3(out + 3 * t, acc);
    }
}

// Same sum with the work spread evenly over exactly the resident workgroups
// (McNaughton's wrap-around rule).  Job j = targets [256 j, 256 j + 256) over
// all L panels; the concatenated jobs are cut into equal ranges of C > L
// panel units, one per workgroup.  A job cut between workgroups b and b+1 runs
// its first panels at the start of b+1's range (saving the partial sums) and
// its last panels at the end of b's range, continuing from those partial sums:
// the per-target fold keeps its ascending order, so the result is bit-identical
// to k_targets.  b reaches the job no earlier than b+1 finishes its part; if
// the partial sums are not published within `spin_limit` polls, b recomputes
// the first panels itself (same bits), so every workgroup terminates.
struct McPlan {
    long C, n_jobs, L;
    double* partial;     // [n_jobs][256][3]
    long long* flags;    // [n_jobs], == epoch once the partial sums are published
    long long epoch;
    int spin_limit;
    const int* tmap;     // target map (McBuffers::tmap) or null for flat order
};

template <int MODE, int SYM, bool NB = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_targets_mc(Inst d, const double* __restrict__ targets, long n_targets,
                                                    double* __restrict__ out, int add, McPlan pl) {
    constexpr int TL = NB ? NB_TILE : TILE;
    __shared__ PanelRec tile[MODE == MODE_FP32_KAHAN ? TL : 1];
    __shared__ double nbuf[NB ? NB_TILE * NEARD : 1];
    __shared__ double ncore[NB ? NB_TILE : 1];
    __shared__ NearShared nsh;
    __shared__ int ready;
    NearCtx nc;
    nc.nbuf = nbuf;
    nc.ncore = ncore;
    nc.sh = &nsh;
    const long L = pl.L;
    const long g0 = (long)blockIdx.x * pl.C;
    const long gend = pl.n_jobs * L;
    const long g1 = g0 + pl.C < gend ? g0 + pl.C : gend;
    for (long j = g0 / L; j * L < g1; ++j) {
        const long js = j * L, je = js + L;
        const long a = js > g0 ? js : g0;
        const long b = je < g1 ? je : g1;
        const bool head = a > js;  // first panels of job j (its last panels end the previous range)
        const bool tail = b < je;  // last panels of job j (its first panels start the next range)
        const long p0 = tail ? L - (b - a) : 0;
        const long p1 = head ? (b - a) : L;
        const long slot = j * 256 + threadIdx.x;
        const long t = target_of_slot(pl.tmap, slot, n_targets);
        V3 x = load3(targets + 3 * (t >= 0 ? t : -t - 1));
        if (NB) block_box(d, x, nc);
        V3 acc{0.0, 0.0, 0.0};
        double* part = pl.partial + 3 * slot;
        long start = p0;
        if (tail) {
            if (threadIdx.x == 0) {
                int ok = 0;
                for (int it = 0; it < pl.spin_limit; ++it) {
                    if (__hip_atomic_load(pl.flags + j, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == pl.epoch) {
                        ok = 1;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(8);
                }
                ready = ok;
            }
            __syncthreads();
            if (ready) {
                __threadfence();
                acc = load3(part);
            } else {
                start = 0;  // recompute the first panels here: same order, same bits
            }
        }
        acc = sweep_targets<MODE, SYM, TL, NB>(d, tile, x, acc, start, p1, nc);
        if (head) {
            store3(part, acc);
            __threadfence();
            __syncthreads();
            if (threadIdx.x == 0) __hip_atomic_store(pl.flags + j, pl.epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else if (t >= 0) {
            if (add) store3(out + 3 * t, load3(out + 3 * t) + acc);
            else store3(out + 3 * t, acc);
        }
    }
}

// Pair census for the roofline: per block, the number of (target, panel)
// pairs with nonzero strength that take the near-field (vortex ring) branch
// and the total; one partial per block, summed on the host.
__global__ __launch_bounds__(256) void k_count_pairs(Inst d, const double* __restrict__ targets, long n_targets,
                                                     long n_panels, double* __restrict__ partial) {
    __shared__ PanelRec tile[TILE];
    __shared__ double red[2][256];
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long tc = t < n_targets ? t : n_targets - 1;
    V3 x = load3(targets + 3 * tc);
    double near = 0.0, total = 0.0;
    for (long base = 0; base < n_panels; base += TILE) {
        int cnt = (int)((n_panels - base) < TILE ? (n_panels - base) : TILE);
        __syncthreads();
        const double* src = reinterpret_cast<const double*>(d.rec + base);
        double* dst = reinterpret_cast<double*>(tile);
        for (int q = threadIdx.x; q < cnt * RECD; q += blockDim.x) dst[q] = src[q];
        __syncthreads();
        for (int p = 0; p < cnt; ++p) {
            PanelRec r = tile[p];
            if (r.strength == 0.0) continue;
            double tx = x.x - r.cx, ty = x.y - r.cy, tz = x.z - r.cz;
            double d2 = tx * tx + ty * ty + tz * tz;
            total += 1.0;
            if (d2 <= r.far2) near += 1.0;
        }
    }
    if (t >= n_targets) near = total = 0.0;
    red[0][threadIdx.x] = near;
    red[1][threadIdx.x] = total;
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int q = 0; q < (int)blockDim.x; ++q) {
            a += red[0][q];
            b += red[1][q];
        }
        partial[2 * blockIdx.x] = a;
        partial[2 * blockIdx.x + 1] = b;
    }
}

// One wave per target, lanes over panels: each lane evaluates one panel's
// term, then the wave folds the 64 terms in ascending panel order.  Used when
// there are few targets (the N control points): it spreads one target's work
// over 64 lanes while keeping the sequential per-target summation order.
// Panels of the target's own wing are skipped in neglect-self mode
// (induced_velocity_calc.rs:128-150).
template <int MODE>
__global__ __launch_bounds__(256) void k_fold(Inst d, const double* __restrict__ targets, int n_targets,
                                              int p_begin, int p_end, int neglect, double* __restrict__ out) {
    __shared__ double terms[4][WAVE * 3];
    __shared__ int valid[4][WAVE];
    int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    int t = blockIdx.x * 4 + wave;
    int tc = t < n_targets ? t : n_targets - 1;
    V3 x = load3(targets + 3 * tc);
    int wt = neglect ? d.wing_of[tc] : -1;
    V3 acc{0.0, 0.0, 0.0};
    for (int base = p_begin; base < p_end; base += WAVE) {
        int k = base + lane;
        V3 term{0.0, 0.0, 0.0};
        int ok = 0;
        if (k < p_end) {
            ok = 1;
            if (neglect && d.wing_of[k % d.N] == wt) {
                ok = 0;
            } else {
                PanelRec r = d.rec[k];
                if (r.strength != 0.0) term = unit_velocity_sym<MODE>(d, r, k, x) * r.strength;
            }
        }
        terms[wave][3 * lane + 0] = term.x;
        terms[wave][3 * lane + 1] = term.y;
        terms[wave][3 * lane + 2] = term.z;
        valid[wave][lane] = ok;
        __syncthreads();
        int cnt = p_end - base < WAVE ? p_end - base : WAVE;
        for (int q = 0; q < cnt; ++q) {
            if (valid[wave][q]) acc = acc + V3{terms[wave][3 * q], terms[wave][3 * q + 1], terms[wave][3 * q + 2]};
        }
        __syncthreads();
    }
    if (t < n_targets && lane == 0) store3(out + 3 * t, acc);
}

// Few targets, many panels, in two passes over a chunk of panels
// [p_begin, p_begin + cnt):
//  k_terms      every (target, panel) term strength * u, one wave per 64
//               targets x tp panels (tp <= 16: fewer when the chunk would give
//               too few waves to fill the GPU), written to terms[c][q][t]
//               (ld >= n_targets);
//               a skipped panel (zero strength, or the target's own wing in
//               neglect-self mode) writes +0;
//  k_fold_terms one lane per (component, target) chain adds the chunk's terms
//               in ascending panel order to the running sum acc[c][t] (fed
//               through LDS by a whole workgroup of loaders).
// Adding +0 leaves a sum unchanged unless the sum is -0, and a sum that starts
// at +0 never becomes -0, so the result equals the reference's fold that
// skips those panels (induced_velocity_calc.rs:111-150, 202-210).
constexpr int TERM_PANELS = 16;

template <int MODE>
__global__ __launch_bounds__(256) void k_terms(Inst d, const double* __restrict__ targets, int n_targets,
                                               long p_begin, long cnt, int neglect, double* __restrict__ terms,
                                               int ld, int tp) {
    int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    int t = blockIdx.y * WAVE + lane;
    int tc = t < n_targets ? t : n_targets - 1;
    V3 x = load3(targets + 3 * tc);
    int wt = neglect ? d.wing_of[tc] : -1;
    long q0 = ((long)blockIdx.x * 4 + wave) * tp;
    for (int j = 0; j < tp; ++j) {
        long q = q0 + j;
        if (q >= cnt) break;
        int k = __builtin_amdgcn_readfirstlane((int)(p_begin + q));
        PanelRec r = d.rec[k];
        V3 term{0.0, 0.0, 0.0};
        bool skip = r.strength == 0.0 || (neglect && d.wing_of[k % d.N] == wt);
        if (!skip) term = unit_velocity_sym<MODE>(d, r, k, x) * r.strength;
        if (t < n_targets) {
            terms[(0 * cnt + q) * ld + t] = term.x;
            terms[(1 * cnt + q) * ld + t] = term.y;
            terms[(2 * cnt + q) * ld + t] = term.z;
        }
    }
}

// One workgroup (16 waves) per 16 chains of one component: every thread keeps
// FOLD_K loads of the next 512-row block in flight while lanes 0-15 of wave 0
// add the current block (double-buffered in LDS) in ascending row order.
constexpr int FOLD_CH = 16;
constexpr int FOLD_K = 8;
constexpr int FOLD_THREADS = 1024;
constexpr int FOLD_ROWS = FOLD_THREADS * FOLD_K / FOLD_CH;  // 512

__global__ __launch_bounds__(1024) void k_fold_terms(const double* __restrict__ terms, int ld, int n_targets,
                                                     long cnt, int first, double* __restrict__ acc,
                                                     double* __restrict__ out) {
    __shared__ double buf[2][FOLD_ROWS][FOLD_CH];
    const int groups = (n_targets + FOLD_CH - 1) / FOLD_CH;
    const int c = blockIdx.x / groups;
    const int t0 = (blockIdx.x % groups) * FOLD_CH;
    const double* src = terms + (long)c * cnt * ld + t0;  // t0 + FOLD_CH <= ld: loads stay in the chunk
    const int lc = threadIdx.x % FOLD_CH, r0 = threadIdx.x / FOLD_CH;
    constexpr int RSTEP = FOLD_THREADS / FOLD_CH;  // 64
    const long nb = (cnt + FOLD_ROWS - 1) / FOLD_ROWS;
    double r[FOLD_K];
    auto load = [&](long b) {
#pragma unroll
        for (int k = 0; k < FOLD_K; ++k) {
            long q = b * FOLD_ROWS + r0 + RSTEP * k;
            r[k] = q < cnt ? src[q * ld + lc] : 0.0;
        }
    };
    auto store = [&](int slot) {
#pragma unroll
        for (int k = 0; k < FOLD_K; ++k) buf[slot][r0 + RSTEP * k][lc] = r[k];
    };
    const int t = t0 + (int)threadIdx.x;
    const bool folder = threadIdx.x < FOLD_CH && t < n_targets;
    double s = (folder && !first) ? acc[(long)c * n_targets + t] : 0.0;
    if (nb > 0) {
        load(0);
        store(0);
    }
    __syncthreads();
    for (long b = 0; b < nb; ++b) {
        if (b + 1 < nb) load(b + 1);
        if (folder) {
            int rows = cnt - b * FOLD_ROWS < FOLD_ROWS ? (int)(cnt - b * FOLD_ROWS) : FOLD_ROWS;
            const double(*cur)[FOLD_CH] = buf[b & 1];
            for (int q = 0; q < rows; ++q) s = s + cur[q][threadIdx.x];
        }
        if (b + 1 < nb) store((int)((b + 1) & 1));
        __syncthreads();
    }
    if (folder) {
        acc[(long)c * n_targets + t] = s;
        if (out) out[3 * t + c] = s;
    }
}

// frozen_wake.rs:198-261: F[j][i] = unit velocity of first-row panel j at ctrl point i.
__device__ inline void factors_body(const Inst& d, long idx) {
    long NN = (long)d.N * d.N;
    if (idx >= NN) return;
    int j = (int)(idx / d.N);  // panel (column)
    int i = (int)(idx - (long)j * d.N);  // ctrl point (row)
    V3 u{0.0, 0.0, 0.0};
    if (!(d.neglect_self && d.wing_of[i] == d.wing_of[j])) {
        V3 x = load3(d.geo + d.gl.ctrl + 3 * i);
        PanelRec r = d.rec[j];
        u = unit_velocity_sym<MODE_EXACT>(d, r, j, x);
    }
    store3(d.F + 3 * idx, u);
}
__global__ void k_factors(Inst d) { factors_body(d, (long)blockIdx.x * blockDim.x + threadIdx.x); }

// ---------------------------------------------------------------------------
// Wake shedding and convection
// ---------------------------------------------------------------------------

// Felt ctrl-point velocity (force_calculations.rs:50-69), its span-point map
// (value_mapping.rs:87-137) and, at the first step, the mean ctrl freestream
// and flow-derivative history (simulation.rs:65-94).
__device__ inline void prepare_body(const Inst& d, int first_step) {
    int t = threadIdx.x;
    const double* rb = d.geo + d.gl.rbm;
    V3 trans = load3(rb), vlin = load3(rb + 6), vang = load3(rb + 9);
    for (int i = t; i < d.N; i += blockDim.x) {
        V3 c = load3(d.geo + d.gl.ctrl + 3 * i);
        V3 motion = vlin + cross(vang, c - trans);
        store3(d.felt + 3 * i, load3(d.fs + 3 * i) - motion);
    }
    __syncthreads();
    for (int j = t; j < d.P; j += blockDim.x) {
        int a = d.span_src[2 * j], b = d.span_src[2 * j + 1];
        V3 v = a == b ? load3(d.felt + 3 * a) : (load3(d.felt + 3 * a) + load3(d.felt + 3 * b)) * 0.5;
        store3(d.felt_span + 3 * j, v);
    }
    if (first_step) {
        if (t == 0) {
            V3 s{0.0, 0.0, 0.0};
            for (int i = 0; i < d.N; ++i) s = s + load3(d.fs + 3 * i);
            store3(d.scal, s / (double)d.N);
            d.scal[3] = 0.0;
        }
        for (int i = t; i < d.N; i += blockDim.x) {
            V3 v = load3(d.fs + 3 * i);
            store3(d.hist + 3 * i, v);
            store3(d.hist + 3 * (d.N + i), v);
            d.prev_gamma[i] = 0.0;
        }
    }
}
__global__ void k_prepare(Inst d, int first_step) { prepare_body(d, first_step); }

// initialization.rs:12-41: M+1 sweeps of update_wake_points_before_solving
// with a uniform velocity v and freestream v.  Each column evolves on its
// own.  Without shape damping the sweeps converge to
//   p1 = p0 + dir*first_len, p_s = p_{s-1} + dt*v (2 <= s < M),
//   p_M = p_{M-1} + last_len * normalize(p_{M-1} - p_{M-2})
// bit for bit (row s is final after sweep s); with damping the sweeps are
// replayed literally.
__device__ inline void init_wake_body(const Inst& d, double dt, int use_mean, V3 vin, int fill, long g, long stride) {
    long P = d.P, M = d.M;
    long npts = P * (M + 1);
    V3 v = use_mean ? load3(d.scal) : vin;
    for (long q = g; q < npts; q += stride) store3(d.vel + 3 * q, v);
    for (long q = g; q < (long)d.N * M; q += stride) d.str[q] = 0.0;
    if (g >= P) return;
    int j = (int)g;
    double first_len = d.rep_chord * d.first_rel;
    double last_len = d.last_rel * d.rep_chord;
    V3 sp = load3(d.geo + d.gl.span_pts + 3 * j);
    V3 dir = d.first_dir == 0 ? normalize(load3(d.geo + d.gl.chord_span + 3 * j)) : normalize(v);
    double* col = d.pts;  // column j: row s at col + 3*(s*P + j)
    auto at = [&](long s) { return col + 3 * (s * P + j); };
    if (fill)  // DynamicWakeBuilder::build: every row starts at the span points (builder.rs:127-135)
        for (long s = 0; s <= M; ++s) store3(at(s), sp);
    if (d.damping <= 0.0) {
        store3(at(0), sp);
        V3 prev = sp + dir * first_len;
        store3(at(1), prev);
        for (long s = 2; s < M; ++s) {
            prev = prev + v * dt;
            store3(at(s), prev);
        }
        if (M > 1) {
            V3 a = load3(at(M - 1)), b = load3(at(M - 2));
            store3(at(M), a + (last_len * normalize(a - b)));
        } else {
            store3(at(1), sp + (last_len * normalize(v)));
        }
    } else {
        double dm = d.damping;
        for (long sweep = 0; sweep <= M; ++sweep) {
            store3(at(0), sp);
            for (long s = M; s >= 2; --s) {
                V3 integ = load3(at(s - 1)) + v * dt;
                store3(at(s), load3(at(s)) * dm + integ * (1.0 - dm));
            }
            V3 est = sp + dir * first_len;
            store3(at(1), load3(at(1)) * dm + est * (1.0 - dm));
            if (M > 1) {
                V3 a = load3(at(M - 1)), b = load3(at(M - 2));
                store3(at(M), a + (last_len * normalize(a - b)));
            } else {
                store3(at(1), sp + (last_len * normalize(v)));
            }
        }
    }
}
__global__ void k_init_wake(Inst d, double dt, int use_mean, double vx, double vy, double vz, int fill) {
    init_wake_body(d, dt, use_mean, V3{vx, vy, vz}, fill, (long)blockIdx.x * blockDim.x + threadIdx.x,
                   (long)gridDim.x * blockDim.x);
}

// update_data.rs:136-150 + 200-252 + 295-316: rows 0..M of the next points.
__device__ inline void wake_step_body(const Inst& d, double dt, long q) {
    long P = d.P, M = d.M;
    if (q >= P * (M + 1)) return;
    long s = q / P;
    int j = (int)(q - s * P);
    V3 sp = load3(d.geo + d.gl.span_pts + 3 * j);
    V3 out;
    if (s == 0) {
        out = sp;
    } else if (s == 1) {
        V3 dir;
        if (d.first_dir == 0) dir = normalize(load3(d.geo + d.gl.chord_span + 3 * j));
        else if (d.first_dir == 2 && d.steps_completed > 2)
            dir = normalize((load3(d.vel + 3 * j) + load3(d.vel + 3 * (P + j))) * 0.5);
        else dir = normalize(load3(d.felt_span + 3 * j));
        V3 est = sp + dir * (d.rep_chord * d.first_rel);
        out = d.damping > 0.0 ? load3(d.pts + 3 * q) * d.damping + est * (1.0 - d.damping) : est;
    } else {
        V3 integ = load3(d.pts + 3 * (q - P)) + load3(d.vel + 3 * (q - P)) * dt;
        out = d.damping > 0.0 ? load3(d.pts + 3 * q) * d.damping + integ * (1.0 - d.damping) : integ;
    }
    store3(d.pts_next + 3 * q, out);
}
__global__ void k_wake_step(Inst d, double dt) { wake_step_body(d, dt, (long)blockIdx.x * blockDim.x + threadIdx.x); }

// update_data.rs:262-292: last row from the two rows before it.
__device__ inline void wake_last_body(const Inst& d, long jj) {
    if (jj >= d.P) return;
    int j = (int)jj;
    long P = d.P, M = d.M;
    double last_len = d.last_rel * d.rep_chord;
    if (M > 1) {
        V3 a = load3(d.pts_next + 3 * ((M - 1) * P + j));
        V3 b = load3(d.pts_next + 3 * ((M - 2) * P + j));
        store3(d.pts_next + 3 * (M * P + j), a + (last_len * normalize(a - b)));
    } else {
        V3 a = load3(d.pts_next + 3 * j);
        store3(d.pts_next + 3 * (P + j), a + (last_len * normalize(load3(d.felt_span + 3 * j))));
    }
}
__global__ void k_wake_last(Inst d) { wake_last_body(d, (long)blockIdx.x * blockDim.x + threadIdx.x); }

// update_data.rs:102-115 + panel.rs:26-50 + geometry_functions.rs:11-30, and
// the strength shift of update_data.rs:319-328 (into str_next).
__device__ inline void panels_body(const Inst& d, int shift, long k) {
    long np = (long)d.N * d.M;
    if (k >= np) return;
    long s_row = k / d.N;
    int i = (int)(k - s_row * d.N);
    long a0 = s_row * d.P + d.col[i];
    V3 c[4];
    c[0] = load3(d.pts + 3 * a0);
    c[1] = load3(d.pts + 3 * (a0 + 1));
    c[2] = load3(d.pts + 3 * (a0 + d.P + 1));
    c[3] = load3(d.pts + 3 * (a0 + d.P));
    PanelRec r;
    r.cbase = (double)a0;
    V3 center = 0.25 * (c[0] + c[1] + c[2] + c[3]);
    double a = len(c[1] - c[0]);
    double b = len(c[2] - c[1]);
    double cc = len(c[3] - c[2]);
    double dd = len(c[0] - c[3]);
    double s = (a + b + cc + dd) * 0.5;
    double area = sqrt((s - a) * (s - b) * (s - cc) * (s - dd));
    V3 n = normalize(cross(c[2] - c[0], c[3] - c[1]));
    double rep = fmax_(a, b);
    double f = rep * d.far_ratio;
    r.cx = center.x;
    r.cy = center.y;
    r.cz = center.z;
    r.nx = n.x;
    r.ny = n.y;
    r.nz = n.z;
    r.far2 = f * f;
    r.area_term = area * FOUR_PI_INVERSE;
    double g;
    if (shift) {
        g = k >= d.N ? d.str[k - d.N] : d.str[k];
        d.str_next[k] = g;
    } else {
        g = d.str[k];
    }
    r.strength = g;
    d.rec[k] = r;
    if (d.segs) {  // the corners, then viscous_core_term's rel_line length and unit direction per edge
        double* q = d.segs + (long)NEARD * k;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            q[3 * e] = c[e].x;
            q[3 * e + 1] = c[e].y;
            q[3 * e + 2] = c[e].z;
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            V3 rel_line = c[(e + 1) & 3] - c[e];
            double line_len = sqrt(len2(rel_line));
            V3 dir = rel_line / line_len;
            q[12 + 4 * e] = dir.x;
            q[12 + 4 * e + 1] = dir.y;
            q[12 + 4 * e + 2] = dir.z;
            q[12 + 4 * e + 3] = line_len;
        }
    }
}
__global__ void k_panels(Inst d, int shift) { panels_body(d, shift, (long)blockIdx.x * blockDim.x + threadIdx.x); }

// ---------------------------------------------------------------------------
// Fixed-point solve (simple_iterative.rs:86-212), one workgroup
// ---------------------------------------------------------------------------
// The foil stays in memory (a pointer, not a copy): its 18 fields held in
// registers for every strip push the one-workgroup kernels into spills.
struct Strip {
    V3 felt, fixed, rel_g, dir_g, chord_g, rel_out;
    double c, local_len;
    const FoilParams* foil;
};

__device__ inline double strip_cl_pre(const Strip& s, double a) { return foil_cl_pre_dropoff(*s.foil, a); }

// force_calculations.rs:552-627 (one strip's |CL(Gamma) - CL(alpha)|)
__device__ inline double residual_term(const Strip& s, double rho, double g, double a, V3 u) {
    V3 F = len(u) == 0.0 ? V3{0.0, 0.0, 0.0} : (cross(u, s.rel_out) * g) * rho;
    double sign = -1.0 * signum(g);
    double area = s.c * s.local_len;
    double ff = 0.5 * area * rho * len2(u);
    double clg = ff == 0.0 ? 0.0 : len(F) * sign / ff;
    return fabs(clg - strip_cl_pre(s, a));
}

// force_calculations.rs:271-292
__device__ inline double circulation(const Strip& s, int viscous, double a, V3 u) {
    double cl = strip_cl_pre(s, a);
    if (viscous) cl += foil_cl_post_weighted(*s.foil, a);
    return -0.5 * s.c * len(u) * cl;
}

// Strip ic; its foil from `foils` (a shared-memory copy) when given.
__device__ inline Strip load_strip(const Inst& d, int ic, const FoilParams* foils = nullptr) {
    const double* geo = d.geo;
    Strip s;
    s.felt = load3(d.felt + 3 * ic);
    s.fixed = load3(d.fixed + 3 * ic);
    s.rel_g = load3(geo + d.gl.rel_g + 3 * ic);
    s.dir_g = load3(geo + d.gl.dir_g + 3 * ic);
    s.chord_g = load3(geo + d.gl.chord_g + 3 * ic);
    s.rel_out = load3(geo + d.gl.rel_out + 3 * ic);
    s.c = geo[d.gl.chord_len + ic];
    s.local_len = geo[d.gl.local_len + ic];
    s.foil = (foils ? foils : d.foil) + d.wing_of[ic];
    return s;
}

// Copies the instance's foils into shared memory (the caller synchronizes).
// Loop-invariant loads from global memory would be hoisted out of the solver's
// iteration and held in registers; shared-memory copies are read where used.
constexpr int FOIL_DOUBLES = (int)(sizeof(FoilParams) / sizeof(double));
__device__ inline void stage_foils(const Inst& d, FoilParams* dst) {
    const double* src = (const double*)d.foil;
    double* o = (double*)dst;
    for (int q = threadIdx.x; q < d.W * FOIL_DOUBLES; q += blockDim.x) o[q] = src[q];
}

// frozen_wake.rs:268-301: fixed velocity plus the factor row times G (three
// running sums in column order, added to the fixed velocity at the end).  The
// factors of MATVEC_U columns are loaded before any of them is used, so that
// many loads are in flight (the sums keep their order).
constexpr int MATVEC_U = 8;
__device__ inline V3 induced_at(const Inst& d, V3 fixed, const double* G, int i) {
    const int N = d.N;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    const double* Fi = d.F + 3 * i;
    int j = 0;
    for (; j + MATVEC_U <= N; j += MATVEC_U) {
        double fx[MATVEC_U], fy[MATVEC_U], fz[MATVEC_U], g[MATVEC_U];
#pragma unroll
        for (int u = 0; u < MATVEC_U; ++u) {
            const double* f = Fi + 3L * N * (j + u);
            fx[u] = f[0];
            fy[u] = f[1];
            fz[u] = f[2];
        }
#pragma unroll
        for (int u = 0; u < MATVEC_U; ++u) g[u] = G[j + u];
#pragma unroll
        for (int u = 0; u < MATVEC_U; ++u) {
            sx += fx[u] * g[u];
            sy += fy[u] * g[u];
            sz += fz[u] * g[u];
        }
    }
    for (; j < N; ++j) {
        const double* f = Fi + 3L * N * j;
        double gj = G[j];
        sx += f[0] * gj;
        sy += f[1] * gj;
        sz += f[2] * gj;
    }
    V3 ind = fixed;
    ind.x += sx;
    ind.y += sy;
    ind.z += sz;
    return ind;
}

// velocity_corrections.rs:18-100: felt velocity plus the (corrected) induced velocity.
__device__ inline V3 corrected_velocity(int vc_mode, double vc_ratio, V3 felt, V3 ind) {
    if (vc_mode == 1) {
        double um = len(ind), fm = len(felt);
        double mx = vc_ratio * fm;
        V3 cvel = um > mx ? normalize(ind) * mx : ind;
        return felt + cvel;
    }
    if (vc_mode == 2) {
        V3 tot = felt + ind;
        return normalize(tot) * len(felt);
    }
    return felt + ind;
}

__device__ inline double angle_of_attack(const Strip& s, V3 u) {  // force_calculations.rs:79-107
    return len(u) > MIN_POSITIVE ? signed_angle_between(s.chord_g, u, s.dir_g) : 0.0;
}

// One evaluation of the fixed-point map for strip i from the circulations in G
// (simple_iterative.rs:110-160): induced velocity, velocity correction,
// spanwise part removed, angle of attack, circulation estimate.
__device__ inline double estimate_circulation(const Inst& d, const Strip& s, const double* G, int i, V3& u,
                                              double& a) {
    V3 ind = induced_at(d, s.fixed, G, i);
    u = corrected_velocity(d.vc_mode, d.vc_ratio, s.felt, ind);
    u = u - project(u, s.rel_g);  // line_force_model/mod.rs:235-254
    a = angle_of_attack(s, u);
    return circulation(s, d.viscous_lift_in_circulation, a, u);
}

__device__ inline void solve_body(const Inst& d, double* sh) {
    int N = d.N;
    double* G = sh;             // [N] current circulation (broadcast for the matvec)
    double* R = sh + N;         // [N] residual terms
    double* wm = R + N;         // [16] per-wave maxima
    double* ctl = wm + 16;      // [0] residual, [1] converged flag
    FoilParams* foils = (FoilParams*)(ctl + 4);  // [W]
    int i = threadIdx.x;
    bool act = i < N;
    int ic = act ? i : N - 1;
    stage_foils(d, foils);
    __syncthreads();
    Strip s = load_strip(d, ic, foils);
    double rho = d.density;

    double g = d.prev_gamma[ic];
    if (fabs(d.correction_factor - 1.0) > MIN_POSITIVE) g *= d.correction_factor;
    if (d.start_linearized) g = d.lin_gamma[ic];  // simple_iterative.rs:103-113
    V3 u{0.0, 0.0, 0.0};
    double a = 0.0;
    if (act) R[i] = residual_term(s, rho, g, a, u);
    __syncthreads();
    if (i == 0) {
        double sum = 0.0;
        for (int q = 0; q < N; ++q) sum += R[q];
        ctl[0] = sum / (double)N;
        ctl[1] = 0.0;
    }
    __syncthreads();
    int it = 0;
    bool conv = false;
    double est = 0.0;
    while (it < d.max_iterations && !conv) {
        it += 1;
        if (act) G[i] = g;
        __syncthreads();
        double ad = 0.0;
        if (act) {
            est = estimate_circulation(d, s, G, i, u, a);
            R[i] = residual_term(s, rho, g, a, u);
            double diff = est - g;
            ad = fabs(diff);
            g += d.solver_damping * diff;
        }
        // max |est - g| over strips (order-free, exact)
        double m = ad;
        for (int off = WAVE / 2; off > 0; off >>= 1) m = fmax_(m, __shfl_xor(m, off, WAVE));
        if ((i % WAVE) == 0) wm[i / WAVE] = m;
        __syncthreads();
        if (i == 0) {
            double sum = 0.0;
            for (int q = 0; q < N; ++q) sum += R[q];
            double residual = sum / (double)N;
            double mx = 0.0;
            for (int w = 0; w < (int)((blockDim.x + WAVE - 1) / WAVE); ++w) mx = fmax_(mx, wm[w]);
            ctl[0] = residual;
            ctl[1] = (residual < d.residual_tol || mx < d.strength_tol) ? 1.0 : 0.0;
        }
        __syncthreads();
        conv = ctl[1] != 0.0;
    }
    if (act) {
        double gamma = circulation(s, d.viscous_lift_in_circulation, a, u);
        double* o = d.out;
        o[OutLayout::gamma(N) + i] = gamma;
        o[OutLayout::alpha(N) + i] = angle_of_attack(s, u);
        store3(o + OutLayout::u(N) + 3 * i, u);
        store3(o + OutLayout::felt(N) + 3 * i, s.felt);
        // update_wing_strength (update_data.rs:146-150): first-row strengths
        if (d.M > 0) {
            d.str[i] = gamma;
            d.rec[i].strength = gamma;
        }
        d.prev_gamma[i] = gamma;
    }
    if (i == 0) {
        d.out[OutLayout::iters(N, d.W)] = (double)it;
        d.out[OutLayout::residual(N, d.W)] = ctl[0];
    }
}
// MAXT = 512 covers N <= 512 with twice the registers per thread (no spills).
template <int MAXT>
__global__ __launch_bounds__(MAXT) void k_solve(Inst d) {
    extern __shared__ double sh[];
    solve_body(d, sh);
}

// ---------------------------------------------------------------------------
// Quasi-steady frozen wake and the Linearized solver
// ---------------------------------------------------------------------------

// horseshoe_vortex.rs:23-51: bound segment, then the trailing legs
// (start + w0 -> start) and (end -> end + w1), summed from zero.
__device__ inline V3 horseshoe_unit(V3 b0, V3 b1, V3 w0, V3 w1, V3 x, double rc) {
    V3 u{0.0, 0.0, 0.0};
    u = u + segment_unit(b0, b1, x, rc);
    u = u + segment_unit(b0 + w0, b0, x, rc);
    u = u + segment_unit(b1, b1 + w1, x, rc);
    return u;
}

// frozen_wake.rs:68-129 with horseshoe_vortex.rs:55-103: F[j][i] = unit velocity
// of strip j's horseshoe at ctrl point i, with the symmetry image.  The legs of
// strip j run along the normalized felt velocity of span points j + wing and
// j + wing + 1, hs_wake_length long.
__device__ inline void horseshoe_factors_body(const Inst& d, long idx) {
    long NN = (long)d.N * d.N;
    if (idx >= NN) return;
    int j = (int)(idx / d.N);            // horseshoe (column)
    int i = (int)(idx - (long)j * d.N);  // ctrl point (row)
    int c = d.col[j];
    V3 b0 = load3(d.geo + d.gl.span_pts + 3 * c);
    V3 b1 = load3(d.geo + d.gl.end_g + 3 * j);
    V3 w0 = normalize(load3(d.felt_span + 3 * c)) * d.hs_wake_length;
    V3 w1 = normalize(load3(d.felt_span + 3 * (c + 1))) * d.hs_wake_length;
    V3 x = load3(d.geo + d.gl.ctrl + 3 * i);
    int nq = d.symmetry == 0 ? 1 : 2;
    V3 s{d.symmetry == 1 ? -1.0 : 1.0, d.symmetry == 2 ? -1.0 : 1.0, d.symmetry == 3 ? -1.0 : 1.0};
    V3 out{0.0, 0.0, 0.0};
#pragma nounroll
    for (int q = 0; q < nq; ++q) {  // symmetry_condition.rs:27-43, as unit_velocity_sym
        V3 p = q == 0 ? x : V3{x.x * s.x, x.y * s.y, x.z * s.z};
        V3 v = horseshoe_unit(b0, b1, w0, w1, p, d.hs_core);
        out = q == 0 ? v : V3{out.x + s.x * v.x, out.y + s.y * v.y, out.z + s.z * v.z};
    }
    store3(d.F + 3 * idx, out);
}
__global__ void k_horseshoe_factors(Inst d) {
    horseshoe_factors_body(d, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// Linearized::solve (linearized.rs:54-222).  The augmented matrix [A | b] is
// kept column-major (column j at aug + j*N), so a pivot column is contiguous
// and trailing updates are coalesced.  Gaussian elimination with partial
// pivoting follows linalg.rs:41-104 element by element: every entry receives
// the same sequence of a - f*p updates (f and p rounded as there), the pivot is
// the first row of largest magnitude (a NaN diagonal is kept), and back
// substitution sums in ascending column order.  Entries below the diagonal of
// an eliminated column are never read again; they are not updated (or hold the
// factors, in the blocked form).  A pivot below 1e-12 in magnitude (where the
// reference's unwrap panics) records its row + 1 in scal[4] and the circulation
// is then zero.  as_start = 1: the default settings, and the circulation goes to
// lin_gamma only (the start of a SimpleIterative solve).

// linearized.rs:75-131: what row i of the system needs from strip i.
struct LinRow {
    V3 ndir;   // normal direction: velocity direction x span direction
    double ut, dcl, cll, c;
};
__device__ inline LinRow lin_row(const Strip& s) {
    V3 u = s.felt + s.fixed;
    u = u - project(u, s.rel_g);  // remove_span_velocity
    V3 vdir = normalize(u);
    LinRow q;
    q.ndir = cross(vdir, normalize(s.rel_g));
    double a = angle_of_attack(s, u);
    q.cll = foil_cl_linear(*s.foil, a);
    q.dcl = s.foil->cl_initial_slope;
    q.ut = dot(u, vdir);
    q.c = s.c;
    return q;
}
__device__ inline double lin_entry(const Inst& d, const LinRow& q, int r, int c) {
    V3 f = load3(d.F + 3L * ((long)c * d.N + r));
    double v = -q.dcl * dot(f, q.ndir) / q.ut;
    if (c == r) v += 2.0 / (q.c * q.ut);
    return v;
}


// linalg.rs:92-101 by one wave over the factored matrix in global memory (A
// column-major [N + 1][N]: a[i][j] = A[j N + i], column N the right-hand side),
// N <= 1024.  Every lane runs the same ordered sum
// sum = ((0 + a[i][i+1] x[i+1]) + a[i][i+2] x[i+2]) + ..., reading the products
// a[i][j] x[j], j >= i + 2, from shared memory in groups of 16 (pb, two rows of
// pstride >= N + 64; x in shared memory too): lane t forms the products of row
// i - 1 while row i is summed, from matrix entries loaded one row earlier, every
// slot from a valid index and +0.0 past the row's end (the sum starts at +0.0 and
// cannot become -0.0).
__device__ inline void back_substitute_wave_cm(const double* A, int N, double* x, double* pb, int pstride) {
    constexpr int PS = 16;  // product slots per lane: rows of up to 1024 products
    const int lane = threadIdx.x & (WAVE - 1);
    const int nsl = (N + WAVE - 1) / WAVE;  // slots in use
    const long n = N;
    double xprev = (A[n * n + N - 1] - 0.0) / A[(n - 1) * n + N - 1];
    x[N - 1] = xprev;
    double cu[PS];  // a[i-1][i+1+t] for the products formed at step i
#pragma unroll
    for (int s = 0; s < PS; ++s) {
        const int r = N >= 3 ? N - 3 : 0;
        cu[s] = s < nsl ? A[(long)min(N - 1 + lane + WAVE * s, N - 1) * n + r] : 0.0;
    }
    for (int i = N - 2; i >= 0; --i) {
        const double di = A[(long)i * n + i], bi = A[n * n + i], u1 = A[(long)(i + 1) * n + i];
        double nu[PS];  // a[i-2][i+t], t < N - i: the products formed at step i - 1
        const int r2 = i >= 2 ? i - 2 : 0;
#pragma unroll
        for (int s = 0; s < PS; ++s) nu[s] = s < nsl ? A[(long)min(i + lane + WAVE * s, N - 1) * n + r2] : 0.0;
        const int lenp = i >= 1 ? N - i - 1 : 0;
        double sum = 0.0 + u1 * xprev;
        const int len = N - i - 2;
        if (len > 0) {
            const double2* p2 = (const double2*)(pb + (long)(i & 1) * pstride);
            double2 c0 = p2[0], c1 = p2[1], c2 = p2[2], c3 = p2[3];
            for (int g = 0; g < len; g += 16) {
                const double2* nx = p2 + (g >> 1);
                const double2 n0 = nx[4], n1 = nx[5], n2 = nx[6], n3 = nx[7];
                sum += c0.x;
                sum += c0.y;
                sum += c1.x;
                sum += c1.y;
                sum += c2.x;
                sum += c2.y;
                sum += c3.x;
                sum += c3.y;
                c0 = nx[8];
                c1 = nx[9];
                c2 = nx[10];
                c3 = nx[11];
                sum += n0.x;
                sum += n0.y;
                sum += n1.x;
                sum += n1.y;
                sum += n2.x;
                sum += n2.y;
                sum += n3.x;
                sum += n3.y;
            }
        }
        double* po = pb + (long)((i - 1) & 1) * pstride;
#pragma unroll
        for (int s = 0; s < PS; ++s) {
            if (s < nsl) {
                const int t = lane + WAVE * s;
                const double p = cu[s] * x[min(i + 1 + t, N - 1)];
                po[t] = t < lenp ? p : 0.0;
            }
        }
        xprev = (bi - sum) / di;
        x[i] = xprev;
#pragma unroll
        for (int s = 0; s < PS; ++s) cu[s] = nu[s];
    }
}

// Everything after the linear solve (linearized.rs:141-221): induced velocity,
// viscous correction of the circulation, induced velocity again, residual and
// outputs.  G holds the solution; R is [N] shared scratch.
__device__ inline void lin_finish(const Inst& d, const Strip& s, double* G, double* R, int as_start) {
    const int N = d.N;
    const int t = threadIdx.x;
    const bool act = t < N;
    const int ic = act ? t : N - 1;
    const int vc = as_start ? 0 : d.vc_mode;
    const double vc_ratio = d.vc_ratio;
    const int viscous = as_start ? 1 : !d.lin_disable_viscous;
    const int ivc = as_start ? 1 : d.lin_ivc;
    double g = G[ic];
    const V3 ind1 = induced_at(d, s.fixed, G, ic);
    V3 u = corrected_velocity(vc, vc_ratio, s.felt, ind1);
    double a;
    if (viscous) {  // linearized.rs:156-197
        a = angle_of_attack(s, u);
        double clf = foil_cl_pre_dropoff(*s.foil, a);
        double cll = foil_cl_linear(*s.foil, a);
        if (fabs(cll) > MIN_POSITIVE) g *= clf / cll;
        __syncthreads();
        if (act) G[t] = g;
        __syncthreads();
        // NoCorrection keeps the induced velocity of the linear solution
        u = corrected_velocity(vc, vc_ratio, s.felt, ivc ? induced_at(d, s.fixed, G, ic) : ind1);
    }
    a = angle_of_attack(s, u);
    if (as_start) {
        if (act) d.lin_gamma[t] = g;
        return;
    }
    if (act) R[t] = residual_term(s, d.density, g, a, u);
    __syncthreads();
    if (act) {
        double* o = d.out;
        o[OutLayout::gamma(N) + t] = g;
        o[OutLayout::alpha(N) + t] = a;
        store3(o + OutLayout::u(N) + 3 * t, u);
        store3(o + OutLayout::felt(N) + 3 * t, s.felt);
        if (d.M > 0) {  // update_wing_strength (update_data.rs:146-150)
            d.str[t] = g;
            d.rec[t].strength = g;
        }
        d.prev_gamma[t] = g;
    }
    if (t == 0) {
        double sum = 0.0;
        for (int q = 0; q < N; ++q) sum += R[q];
        d.out[OutLayout::iters(N, d.W)] = 1.0;
        d.out[OutLayout::residual(N, d.W)] = sum / (double)N;
    }
}

// ---- one-workgroup form (small N) -------------------------------------------
// T threads in 8 column groups x T/8 row groups: thread (rg, cg) holds rows
// r = rg + s T/8 (s < R) and columns j = cg + 8 c (c < C = 4 CB; the right-hand
// side is column N) of the augmented matrix in registers for the whole
// elimination, so a pivot-row value read from shared memory serves R rows.  Row
// exchanges are not carried out: every wave keeps the logical position of every
// row (lanes hold rows l, l + 64, ...) and which rows have been pivots.  Per pivot k:
//   - column k of every row is read from shared memory (cst, published at step k - 1);
//   - every wave reduces it on its own: the largest magnitude among the rows not
//     yet used, NaN excluded, ties to the lower logical position; the row at
//     position k when that row holds NaN (linalg.rs:58-64 scans positions
//     k + 1.. with a strict >, so a NaN at position k is never replaced);
//   - the pivot row's threads publish its columns >= k (prow, [8][C + 2]), and
//     each thread forms its factors f = a[r][k] / pivot; barrier;
//   - every row subtracts f * prow[j] entry by entry in the blocks of 4 register
//     columns that hold a column >= k (the rows already used as pivots and the
//     entries left of column k + 1 take harmless updates: they are never read
//     again), wave 0 copies the pivot row into U (row-major, logical position k),
//     and the owners of column k + 1 publish it; barrier.
// Wave 0 then back-substitutes.
template <int V>
struct LinrV;  // <threads, rows per thread, blocks of 4 columns per thread, 64-row slices of the pivot search>
template <>
struct LinrV<1> {
    static constexpr int T = 256, R = 2, CB = 3, MS = 1;  // N <= 64
};
template <>
struct LinrV<2> {
    static constexpr int T = 256, R = 4, CB = 4, MS = 2;  // N <= 127
};
template <>
struct LinrV<3> {
    static constexpr int T = 256, R = 4, CB = 5, MS = 2;  // N <= 128
};
template <>
struct LinrV<4> {
    static constexpr int T = 512, R = 3, CB = 5, MS = 3;  // N <= 159 (160 KB of shared memory: N <= 139)
};
template <int V>
SB_HD inline bool linr_fits(int N) {
    using P = LinrV<V>;
    return N >= 1 && N <= P::T / 8 * P::R && N + 1 <= 32 * P::CB && N <= 64 * P::MS;
}
// Shared memory (doubles): U [N][N + 1] (+8), G [N] (+8), R [N], cst (column k of
// every row, [N + 1] with a spare slot; later the back substitution's two product
// rows), prow [8][C + 2].  Every region starts on a 16-byte boundary.
struct LinrLayout {
    long G, R, cst, prow, total;
    int pstride;  // product row stride of the back substitution
};
SB_HD inline long linr_even(long v) { return (v + 1) & ~1L; }
constexpr int LINR_PS = 3;  // product slots per lane of the back substitution (N <= 192)
template <int V>
SB_HD inline LinrLayout linr_layout(int N) {
    LinrLayout L;
    L.pstride = (int)(N + 40 > 64 * LINR_PS ? linr_even(N + 40) : 64 * LINR_PS);
    L.G = (long)N * (N + 1) + 8;
    L.R = L.G + linr_even(N + 8);
    L.cst = L.R + linr_even(N);
    const long cst = N + 1 > 2L * L.pstride + 16 ? N + 1 : 2L * L.pstride + 16;
    L.prow = L.cst + linr_even(cst);
    L.total = L.prow + 8L * (4 * LinrV<V>::CB + 2);
    return L;
}
template <int V>
SB_HD inline long linr_shared_doubles(int N) { return linr_layout<V>(N).total; }

__device__ inline double readlane_f64(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}
template <int CTRL>
__device__ inline double dpp_f64(double v) {
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(v), __double2loint(v), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(v), __double2hiint(v), CTRL, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}
// Largest value over the wave (fmax: a NaN operand loses), the same on every lane.
// Four DPP exchange levels inside each row of 16 lanes, then the four row results.
__device__ inline double wave_max_f64(double v) {
    v = fmax(v, dpp_f64<0xB1>(v));   // quad_perm [1,0,3,2]
    v = fmax(v, dpp_f64<0x4E>(v));   // quad_perm [2,3,0,1]
    v = fmax(v, dpp_f64<0x141>(v));  // row_half_mirror
    v = fmax(v, dpp_f64<0x140>(v));  // row_mirror
    return fmax(fmax(readlane_f64(v, 0), readlane_f64(v, 16)), fmax(readlane_f64(v, 32), readlane_f64(v, 48)));
}
__device__ inline int wave_min_i32(int v) {
    v = min(v, __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xf, 0xf, false));
    v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xf, 0xf, false));
    v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x141, 0xf, 0xf, false));
    v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x140, 0xf, 0xf, false));
    return min(min(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
               min(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}

// linalg.rs:92-101 by one wave over the row-major upper triangle U (logical
// rows, leading dimension N + 1), N <= 64 PS.  Every lane runs the same ordered
// sum sum = ((0 + u[i][i+1] x[i+1]) + u[i][i+2] x[i+2]) + ..., reading the
// products u[i][j] x[j], j >= i + 2, from shared memory (pb, two rows of pstride
// >= max(N + 40, 64 PS)) in groups of 16: lane t forms the products of row i - 1
// while row i is summed (its x[j], j > i, are known by then), every slot from a
// valid index and +0.0 past the row's end.  The padding leaves the sum unchanged:
// it starts at +0.0 and cannot become -0.0.  (Fewer instructions per row, not
// fewer latencies, is what shortens this loop: one wave issues all of it.)
template <int PS>
__device__ inline void back_substitute_wave(const double* U, int N, double* x, double* pb, int pstride) {
    const int lane = threadIdx.x & (WAVE - 1);
    const long ld = N + 1;
    double xprev = (U[(N - 1) * ld + N] - 0.0) / U[(N - 1) * ld + N - 1];
    x[N - 1] = xprev;
    for (int i = N - 2; i >= 0; --i) {
        const double* row = U + i * ld;
        const double di = row[i], bi = row[N], u1 = row[i + 1];
        const int lenp = i >= 1 ? N - i - 1 : 0;  // products of row i - 1
        const double* rowp = row - (i >= 1 ? ld : 0);
        double pr[PS];
#pragma unroll
        for (int s = 0; s < PS; ++s) {
            const int t = lane + WAVE * s;
            const int j = min(i + 1 + t, N - 1);
            const double p = rowp[j] * x[j];
            pr[s] = t < lenp ? p : 0.0;
        }
        double sum = 0.0 + u1 * xprev;
        const int len = N - i - 2;
        if (len > 0) {
            const double2* p2 = (const double2*)(pb + (long)(i & 1) * pstride);
            double2 c0 = p2[0], c1 = p2[1], c2 = p2[2], c3 = p2[3];
            for (int g = 0; g < len; g += 16) {
                const double2* nx = p2 + (g >> 1);
                const double2 n0 = nx[4], n1 = nx[5], n2 = nx[6], n3 = nx[7];
                sum += c0.x;
                sum += c0.y;
                sum += c1.x;
                sum += c1.y;
                sum += c2.x;
                sum += c2.y;
                sum += c3.x;
                sum += c3.y;
                c0 = nx[8];
                c1 = nx[9];
                c2 = nx[10];
                c3 = nx[11];
                sum += n0.x;
                sum += n0.y;
                sum += n1.x;
                sum += n1.y;
                sum += n2.x;
                sum += n2.y;
                sum += n3.x;
                sum += n3.y;
            }
        }
        double* po = pb + (long)((i - 1) & 1) * pstride;
#pragma unroll
        for (int s = 0; s < PS; ++s) po[lane + WAVE * s] = pr[s];
        xprev = (bi - sum) / di;
        x[i] = xprev;
    }
}

// Where the one-workgroup elimination takes its augmented matrix from: row(r)
// returns what every entry of row r needs, entry(row, r, j) the entry in column
// j (j = N: the right-hand side).
struct LinSystemSrc {  // linearized.rs:75-131
    const Inst* d;
    __device__ LinRow row(int r) const { return lin_row(load_strip(*d, r)); }
    __device__ double entry(const LinRow& q, int r, int j) const { return j < d->N ? lin_entry(*d, q, r, j) : -q.cll; }
};
struct DenseSrc {  // a row-major N x N matrix and its right-hand side
    const double* A;
    const double* b;
    int N;
    struct Row {};
    __device__ Row row(int) const { return Row{}; }
    __device__ double entry(Row, int r, int j) const { return j < N ? A[(long)r * N + j] : b[r]; }
};

// Solves the system on one workgroup of LinrV<V>::T threads; the solution goes to
// G of the layout.  Returns 0, or k + 1 when the pivot of step k is below 1e-12 in
// magnitude (linalg.rs:67-71; G is then not written).  Ends with a barrier.
template <int V, class Src>
__device__ inline int gauss_regs(int N, const Src& src, double* sh) {
    using P = LinrV<V>;
    constexpr int T = P::T, R = P::R, CB = P::CB, MS = P::MS;
    constexpr int CG = 8, RG = T / CG, C = 4 * CB;
    constexpr int US = (8 * C + WAVE - 1) / WAVE;  // wave 0's slots for the columns <= N
    const LinrLayout L = linr_layout<V>(N);
    const int t = threadIdx.x, lane = t & (WAVE - 1);
    const int rg = t >> 3, cg = t & 7;
    double* U = sh;
    double* G = sh + L.G;
    double* cst = sh + L.cst;
    double* prow = sh + L.prow;
    double* prq = prow + cg * (C + 2);
    const long ld = N + 1;
    const int nbu = (N + 1 + 31) >> 5;  // blocks of 4 register columns that hold a column <= N
    int rs[R];                          // this thread's rows; N (a spare slot of cst) past the end
#pragma unroll
    for (int s = 0; s < R; ++s) rs[s] = rg + RG * s < N ? rg + RG * s : N;
    double a[R][C];
#pragma unroll
    for (int s = 0; s < R; ++s) {
        const int r = rg + RG * s;
        const bool own = r < N;
        const auto qr = src.row(own ? r : N - 1);
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const int j = cg + CG * c;
            a[s][c] = own && j <= N ? src.entry(qr, r, j) : 0.0;
        }
    }
    if (cg == 0) {
#pragma unroll
        for (int s = 0; s < R; ++s) cst[rs[s]] = a[s][0];
    }
    int wpos[MS];
    bool wdone[MS];
#pragma unroll
    for (int sl = 0; sl < MS; ++sl) {
        wpos[sl] = lane + WAVE * sl;
        wdone[sl] = false;
    }
    int singular = 0;
    __syncthreads();
    for (int k = 0; k < N; ++k) {
        double v[MS];
#pragma unroll
        for (int sl = 0; sl < MS; ++sl) {
            const int row = lane + WAVE * sl;
            v[sl] = cst[row < N ? row : N - 1];
        }
        double fa[R];
#pragma unroll
        for (int s = 0; s < R; ++s) fa[s] = cst[rs[s]];
        double mm = -1.0;
#pragma unroll
        for (int sl = 0; sl < MS; ++sl) {
            const bool cand = lane + WAVE * sl < N && !wdone[sl];
            mm = fmax(mm, cand ? fabs(v[sl]) : -1.0);
        }
        const double M = wave_max_f64(mm);
        int key = 0x7fffffff;  // logical position * 1024 + row
        int rk = -1;           // row at logical position k
        bool nank = false;
#pragma unroll
        for (int sl = 0; sl < MS; ++sl) {
            const int row = lane + WAVE * sl;
            const bool cand = row < N && !wdone[sl];
            const int kk = wpos[sl] * 1024 + row;
            if (cand && fabs(v[sl]) == M && kk < key) key = kk;
            if (cand && wpos[sl] == k) {
                rk = row;
                nank = v[sl] != v[sl];
            }
        }
        const int K = wave_min_i32(key);
        {
            const unsigned long long hit = __ballot(rk >= 0);
            const int src_lane = (int)__builtin_ctzll(hit);
            rk = __builtin_amdgcn_readlane(rk, src_lane);
            nank = __builtin_amdgcn_readlane((int)nank, src_lane) != 0;
        }
        int pr = K & 1023, ppos = K >> 10;
        if (nank) {
            pr = rk;
            ppos = k;
        }
        pr = __builtin_amdgcn_readfirstlane(pr);
        const int ps = pr >> 6;
        double vsel = v[0];
#pragma unroll
        for (int sl = 1; sl < MS; ++sl)
            if (ps == sl) vsel = v[sl];
        const double piv = readlane_f64(vsel, pr & (WAVE - 1));
        // a singular pivot is recorded and the remaining steps run on (their results
        // are not used): the loop keeps one uniform exit
        if (fabs(piv) < 1e-12 && !singular) singular = k + 1;
#pragma unroll
        for (int sl = 0; sl < MS; ++sl) {
            const int row = lane + WAVE * sl;
            if (row == pr) {
                wpos[sl] = k;
                wdone[sl] = true;
            } else if (row == rk) {
                wpos[sl] = ppos;
            }
        }
        const int b0 = k >> 5;  // blocks that hold a column >= k
        if (rg == pr % RG) {    // the pivot row's threads
            const int sp = pr / RG;
#pragma unroll
            for (int s = 0; s < R; ++s) {
                if (s == sp) {
#pragma unroll
                    for (int b = 0; b < CB; ++b) {
                        if (b >= b0 && b < nbu) {
                            double2* o = (double2*)(prq + 4 * b);
                            o[0] = make_double2(a[s][4 * b], a[s][4 * b + 1]);
                            o[1] = make_double2(a[s][4 * b + 2], a[s][4 * b + 3]);
                        }
                    }
                }
            }
        }
        // the factors, formed after the pivot row's stores are issued and before the barrier
        double f[R];
#pragma unroll
        for (int s = 0; s < R; ++s) {
            f[s] = fa[s] / piv;
            asm volatile("" ::"v"(f[s]));
        }
        __syncthreads();
        double2 w[C / 2];
#pragma unroll
        for (int b = 0; b < CB; ++b) {
            if (b < nbu) {
                const double2* s2 = (const double2*)(prq + 4 * b);
                w[2 * b] = s2[0];
                w[2 * b + 1] = s2[1];
            }
        }
        double uc[US];
        if (t < WAVE) {  // wave 0 copies the pivot row into U (logical row k)
#pragma unroll
            for (int sl = 0; sl < US; ++sl) {
                const int j = lane + WAVE * sl;
                uc[sl] = j <= N ? prow[(j & 7) * (C + 2) + (j >> 3)] : 0.0;
            }
        }
#pragma unroll
        for (int b = 0; b < CB; ++b) {
            if (b >= b0 && b < nbu) {
#pragma unroll
                for (int s = 0; s < R; ++s) {
                    a[s][4 * b] = a[s][4 * b] - f[s] * w[2 * b].x;
                    a[s][4 * b + 1] = a[s][4 * b + 1] - f[s] * w[2 * b].y;
                    a[s][4 * b + 2] = a[s][4 * b + 2] - f[s] * w[2 * b + 1].x;
                    a[s][4 * b + 3] = a[s][4 * b + 3] - f[s] * w[2 * b + 1].y;
                }
            }
        }
        if (k + 1 < N && cg == ((k + 1) & 7)) {  // publish column k + 1
            const int c1 = (k + 1) >> 3;
#pragma unroll
            for (int c = 0; c < C; ++c) {
                if (c == c1) {
#pragma unroll
                    for (int s = 0; s < R; ++s) cst[rs[s]] = a[s][c];
                }
            }
        }
        if (t < WAVE) {
#pragma unroll
            for (int sl = 0; sl < US; ++sl) {
                const int j = lane + WAVE * sl;
                if (j <= N) U[k * ld + j] = uc[sl];
            }
        }
        __syncthreads();
    }
    if (singular) return singular;
    if (t < WAVE) {
        if (N <= 128) back_substitute_wave<2>(U, N, G, cst, L.pstride);
        else back_substitute_wave<LINR_PS>(U, N, G, cst, L.pstride);
    }
    __syncthreads();
    return 0;
}

template <int V>
__device__ inline void linearized_reg_body(const Inst& d, double* sh, int as_start) {
    const int N = d.N;
    const int t = threadIdx.x;
    const LinrLayout L = linr_layout<V>(N);
    double* G = sh + L.G;
    double* R = sh + L.R;
    const int singular = gauss_regs<V>(N, LinSystemSrc{&d}, sh);
    if (singular) {
        if (t == 0) {
            for (int qq = 0; qq < N; ++qq) G[qq] = 0.0;
            d.scal[4] = (double)singular;
        }
        __syncthreads();
    }
    lin_finish(d, load_strip(d, t < N ? t : N - 1), G, R, as_start);
}
template <int V>
__global__ __launch_bounds__(LinrV<V>::T) void k_linearized(Inst d, int as_start) {
    extern __shared__ double sh[];
    linearized_reg_body<V>(d, sh, as_start);
}

// The one-workgroup solve on a given system (checks of the elimination itself):
// status = 0, or the singular step + 1 (x is then zero).
template <int V>
__global__ __launch_bounds__(LinrV<V>::T) void k_gauss_dense(const double* A, const double* b, int N, double* x, int* status) {
    extern __shared__ double sh[];
    const int s = gauss_regs<V>(N, DenseSrc{A, b, N}, sh);
    const double* G = sh + linr_layout<V>(N).G;
    if (threadIdx.x == 0) *status = s;
    for (int i = threadIdx.x; i < N; i += LinrV<V>::T) x[i] = s ? 0.0 : G[i];
}

// ---- blocked form (matrix in global memory, d.lin) -------------------------
// d.lin: [N + 1][N] matrix, then [N][8] row data (LinRow), then [N] ints: the
// source position of every row after the current panel's exchanges.
__device__ inline double* lin_rows(const Inst& d) { return d.lin + (long)d.N * (d.N + 1); }
__device__ inline int* lin_perm(const Inst& d) { return (int*)(lin_rows(d) + 8L * d.N); }

__global__ void k_lin_rows(Inst d) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.N) return;
    LinRow q = lin_row(load_strip(d, i));
    double* o = lin_rows(d) + 8L * i;
    o[0] = q.ndir.x;
    o[1] = q.ndir.y;
    o[2] = q.ndir.z;
    o[3] = q.ut;
    o[4] = q.dcl;
    o[5] = q.cll;
    o[6] = q.c;
}

__global__ void k_lin_assemble(Inst d) {
    const long N = d.N;
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= N * (N + 1)) return;
    int c = (int)(idx / N), r = (int)(idx - (long)c * N);
    const double* o = lin_rows(d) + 8L * r;
    LinRow q;
    q.ndir = V3{o[0], o[1], o[2]};
    q.ut = o[3];
    q.dcl = o[4];
    q.cll = o[5];
    q.c = o[6];
    d.lin[idx] = c < N ? lin_entry(d, q, r, c) : -q.cll;
}

// Panel factorization of columns [k0, k0 + b), b <= 16: one workgroup of
// LINP_T threads, thread t holding rows k0 + t + LINP_T s (s < 4) of the panel in
// registers.  Rows are not exchanged while the panel is factored: every row keeps
// its logical position.  Per pivot, each wave reduces its rows (the largest
// magnitude, NaN excluded, ties to the lower position) and its best row publishes
// its entries with the wave's result (LINP_SLOT doubles per wave, two sets by step
// parity), and the row at position kk publishes its entries too (the pivot when it
// holds NaN: linalg.rs:58-64 never replaces a NaN at the pivot position); after
// one barrier every wave reduces the wave results and takes the pivot row from its
// slot.  The factors replace the entries below the diagonal; rows are written back
// at their logical positions, and the source position of every row goes to
// lin_perm for the trailing update.
constexpr int LINP_T = 256;                // threads of the panel kernel (4 rows each, N <= 1024)
constexpr int LINP_W = LINP_T / 64;        // its waves
constexpr int LINP_R = 4;                  // rows per thread
constexpr int LINP_SLOT = 20;              // a candidate row (16 entries), |v|, key, row at position kk, NaN flag
struct LinPanelState {
    double a[LINP_R][16];
    int pos[LINP_R];  // logical position of each row in the panel
    bool done[LINP_R];
};
// Pivot KK of the panel (a template, so that every register index is fixed).
// Returns false at a singular pivot (recorded in scal[4]).
template <int KK>
__device__ __forceinline__ bool lin_panel_step(const Inst& d, int k0, int b, double* slots, double* pks, LinPanelState& st) {
    constexpr int R = LINP_R;
    const int t = threadIdx.x, lane = t & (WAVE - 1), w = t >> 6;
    double m = -1.0;
#pragma unroll
    for (int s = 0; s < R; ++s) m = fmax(m, st.done[s] ? -1.0 : fabs(st.a[s][KK]));  // NaN |v| never wins
    const double Mw = wave_max_f64(m);
    int key = 0x7fffffff;
#pragma unroll
    for (int s = 0; s < R; ++s) {
        const int kk2 = st.pos[s] * 1024 + t + LINP_T * s;
        if (!st.done[s] && fabs(st.a[s][KK]) == Mw && kk2 < key) key = kk2;
    }
    const int Kw = wave_min_i32(key);
    double* sl = slots + ((KK & 1) * LINP_W + w) * LINP_SLOT;
    double* pk = pks + (KK & 1) * LINP_SLOT;
    if (lane == 0) {
        sl[16] = Mw;
        sl[17] = (double)Kw;
    }
#pragma unroll
    for (int s = 0; s < R; ++s) {
        const int r = t + LINP_T * s;
        if (Kw != 0x7fffffff && r == (Kw & 1023)) {
#pragma unroll
            for (int c = KK; c < 16; ++c)
                if (c < b) sl[c] = st.a[s][c];
        }
        if (!st.done[s] && st.pos[s] == KK) {
#pragma unroll
            for (int c = KK; c < 16; ++c)
                if (c < b) pk[c] = st.a[s][c];
            pk[16] = (double)r;
            pk[17] = st.a[s][KK] != st.a[s][KK] ? 1.0 : 0.0;
        }
    }
    __syncthreads();
    double mr = -1.0;
    int kr = 0x7fffffff;
    if (lane < LINP_W) {
        mr = slots[((KK & 1) * LINP_W + lane) * LINP_SLOT + 16];
        kr = (int)slots[((KK & 1) * LINP_W + lane) * LINP_SLOT + 17];
    }
    const double M = wave_max_f64(mr);
    const int K = wave_min_i32(mr == M ? kr : 0x7fffffff);
    int pr, ppos;
    const double* prow;
    if (pk[17] != 0.0) {  // NaN at the pivot position: that row
        pr = (int)pk[16];
        ppos = KK;
        prow = pk;
    } else {
        pr = K & 1023;
        ppos = K >> 10;
        prow = slots + ((KK & 1) * LINP_W + ((pr % LINP_T) >> 6)) * LINP_SLOT;
    }
    const double piv = prow[KK];
    if (fabs(piv) < 1e-12) {  // uniform: every thread read the same slots
        if (t == 0) d.scal[4] = (double)(k0 + KK + 1);
        return false;
    }
#pragma unroll
    for (int s = 0; s < R; ++s) {
        const int r = t + LINP_T * s;
        if (r == pr) {
            st.pos[s] = KK;
            st.done[s] = true;
        } else if (!st.done[s] && st.pos[s] == KK) {
            st.pos[s] = ppos;
        }
        if (!st.done[s]) {
            const double f = st.a[s][KK] / piv;
#pragma unroll
            for (int c = KK + 1; c < 16; ++c)
                if (c < b) st.a[s][c] = st.a[s][c] - f * prow[c];
            st.a[s][KK] = f;
        }
    }
    return true;
}
template <int KK>
struct LinPanelSteps {
    __device__ __forceinline__ static bool run(const Inst& d, int k0, int b, double* slots, double* pks, LinPanelState& st) {
        if (KK >= b) return true;
        if (!lin_panel_step<KK>(d, k0, b, slots, pks, st)) return false;
        return LinPanelSteps<KK + 1>::run(d, k0, b, slots, pks, st);
    }
};
template <>
struct LinPanelSteps<16> {
    __device__ __forceinline__ static bool run(const Inst&, int, int, double*, double*, LinPanelState&) { return true; }
};

__global__ __launch_bounds__(LINP_T) void k_lin_panel(Inst d, int k0, int b) {
    extern __shared__ double sh[];
    if (d.scal[4] != 0.0) return;
    const int N = d.N, t = threadIdx.x;
    const int nr = N - k0;
    double* slots = sh;                              // [2][LINP_W][LINP_SLOT]
    double* pks = slots + 2 * LINP_W * LINP_SLOT;    // [2][LINP_SLOT]: the row at position kk
    LinPanelState st;
#pragma unroll
    for (int s = 0; s < LINP_R; ++s) {
        const int r = t + LINP_T * s;
        st.pos[s] = r;
        st.done[s] = r >= nr;  // past the panel: never a candidate, never updated
#pragma unroll
        for (int c = 0; c < 16; ++c) st.a[s][c] = r < nr && c < b ? d.lin[(long)(k0 + c) * N + k0 + r] : 0.0;
    }
    if (!LinPanelSteps<0>::run(d, k0, b, slots, pks, st)) return;
#pragma unroll
    for (int s = 0; s < LINP_R; ++s) {
        const int r = t + LINP_T * s;
        if (r < nr) {
#pragma unroll
            for (int c = 0; c < 16; ++c)
                if (c < b) d.lin[(long)(k0 + c) * N + k0 + st.pos[s]] = st.a[s][c];
            lin_perm(d)[st.pos[s]] = r;
        }
    }
}

// Trailing update after the panel [k0, k0 + b): one workgroup per column j >=
// k0 + b (the right-hand side included).  The panel's exchanges are applied as
// one gather; the panel rows' entries take their updates from the earlier
// pivots in order (these are the pivot rows' values), then every lower entry
// takes the b updates a - l*u in pivot order.
constexpr int LIN_TRAIL_THREADS = 256;
__global__ __launch_bounds__(LIN_TRAIL_THREADS) void k_lin_trailing(Inst d, int k0, int b) {
    __shared__ double u[16];
    if (d.scal[4] != 0.0) return;
    const int N = d.N, t = threadIdx.x;
    const int nr = N - k0;
    const int j = k0 + b + blockIdx.x;
    if (j > N) return;
    double* cj = d.lin + (long)j * N + k0;  // rows from k0
    const int* src = lin_perm(d);
    constexpr int PER = 1024 / LIN_TRAIL_THREADS;
    double v[PER];
#pragma unroll
    for (int m = 0; m < PER; ++m) {
        int q = t + m * LIN_TRAIL_THREADS;
        v[m] = (q < nr && src[q] != q) ? cj[src[q]] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < PER; ++m) {
        int q = t + m * LIN_TRAIL_THREADS;
        if (q < nr && src[q] != q) cj[q] = v[m];
    }
    __syncthreads();
    if (t == 0) {
        double uu[16];
        for (int kk = 0; kk < b; ++kk) uu[kk] = cj[kk];
        for (int kk = 1; kk < b; ++kk) {
            double a = uu[kk];
            for (int m = 0; m < kk; ++m) a = a - d.lin[(long)(k0 + m) * N + k0 + kk] * uu[m];
            uu[kk] = a;
        }
        for (int kk = 0; kk < b; ++kk) {
            u[kk] = uu[kk];
            cj[kk] = uu[kk];
        }
    }
    __syncthreads();
    for (int q = b + t; q < nr; q += LIN_TRAIL_THREADS) {
        double a = cj[q];
        for (int kk = 0; kk < b; ++kk) a = a - d.lin[(long)(k0 + kk) * N + k0 + q] * u[kk];
        cj[q] = a;
    }
}

// Back substitution and the rest of the solve after the blocked factorization.
// Shared memory: G [N] (+8), R [N], two product rows of linr_even(N + 64) (+16).
__global__ __launch_bounds__(1024) void k_lin_finish(Inst d, int as_start) {
    extern __shared__ double sh[];
    const int N = d.N, t = threadIdx.x;
    double* G = sh;
    double* R = G + linr_even(N + 8);
    double* pb = R + linr_even(N);
    if (d.scal[4] != 0.0) {
        if (t == 0)
            for (int q = 0; q < N; ++q) G[q] = 0.0;
    } else if (t < WAVE) {
        back_substitute_wave_cm(d.lin, N, G, pb, (int)linr_even(N + 64));
    }
    __syncthreads();
    const Strip s = load_strip(d, t < N ? t : N - 1);
    lin_finish(d, s, G, R, as_start);
}

// The blocked solve on a given system (checks of the elimination itself): the
// back substitution of the factored d.lin into x; status = the singular step + 1
// or 0.  Shared memory as k_lin_finish.
__global__ __launch_bounds__(1024) void k_gauss_blocked_finish(Inst d, double* x, int* status) {
    extern __shared__ double sh[];
    const int N = d.N, t = threadIdx.x;
    double* G = sh;
    double* pb = G + linr_even(N + 8) + linr_even(N);
    const int s = (int)d.scal[4];
    if (s == 0 && t < WAVE) back_substitute_wave_cm(d.lin, N, G, pb, (int)linr_even(N + 64));
    __syncthreads();
    if (t == 0) *status = s;
    for (int i = t; i < N; i += 1024) x[i] = s ? 0.0 : G[i];
}

// ---------------------------------------------------------------------------
// Forces (force_calculations.rs:322-548, forces_and_moments.rs:163-245,
// flow_derivatives.rs:31-62), one workgroup
// ---------------------------------------------------------------------------
__device__ inline void forces_body(const Inst& d, double dt) {
    int N = d.N, W = d.W;
    const double* geo = d.geo;
    double* o = d.out;
    int i = threadIdx.x;
    double count = d.scal[3];
    if (i < N) {
        double gamma = o[OutLayout::gamma(N) + i];
        double alpha = o[OutLayout::alpha(N) + i];
        V3 ug = load3(o + OutLayout::u(N) + 3 * i);
        V3 acc{0.0, 0.0, 0.0};
        if (count >= 2.0) {  // first_derivative_second_order_backward
            V3 h0 = load3(d.hist + 3 * i), h1 = load3(d.hist + 3 * (N + i));
            acc = (ug * 3.0 - h1 * 4.0 + h0) / (2.0 * dt);
        }
        V3 vel = ug;
        if (d.output_body) {
            const double* ax = geo + d.gl.body_axes;
            V3 X = load3(ax), Y = load3(ax + 3), Z = load3(ax + 6);
            vel = V3{dot(ug, X), dot(ug, Y), dot(ug, Z)};
            acc = V3{dot(acc, X), dot(acc, Y), dot(acc, Z)};
        }
        const FoilParams& foil = d.foil[d.wing_of[i]];
        double rho = d.density;
        double c = geo[d.gl.chord_len + i];
        double llen = geo[d.gl.local_len + i];
        V3 rel_out = load3(geo + d.gl.rel_out + 3 * i);
        V3 rel_local = load3(geo + d.gl.rel_local + 3 * i);
        V3 F[6];
        F[0] = len(vel) == 0.0 ? V3{0.0, 0.0, 0.0} : (cross(vel, rel_out) * gamma) * rho;
        F[1] = V3{0.0, 0.0, 0.0};
        if (!d.viscous_lift_in_circulation) {
            double clv = foil_cl_post_weighted(foil, alpha);
            V3 dir = normalize(cross(rel_local, vel));
            double area = c * llen;
            double ff = 0.5 * area * rho * len2(vel);
            F[1] = dir * clv * ff;
        }
        {
            double cd = foil_cd(foil, alpha);
            V3 dir = normalize(vel);
            double area = c * llen;
            double ff = 0.5 * area * rho * len2(vel);
            F[2] = dir * cd * ff;
        }
        {
            V3 dir_out = load3(geo + d.gl.dir_out + 3 * i);
            V3 chord_out = load3(geo + d.gl.chord_out + 3 * i);
            double area = c * len(rel_out);
            V3 ra = acc;
            ra = ra - project(ra, dir_out);
            ra = ra - project(ra, chord_out);
            if (len(ra) == 0.0) F[3] = V3{0.0, 0.0, 0.0};
            else {
                double coef = foil.added_mass_factor * len(ra);
                F[3] = (coef * rho * area) * normalize(ra);
            }
        }
        F[4] = V3{0.0, 0.0, 0.0};
        V3 tot = F[0];
        tot = tot + F[1];
        tot = tot + F[2];
        tot = tot + F[3];
        tot = tot + F[4];
        F[5] = tot;
        for (int k = 0; k < 6; ++k) store3(o + OutLayout::sect(N) + 3 * ((long)k * N + i), F[k]);
        store3(o + OutLayout::ctrl(N) + 3 * i, load3(geo + d.gl.ctrl + 3 * i));
        store3(o + OutLayout::u(N) + 3 * i, vel);  // force_input.velocity is in the output frame
        // flow-derivative history update (global-frame velocity)
        store3(d.hist + 3 * i, load3(d.hist + 3 * (N + i)));
        store3(d.hist + 3 * (N + i), ug);
    }
    __syncthreads();
    for (int idx = i; idx < 6 * W; idx += blockDim.x) {
        int k = idx / W, w = idx - k * W;
        int a = -1, b = -1;
        for (int q = 0; q < N; ++q) {
            if (d.wing_of[q] == w) {
                if (a < 0) a = q;
                b = q + 1;
            }
        }
        V3 fs{0.0, 0.0, 0.0}, ms{0.0, 0.0, 0.0};
        for (int q = a; a >= 0 && q < b; ++q) {
            if (geo[d.gl.is_virtual + q] != 0.0) continue;
            V3 f = load3(o + OutLayout::sect(N) + 3 * ((long)k * N + q));
            fs = fs + f;
            ms = ms + cross(load3(geo + d.gl.ctrl_out + 3 * q), f);
        }
        store3(o + OutLayout::integ_f(N) + 3 * ((long)k * W + w), fs);
        store3(o + OutLayout::integ_m(N, W) + 3 * ((long)k * W + w), ms);
    }
    if (i == 0) {
        d.scal[3] = count + 1.0;
        d.out[OutLayout::singular(d.N, d.W)] = d.scal[4];  // with the results, one download
    }
}
__global__ __launch_bounds__(1024) void k_forces(Inst d, double dt) { forces_body(d, dt); }

__global__ void k_copy3(const double* __restrict__ src, double* __restrict__ dst, long n) {
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < 3 * n) dst[q] = src[q];
}

// Freestream of the wind environment at every wake point, written after the
// control points (environment.rs:182-190 per point; complete_sail_model/mod.rs:217-255).
__global__ void k_wind_points(Inst d, WindField w) {
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long)d.P * (d.M + 1)) return;
    store3(const_cast<double*>(d.fs) + 3L * (d.N + q), apparent_wind_at(w, load3(d.pts + 3 * q)));
}

__global__ void k_libm(int fn, const double* x, const double* y, double* out, long n) {
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    double r;
    switch (fn) {
        case 0: r = sbm::sin(x[q]); break;
        case 1: r = sbm::cos(x[q]); break;
        case 2: r = sbm::exp(x[q]); break;
        case 3: r = sbm::acos(x[q]); break;
        case 4: r = sbm::pow(x[q], y[q]); break;
        case 6: r = sbm::atan(x[q]); break;
        default: r = sbm::log(x[q]); break;
    }
    out[q] = r;
}


// ---------------------------------------------------------------------------
// Batched instances
// ---------------------------------------------------------------------------

// Owner of flat block b: the last instance i with off[i] <= b.
__device__ inline int owner_of(const long* off, int n, long b) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Freestream of batched instances from their wind fields (environment.rs:182-190 at
// every point): one 256-point block per thread block of the flat range.
__global__ void kb_wind(const WindJob* J, const long* blk_off, int n) {
    const int i = owner_of(blk_off, n, (long)blockIdx.x);
    const WindJob& j = J[i];
    const long q = (long)(blockIdx.x - blk_off[i]) * 256 + threadIdx.x;
    if (q < j.n_ctrl) store3(j.dst + 3 * q, load3(j.ctrl + 3 * q));
    else if (q < j.n_ctrl + j.n_pts) store3(j.dst + 3 * q, apparent_wind_at(j.w, load3(j.pts + 3 * (q - j.n_ctrl))));
}


__global__ void kb_prepare(const Inst* I, int first_step) {
    const Inst d = I[blockIdx.x];
    prepare_body(d, first_step);
}

__global__ void kb_init_wake(const Inst* I, const long* off, int n) {
    long b = blockIdx.x;
    int i = owner_of(off, n, b);
    const Inst d = I[i];
    long nb = off[i + 1] - off[i];
    init_wake_body(d, d.dt, 1, V3{0.0, 0.0, 0.0}, 0, (b - off[i]) * blockDim.x + threadIdx.x, nb * blockDim.x);
}

__global__ void kb_wake_step(const Inst* I, const long* off, int n) {
    long b = blockIdx.x;
    int i = owner_of(off, n, b);
    const Inst d = I[i];
    wake_step_body(d, d.dt, (b - off[i]) * blockDim.x + threadIdx.x);
}

__global__ void kb_wake_last(const Inst* I, const long* off, int n) {
    long b = blockIdx.x;
    int i = owner_of(off, n, b);
    const Inst d = I[i];
    wake_last_body(d, (b - off[i]) * blockDim.x + threadIdx.x);
}

__global__ void kb_panels(const Inst* I, const long* off, int n, int shift) {
    long b = blockIdx.x;
    int i = owner_of(off, n, b);
    const Inst d = I[i];
    panels_body(d, shift, (b - off[i]) * blockDim.x + threadIdx.x);
}

__global__ void kb_factors(const Inst* I, const long* off, int n) {
    long b = blockIdx.x;
    int i = owner_of(off, n, b);
    const Inst d = I[i];
    factors_body(d, (b - off[i]) * blockDim.x + threadIdx.x);
}

__global__ void kb_horseshoe_factors(const Inst* I, const long* off, int n) {
    long b = blockIdx.x;
    int i = owner_of(off, n, b);
    const Inst d = I[i];
    horseshoe_factors_body(d, (b - off[i]) * blockDim.x + threadIdx.x);
}

// One workgroup per instance (the one-workgroup Linearized form).
template <int V>
__global__ __launch_bounds__(LinrV<V>::T) void kb_linearized(const Inst* I, int as_start) {
    extern __shared__ double sh[];
    const Inst d = I[blockIdx.x];
    linearized_reg_body<V>(d, sh, as_start);
}

// Singular-pivot flags of every instance (scal[4]) into flags[n], then cleared.
__global__ void kb_take_singular(const Inst* I, int n, double* flags) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double* sc = I[i].scal;
    flags[i] = sc[4];
    sc[4] = 0.0;
}

__global__ void kb_velocity_copy(const Inst* I, const long* off, int n) {
    long b = blockIdx.x;
    int i = owner_of(off, n, b);
    const Inst d = I[i];
    long q = (b - off[i]) * blockDim.x + threadIdx.x;
    if (q < 3L * d.P * (d.M + 1)) d.vel[q] = d.fs[3L * d.N + q];
}

template <int MAXT>
__global__ __launch_bounds__(MAXT) void kb_solve(const Inst* I) {
    extern __shared__ double sh[];
    const Inst d = I[blockIdx.x];
    solve_body(d, sh);
}

__global__ __launch_bounds__(1024) void kb_forces(const Inst* I) {
    const Inst d = I[blockIdx.x];
    forces_body(d, d.dt);
}

__global__ void kb_zero_counters(unsigned* c) {
    if (threadIdx.x < 2) c[threadIdx.x] = 0u;
}

__device__ inline int next_job(unsigned* head, int* slot) {
    __syncthreads();  // the previous job's readers of *slot are done
    if (threadIdx.x == 0) *slot = (int)__hip_atomic_fetch_add(head, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(*slot);
}

// Fixed velocities of many instances (frozen_wake.rs:189-261).  Job = one
// instance's 64 control points (lane = target) over its panels of rows >= 1,
// run by one 16-wave workgroup (a whole CU).  Fifteen compute waves evaluate
// the panels of a 45-panel block (three consecutive panels each) into one
// half of a double-buffered LDS term array while the last wave adds the
// previous block's terms, component by component, in ascending panel order:
// one barrier per block, and every target sees the same sequence of additions
// as the one-target fold (a skipped panel adds +0, which leaves a sum that
// started at +0 unchanged).  Panel records are read with scalar loads (the
// address is wave-uniform), the next block's across the barrier, so the
// compute waves use no LDS for them.  Jobs are taken longest first from a
// queue by persistent workgroups.
constexpr int COOP_WAVES = 16;
constexpr int FX_COMPUTE = 15;  // compute waves; wave 15 folds

// PER panels per compute wave per block (1, 2 or 3; the autotuner picks it):
// blocks of 15 PER panels, a double-buffered term array of 2 x 15 PER x 1.5 KiB.
template <int MODE, int SYM, int PER>
__global__ __launch_bounds__(1024) void kb_fixed(const Inst* I, const int2* __restrict__ jobs, int n_jobs,
                                                 unsigned* head) {
    constexpr int FX_BLOCK = FX_COMPUTE * PER;
    constexpr int FOLD_CHUNK = FX_BLOCK / 5;  // the fold's loads run one chunk ahead of its adds
    static_assert(FX_BLOCK % FOLD_CHUNK == 0, "the fold reads full blocks in chunks");
    __shared__ double T[2][FX_BLOCK][3][WAVE];  // 46,080 B per panel per wave
    __shared__ int slot;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
    const int lane = threadIdx.x % WAVE;
    for (;;) {
        int j = next_job(head, &slot);
        if (j >= n_jobs) break;
        int2 jb = jobs[j];
        const Inst d = I[jb.x];
        int t = jb.y * WAVE + lane;
        int tc = t < d.N ? t : d.N - 1;
        V3 x = load3(d.geo + d.gl.ctrl + 3 * tc);
        int wt = d.neglect_self ? d.wing_of[tc] : -1;
        const long p_begin = d.N, p_end = (long)d.N * d.nz_rows;
        const long nblk = p_end > p_begin ? (p_end - p_begin + FX_BLOCK - 1) / FX_BLOCK : 0;
        double acc = 0.0;
        if (wave < FX_COMPUTE) {
            ScalarRecs recs = (ScalarRecs)d.rec;
            const int pb = (int)p_begin, pe = (int)p_end;
            const int mine = pb + wave * PER;  // first panel of this wave in block 0
            PanelRec r[PER];
            if (nblk > 0) {
#pragma unroll
                for (int q = 0; q < PER; ++q) r[q] = scalar_rec(recs, mine + q < pe ? mine + q : pe - 1);
            }
            for (int b = 0; b <= (int)nblk; ++b) {
                if (b < (int)nblk) {
                    const int buf = b & 1;
                    const int k0 = mine + b * FX_BLOCK;
                    V3 term[PER];
                    if constexpr (SYM == 0 && MODE != MODE_FP32_KAHAN) {
                        // far-field terms of the three panels, interleaved; then the
                        // lanes that need the general evaluation
                        bool ok[PER];
#pragma unroll
                        for (int q = 0; q < PER; ++q) ok[q] = far_fast<MODE>(r[q], x, term[q]);
#pragma unroll
                        for (int q = 0; q < PER; ++q)
                            if (!ok[q]) term[q] = panel_unit<MODE>(d, r[q], k0 + q < pe ? k0 + q : pe - 1, x);
                    } else {
#pragma unroll
                        for (int q = 0; q < PER; ++q)
                            term[q] = pair_unit<MODE, SYM>(d, r[q], k0 + q < pe ? k0 + q : pe - 1, x);
                    }
#pragma unroll
                    for (int q = 0; q < PER; ++q) {
                        const int k = k0 + q;
                        bool skip = k >= pe || r[q].strength == 0.0 ||
                                    (d.neglect_self && d.wing_of[k % d.N] == wt);
                        V3 u = skip ? V3{0.0, 0.0, 0.0} : term[q] * r[q].strength;
                        const int kl = wave * PER + q;
                        T[buf][kl][0][lane] = u.x;
                        T[buf][kl][1][lane] = u.y;
                        T[buf][kl][2][lane] = u.z;
                    }
                    // records of the next block, loaded across the barrier
                    const int kn = k0 + FX_BLOCK;
#pragma unroll
                    for (int q = 0; q < PER; ++q) r[q] = scalar_rec(recs, kn + q < pe ? kn + q : pe - 1);
                }
                __syncthreads();
            }
        } else {
            double ay = 0.0, az = 0.0;
            for (long b = 0; b <= nblk; ++b) {
                if (b > 0) {
                    const int buf = (int)((b - 1) & 1);
                    const long base = p_begin + (b - 1) * FX_BLOCK;
                    const int cnt = (int)(p_end - base < FX_BLOCK ? p_end - base : FX_BLOCK);
                    if (cnt == FX_BLOCK) {
                        // a full block: loads issued a chunk ahead of the adds
#pragma unroll
                        for (int q0 = 0; q0 < FX_BLOCK; q0 += FOLD_CHUNK) {
                            double vx[FOLD_CHUNK], vy[FOLD_CHUNK], vz[FOLD_CHUNK];
#pragma unroll
                            for (int i = 0; i < FOLD_CHUNK; ++i) {
                                vx[i] = T[buf][q0 + i][0][lane];
                                vy[i] = T[buf][q0 + i][1][lane];
                                vz[i] = T[buf][q0 + i][2][lane];
                            }
#pragma unroll
                            for (int i = 0; i < FOLD_CHUNK; ++i) {
                                acc = acc + vx[i];
                                ay = ay + vy[i];
                                az = az + vz[i];
                            }
                        }
                    } else {
                        for (int q = 0; q < cnt; ++q) {
                            acc = acc + T[buf][q][0][lane];
                            ay = ay + T[buf][q][1][lane];
                            az = az + T[buf][q][2][lane];
                        }
                    }
                }
                __syncthreads();
            }
            if (t < d.N) store3(d.fixed + 3 * t, V3{acc, ay, az});
        }
    }
}

// Free-wake self-induction of many instances (update_data.rs:125-141): job =
// one instance's 256 wake points over all its panels, one target per lane,
// panel records staged through LDS as in k_targets.
// TL panel records per LDS tile (64, 128 or 256; the autotuner picks it).
template <int MODE, int SYM, int TL, bool NB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void kb_targets(
    const Inst* I, const int2* __restrict__ jobs, int n_jobs, unsigned* head) {
    __shared__ PanelRec tile[TL];
    __shared__ double nbuf[NB ? TL * NEARD : 1];
    __shared__ double ncore[NB ? TL : 1];
    __shared__ NearShared nsh;
    __shared__ int slot;
    NearCtx nc;
    nc.nbuf = nbuf;
    nc.ncore = ncore;
    nc.sh = &nsh;
    for (;;) {
        int j = next_job(head, &slot);
        if (j >= n_jobs) break;
        int2 jb = jobs[j];
        const Inst d = I[jb.x];
        long np = (long)d.P * (d.M + 1);
        long end = d.end_index < np ? d.end_index : np;
        long t = (long)jb.y * 256 + threadIdx.x;
        long tc = t < end ? t : end - 1;
        V3 x = load3(d.pts + 3 * tc);
        if (NB) block_box(d, x, nc);
        V3 acc = sweep<MODE, SYM, TL, NB>(d, tile, x, V3{0.0, 0.0, 0.0}, 0, (long)d.N * d.nz_rows, nc);
        if (t < end) store3(d.vel + 3 * t, load3(d.vel + 3 * t) + acc);
    }
}


// ---- learned warm start: per-strip features and the initial guess ----
// Features of strip i (wing strips [a, b)): for o = -4..4, strip j = i + o
// inside the wing gives [c_j / mean chord, (j - a + 1/2)/(b - a), twist_j / 0.25,
// alpha_j / 0.25, |u_j| / mean |u|] (zeros outside the wing), then the wing's
// aspect ratio / 6, the number of wings / 6 and its place among the wings.  u = felt ctrl-point velocity
// without its spanwise part and alpha its angle of attack (the solver's first
// evaluation with no induced velocity, line_force_model/mod.rs:235-254,
// force_calculations.rs:50-69); twist = chord angle about the span axis
// relative to the wing's first strip; then the wing's centre relative to the
// centre of all strips along and across the mean felt velocity (in 4 mean
// chords); then the solver's default initial guess (the previous step's
// circulation, zero at the first step) of strips i - 4 .. i + 4 over their
// scales.  scale = c |u| / 2, so that circulation = prediction * scale.
constexpr int WS_FEATURES = 64;
constexpr int WS_BASE = 5;

__device__ inline void features_body(const Inst& d, float* X, float* scale) {
    __shared__ float base[1024][WS_BASE];
    __shared__ int wa[1024], wb[1024];
    const int N = d.N;
    const double* geo = d.geo;
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        int w = d.wing_of[i];
        int a = i, b = i + 1;
        while (a > 0 && d.wing_of[a - 1] == w) --a;
        while (b < N && d.wing_of[b] == w) ++b;
        wa[i] = a;
        wb[i] = b;
        V3 felt = load3(d.felt + 3 * i);
        V3 rel = load3(geo + d.gl.rel_g + 3 * i);
        V3 dir = load3(geo + d.gl.dir_g + 3 * i);
        V3 ch = load3(geo + d.gl.chord_g + 3 * i);
        V3 u = felt - project(felt, rel);
        double sp = len(u);
        double alpha = sp > MIN_POSITIVE ? signed_angle_between(ch, u, dir) : 0.0;
        V3 ch0 = load3(geo + d.gl.chord_g + 3 * a);
        double twist = signed_angle_between(ch, ch0, dir);
        double c = geo[d.gl.chord_len + i];
        base[i][0] = (float)c;
        base[i][1] = (float)(((double)(i - a) + 0.5) / (double)(b - a));
        base[i][2] = (float)(twist / 0.25);
        base[i][3] = (float)(alpha / 0.25);
        base[i][4] = (float)sp;
        scale[i] = (float)(0.5 * c * sp);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        int a = wa[i], b = wb[i];
        double cs = 0.0, us = 0.0, span = 0.0;
        for (int j = a; j < b; ++j) {
            cs += base[j][0];
            us += base[j][4];
            span += geo[d.gl.local_len + j];
        }
        double cm = cs / (b - a), um = us / (b - a);
        float* x = X + (long)i * WS_FEATURES;
        for (int o = -4; o <= 4; ++o) {
            int j = i + o;
            float* f = x + (o + 4) * WS_BASE;
            if (j < a || j >= b) {
                for (int q = 0; q < WS_BASE; ++q) f[q] = 0.0f;
                continue;
            }
            f[0] = (float)(base[j][0] / cm);
            f[1] = base[j][1];
            f[2] = base[j][2];
            f[3] = base[j][3];
            f[4] = um > 0.0 ? (float)(base[j][4] / um) : 0.0f;
        }
        x[45] = (float)(span / cm / 6.0);
        x[46] = (float)(d.W / 6.0);
        // the wing's place among the wings: its centre relative to the centre of
        // all control points, along and across the mean felt velocity, in chords
        V3 cw{0.0, 0.0, 0.0}, ca{0.0, 0.0, 0.0}, ua{0.0, 0.0, 0.0};
        for (int j = 0; j < N; ++j) {
            V3 cj = load3(geo + d.gl.ctrl + 3 * j);
            ca = ca + cj;
            ua = ua + load3(d.felt + 3 * j);
            if (j >= a && j < b) cw = cw + cj;
        }
        V3 off = cw / (double)(b - a) - ca / (double)N;
        V3 along = normalize(ua);
        V3 across = normalize(cross(load3(geo + d.gl.dir_g + 3 * i), along));
        x[47] = (float)(dot(off, along) / cm / 4.0);
        x[48] = (float)(dot(off, across) / cm / 4.0);
        // the solver's default initial guess (the previous step's circulation; zero at the
        // first step) of strips i - 4 .. i + 4 of the wing, in units of this step's scale
        for (int o = -4; o <= 4; ++o) {
            int j = i + o;
            double sc = j >= a && j < b ? (double)scale[j] : 0.0;
            x[53 + o] = sc > 0.0 ? (float)(d.prev_gamma[j] / sc) : 0.0f;
        }
        for (int q = 58; q < WS_FEATURES; ++q) x[q] = 0.0f;
    }
}

__global__ void kb_features(const Inst* I, const long* strip_off, float* X, float* scale) {
    const Inst d = I[blockIdx.x];
    long o = strip_off[blockIdx.x];
    features_body(d, X + o * WS_FEATURES, scale + o);
}

// prev_gamma = prediction * scale for every strip of the instance, unless a
// prediction is not finite or exceeds 10 in magnitude (|CL| > 10): then the
// instance keeps its default initial guess and counts as a fallback.
// residual: the prediction is the change from the default guess (the previous
// circulation), in units of the scale.
__global__ void kb_warm_apply(const Inst* I, const long* strip_off, const float* pred, const float* scale,
                              unsigned* fallbacks, int residual) {
    __shared__ int bad;
    const Inst d = I[blockIdx.x];
    long o = strip_off[blockIdx.x];
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < d.N; i += blockDim.x) {
        float y = pred[o + i];
        if (!(fabsf(y) <= 10.0f)) atomicOr(&bad, 1);
    }
    __syncthreads();
    if (bad) {
        if (threadIdx.x == 0) __hip_atomic_fetch_add(fallbacks, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    for (int i = threadIdx.x; i < d.N; i += blockDim.x) {
        const double g = (double)pred[o + i] * (double)scale[o + i];
        d.prev_gamma[i] = residual ? d.prev_gamma[i] + g : g;
    }
}

// Judged warm start: the solver's own first evaluation decides between the
// default initial guess (previous circulation) and the network's guess, per
// instance: the one with the smaller max |Gamma_est - Gamma| is kept.  A guess
// that is not finite or exceeds 10 in magnitude is never taken.  Runs after
// the fixed velocities and factors, before the solve; `taken` counts the
// instances that start from the network's guess.
__global__ __launch_bounds__(1024) void kb_warm_judge(const Inst* I, const long* strip_off, const float* pred,
                                                      const float* scale, unsigned* taken, unsigned* fallbacks,
                                                      int residual) {
    extern __shared__ double sh[];
    const Inst d = I[blockIdx.x];
    const int N = d.N;
    double* G = sh;        // [N]
    double* wm = sh + N;   // [16]
    __shared__ int bad;
    __shared__ double best[2];
    long o = strip_off[blockIdx.x];
    int i = threadIdx.x;
    bool act = i < N;
    int ic = act ? i : N - 1;
    if (i == 0) bad = fabs(d.correction_factor - 1.0) > MIN_POSITIVE ? 1 : 0;
    __syncthreads();
    if (act && !(fabsf(pred[o + i]) <= 10.0f)) atomicOr(&bad, 1);
    __syncthreads();
    if (bad) {
        if (i == 0) __hip_atomic_fetch_add(fallbacks, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    Strip s = load_strip(d, ic);
    const double g = (double)pred[o + ic] * (double)scale[o + ic];
    double cand[2] = {d.prev_gamma[ic], residual ? d.prev_gamma[ic] + g : g};
    for (int c = 0; c < 2; ++c) {
        if (act) G[i] = cand[c];
        __syncthreads();
        double ad = 0.0;
        if (act) {
            V3 u{0.0, 0.0, 0.0};
            double a = 0.0;
            ad = fabs(estimate_circulation(d, s, G, i, u, a) - cand[c]);
        }
        double m = ad;
        for (int off = WAVE / 2; off > 0; off >>= 1) m = fmax_(m, __shfl_xor(m, off, WAVE));
        if ((i % WAVE) == 0) wm[i / WAVE] = m;
        __syncthreads();
        if (i == 0) {
            double mx = 0.0;
            for (int w = 0; w < (int)((blockDim.x + WAVE - 1) / WAVE); ++w) mx = fmax_(mx, wm[w]);
            best[c] = mx;
        }
        __syncthreads();
    }
    if (best[1] < best[0]) {
        if (act) d.prev_gamma[i] = cand[1];
        if (i == 0) __hip_atomic_fetch_add(taken, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__global__ void kb_zero_u32(unsigned* c) {
    if (threadIdx.x < 2) c[threadIdx.x] = 0u;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline unsigned blocks_for(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

hipError_t launch_prepare(const Inst& d, int first_step, hipStream_t st) {
    hipLaunchKernelGGL(k_prepare, dim3(1), dim3(1024), 0, st, d, first_step);
    return hipGetLastError();
}

hipError_t launch_init_wake(const Inst& d, double dt, int use_mean, const double* v3p, hipStream_t st) {
    long n = (long)d.P * (d.M + 1);
    unsigned nb = blocks_for(n, 256);
    double vx = v3p ? v3p[0] : 0.0, vy = v3p ? v3p[1] : 0.0, vz = v3p ? v3p[2] : 0.0;
    int fill = use_mean ? 0 : 1;
    hipLaunchKernelGGL(k_init_wake, dim3(nb), dim3(256), 0, st, d, dt, use_mean, vx, vy, vz, fill);
    return hipGetLastError();
}

hipError_t launch_wake_step(const Inst& d, double dt, hipStream_t st) {
    long n = (long)d.P * (d.M + 1);
    hipLaunchKernelGGL(k_wake_step, dim3(blocks_for(n, 256)), dim3(256), 0, st, d, dt);
    hipLaunchKernelGGL(k_wake_last, dim3(blocks_for(d.P, 256)), dim3(256), 0, st, d);
    return hipGetLastError();
}

hipError_t launch_panels(const Inst& d, const double*, int shift, hipStream_t st) {
    long n = (long)d.N * d.M;
    hipLaunchKernelGGL(k_panels, dim3(blocks_for(n, 256)), dim3(256), 0, st, d, shift);
    return hipGetLastError();
}

hipError_t launch_fixed(const Inst& d, int mode, FoldBuffers* fb, hipStream_t st) {
    int p_begin = d.N;
    int p_end = d.N * d.nz_rows;
    if (p_end < p_begin) p_end = p_begin;
    const double* ctrl = d.geo + d.gl.ctrl;
    const int ld = (d.N + WAVE - 1) / WAVE * WAVE;
    const long chunk = fb && fb->cap_targets >= d.N ? fb->cap / (3L * ld) / (4 * TERM_PANELS) * (4 * TERM_PANELS) : 0;
    if (chunk > 0 && p_end > p_begin) {
        int first = 1;
        for (long b = p_begin; b < p_end; b += chunk) {
            long cnt = p_end - b < chunk ? p_end - b : chunk;
            // panels per wave: 16, or fewer while the chunk gives less than ~16 waves per CU
            int tp = TERM_PANELS;
            while (tp > 1 && (cnt + tp - 1) / tp * (ld / WAVE) < 4096) tp >>= 1;
            dim3 grid(blocks_for(cnt, 4 * tp), (unsigned)(ld / WAVE));
            if (mode == MODE_FAST)
                hipLaunchKernelGGL(k_terms<MODE_FAST>, grid, dim3(256), 0, st, d, ctrl, d.N, b, cnt, d.neglect_self,
                                   fb->terms, ld, tp);
            else if (mode == MODE_FP32_KAHAN)
                hipLaunchKernelGGL(k_terms<MODE_FP32_KAHAN>, grid, dim3(256), 0, st, d, ctrl, d.N, b, cnt,
                                   d.neglect_self, fb->terms, ld, tp);
            else
                hipLaunchKernelGGL(k_terms<MODE_EXACT>, grid, dim3(256), 0, st, d, ctrl, d.N, b, cnt, d.neglect_self,
                                   fb->terms, ld, tp);
            bool last = b + cnt >= p_end;
            hipLaunchKernelGGL(k_fold_terms, dim3(3 * ((d.N + FOLD_CH - 1) / FOLD_CH)), dim3(FOLD_THREADS), 0, st,
                               fb->terms, ld, d.N, cnt, first, fb->acc, last ? d.fixed : nullptr);
            first = 0;
        }
        return hipGetLastError();
    }
    unsigned nb = blocks_for(d.N, 4);
    if (mode == MODE_FAST)
        hipLaunchKernelGGL(k_fold<MODE_FAST>, dim3(nb), dim3(256), 0, st, d, ctrl, d.N, p_begin, p_end,
                           d.neglect_self, d.fixed);
    else if (mode == MODE_FP32_KAHAN)
        hipLaunchKernelGGL(k_fold<MODE_FP32_KAHAN>, dim3(nb), dim3(256), 0, st, d, ctrl, d.N, p_begin, p_end,
                           d.neglect_self, d.fixed);
    else
        hipLaunchKernelGGL(k_fold<MODE_EXACT>, dim3(nb), dim3(256), 0, st, d, ctrl, d.N, p_begin, p_end,
                           d.neglect_self, d.fixed);
    return hipGetLastError();
}

hipError_t launch_horseshoe_factors(const Inst& d, hipStream_t st) {
    long n = (long)d.N * d.N;
    hipLaunchKernelGGL(k_horseshoe_factors, dim3(blocks_for(n, 256)), dim3(256), 0, st, d);
    return hipGetLastError();
}

static long lds_limit_bytes() {
    static long cached = -1;
    if (cached < 0) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess)
            cached = v;
        else
            cached = 65536;
    }
    return cached;
}

// One-workgroup variants: 256 threads x 72 entries (N <= 128), 512 x 48.
static hipError_t allow_shared(const void* fn, size_t bytes) {
    if (bytes <= 65536) return hipSuccess;
    return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

// One-workgroup variants (LinrV): a variant applies when the rows and columns
// fit its threads' registers and its shared memory fits the device.
template <int V>
static bool linr_ok(int N) {
    return linr_fits<V>(N) && (long)sizeof(double) * linr_shared_doubles<V>(N) <= lds_limit_bytes();
}
static bool linr_variant_ok(int v, int N) {
    switch (v) {
        case 1: return linr_ok<1>(N);
        case 2: return linr_ok<2>(N);
        case 3: return linr_ok<3>(N);
        case 4: return linr_ok<4>(N);
        default: return false;
    }
}
static size_t linr_variant_shared(int v, int N) {
    long n = v == 1 ? linr_shared_doubles<1>(N) : v == 2 ? linr_shared_doubles<2>(N)
           : v == 3 ? linr_shared_doubles<3>(N) : linr_shared_doubles<4>(N);
    return sizeof(double) * (size_t)n;
}
static int linr_variant(int N) {
    for (int v = 1; v <= 4; ++v)
        if (linr_variant_ok(v, N)) return v;
    return 0;
}
template <int V>
static hipError_t launch_linr(const Inst& d, int as_start, size_t shm, hipStream_t st) {
    hipError_t e = allow_shared((const void*)k_linearized<V>, shm);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_linearized<V>), dim3(1), dim3(LinrV<V>::T), shm, st, d, as_start);
    return hipGetLastError();
}
template <int V>
static hipError_t launch_linr_batch(const BatchDev& b, int as_start, size_t shm, hipStream_t st) {
    hipError_t e = allow_shared((const void*)kb_linearized<V>, shm);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((kb_linearized<V>), dim3(b.n), dim3(LinrV<V>::T), shm, st, b.inst, as_start);
    return hipGetLastError();
}
template <int V>
static hipError_t launch_linr_dense(const double* A, const double* b, int N, double* x, int* status, size_t shm,
                                    hipStream_t st) {
    hipError_t e = allow_shared((const void*)k_gauss_dense<V>, shm);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_gauss_dense<V>), dim3(1), dim3(LinrV<V>::T), shm, st, A, b, N, x, status);
    return hipGetLastError();
}

bool linearized_in_lds(int N) { return linr_variant(N) != 0; }


// Panel width of the blocked form: 16 columns (the panel kernel's registers per row).
static int linearized_panel_width(int) { return 16; }

// The blocked factorization of d.lin ([N + 1][N] column-major, N <= 1024): panels
// of 16 columns and their trailing updates.
static hipError_t launch_lin_blocked_factor(const Inst& d, hipStream_t st) {
    const int N = d.N;
    const int b = linearized_panel_width(N);
    for (int k0 = 0; k0 < N; k0 += b) {
        const int bb = N - k0 < b ? N - k0 : b;
        const size_t shm = sizeof(double) * (size_t)(2 * LINP_W * LINP_SLOT + 2 * LINP_SLOT);
        hipLaunchKernelGGL(k_lin_panel, dim3(1), dim3(LINP_T), shm, st, d, k0, bb);
        const unsigned cols = (unsigned)(N + 1 - k0 - bb);
        hipLaunchKernelGGL(k_lin_trailing, dim3(cols), dim3(LIN_TRAIL_THREADS), 0, st, d, k0, bb);
    }
    return hipGetLastError();
}

hipError_t launch_linearized(const Inst& d, int as_start, hipStream_t st) {
    const int N = d.N;
    if (N < 1 || N > 1024) return hipErrorInvalidValue;
    const int v = linr_variant(N);
    if (v) {
        const size_t shm = linr_variant_shared(v, N);
        switch (v) {
            case 1: return launch_linr<1>(d, as_start, shm, st);
            case 2: return launch_linr<2>(d, as_start, shm, st);
            case 3: return launch_linr<3>(d, as_start, shm, st);
            default: return launch_linr<4>(d, as_start, shm, st);
        }
    }
    if (!d.lin) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_lin_rows, dim3(blocks_for(N, 256)), dim3(256), 0, st, d);
    hipLaunchKernelGGL(k_lin_assemble, dim3(blocks_for((long)N * (N + 1), 256)), dim3(256), 0, st, d);
    hipError_t eb = launch_lin_blocked_factor(d, st);
    if (eb != hipSuccess) return eb;
    const size_t fin = sizeof(double) * (size_t)(linr_even(N + 8) + linr_even(N) + 2 * linr_even(N + 64) + 16);
    hipError_t ef = allow_shared((const void*)k_lin_finish, fin);
    if (ef != hipSuccess) return ef;
    hipLaunchKernelGGL(k_lin_finish, dim3(1), dim3(1024), fin, st, d, as_start);
    return hipGetLastError();
}

hipError_t launch_gauss_dense(const double* A, const double* b, int N, double* x, int* status, int variant, hipStream_t st) {
    if (!linr_variant_ok(variant, N)) return hipErrorInvalidValue;
    const size_t shm = linr_variant_shared(variant, N);
    switch (variant) {
        case 1: return launch_linr_dense<1>(A, b, N, x, status, shm, st);
        case 2: return launch_linr_dense<2>(A, b, N, x, status, shm, st);
        case 3: return launch_linr_dense<3>(A, b, N, x, status, shm, st);
        default: return launch_linr_dense<4>(A, b, N, x, status, shm, st);
    }
}

hipError_t launch_gauss_blocked(double* lin, double* scal, int N, double* x, int* status, hipStream_t st) {
    if (N < 1 || N > 1024) return hipErrorInvalidValue;
    Inst d{};
    d.N = N;
    d.lin = lin;
    d.scal = scal;
    hipError_t e = launch_lin_blocked_factor(d, st);
    if (e != hipSuccess) return e;
    const size_t fin = sizeof(double) * (size_t)(linr_even(N + 8) + linr_even(N) + 2 * linr_even(N + 64) + 16);
    e = allow_shared((const void*)k_gauss_blocked_finish, fin);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_gauss_blocked_finish, dim3(1), dim3(1024), fin, st, d, x, status);
    return hipGetLastError();
}

// Doubles of d.lin for the blocked form: matrix, row data, exchange sources.
long linearized_global_doubles(int N) { return (long)N * (N + 1) + 8L * N + N; }

hipError_t launch_factors(const Inst& d, hipStream_t st) {
    long n = (long)d.N * d.N;
    hipLaunchKernelGGL(k_factors, dim3(blocks_for(n, 256)), dim3(256), 0, st, d);
    return hipGetLastError();
}

// Shared memory of the one-workgroup fixed-point solve (solve_body).
size_t solve_shared_bytes(int N, int W) {
    return sizeof(double) * (2 * (size_t)N + 16 + 4) + sizeof(FoilParams) * (size_t)W;
}

hipError_t launch_solve(const Inst& d, hipStream_t st) {
    int threads = ((d.N + WAVE - 1) / WAVE) * WAVE;
    size_t shm = solve_shared_bytes(d.N, d.W);
    if (threads <= 512) hipLaunchKernelGGL(k_solve<512>, dim3(1), dim3(threads), shm, st, d);
    else hipLaunchKernelGGL(k_solve<1024>, dim3(1), dim3(threads), shm, st, d);
    return hipGetLastError();
}

hipError_t launch_set_row0(const Inst&, hipStream_t) { return hipSuccess; }

hipError_t launch_copy3(const double* src, double* dst, long n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_copy3, dim3(blocks_for(3 * n, 256)), dim3(256), 0, st, src, dst, n);
    return hipGetLastError();
}

__global__ void k_copy_doubles(const double* __restrict__ src, double* __restrict__ dst, long count) {
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < count) dst[q] = src[q];
}

hipError_t launch_copy_doubles(const double* src, double* dst, long count, hipStream_t st) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_copy_doubles, dim3(blocks_for(count, 256)), dim3(256), 0, st, src, dst, count);
    return hipGetLastError();
}

hipError_t launch_velocity_copy(const Inst& d, hipStream_t st) {
    long n = (long)d.P * (d.M + 1);
    hipLaunchKernelGGL(k_copy3, dim3(blocks_for(3 * n, 256)), dim3(256), 0, st, d.fs + 3L * d.N, d.vel, n);
    return hipGetLastError();
}

template <int MODE, int SYM, bool NB>
static int mc_slots_for() {
    int per_cu = 0, cus = 0, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_targets_mc<MODE, SYM, NB>, 256, 0) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    return per_cu * cus;
}

template <int MODE, int SYM>
static hipError_t launch_targets_t(const Inst& d, const double* targets, long n_targets, long n_panels, double* out,
                                   int add, McBuffers* mc, hipStream_t st, const int* tmap, long map_jobs) {
    long n_jobs = tmap ? map_jobs : (n_targets + 255) / 256;
    // near-field data staged in shared memory when the panels have ring-edge geometry
    const bool nb = d.segs && !d.corners;
    int slots = 0;
    if (mc) {
        if (mc->slots[MODE][SYM][nb] == 0)
            mc->slots[MODE][SYM][nb] = nb ? mc_slots_for<MODE, SYM, true>() : mc_slots_for<MODE, SYM, false>();
        slots = mc->slots[MODE][SYM][nb];
    }
    if (mc && slots > 0 && n_jobs > slots && mc->cap_jobs >= n_jobs && n_panels > 0) {
        McPlan pl;
        pl.L = n_panels;
        pl.n_jobs = n_jobs;
        pl.C = (n_jobs * n_panels + slots - 1) / slots;
        pl.partial = mc->partial;
        pl.flags = mc->flags;
        pl.epoch = ++mc->epoch;
        pl.spin_limit = 200000;
        pl.tmap = tmap;
        long grid = (n_jobs * n_panels + pl.C - 1) / pl.C;
        if (nb)
            hipLaunchKernelGGL((k_targets_mc<MODE, SYM, true>), dim3((unsigned)grid), dim3(256), 0, st, d, targets,
                               n_targets, out, add, pl);
        else
            hipLaunchKernelGGL((k_targets_mc<MODE, SYM, false>), dim3((unsigned)grid), dim3(256), 0, st, d, targets,
                               n_targets, out, add, pl);
    } else {
        if (nb)
            hipLaunchKernelGGL((k_targets<MODE, SYM, true>), dim3((unsigned)n_jobs), dim3(256), 0, st, d, targets,
                               n_targets, n_panels, out, add, tmap);
        else
            hipLaunchKernelGGL((k_targets<MODE, SYM, false>), dim3((unsigned)n_jobs), dim3(256), 0, st, d, targets,
                               n_targets, n_panels, out, add, tmap);
    }
    return hipGetLastError();
}

hipError_t launch_targets(const Inst& d, const double* targets, long n_targets, long n_panels, int mode, double* out,
                          int add, McBuffers* mc, hipStream_t st, const int* tmap, long map_jobs) {
    if (n_targets <= 0) return hipSuccess;
    bool sym = d.symmetry != 0;
    if (mode == MODE_FAST)
        return sym ? launch_targets_t<MODE_FAST, 1>(d, targets, n_targets, n_panels, out, add, mc, st, tmap, map_jobs)
                   : launch_targets_t<MODE_FAST, 0>(d, targets, n_targets, n_panels, out, add, mc, st, tmap, map_jobs);
    if (mode == MODE_FP32_KAHAN)  // one workgroup per 256 targets (the Kahan state is not carried across pieces)
        return sym ? launch_targets_t<MODE_FP32_KAHAN, 1>(d, targets, n_targets, n_panels, out, add, nullptr, st, tmap,
                                                          map_jobs)
                   : launch_targets_t<MODE_FP32_KAHAN, 0>(d, targets, n_targets, n_panels, out, add, nullptr, st, tmap,
                                                          map_jobs);
    return sym ? launch_targets_t<MODE_EXACT, 1>(d, targets, n_targets, n_panels, out, add, mc, st, tmap, map_jobs)
               : launch_targets_t<MODE_EXACT, 0>(d, targets, n_targets, n_panels, out, add, mc, st, tmap, map_jobs);
}

hipError_t launch_self_induction(const Inst& d, int mode, McBuffers* mc, hipStream_t st) {
    long np = (long)d.P * (d.M + 1);
    long end = d.end_index < np ? d.end_index : np;
    long n_panels = (long)d.N * d.nz_rows;
    if (end <= 0 || n_panels <= 0) return hipSuccess;
    const bool map = mc && mc->tmap && mc->tmap_jobs > 0;
    return launch_targets(d, d.pts, end, n_panels, mode, d.vel, 1, mc, st, map ? mc->tmap : nullptr,
                          map ? mc->tmap_jobs : 0);
}

hipError_t launch_forces(const Inst& d, double dt, hipStream_t st) {
    int threads = ((d.N + WAVE - 1) / WAVE) * WAVE;
    hipLaunchKernelGGL(k_forces, dim3(1), dim3(threads), 0, st, d, dt);
    return hipGetLastError();
}

hipError_t launch_flow_update(const Inst&, hipStream_t) { return hipSuccess; }

hipError_t launch_induced_targets(const Inst& d, const double* targets, long n_targets, long n_panels, int mode,
                                  double* out, McBuffers* mc, hipStream_t st) {
    if (n_targets <= 0) return hipSuccess;
    if (n_panels <= 0) return hipMemsetAsync(out, 0, sizeof(double) * 3 * n_targets, st);
    if (n_targets < 2048) {
        unsigned nb = blocks_for(n_targets, 4);
        if (mode == MODE_FAST)
            hipLaunchKernelGGL(k_fold<MODE_FAST>, dim3(nb), dim3(256), 0, st, d, targets, (int)n_targets, 0,
                               (int)n_panels, 0, out);
        else if (mode == MODE_FP32_KAHAN)
            hipLaunchKernelGGL(k_fold<MODE_FP32_KAHAN>, dim3(nb), dim3(256), 0, st, d, targets, (int)n_targets, 0,
                               (int)n_panels, 0, out);
        else
            hipLaunchKernelGGL(k_fold<MODE_EXACT>, dim3(nb), dim3(256), 0, st, d, targets, (int)n_targets, 0,
                               (int)n_panels, 0, out);
        return hipGetLastError();
    }
    return launch_targets(d, targets, n_targets, n_panels, mode, out, 0, mc, st);
}

hipError_t launch_count_pairs(const Inst& d, const double* targets, long n_targets, long n_panels,
                              double* partial, long* n_blocks, hipStream_t st) {
    *n_blocks = (long)blocks_for(n_targets, 256);
    if (n_targets <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_count_pairs, dim3((unsigned)*n_blocks), dim3(256), 0, st, d, targets, n_targets, n_panels,
                       partial);
    return hipGetLastError();
}

hipError_t launch_wind_points(const Inst& d, const WindField& w, hipStream_t st) {
    long n = (long)d.P * (d.M + 1);
    if (d.quasi || n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_wind_points, dim3(blocks_for(n, 256)), dim3(256), 0, st, d, w);
    return hipGetLastError();
}

hipError_t launch_batch_wind(const WindJob* jobs, const long* blk_off, int n, long blocks, hipStream_t st) {
    if (blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(kb_wind, dim3((unsigned)blocks), dim3(256), 0, st, jobs, blk_off, n);
    return hipGetLastError();
}

hipError_t launch_libm(int fn, const double* x, const double* y, double* out, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_libm, dim3(blocks_for(n, 256)), dim3(256), 0, st, fn, x, y, out, n);
    return hipGetLastError();
}


// ---- batched launchers ----
hipError_t launch_batch_warm_start(const BatchDev& b, sb_mlp* mlp, int apply, hipStream_t st) {
    hipLaunchKernelGGL(kb_features, dim3(b.n), dim3(b.max_threads), 0, st, b.inst, b.strip_off, b.feat, b.scale);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !apply) return e;
    hipLaunchKernelGGL(kb_zero_u32, dim3(1), dim3(64), 0, st, b.fallbacks);
    e = mlp_forward_on(mlp, b.feat, b.n_strips, b.pred, st);
    if (e != hipSuccess) return e;
    if (apply == 1)
        hipLaunchKernelGGL(kb_warm_apply, dim3(b.n), dim3(256), 0, st, b.inst, b.strip_off, b.pred, b.scale,
                           b.fallbacks, b.warm_residual);
    return hipGetLastError();
}

hipError_t launch_batch_warm_judge(const BatchDev& b, hipStream_t st) {
    size_t shm = sizeof(double) * ((size_t)b.max_threads + 16);
    hipLaunchKernelGGL(kb_warm_judge, dim3(b.n), dim3(b.max_threads), shm, st, b.inst, b.strip_off, b.pred, b.scale,
                       b.fallbacks + 1, b.fallbacks, b.warm_residual);
    return hipGetLastError();
}

hipError_t launch_batch_prepare(const BatchDev& b, int first_step, hipStream_t st) {
    hipLaunchKernelGGL(kb_prepare, dim3(b.n), dim3(256), 0, st, b.inst, first_step);
    return hipGetLastError();
}

static const long* boff(const BatchDev& b, int kind) { return b.off + (long)kind * (b.n + 1); }

hipError_t launch_batch_init_wake(const BatchDev& b, hipStream_t st) {
    if (b.blocks[BK_POINTS] > 0)
        hipLaunchKernelGGL(kb_init_wake, dim3((unsigned)b.blocks[BK_POINTS]), dim3(256), 0, st, b.inst,
                           boff(b, BK_POINTS), b.n);
    return hipGetLastError();
}

hipError_t launch_batch_wake_step(const BatchDev& b, hipStream_t st) {
    hipLaunchKernelGGL(kb_wake_step, dim3((unsigned)b.blocks[BK_POINTS]), dim3(256), 0, st, b.inst,
                       boff(b, BK_POINTS), b.n);
    hipLaunchKernelGGL(kb_wake_last, dim3((unsigned)b.blocks[BK_SPAN]), dim3(256), 0, st, b.inst, boff(b, BK_SPAN),
                       b.n);
    return hipGetLastError();
}

hipError_t launch_batch_panels(const BatchDev& b, int shift, hipStream_t st) {
    hipLaunchKernelGGL(kb_panels, dim3((unsigned)b.blocks[BK_PANELS]), dim3(256), 0, st, b.inst,
                       boff(b, BK_PANELS), b.n, shift);
    return hipGetLastError();
}

template <int MODE, int SYM, int PER>
static void launch_kb_fixed(const BatchDev& b, hipStream_t st) {
    hipLaunchKernelGGL((kb_fixed<MODE, SYM, PER>), dim3(b.fixed_grid), dim3(COOP_WAVES * WAVE), 0, st, b.inst,
                       b.fixed_jobs, b.n_fixed_jobs, b.counters);
}

template <int MODE, int SYM>
static void launch_kb_fixed_per(const BatchDev& b, hipStream_t st) {
    if (b.fixed_per == 1) launch_kb_fixed<MODE, SYM, 1>(b, st);
    else if (b.fixed_per == 2) launch_kb_fixed<MODE, SYM, 2>(b, st);
    else launch_kb_fixed<MODE, SYM, 3>(b, st);
}

hipError_t launch_batch_fixed(const BatchDev& b, int mode, hipStream_t st) {
    if (b.n_fixed_jobs <= 0) return hipSuccess;
    if (mode == MODE_FAST) b.sym ? launch_kb_fixed_per<MODE_FAST, 1>(b, st) : launch_kb_fixed_per<MODE_FAST, 0>(b, st);
    else b.sym ? launch_kb_fixed_per<MODE_EXACT, 1>(b, st) : launch_kb_fixed_per<MODE_EXACT, 0>(b, st);
    hipLaunchKernelGGL(kb_zero_counters, dim3(1), dim3(64), 0, st, b.counters);
    return hipGetLastError();
}

hipError_t launch_batch_factors(const BatchDev& b, hipStream_t st) {
    hipLaunchKernelGGL(kb_factors, dim3((unsigned)b.blocks[BK_FACTORS]), dim3(256), 0, st, b.inst,
                       boff(b, BK_FACTORS), b.n);
    return hipGetLastError();
}

hipError_t launch_batch_horseshoe_factors(const BatchDev& b, hipStream_t st) {
    hipLaunchKernelGGL(kb_horseshoe_factors, dim3((unsigned)b.blocks[BK_FACTORS]), dim3(256), 0, st, b.inst,
                       boff(b, BK_FACTORS), b.n);
    return hipGetLastError();
}

int batch_linearized_variant(int max_n) {
    const int v = linr_variant(max_n);
    return v;
}

size_t batch_linearized_shared(const int* ns, int n, int variant) {
    size_t m = 0;
    for (int i = 0; i < n; ++i) {
        if (!linr_variant_ok(variant, ns[i])) return 0;
        const size_t x = linr_variant_shared(variant, ns[i]);
        m = x > m ? x : m;
    }
    return m;
}

hipError_t launch_batch_linearized(const BatchDev& b, int variant, size_t shm, int as_start, hipStream_t st) {
    switch (variant) {
        case 1: return launch_linr_batch<1>(b, as_start, shm, st);
        case 2: return launch_linr_batch<2>(b, as_start, shm, st);
        case 3: return launch_linr_batch<3>(b, as_start, shm, st);
        case 4: return launch_linr_batch<4>(b, as_start, shm, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_batch_take_singular(const BatchDev& b, double* flags, hipStream_t st) {
    hipLaunchKernelGGL(kb_take_singular, dim3(blocks_for(b.n, 256)), dim3(256), 0, st, b.inst, b.n, flags);
    return hipGetLastError();
}

hipError_t launch_batch_solve(const BatchDev& b, hipStream_t st) {
    if (b.max_threads <= 512) hipLaunchKernelGGL(kb_solve<512>, dim3(b.n), dim3(b.max_threads), b.solve_shm, st, b.inst);
    else hipLaunchKernelGGL(kb_solve<1024>, dim3(b.n), dim3(b.max_threads), b.solve_shm, st, b.inst);
    return hipGetLastError();
}

hipError_t launch_batch_velocity_copy(const BatchDev& b, hipStream_t st) {
    hipLaunchKernelGGL(kb_velocity_copy, dim3((unsigned)b.blocks[BK_VEL]), dim3(256), 0, st, b.inst,
                       boff(b, BK_VEL), b.n);
    return hipGetLastError();
}

template <int MODE, int SYM, int TL>
static void launch_kb_targets(const BatchDev& b, hipStream_t st) {
    if (b.stage_near)
        hipLaunchKernelGGL((kb_targets<MODE, SYM, TL, true>), dim3(b.target_grid), dim3(256), 0, st, b.inst,
                           b.target_jobs, b.n_target_jobs, b.counters + 1);
    else
        hipLaunchKernelGGL((kb_targets<MODE, SYM, TL, false>), dim3(b.target_grid), dim3(256), 0, st, b.inst,
                           b.target_jobs, b.n_target_jobs, b.counters + 1);
}

template <int MODE, int SYM>
static void launch_kb_targets_tile(const BatchDev& b, hipStream_t st) {
    if (b.targets_tile == 64) launch_kb_targets<MODE, SYM, 64>(b, st);
    else if (b.targets_tile == 128) launch_kb_targets<MODE, SYM, 128>(b, st);
    else launch_kb_targets<MODE, SYM, 256>(b, st);
}

hipError_t launch_batch_self_induction(const BatchDev& b, int mode, hipStream_t st) {
    if (b.n_target_jobs <= 0) return hipSuccess;
    if (mode == MODE_FAST)
        b.sym ? launch_kb_targets_tile<MODE_FAST, 1>(b, st) : launch_kb_targets_tile<MODE_FAST, 0>(b, st);
    else b.sym ? launch_kb_targets_tile<MODE_EXACT, 1>(b, st) : launch_kb_targets_tile<MODE_EXACT, 0>(b, st);
    hipLaunchKernelGGL(kb_zero_counters, dim3(1), dim3(64), 0, st, b.counters);
    return hipGetLastError();
}

hipError_t launch_batch_forces(const BatchDev& b, hipStream_t st) {
    hipLaunchKernelGGL(kb_forces, dim3(b.n), dim3(b.max_threads), 0, st, b.inst);
    return hipGetLastError();
}

template <class K>
static int resident_groups(K kernel, int threads) {
    int per_cu = 0, cus = 0, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    return per_cu * cus;
}

template <int MODE, int SYM>
static int fixed_groups(int per) {
    const int th = COOP_WAVES * WAVE;
    if (per == 1) return resident_groups(kb_fixed<MODE, SYM, 1>, th);
    if (per == 2) return resident_groups(kb_fixed<MODE, SYM, 2>, th);
    return resident_groups(kb_fixed<MODE, SYM, 3>, th);
}

template <int MODE, int SYM, bool NB>
static int target_groups(int tile) {
    if (tile == 64) return resident_groups(kb_targets<MODE, SYM, 64, NB>, 256);
    if (tile == 128) return resident_groups(kb_targets<MODE, SYM, 128, NB>, 256);
    return resident_groups(kb_targets<MODE, SYM, 256, NB>, 256);
}

template <int MODE, int SYM>
static int target_groups_staged(int tile, int stage_near) {
    return stage_near ? target_groups<MODE, SYM, true>(tile) : target_groups<MODE, SYM, false>(tile);
}

int batch_resident_groups(int which, int mode, int sym, int variant, int stage_near) {
    if (which == 0) {
        if (mode == MODE_FAST) return sym ? fixed_groups<MODE_FAST, 1>(variant) : fixed_groups<MODE_FAST, 0>(variant);
        return sym ? fixed_groups<MODE_EXACT, 1>(variant) : fixed_groups<MODE_EXACT, 0>(variant);
    }
    if (mode == MODE_FAST)
        return sym ? target_groups_staged<MODE_FAST, 1>(variant, stage_near)
                   : target_groups_staged<MODE_FAST, 0>(variant, stage_near);
    return sym ? target_groups_staged<MODE_EXACT, 1>(variant, stage_near)
               : target_groups_staged<MODE_EXACT, 0>(variant, stage_near);
}

}  // namespace sb
