/*
 * oracle.c -- FP64 CPU restatement of gbkfit's model-evaluation hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Written from the semantics of the
 * reference's header-only core; each function cites the reference file:line
 * it follows.  Reference paths are relative to src/gbkfit/driver/native/:
 *   core/ = gbkfit/include/gbkfit/
 */
#include "oracle.h"

#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#define OR_PI 3.14159265358979323846
#define OR_D2R (OR_PI / 180.0)
#define LWM 5.0  /* LINE_WIDTH_MULTIPLIER, core/constants.hpp:5-6 */

/* ------------------------------------------------------------------------ */
/* 1-D profiles: core/math/math.hpp:126-498                                 */
/* ------------------------------------------------------------------------ */

static double sgn(double x) { return (double)((0.0 < x) - (x < 0.0)); }

static double wrap_angle(double a)                 /* math.hpp:28-36 */
{
    while (a < -OR_PI) a += 2 * OR_PI;
    while (a > +OR_PI) a -= 2 * OR_PI;
    return a;
}

static double f_uniform(double x, double a, double b, double c)  /* :126 */
{ return (x >= b && x <= c) ? a : 0.0; }
static double f_uniform_wm(double x, double a, double b, double c)
{ return f_uniform(x, a, b - c, b + c); }
static double f_expon(double x, double a, double b, double c)    /* :219 */
{ return a * exp(-fabs(x - b) / c); }
static double f_gauss(double x, double a, double b, double c)    /* :264 */
{ return a * exp(-(x - b) * (x - b) / (2 * c * c)); }
static double f_ggauss(double x, double a, double b, double c, double d)
{ return a * exp(-pow(fabs(x - b) / c, d)); }                     /* :310 */
static double f_lorentz(double x, double a, double b, double c)  /* :358 */
{ return a * c * c / ((x - b) * (x - b) + c * c); }
static double f_moffat(double x, double a, double b, double c, double d)
{ return a / pow(1 + ((x - b) / c) * ((x - b) / c), d); }         /* :403 */
static double f_sech2(double x, double a, double b, double c)    /* :455 */
{ double s = 1 / cosh((x - b) / c); return a * s * s; }

static double cdf_uniform(double x, double b, double c)          /* :132 */
{ return x < b ? 0 : (x > c ? 1 : (x - b) / (c - b)); }
static double cdf_expon(double x, double b, double c)            /* :232 */
{ return 0.5 + 0.5 * sgn(x - b) * (1 - exp(-fabs(x - b) / c)); }
static double cdf_gauss(double x, double b, double c)            /* :276 */
{ return 0.5 * (1 + erf((x - b) / (c * sqrt(2.0)))); }
static double cdf_lorentz(double x, double b, double c)          /* :370 */
{ return 0.5 + atan((x - b) / c) / OR_PI; }
static double cdf_sech2(double x, double b, double c)            /* :467 */
{ return 0.5 * (1 + tanh((x - b) / c)); }
/* ggauss and moffat cdfs are stubs returning 0 in the reference (:322,:417) */

static double pdf_uniform(double x, double b, double c)
{ return f_uniform(x, 1 / (c - b), b, c); }
static double pdf_expon(double x, double b, double c)
{ return f_expon(x, 1 / (2 * c), b, c); }
static double pdf_gauss(double x, double b, double c)
{ return f_gauss(x, 1 / (c * sqrt(2 * OR_PI)), b, c); }
static double pdf_ggauss(double x, double b, double c, double d)
{ return f_ggauss(x, d / (2 * c * tgamma(1 / d)), b, c, d); }
static double pdf_lorentz(double x, double b, double c)
{ return f_lorentz(x, 1 / (OR_PI * c), b, c); }
static double pdf_moffat(double x, double b, double c, double d)
{ (void)x; (void)b; (void)c; (void)d; return 0.0; }  /* stub, :426-434 */
static double pdf_sech2(double x, double b, double c)
{ return f_sech2(x, 1 / (2 * c), b, c); }

/* Truncated pdf: math.hpp:108-114.  cdfs of ggauss/moffat are 0, so their
 * truncated pdfs divide by zero exactly as the reference does. */
static double trunc_norm(int kind, double lo, double hi, double b, double c,
                         double d)
{
    (void)d;
    switch (kind) {
    case 1: return cdf_uniform(hi, b - c, b + c) - cdf_uniform(lo, b - c, b + c);
    case 2: return cdf_expon(hi, b, c) - cdf_expon(lo, b, c);
    case 3: return cdf_gauss(hi, b, c) - cdf_gauss(lo, b, c);
    case 5: return cdf_lorentz(hi, b, c) - cdf_lorentz(lo, b, c);
    case 7: return cdf_sech2(hi, b, c) - cdf_sech2(lo, b, c);
    default: return 0.0;
    }
}

static double lerp_nw(double r, int idx, const double* xs, const double* ys,
                      int offset)                                /* :500-517 */
{
    const double x1 = xs[idx - 1], x2 = xs[idx];
    const double y1 = ys[offset + idx - 1], y2 = ys[offset + idx];
    const double t = (r - x1) / (x2 - x1);
    return y1 + t * (y2 - y1);
}

/* ------------------------------------------------------------------------ */
/* Polar / height traits: core/gmodel/traits.hpp                            */
/* ------------------------------------------------------------------------ */

typedef struct {
    int rnidx, nrnodes;
    const double* rnodes;
    double x, y, r, theta, incl;
} pctx;

static void rot_lh_z(double* ox, double* oy, double x, double y, double t)
{   /* math.hpp:62-71 */
    const double s = sin(t), c = cos(t);
    *ox = -x * s + y * c;
    *oy = -x * c - y * s;
}

static void rot_lh_x(double* oy, double* oz, double y, double z, double t)
{   /* math.hpp:44-51 */
    const double s = sin(t), c = cos(t);
    *oy = y * c + z * s;
    *oz = -y * s + z * c;
}

/* Blob mixtures: traits.hpp:199-246 (np = 6 or 7 parameters per blob). */
static double mixture(int fun, int np, const pctx* p, const double* consts,
                      const double* params)
{
    const int nb = (int)rint(consts[0]);
    double res = 0;
    for (int i = 0; i < nb; ++i) {
        const double br = params[i];
        const double bt = params[nb + i] * OR_D2R;
        const double ba = params[2 * nb + i];
        const double bs = params[3 * nb + i];
        const double bb = np == 7 ? params[4 * nb + i] : 0;
        const double bq = params[(np == 7 ? 5 : 4) * nb + i];
        const double bp = params[(np == 7 ? 6 : 5) * nb + i] * OR_D2R;
        double xn = p->r * cos(p->theta) - br * cos(bt);
        double yn = p->r * sin(p->theta) - br * sin(bt);
        rot_lh_z(&xn, &yn, xn, yn, bt + bp);
        const double rn = sqrt(xn * xn + (yn * yn) / (bq * bq));
        switch (fun) {
        case 2: res += f_expon(rn, ba, 0, bs); break;
        case 3: res += f_gauss(rn, ba, 0, bs); break;
        case 4: res += f_ggauss(rn, ba, 0, bs, bb); break;
        case 6: res += f_moffat(rn, ba, 0, bs, bb); break;
        default: break;
        }
    }
    return res;
}

static double nw_harmonic(const pctx* p, const double* consts,
                          const double* params)              /* :285-295 */
{
    const int k = (int)rint(consts[0]);
    const double a = lerp_nw(p->r, p->rnidx, p->rnodes, params, 0);
    const double ph = k ? lerp_nw(p->r, p->rnidx, p->rnodes, params,
                                  p->nrnodes) * OR_D2R : 0;
    return a * cos(k * (p->theta - ph));
}

static double nw_distortion(const pctx* p, const double* params) /* :297 */
{
    const double a = lerp_nw(p->r, p->rnidx, p->rnodes, params, 0);
    const double ph = lerp_nw(p->r, p->rnidx, p->rnodes, params, p->nrnodes)
                      * OR_D2R;
    const double s = lerp_nw(p->r, p->rnidx, p->rnodes, params,
                             2 * p->nrnodes);
    const double t = wrap_angle(p->theta - ph);
    return a * exp(-(t * t * p->r * p->r) / (2 * s * s));
}

/* Density (and opacity, uid + 1000) polar traits: traits.hpp:1242-1331. */
static double rp_trait(int uid, const double* c, const double* q,
                       const pctx* p)
{
    const int u = uid >= 1000 ? uid - 1000 : uid;
    switch (u) {
    case 1: return q[0];
    case 2: return f_expon(p->r, q[0], 0, q[1]);
    case 3: return f_gauss(p->r, q[0], 0, q[1]);
    case 4: return f_ggauss(p->r, q[0], 0, q[1], q[2]);
    case 5: return f_lorentz(p->r, q[0], 0, q[1]);
    case 6: return f_moffat(p->r, q[0], 0, q[1], q[2]);
    case 7: return f_sech2(p->r, q[0], 0, q[1]);
    case 51: return mixture(2, 6, p, c, q);
    case 52: return mixture(3, 6, p, c, q);
    case 53: return mixture(4, 7, p, c, q);
    case 54: return mixture(6, 7, p, c, q);
    case 101: return lerp_nw(p->r, p->rnidx, p->rnodes, q, 0);
    case 102: return nw_harmonic(p, c, q);
    case 103: return nw_distortion(p, q);
    default: return NAN;
    }
}

/* Density height traits: traits.hpp:561-896,1437-1514.  Surface-brightness
 * uids 1..7 are pdfs; opacity uids 1001..1007 are amplitude functions. */
static double rh_trait(int uid, const double* c, const double* q,
                       const pctx* p, double z)
{
    const double trunc = c[0];
    const int use_nodes = c[1] != 0;
    const int n = p->nrnodes;
#define NWP(k) (use_nodes ? lerp_nw(p->r, p->rnidx, p->rnodes, q, (k) * n) \
                          : q[k])
    if (uid < 1000) {
        const double p1 = NWP(0);
        double p2 = 0;
        if (uid == 4 || uid == 6) p2 = NWP(1);
        double v;
        switch (uid) {
        case 1: v = pdf_uniform(z, -p1, p1); break;
        case 2: v = pdf_expon(z, 0, p1); break;
        case 3: v = pdf_gauss(z, 0, p1); break;
        case 4: v = pdf_ggauss(z, 0, p1, p2); break;
        case 5: v = pdf_lorentz(z, 0, p1); break;
        case 6: v = pdf_moffat(z, 0, p1, p2); break;
        case 7: v = pdf_sech2(z, 0, p1); break;
        default: return NAN;
        }
        if (trunc == 0) return v;
        {
            const double lo = -trunc * p1, hi = trunc * p1;
            if (!(z >= lo && z <= hi)) return 0;
            /* uniform_wm truncation uses the plain uniform pdf/cdf over
             * (b - c, b + c): math.hpp:188-193 */
            return v / trunc_norm(uid, lo, hi, 0, p1, p2);
        }
    } else {
        const int u = uid - 1000;
        const double p1 = NWP(0);               /* amplitude */
        const double p2 = NWP(1);               /* scale */
        /* 3-parameter forms read params[1] for the shape when not nodewise:
         * traits.hpp:666-668 (reproduced). */
        const double p3 = (u == 4 || u == 6)
            ? (use_nodes ? lerp_nw(p->r, p->rnidx, p->rnodes, q, 2 * n) : q[1])
            : 0;
        double v;
        switch (u) {
        case 1: v = f_uniform_wm(z, p1, 0, p2); break;
        case 2: v = f_expon(z, p1, 0, p2); break;
        case 3: v = f_gauss(z, p1, 0, p2); break;
        case 4: v = f_ggauss(z, p1, 0, p2, p3); break;
        case 5: v = f_lorentz(z, p1, 0, p2); break;
        case 6: v = f_moffat(z, p1, 0, p2, p3); break;
        case 7: v = f_sech2(z, p1, 0, p2); break;
        default: return NAN;
        }
        if (trunc == 0) return v;
        return (z >= -trunc * p2 && z <= trunc * p2) ? v : 0;
    }
#undef NWP
}

/* Velocity polar traits: traits.hpp:898-1078,1592-1673. */
static double vp_trait(int uid, const double* c, const double* q,
                       const pctx* p)
{
    const double r = p->r, ct = cos(p->theta), st = sin(p->theta);
    const double si = sin(p->incl), ci = cos(p->incl);
    double v;
    switch (uid) {
    case 1: return q[0] * ct * si;
    case 2: v = q[1] * (2 / OR_PI) * atan(r / q[0]); return v * ct * si;
    case 3: v = q[1] * (1 - exp(-r / q[0])); return v * ct * si;
    case 4: v = q[1] * pow(r / q[0], q[3]) / (1 + pow(r / q[0], q[2]));
            return v * ct * si;
    case 5: v = r < q[0] ? q[1] * r / q[0] : q[1]; return v * ct * si;
    case 6: v = q[1] * tanh(r / q[0]); return v * ct * si;
    case 7: v = q[1] * (1 - exp(-r / q[0])) * (1 + q[2] * r / q[0]);
            return v * ct * si;
    case 8: v = q[1] * pow(1 + r / q[0], q[2]) *
                pow(1 + pow(r / q[0], -q[3]), -1 / q[3]);
            return v * ct * si;
    case 101: if (r == 0) return 0;
              return lerp_nw(r, p->rnidx, p->rnodes, q, 0) * ct * si;
    case 102: if (r == 0) return 0; return nw_harmonic(p, c, q) * ct * si;
    case 103: if (r == 0) return 0;
              return lerp_nw(r, p->rnidx, p->rnodes, q, 0) * st * si;
    case 104: if (r == 0) return 0; return nw_harmonic(p, c, q) * st * si;
    case 105: return lerp_nw(r, p->rnidx, p->rnodes, q, 0) * ci;
    case 106: return nw_harmonic(p, c, q) * ci;
    case 107: return lerp_nw(r, p->rnidx, p->rnodes, q, 0);
    case 108: return nw_harmonic(p, c, q);
    default: return NAN;
    }
}

/* Dispersion polar traits: traits.hpp:1080-1175,1702-1775. */
static double dp_trait(int uid, const double* c, const double* q,
                       const pctx* p)
{
    if (uid == 1 || (uid >= 2 && uid <= 7) || (uid >= 51 && uid <= 54) ||
        uid == 101 || uid == 102 || uid == 103)
        return rp_trait(uid, c, q, p);
    return NAN;
}

static double zp_trait(int uid, const double* c, const double* q,
                       const pctx* p)                        /* :1804-1829 */
{
    if (uid == 101) return lerp_nw(p->r, p->rnidx, p->rnodes, q, 0);
    if (uid == 102) return nw_harmonic(p, c, q);
    return NAN;
}

static double azrange(double theta, double ph, double s)  /* :1195-1208 */
{
    double pmin = ph - s * 0.5, pmax = ph + s * 0.5;
    pmin = wrap_angle(pmin);
    pmax = wrap_angle(pmax);
    if (pmin < pmax) return (theta > pmin && theta < pmax) ? 1 : 0;
    return (theta < pmin && theta > pmax) ? 0 : 1;
}

static double sp_trait(int uid, const double* c, const double* q,
                       const pctx* p)                        /* :1831-1857 */
{
    (void)c;
    if (uid == 1) return azrange(p->theta, q[0] * OR_D2R, q[1] * OR_D2R);
    if (uid == 101)
        return azrange(p->theta,
                       lerp_nw(p->r, p->rnidx, p->rnodes, q, 0) * OR_D2R,
                       lerp_nw(p->r, p->rnidx, p->rnodes, q, p->nrnodes)
                           * OR_D2R);
    return NAN;
}

/* Sums (or products) over a trait list, advancing the packed consts/params
 * by ccounts/pcounts: traits.hpp:1887-1919. */
enum { K_RP, K_VP, K_DP, K_ZP, K_SP, K_WP };

static void p_traits(int kind, const or_traits* t, const pctx* p, double* out)
{
    const double* c = t->cvalues;
    const double* q = t->pvalues;
    for (int i = 0; i < t->n; ++i) {
        const int uid = t->uids[i];
        switch (kind) {
        case K_RP: out[i] = rp_trait(uid, c, q, p); break;
        case K_VP: out[i] = vp_trait(uid, c, q, p); break;
        case K_DP: out[i] = dp_trait(uid, c, q, p); break;
        case K_ZP: out[i] = zp_trait(uid, c, q, p); break;
        case K_SP: out[i] = sp_trait(uid, c, q, p); break;
        case K_WP: out[i] = uid == 1 ? 1.0 : NAN; break;   /* :1229-1240 */
        }
        c += t->ccounts[i];
        q += t->pcounts[i];
    }
}

static void rh_traits(const or_traits* t, const pctx* p, double z, double* out)
{
    const double* c = t->cvalues;
    const double* q = t->pvalues;
    for (int i = 0; i < t->n; ++i) {
        out[i] = rh_trait(t->uids[i], c, q, p, z);
        c += t->ccounts[i];
        q += t->pcounts[i];
    }
}

/* vh / dh traits: uid 1 ("one") is the only one (traits.hpp:1675-1700,
 * 1777-1802). */
static void one_traits(const or_traits* t, double* out)
{
    for (int i = 0; i < t->n; ++i) out[i] = t->uids[i] == 1 ? 1.0 : NAN;
}

/* ------------------------------------------------------------------------ */
/* Geometry: core/gmodel/gmodels.hpp:118-323                                 */
/* ------------------------------------------------------------------------ */

/* disk_info: gmodels.hpp:197-215.  Sets *amb when the radius lies within the
 * relative margin of either edge test. */
static int disk_info(int* index, double r, int n, const double* nodes,
                     int* amb)
{
    const double m0 = OR_AMB_EPS_R * fabs(nodes[0]);
    const double m1 = OR_AMB_EPS_R * fabs(nodes[n - 1]);
    if (amb) {
        if (m0 > 0 && fabs(r - nodes[0]) <= m0) *amb = 1;
        if (fabs(r - nodes[n - 1]) <= m1) *amb = 1;
    }
    int i;
    for (i = 1; i < n; ++i)
        if (r < nodes[i]) break;
    *index = i < n ? i : n - 1;
    if (r < nodes[0]) return 0;
    if (r >= nodes[n - 1]) return 0;
    return 1;
}

typedef struct {
    int loose, tilted;
    const double *xpos, *ypos, *posa, *incl;
} geom;

/* rnode_radius: gmodels.hpp:167-195 (thick when has_z). */
static double rnode_radius(const geom* G, int idx, double x, double y,
                           double z, int has_z)
{
    const double xp = G->loose ? G->xpos[idx] : G->xpos[0];
    const double yp = G->loose ? G->ypos[idx] : G->ypos[0];
    const double pa = (G->tilted ? G->posa[idx] : G->posa[0]) * OR_D2R;
    const double in = (G->tilted ? G->incl[idx] : G->incl[0]) * OR_D2R;
    x -= xp;
    y -= yp;
    rot_lh_z(&x, &y, x, y, pa);
    if (has_z) rot_lh_x(&y, &z, y, z, in);
    else y /= cos(in);
    return sqrt(x * x + y * y);
}

/* ring_info: gmodels.hpp:217-323 (linear search + "velfi" radius). */
static int ring_info(int* index, double* radius, const geom* G, int n,
                     const double* nodes, double x, double y, double z,
                     int has_z, int* amb)
{
    const double rmin = rnode_radius(G, 0, x, y, z, has_z);
    if (amb && nodes[0] != 0 &&
        fabs(rmin - nodes[0]) <= OR_AMB_EPS_R * fabs(nodes[0])) *amb = 1;
    if (rmin < nodes[0]) return 0;
    const double rmax = rnode_radius(G, n - 1, x, y, z, has_z);
    if (amb && fabs(rmax - nodes[n - 1]) <= OR_AMB_EPS_R * fabs(nodes[n - 1]))
        *amb = 1;
    if (rmax >= nodes[n - 1]) return 0;
    int i;
    for (i = 1; i < n; ++i)
        if (rnode_radius(G, i, x, y, z, has_z) < nodes[i]) break;
    if (i >= n) i = n - 1;
    const double d1 = nodes[i - 1] - rnode_radius(G, i - 1, x, y, z, has_z);
    const double d2 = nodes[i] - rnode_radius(G, i, x, y, z, has_z);
    *radius = (nodes[i] * d1 - nodes[i - 1] * d2) / (d1 - d2);
    *index = i;
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Spectral line deposit: gmodels.hpp:91-116                                */
/* ------------------------------------------------------------------------ */

static int near_half(double a)
{
    return fabs(a - floor(a) - 0.5) < OR_AMB_EPS_CH;
}

/* Adds flux * N(v_k; v, s) over the rint window.  With `only_amb` the line
 * goes to the ambiguity bound instead of the cube.  Returns the window size. */
static int deposit_line(const or_grid* g, double* cube, double* amb,
                        long long base, long long plane, double flux,
                        double v, double s, int only_amb)
{
    const int n = g->spec_size;
    const double a0 = (v - s * LWM - g->spec_zero) / g->spec_step;
    const double a1 = (v + s * LWM - g->spec_zero) / g->spec_step;
    double r0 = rint(a0), r1 = rint(a1);
    if (r0 < 0) r0 = 0;
    if (r1 > n - 1) r1 = n - 1;
    const int k0 = (int)r0, k1 = (int)r1;
    for (int k = k0; k <= k1; ++k) {
        const double val = flux * pdf_gauss(g->spec_zero + k * g->spec_step,
                                            v, s);
        if (only_amb) amb[base + k * plane] += fabs(val);
        else cube[base + k * plane] += val;
    }
    if (amb && !only_amb) {
        /* channels whose inclusion hinges on a rint near-tie */
        /* lower edge: channel floor(a0) is in for one rounding, out for
         * the other; upper edge: channel floor(a1) + 1 likewise */
        if (near_half(a0) && floor(a0) >= 0 && floor(a0) <= n - 1) {
            const int c = (int)floor(a0);
            amb[base + c * plane] += fabs(flux * pdf_gauss(
                g->spec_zero + c * g->spec_step, v, s));
        }
        if (near_half(a1) && floor(a1) + 1 >= 0 && floor(a1) + 1 <= n - 1) {
            const int c = (int)floor(a1) + 1;
            amb[base + c * plane] += fabs(flux * pdf_gauss(
                g->spec_zero + c * g->spec_step, v, s));
        }
    }
    return k1 >= k0 ? k1 - k0 + 1 : 0;
}

/* ------------------------------------------------------------------------ */
/* smdisk: gmodels.hpp:645-958                                              */
/* ------------------------------------------------------------------------ */

#define TMAX 16

static void smdisk_voxel(const or_disk* d, const or_grid* g, const or_out* o,
                         int x, int y, int z, or_stats* st)
{
    const int is_thin = d->rh.uids == NULL;
    const int nx = g->spat_size[0], ny = g->spat_size[1], nz = g->spat_size[2];
    const geom G = {d->loose, d->tilted, d->xpos, d->ypos, d->posa, d->incl};
    int amb = 0, rnidx = -1, in = 1;
    double rn = 0;

    double xn = g->spat_zero[0] + x * g->spat_step[0];
    double yn = g->spat_zero[1] + y * g->spat_step[1];
    double zn = g->spat_zero[2] + z * g->spat_step[2];

    if (d->loose || d->tilted) {
        in = ring_info(&rnidx, &rn, &G, d->nrnodes, d->rnodes, xn, yn, zn,
                       !is_thin, o->amb ? &amb : NULL);
        if (!in && !amb) return;
        if (!in) {  /* evaluate an ambiguous miss at the outermost ring */
            rnidx = d->nrnodes - 1;
            rn = d->rnodes[d->nrnodes - 1];
        }
    }

    /* vsys follows xpos for loose disks: gmodels.hpp:726 (reproduced) */
    const double vsysi = d->loose ? lerp_nw(rn, rnidx, d->rnodes, d->xpos, 0)
                                  : (d->vsys ? d->vsys[0] : 0);
    const double xposi = d->loose ? lerp_nw(rn, rnidx, d->rnodes, d->xpos, 0)
                                  : d->xpos[0];
    const double yposi = d->loose ? lerp_nw(rn, rnidx, d->rnodes, d->ypos, 0)
                                  : d->ypos[0];
    const double posai = (d->tilted ? lerp_nw(rn, rnidx, d->rnodes, d->posa, 0)
                                    : d->posa[0]) * OR_D2R;
    const double incli = (d->tilted ? lerp_nw(rn, rnidx, d->rnodes, d->incl, 0)
                                    : d->incl[0]) * OR_D2R;

    xn -= xposi;
    yn -= yposi;
    rot_lh_z(&xn, &yn, xn, yn, posai);
    if (is_thin) yn /= cos(incli);
    else rot_lh_x(&yn, &zn, yn, zn, incli);
    const double theta = atan2(yn, xn);

    if (!(d->loose || d->tilted)) {
        rn = sqrt(xn * xn + yn * yn);
        in = disk_info(&rnidx, rn, d->nrnodes, d->rnodes,
                       o->amb ? &amb : NULL);
        if (!in && !amb) return;
    }

    double pt[TMAX] = {0}, ht[TMAX] = {0};
    double rvalue = 0, vvalue = 0, dvalue = 0, zvalue = 0, svalue = 0,
           wvalue = 1;
    pctx P = {rnidx, d->nrnodes, d->rnodes, xn, yn, rn, theta, incli};

    if (d->sp.uids) {
        p_traits(K_SP, &d->sp, &P, pt);
        for (int i = 0; i < d->sp.n; ++i) svalue += pt[i];
        if (!svalue) return;
    }
    if (d->zp.uids) {
        p_traits(K_ZP, &d->zp, &P, pt);
        for (int i = 0; i < d->zp.n; ++i) zvalue += pt[i];
        zn += zvalue;
    }
    if (d->rp.uids) p_traits(K_RP, &d->rp, &P, pt);
    if (d->rh.uids) rh_traits(&d->rh, &P, fabs(zn), ht);
    for (int i = 0; i < d->rp.n; ++i) rvalue += pt[i] * (is_thin ? 1 : ht[i]);
    if (!rvalue) return;
    if (is_thin) rvalue /= cos(incli);
    else rvalue *= g->spat_step[2];

    if (d->vp.uids) p_traits(K_VP, &d->vp, &P, pt);
    if (d->vh.uids) one_traits(&d->vh, ht);
    for (int i = 0; i < d->vp.n; ++i) vvalue += pt[i] * (is_thin ? 1 : ht[i]);
    vvalue += vsysi;

    if (d->dp.uids) p_traits(K_DP, &d->dp, &P, pt);
    if (d->dh.uids) one_traits(&d->dh, ht);
    for (int i = 0; i < d->dp.n; ++i) dvalue += pt[i] * (is_thin ? 1 : ht[i]);
    dvalue = fabs(dvalue);

    if (d->wp.uids && o->wdata) {
        p_traits(K_WP, &d->wp, &P, pt);
        for (int i = 0; i < d->wp.n; ++i) wvalue *= pt[i];
    }

    double orvalue = rvalue;
    if (d->opacity)
        for (int oz = z + 1; oz < nz; ++oz)
            orvalue -= orvalue * d->opacity[x + (long long)y * nx +
                                            (long long)oz * nx * ny];

    const long long plane = (long long)nx * ny;
    const long long col = x + (long long)y * nx;
    /* with no spectral cube wanted, amb bounds the image (its first plane) */
    const int image_amb = o->amb && o->image && !o->scube;
    if (!in) {  /* ambiguous miss: the line may appear in an FP32 run */
        if (o->scube && o->amb)
            deposit_line(g, o->scube, o->amb, col, plane, orvalue, vvalue,
                         dvalue, 1);
        if (image_amb) o->amb[col] += fabs(orvalue);
        return;
    }
    if (o->image) o->image[col] += orvalue;
    if (image_amb && amb) o->amb[col] += fabs(orvalue);
    if (o->scube) {
        const int nch = deposit_line(g, o->scube, o->amb, col, plane, orvalue,
                                     vvalue, dvalue, 0);
        if (amb && o->amb)  /* ambiguous hit: the line may vanish */
            deposit_line(g, o->scube, o->amb, col, plane, orvalue, vvalue,
                         dvalue, 1);
        if (st) { st->voxels_on_disk++; st->channel_deposits += nch; }
    }
    const long long idx = col + (long long)z * plane;
    if (o->wdata) o->wdata[idx] += wvalue;
    if (o->wdata_cmp) o->wdata_cmp[idx] = wvalue;
    if (o->rdata) o->rdata[idx] += rvalue;
    if (o->rdata_cmp) o->rdata_cmp[idx] = rvalue;
    if (o->ordata) o->ordata[idx] += orvalue;
    if (o->ordata_cmp) o->ordata_cmp[idx] = orvalue;
    if (o->vdata_cmp) o->vdata_cmp[idx] = vvalue;
    if (o->ddata_cmp) o->ddata_cmp[idx] = dvalue;
}

void or_smdisk(const or_disk* d, const or_grid* g, const or_out* o,
               or_stats* stats)
{
    for (int y = 0; y < g->spat_size[1]; ++y)
    for (int x = 0; x < g->spat_size[0]; ++x)
    for (int z = 0; z < g->spat_size[2]; ++z)
        smdisk_voxel(d, g, o, x, y, z, stats);
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11 "Parallel random numbers:  */
/* as easy as 1, 2, 3"), multipliers and Weyl key increments as published.   */
/* ------------------------------------------------------------------------ */

void or_philox4x32(const unsigned int ctr_in[4], const unsigned int key_in[2],
                   unsigned int out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int round = 0; round < 10; ++round) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

double or_cloud_uniform(unsigned long long seed, unsigned int cid,
                        unsigned int i)
{
    const unsigned int ctr[4] = {cid, i / 4u, 0u, 0u};
    const unsigned int key[2] = {(unsigned int)seed,
                                 (unsigned int)(seed >> 32)};
    unsigned int w[4];
    or_philox4x32(ctr, key, w);
    return ((double)(w[i % 4u] >> 9) + 0.5) * (1.0 / 8388608.0);
}

typedef struct {
    unsigned long long seed;
    unsigned int cid, next;
} cloud_rng;

static double rng_next(cloud_rng* s)
{
    return or_cloud_uniform(s->seed, s->cid, s->next++);
}

/* Samplers: math.hpp:152-498 (rng() draws in the reference's order). */
static double rnd_uniform(cloud_rng* s, double b, double c)
{ return b + rng_next(s) * (c - b); }
static double rnd_expon(cloud_rng* s, double b, double c)
{ const double u = rng_next(s) * 2 - 1;
  return b - c * sgn(u) * log(1 - 2 * fabs(u)); }
static double rnd_gauss(cloud_rng* s, double b, double c)
{ const double u1 = rng_next(s), u2 = rng_next(s);
  return b + c * sqrt(-2 * log(u1)) * cos(2 * OR_PI * u2); }
static double rnd_lorentz(cloud_rng* s, double b, double c)
{ const double u = rng_next(s) * 2 - 1;
  return b + c * tan(OR_PI * 0.5 * u); }
static double rnd_sech2(cloud_rng* s, double b, double c)
{ const double u = rng_next(s) * 2 - 1; return b + c * atanh(u); }

/* rejection_sampling: math.hpp:206-217, target = ggauss/moffat pdf. */
static double rnd_reject(cloud_rng* s, int uid, double trunc, double b,
                         double c, double d)
{
    double x, yv, zv;
    int guard = 0;
    do {
        x = rnd_uniform(s, -trunc, trunc);
        const double t0 = uid == 4 ? pdf_ggauss(0, b, c, d)
                                   : pdf_moffat(0, b, c, d);
        yv = t0 * rng_next(s);
        zv = uid == 4 ? pdf_ggauss(x, b, c, d) : pdf_moffat(x, b, c, d);
    } while (yv > zv && ++guard < 100000);
    return x;
}

/* rh_trait_rnd: traits.hpp:1516-1590; truncation never rejects (the loop
 * condition of _trunc_1d_rnd, math.hpp:116-124, is never true). */
static double rh_trait_rnd(int uid, const double* c, const double* q,
                           const pctx* p, cloud_rng* s)
{
    const int off = uid >= 1000 ? 1 : 0;
    const int u = uid >= 1000 ? uid - 1000 : uid;
    const int use_nodes = c[1] != 0;
    const int n = p->nrnodes;
    const double p1 = use_nodes ? lerp_nw(p->r, p->rnidx, p->rnodes, q, off * n)
                                : q[off];
    double p2 = 0;
    if (u == 4 || u == 6)
        p2 = use_nodes ? lerp_nw(p->r, p->rnidx, p->rnodes, q, (off + 1) * n)
                       : q[off + 1];
    switch (u) {
    case 1: return rnd_uniform(s, -p1, p1);
    case 2: return rnd_expon(s, 0, p1);
    case 3: return rnd_gauss(s, 0, p1);
    case 4: return rnd_reject(s, 4, 5 * p1, 0, p1, p2);
    case 5: return rnd_lorentz(s, 0, p1);
    case 6: return rnd_reject(s, 6, 5 * p1, 0, p1, p2);
    case 7: return rnd_sech2(s, 0, p1);
    default: return NAN;
    }
}

/* rp_trait_rnd: traits.hpp:310-326,1333-1435. */
static void rp_trait_rnd(int uid, const double* c, const double* q, int ridx,
                         const double* nodes, int n, cloud_rng* s,
                         double* sign, double* r, double* theta)
{
    const int u = uid >= 1000 ? uid - 1000 : uid;
    const int idx = ridx + 1;
    double rmin, rmax;
    switch (u) {
    case 1: rmin = nodes[0]; rmax = nodes[n - 1]; break;
    case 2: case 3: case 4: case 5: case 6: case 7: case 101: case 102: {
        const double hw = nodes[1] - nodes[0];
        rmin = nodes[idx] - hw;
        rmax = nodes[idx] + hw;
        break;
    }
    default:
        return;  /* mixtures and nw_distortion: not implemented (no-op) */
    }
    *r = sqrt(rmax * rmax + (rmin * rmin - rmax * rmax) * rng_next(s));
    *theta = 2 * OR_PI * rng_next(s);
    if (u == 102) {
        const int k = (int)rint(c[0]);
        const double a = q[idx];
        const double ph = q[n + idx] * OR_D2R;
        const double vr = a * (rng_next(s) * 2 - 1);
        const double vf = a * cos(k * (*theta - ph));
        *sign = vr > vf ? 1 : -1;
    }
}

static void mcdisk_cloud(double cflux, int ci, const int* csum,
                         const unsigned char* hoi, const or_disk* d,
                         const or_grid* g, const or_out* o,
                         unsigned long long seed, or_stats* st)
{
    const int nx = g->spat_size[0], ny = g->spat_size[1], nz = g->spat_size[2];
    cloud_rng S = {seed, (unsigned int)ci, 0u};
    int rnidx = 0, tidx, amb = 0;
    double sign = 1, rd = 0, theta = 0;
    const double* rpc = d->rp.cvalues;
    const double* rpp = d->rp.pvalues;
    const double* rhc = d->rh.cvalues;
    const double* rhp = d->rh.pvalues;
    double rvalue = cflux / g->spat_step[2];

    while (ci >= csum[rnidx]) rnidx++;
    for (tidx = 0; tidx < d->rp.n; ++tidx) {
        const int size = hoi[tidx] ? 1 : d->nrnodes - 2;
        if (rnidx < size) break;
        rnidx -= size;
        rpc += d->rp.ccounts[tidx];
        rpp += d->rp.pcounts[tidx];
        rhc += d->rh.ccounts[tidx];
        rhp += d->rh.pcounts[tidx];
    }
    rp_trait_rnd(d->rp.uids[tidx], rpc, rpp, rnidx, d->rnodes, d->nrnodes, &S,
                 &sign, &rd, &theta);
    if (sign < 0) rvalue = -rvalue;
    rvalue *= g->spat_step[2];
    rvalue /= g->spat_step[0] * g->spat_step[1];

    const double xd = rd * cos(theta), yd = rd * sin(theta);
    const int in = disk_info(&rnidx, rd, d->nrnodes, d->rnodes,
                             o->amb ? &amb : NULL);
    if (!in && !amb) return;

    double pt[TMAX] = {0}, ht[TMAX] = {0};
    pctx P = {rnidx, d->nrnodes, d->rnodes, xd, yd, rd, theta, 0};
    if (d->sp.uids) {
        double sv = 0;
        p_traits(K_SP, &d->sp, &P, pt);
        for (int i = 0; i < d->sp.n; ++i) sv += pt[i];
        if (!sv) return;
    }
    double zd = rh_trait_rnd(d->rh.uids[tidx], rhc, rhp, &P, &S);
    if (d->zp.uids) {
        double zv = 0;
        p_traits(K_ZP, &d->zp, &P, pt);
        for (int i = 0; i < d->zp.n; ++i) zv += pt[i];
        zd += zv;
    }
    const double vsysi = d->vsys ? (d->loose ? lerp_nw(rd, rnidx, d->rnodes,
                                                       d->vsys, 0)
                                             : d->vsys[0]) : 0;
    const double xposi = d->loose ? lerp_nw(rd, rnidx, d->rnodes, d->xpos, 0)
                                  : d->xpos[0];
    const double yposi = d->loose ? lerp_nw(rd, rnidx, d->rnodes, d->ypos, 0)
                                  : d->ypos[0];
    const double posai = (d->tilted ? lerp_nw(rd, rnidx, d->rnodes, d->posa, 0)
                                    : d->posa[0]) * OR_D2R;
    const double incli = (d->tilted ? lerp_nw(rd, rnidx, d->rnodes, d->incl, 0)
                                    : d->incl[0]) * OR_D2R;
    P.incl = incli;

    /* disk -> sky: transform_incl_posa_cpos with negated angles/centre,
     * gmodels.hpp:159-165,496 */
    double xn = xd, yn = yd, zn = zd;
    rot_lh_x(&yn, &zn, yn, zn, -incli);
    rot_lh_z(&xn, &yn, xn, yn, -posai);
    xn += xposi;
    yn += yposi;

    /* NaN positions (exponential height sampler, math.hpp:252-255, for
     * |u| > 1/2) land outside the cube on the host driver: rint(NaN) -> int
     * gives INT_MIN on x86.  Made explicit here (the C conversion is UB). */
    if (isnan(xn) || isnan(yn) || isnan(zn)) return;
    const double ax = (xn - g->spat_zero[0]) / g->spat_step[0];
    const double ay = (yn - g->spat_zero[1]) / g->spat_step[1];
    const double az = (zn - g->spat_zero[2]) / g->spat_step[2];
    const int x = (int)rint(ax), y = (int)rint(ay), z = (int)rint(az);
    const int inside = !(x < 0 || x >= nx || y < 0 || y >= ny || z < 0 ||
                         z >= nz);
    const int pos_amb = near_half(ax) || near_half(ay) || near_half(az);
    if (!inside && !(o->amb && pos_amb)) return;

    if (d->vp.uids) p_traits(K_VP, &d->vp, &P, pt);
    if (d->vh.uids) one_traits(&d->vh, ht);
    double vvalue = 0, dvalue = 0, wvalue = 1;
    for (int i = 0; i < d->vp.n; ++i) vvalue += pt[i] * ht[i];
    vvalue += vsysi;
    if (d->dp.uids) p_traits(K_DP, &d->dp, &P, pt);
    if (d->dh.uids) one_traits(&d->dh, ht);
    for (int i = 0; i < d->dp.n; ++i) dvalue += pt[i] * ht[i];
    dvalue = fabs(dvalue);
    if (d->wp.uids && o->wdata) {
        p_traits(K_WP, &d->wp, &P, pt);
        for (int i = 0; i < d->wp.n; ++i) wvalue *= pt[i];
    }

    const long long plane = (long long)nx * ny;
    /* ambiguous position: every candidate voxel gets the line in `amb`
     * (its flux, when amb bounds the image) */
    if (o->amb && (o->scube || o->image) && (pos_amb || amb)) {
        const int xs[2] = {(int)floor(ax), (int)floor(ax) + 1};
        const int ys[2] = {(int)floor(ay), (int)floor(ay) + 1};
        const int zs[2] = {(int)floor(az), (int)floor(az) + 1};
        for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
        for (int c = 0; c < 2; ++c) {
            const int cx = near_half(ax) ? xs[a] : x;
            const int cy = near_half(ay) ? ys[b] : y;
            const int cz = near_half(az) ? zs[c] : z;
            if ((!near_half(ax) && a) || (!near_half(ay) && b) ||
                (!near_half(az) && c)) continue;
            if (cx < 0 || cx >= nx || cy < 0 || cy >= ny || cz < 0 || cz >= nz)
                continue;
            double orv = rvalue;
            if (d->opacity)
                for (int oz = cz + 1; oz < nz; ++oz)
                    orv -= orv * d->opacity[cx + (long long)cy * nx +
                                            (long long)oz * plane];
            if (o->scube)
                deposit_line(g, o->scube, o->amb, cx + (long long)cy * nx,
                             plane, orv, vvalue, dvalue, 1);
            else
                o->amb[cx + (long long)cy * nx] += fabs(orv);
        }
    }
    if (!inside || !in) return;

    double orvalue = rvalue;
    if (d->opacity)
        for (int oz = z + 1; oz < nz; ++oz)
            orvalue -= orvalue * d->opacity[x + (long long)y * nx +
                                            (long long)oz * plane];
    const long long col = x + (long long)y * nx;
    if (o->image) o->image[col] += orvalue;
    if (o->scube) {
        const int nch = deposit_line(g, o->scube, o->amb, col, plane, orvalue,
                                     vvalue, dvalue, 0);
        if (st) { st->voxels_on_disk++; st->channel_deposits += nch; }
    }
    const long long idx = col + (long long)z * plane;
    if (o->wdata) o->wdata[idx] = wvalue;
    if (o->wdata_cmp) o->wdata_cmp[idx] = wvalue;
    if (o->rdata) o->rdata[idx] += rvalue;
    if (o->rdata_cmp) o->rdata_cmp[idx] += rvalue;
    if (o->ordata) o->ordata[idx] += orvalue;
    if (o->ordata_cmp) o->ordata_cmp[idx] += orvalue;
    if (o->vdata_cmp) o->vdata_cmp[idx] = vvalue;
    if (o->ddata_cmp) o->ddata_cmp[idx] = dvalue;
}

void or_mcdisk(double cflux, int nclouds, const int* ncloudscsum,
               int ncloudscsum_len, const unsigned char* hasordint,
               const or_disk* d, const or_grid* g, const or_out* o,
               unsigned long long seed, or_stats* stats)
{
    (void)ncloudscsum_len;
    for (int ci = 0; ci < nclouds; ++ci)
        mcdisk_cloud(cflux, ci, ncloudscsum, hasordint, d, g, o, seed, stats);
}

/* ------------------------------------------------------------------------ */
/* Data model: core/dmodel/dmodels.hpp                                      */
/* ------------------------------------------------------------------------ */

void or_downscale(const int scale[3], const int offset[3],
                  const int src_size[3], const int dst_size[3],
                  const double* src, double* dst)          /* dmodels.hpp:7 */
{
    const double nf = 1.0 / (scale[0] * scale[1] * scale[2]);
    const long long sx = src_size[0], sxy = (long long)src_size[0] * src_size[1];
    for (int z = 0; z < dst_size[2]; ++z)
    for (int y = 0; y < dst_size[1]; ++y)
    for (int x = 0; x < dst_size[0]; ++x) {
        double sum = 0;
        for (int dz = 0; dz < scale[2]; ++dz)
        for (int dy = 0; dy < scale[1]; ++dy)
        for (int dx = 0; dx < scale[0]; ++dx)
            sum += src[(offset[0] + x * scale[0] + dx) +
                       (offset[1] + y * scale[1] + dy) * sx +
                       (offset[2] + z * scale[2] + dz) * sxy];
        dst[x + (long long)y * dst_size[0] +
            (long long)z * dst_size[0] * dst_size[1]] = sum * nf;
    }
}

void or_wcube(const int spat_size[3], int spec_size, const double* spat,
              double* spec)                             /* gmodels.hpp:52-82 */
{
    const int nx = spat_size[0], ny = spat_size[1], nz = spat_size[2];
    const long long plane = (long long)nx * ny;
    for (long long c = 0; c < plane; ++c) {
        double sum = 0, mx = 0;
        for (int z = 0; z < nz; ++z) {
            const double v = spat[c + z * plane];
            sum += v;
            if (v > mx) mx = v;
        }
        const double w = mx > 0 ? (sum / nz) / mx : 0.0;
        for (int z = 0; z < spec_size; ++z) spec[c + z * plane] = w;
    }
}

void or_mask(double cutoff, int apply, const int size[3], double* dcube,
             double* mcube, double* wcube)              /* dmodels.hpp:48-74 */
{
    (void)wcube;
    const long long n = (long long)size[0] * size[1] * size[2];
    for (long long i = 0; i < n; ++i) {
        double m = 1;
        if (fabs(dcube[i]) <= cutoff) {
            m = 0;
            if (apply) dcube[i] = NAN;
        }
        if (mcube) mcube[i] = m;
    }
}

void or_moments(const int size[3], const double step[3], const double zero[3],
                const double* dcube, const double* wcube, double cutoff,
                int norders, const int* orders, double* mmaps_d,
                double* mmaps_m, double* mmaps_w)     /* dmodels.hpp:76-228 */
{
    const int nx = size[0], ny = size[1], nz = size[2];
    const long long plane = (long long)nx * ny;
    const int max_order = orders[norders - 1];
    for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x) {
        const long long c = x + (long long)y * nx;
        int m = 0;
        double m0s = 0;
        for (int z = 0; z < nz; ++z) m0s += dcube[c + z * plane] * step[2];
        const int valid = fabs(m0s) > cutoff;
        mmaps_m[c] = valid;
        const double m0 = valid ? m0s : NAN;
        if (orders[m] == 0) { mmaps_d[c + m * plane] = m0; m++; }
        double ws = 0;
        for (int z = 0; wcube && valid && z < nz; ++z)
            ws += wcube[c + z * plane] * dcube[c + z * plane] * step[2] / m0;
        ws = valid ? ws : NAN;
        if (wcube) mmaps_w[c] = ws;
        if (max_order == 0) continue;
        double m1s = 0;
        for (int z = 0; valid && z < nz; ++z)
            m1s += dcube[c + z * plane] * (zero[2] + z * step[2]) * step[2];
        const double m1 = valid ? m1s / m0 : NAN;
        if (m < norders && orders[m] == 1) { mmaps_d[c + m * plane] = m1; m++; }
        if (max_order == 1) continue;
        double m2s = 0;
        for (int z = 0; valid && z < nz; ++z) {
            const double dv = zero[2] + z * step[2] - m1;
            m2s += dcube[c + z * plane] * dv * dv * step[2];
        }
        const double m2 = valid ? sqrt(m2s / m0) : NAN;
        if (m < norders && orders[m] == 2) { mmaps_d[c + m * plane] = m2; m++; }
        for (; m < norders; ++m) {
            double s = 0;
            for (int z = 0; valid && z < nz; ++z)
                s += dcube[c + z * plane] *
                     pow(zero[2] + z * step[2] - m1, orders[m]);
            mmaps_d[c + m * plane] = valid ? s / m0 : NAN;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Objective: core/objective/objective.hpp:7-33, host kernels.hpp:502-514   */
/* ------------------------------------------------------------------------ */

void or_residual(const double* obs_d, const double* obs_e, const double* obs_m,
                 const double* mdl_d, const double* mdl_w, const double* mdl_m,
                 long long size, double weight, double* res)
{
    for (long long i = 0; i < size; ++i) {
        double r = mdl_d[i] - obs_d[i];
        if (obs_e) r /= obs_e[i];
        if (mdl_w) r *= mdl_w[i];
        /* a zero mask entry gives a zero residual, also where the masked
           value is NaN (the reference multiplies: NaN * 0 = NaN) */
        if (obs_m) r = obs_m[i] != 0.0 ? r * obs_m[i] : 0.0;
        if (mdl_m) r = mdl_m[i] != 0.0 ? r * mdl_m[i] : 0.0;
        res[i] = r * weight;
    }
}

double or_residual_sum(const double* r, long long size, int squared)
{
    double s = 0, comp = 0;
    for (long long i = 0; i < size; ++i) {
        const double v = squared ? r[i] * r[i] : fabs(r[i]);
        const double t = s + v;
        if (fabs(s) >= fabs(v)) comp += (s - t) + v;
        else comp += (v - t) + s;
        s = t;
    }
    return s + comp;
}

void or_circular_convolve(const int size[3], const double* a, const double* k,
                          double* out)
{
    const int nx = size[0], ny = size[1], nz = size[2];
    for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x) {
        double s = 0;
        for (int jz = 0; jz < nz; ++jz)
        for (int jy = 0; jy < ny; ++jy)
        for (int jx = 0; jx < nx; ++jx) {
            const int dx = ((x - jx) % nx + nx) % nx;
            const int dy = ((y - jy) % ny + ny) % ny;
            const int dz = ((z - jz) % nz + nz) % nz;
            s += a[jx + jy * nx + jz * nx * ny] *
                 k[dx + dy * nx + dz * nx * ny];
        }
        out[x + y * nx + z * nx * ny] = s;
    }
}
