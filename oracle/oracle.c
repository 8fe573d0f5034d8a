/* CPU oracle: the reference's sequential single-site chains, restated in C.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE. Only tests/, the smoke check and bench.py's CPU-baseline leg may
 * load it, and only as the checker (or, for the CPU baseline, as the timed CPU path). The product path
 * (libsck.so + generated HIP kernels) never calls it and has no CPU fallback.
 *
 * What it restates:
 *   - the per-site update emitted by SCK/generator.py:150-319 (local field in bond-table order grouped by J
 *     value, extra matrices, field, Philox draw, proposal, dE, cubic anisotropy, Metropolis/Barker rule),
 *     evaluated one site at a time in the order of Sampler.reference_sweep (SCK/sampler.py:281-293):
 *     classes colour-major then sublattice, sites of a class in raster order (z, y, x);
 *   - the synchronous control (SCK/sampler.py:271-279): every active site from the old state;
 *   - the bitwise chains on unpacked +-1 spins (SCK/bitwise.py:194-246): shared uniform per (word, site)
 *     and independent 32-bit U per replica built bit by bit from Philox words (SCK/bitwise.py:150-184);
 *   - the float64 energy and magnetization (SCK/sampler.py:296-345).
 * Additions beyond the reference (its parity is unpinned by the reference; pinned here by closed forms and
 * detailed-balance tests): the uniaxial terms and the cone proposal of width w.
 *
 * Arithmetic: every float operation is IEEE fp32 in the order the reference's generated source writes it
 * (the interpreter evaluates it with numpy float32, no fused multiply-add; build with -ffp-contract=off).
 * Ising exp values come from a caller-supplied function (numpy's float32 exp), memoised per thread, because
 * the reference's CPU chain evaluates tl.exp with numpy; vector models use libm's expf/cosf/sinf/sqrtf,
 * whose last-ulp differences from numpy are inside the decision tolerance their parity uses.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "philox.h"

#define MAXB 8         /* sublattices */
#define MAXR 4         /* colouring group factors */
#define MAXBOND 256

typedef float (*exp_fn)(float);

typedef struct {
    int32_t n, b, D;
    int32_t L[3];            /* L1 (x), L2 (y), L3 (z) */
    int32_t periodic[3];
    int32_t nbond;
    const int32_t *bond;     /* [nbond][6]: alpha, alpha2, dx, dy, dz, shell */
    const double *bondJ;     /* [nbond] J of the bond's shell (0 when absent) */
    const int32_t *has_extra;/* [nbond] */
    const double *extra;     /* [nbond][3][3] */
    int32_t has_h;
    double h[3];
    const int8_t *mask;      /* [nsite] 1 = spin present */
    int32_t has_cubic;
    const float *k1;         /* [nsite] */
    double k2;
    int32_t has_uniaxial;
    double du, du_axis[3];
    int32_t has_site;
    const float *site_ku;    /* [nsite] */
    const float *site_axis;  /* [nsite][n] */
    int32_t barker;
    /* colouring gamma(c, alpha) = K c + o_alpha (mod m), mixed radix */
    int32_t r;
    int32_t m[MAXR];
    int32_t K[MAXR][3];
    int32_t o[MAXB][MAXR];
} oracle_sys;

/* proposal kinds */
enum { PROP_UNIFORM = 0, PROP_CONE = 1 };

typedef struct {
    int32_t kind;
    const float *width;      /* [B] fixed cone half-angle per replica (PROP_CONE, width_table == NULL) */
    /* learned width: w = table[i0][i1][i2], i_d = clamp(int(f_d * s_d), 0, n_d - 1) with the features
       f0 = |v| beta, f1 = (|Du| + |Ku_i| + |K1_i|) beta, f2 = |v . e| / |v| (0 when |v| = 0) */
    const float *width_table;
    int32_t n[3];
    float s[3];
    float axis[3];
} oracle_prop;

/* ------------------------------------------------------------------------------------------ helpers */
static int64_t nsite_of(const oracle_sys *s) { return (int64_t)s->b * s->L[2] * s->L[1] * s->L[0]; }

static int colour_of(const oracle_sys *s, int x, int y, int z, int a) {
    int c[3] = {x, y, z};
    int idx = 0;
    for (int r = 0; r < s->r; ++r) {
        int64_t g = s->o[a][r];
        for (int k = 0; k < s->D; ++k) g += (int64_t)s->K[r][k] * c[k];
        int mr = s->m[r];
        int v = (int)(((g % mr) + mr) % mr);
        idx = idx * mr + v;
    }
    return idx;
}

static int n_colours(const oracle_sys *s) {
    int n = 1;
    for (int r = 0; r < s->r; ++r) n *= s->m[r];
    return n;
}

/* neighbour index of site (a, z, y, x) along bond i, or -1 outside the box */
static int64_t neighbour(const oracle_sys *s, int i, int x, int y, int z) {
    const int32_t *bd = s->bond + 6 * i;
    int c[3] = {x + bd[2], y + bd[3], z + bd[4]};
    for (int k = 0; k < 3; ++k) {
        int Lk = s->L[k];
        if (k < s->D && s->periodic[k]) {
            c[k] = ((c[k] % Lk) + Lk) % Lk;
        } else if (c[k] < 0 || c[k] >= Lk) {
            return -1;
        }
    }
    int64_t pl = (int64_t)s->L[2] * s->L[1] * s->L[0];
    return bd[1] * pl + ((int64_t)c[2] * s->L[1] + c[1]) * s->L[0] + c[0];
}

/* per-thread memo of exp values keyed by the fp32 argument */
typedef struct {
    uint32_t key[1024];
    float val[1024];
    uint8_t used[1024];
} exp_memo;

static float memo_exp(exp_memo *mm, exp_fn fn, float x) {
    uint32_t bits;
    memcpy(&bits, &x, 4);
    uint32_t h = (bits * 2654435761u) >> 22;
    for (int probe = 0; probe < 1024; ++probe) {
        uint32_t j = (h + probe) & 1023;
        if (!mm->used[j]) {
            float v = fn(x);
            mm->used[j] = 1;
            mm->key[j] = bits;
            mm->val[j] = v;
            return v;
        }
        if (mm->key[j] == bits) return mm->val[j];
    }
    return fn(x);
}

/* ----------------------------------------------------------------------------- one site, one replica */
typedef struct {
    int ngroup;
    double gJ[MAXBOND];
    int gstart[MAXBOND + 1];
    int member[MAXBOND];     /* bond indices ordered by group, bond order inside a group */
    int gskip[MAXBOND];
    int nextra;
    int extra_bond[MAXBOND];
} sub_plan;

static void plan_sublattice(const oracle_sys *s, int a, sub_plan *p) {
    memset(p, 0, sizeof(*p));
    double vals[MAXBOND];
    int nv = 0;
    for (int i = 0; i < s->nbond; ++i) {
        if (s->bond[6 * i] != a) continue;
        int g;
        for (g = 0; g < nv; ++g)
            if (vals[g] == s->bondJ[i]) break;
        if (g == nv) vals[nv++] = s->bondJ[i];
        if (s->has_extra[i]) p->extra_bond[p->nextra++] = i;
    }
    p->ngroup = nv;
    int k = 0;
    for (int g = 0; g < nv; ++g) {
        p->gJ[g] = vals[g];
        p->gstart[g] = k;
        int any_extra = 0;
        for (int i = 0; i < s->nbond; ++i) {
            if (s->bond[6 * i] != a || s->bondJ[i] != vals[g]) continue;
            p->member[k++] = i;
            any_extra |= s->has_extra[i];
        }
        p->gskip[g] = (vals[g] == 0.0 && !any_extra);
    }
    p->gstart[nv] = k;
}

typedef struct {
    const oracle_sys *s;
    const sub_plan *plans;
    const oracle_prop *prop;
    exp_fn ising_exp;
    exp_memo *memo;
} ctx_t;

/* Local field (bond part + extras + field) of site `own` of sublattice a, read from spins `src` of one
 * replica laid out [n][nsite]. Writes v[0..n). */
static void local_field(const ctx_t *cx, const float *src, int64_t nsite, int a, int x, int y, int z,
                        float v[3]) {
    const oracle_sys *s = cx->s;
    const sub_plan *p = &cx->plans[a];
    int n = s->n;
    float ext[MAXBOND][3];
    int ext_slot[MAXBOND];
    for (int i = 0; i < s->nbond; ++i) ext_slot[i] = -1;
    for (int e = 0; e < p->nextra; ++e) ext_slot[p->extra_bond[e]] = e;
    for (int k = 0; k < n; ++k) v[k] = 0.0f;
    for (int g = 0; g < p->ngroup; ++g) {
        if (p->gskip[g]) continue;
        float gs[3] = {0.0f, 0.0f, 0.0f};
        for (int j = p->gstart[g]; j < p->gstart[g + 1]; ++j) {
            int i = p->member[j];
            int64_t q = neighbour(s, i, x, y, z);
            for (int k = 0; k < n; ++k) {
                float t = q >= 0 ? src[k * nsite + q] : 0.0f;
                gs[k] = gs[k] + t;
                if (ext_slot[i] >= 0) ext[ext_slot[i]][k] = t;
            }
        }
        if (p->gJ[g] != 0.0) {
            float Jf = (float)p->gJ[g];
            for (int k = 0; k < n; ++k) v[k] = v[k] + Jf * gs[k];
        }
    }
    for (int e = 0; e < p->nextra; ++e) {
        const double *E = s->extra + 9 * p->extra_bond[e];
        for (int k = 0; k < n; ++k) {
            int any = 0;
            float acc = 0.0f;
            for (int l = 0; l < n; ++l) {
                double El = E[3 * k + l];
                if (El == 0.0) continue;
                float term = (float)El * ext[e][l];
                acc = any ? acc + term : term;
                any = 1;
            }
            if (any) v[k] = v[k] + acc;
        }
    }
    if (s->has_h)
        for (int k = 0; k < n; ++k)
            if (s->h[k] != 0.0) v[k] = v[k] + (float)s->h[k];
}

static float cubic_term(float k1, float k2, const float *p) {
    float a = (p[0] * p[1]) * (p[0] * p[1]) + (p[0] * p[2]) * (p[0] * p[2]);
    a = a + (p[1] * p[2]) * (p[1] * p[2]);
    float xyz = (p[0] * p[1]) * p[2];
    return k1 * a + (k2 * xyz) * xyz;
}

static float dot_n(const float *a, const float *b, int n) {
    float d = a[0] * b[0];
    for (int k = 1; k < n; ++k) d = d + a[k] * b[k];
    return d;
}

/* Cone width of one update: fixed per replica, or the learned table looked up from features of the site's
 * environment that its own move leaves unchanged. */
static float cone_width(const ctx_t *cx, int64_t rep, int64_t own, const float *v, float beta) {
    const oracle_prop *pp = cx->prop;
    const oracle_sys *s = cx->s;
    int n = s->n;
    if (!pp->width_table) return pp->width[rep];
    float vv = dot_n(v, v, n);
    float hv = sqrtf(vv);
    float K = s->has_uniaxial ? (float)fabs(s->du) : 0.0f;
    if (s->has_site) K = K + fabsf(s->site_ku[own]);
    if (s->has_cubic) K = K + fabsf(s->k1[own]);
    float f[3];
    f[0] = hv * beta;
    f[1] = K * beta;
    f[2] = hv > 0.0f ? fabsf(dot_n(v, pp->axis, n)) / hv : 0.0f;
    int idx[3];
    for (int d = 0; d < 3; ++d) {
        int i = (int)(f[d] * pp->s[d]);
        idx[d] = i > pp->n[d] - 1 ? pp->n[d] - 1 : (i < 0 ? 0 : i);
    }
    return pp->width_table[(idx[0] * pp->n[1] + idx[1]) * pp->n[2] + idx[2]];
}

/* The Metropolis/Barker update of one site of one replica. `src` is read for the local field, `own` is
 * the site's spin (n components, read from src), the result goes to `dst`. Returns 1 when accepted;
 * margin (optional) receives dE * beta + ln(u0) for Metropolis, in double. */
static int update_site(const ctx_t *cx, const float *src, float *dst, int64_t nsite, int64_t rep, int64_t seed,
                       int64_t counter_base, float beta, int a, int x, int y, int z, double *margin) {
    const oracle_sys *s = cx->s;
    int n = s->n;
    int64_t pl = (int64_t)s->L[2] * s->L[1] * s->L[0];
    int64_t own = a * pl + ((int64_t)z * s->L[1] + y) * s->L[0] + x;
    float sv[3] = {0, 0, 0}, v[3] = {0, 0, 0}, t[3] = {0, 0, 0};
    for (int k = 0; k < n; ++k) sv[k] = src[k * nsite + own];
    local_field(cx, src, nsite, a, x, y, z, v);
    uint32_t r4[4];
    philox_raw4(seed, (uint64_t)(counter_base + own), r4, 10);
    float u0 = philox_u01(r4[0]), u1 = philox_u01(r4[1]), u2 = philox_u01(r4[2]);
    const float TWO_PI = 6.283185307179586f;
    if (n == 1) {
        t[0] = -sv[0];
    } else if (cx->prop->kind == PROP_UNIFORM) {
        if (n == 2) {
            float phi = TWO_PI * u1;
            t[0] = cosf(phi);
            t[1] = sinf(phi);
        } else {
            float cz = 2.0f * u1 - 1.0f;
            float phi = TWO_PI * u2;
            float om = 1.0f - cz * cz;
            float rr = sqrtf(om > 0.0f ? om : 0.0f);
            t[0] = rr * cosf(phi);
            t[1] = rr * sinf(phi);
            t[2] = cz;
        }
    } else {
        float w = cone_width(cx, rep, own, v, beta);
        if (n == 2) {
            float ph = atan2f(sv[1], sv[0]) + w * (2.0f * u1 - 1.0f);
            t[0] = cosf(ph);
            t[1] = sinf(ph);
        } else {
            float c = 1.0f - u1 * (1.0f - cosf(w));
            float om = 1.0f - c * c;
            float st = sqrtf(om > 0.0f ? om : 0.0f);
            float phi = TWO_PI * u2;
            /* orthonormal frame around s: e1 = normalize(a x s), a = z unless s is near z, then x */
            float ax[3] = {0.0f, 0.0f, 1.0f};
            if (fabsf(sv[2]) > 0.9f) {
                ax[0] = 1.0f;
                ax[2] = 0.0f;
            }
            float e1[3] = {ax[1] * sv[2] - ax[2] * sv[1], ax[2] * sv[0] - ax[0] * sv[2], ax[0] * sv[1] - ax[1] * sv[0]};
            float ne = sqrtf(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
            for (int k = 0; k < 3; ++k) e1[k] = e1[k] / ne;
            float e2[3] = {sv[1] * e1[2] - sv[2] * e1[1], sv[2] * e1[0] - sv[0] * e1[2], sv[0] * e1[1] - sv[1] * e1[0]};
            float cp = cosf(phi), sp = sinf(phi);
            for (int k = 0; k < 3; ++k) t[k] = c * sv[k] + st * (cp * e1[k] + sp * e2[k]);
            float nt = sqrtf(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
            for (int k = 0; k < 3; ++k) t[k] = t[k] / nt;
        }
    }
    float dsum = (t[0] - sv[0]) * v[0];
    for (int k = 1; k < n; ++k) dsum = dsum + (t[k] - sv[k]) * v[k];
    float dE = -dsum;
    if (s->has_cubic) {
        float k1 = s->k1[own];
        float k2 = (float)s->k2;
        float at = cubic_term(k1, k2, t), as = cubic_term(k1, k2, sv);
        dE = dE + (at - as);
    }
    if (s->has_uniaxial) {
        float e[3] = {(float)s->du_axis[0], (float)s->du_axis[1], (float)s->du_axis[2]};
        float se = dot_n(sv, e, n), te = dot_n(t, e, n);
        dE = dE + (float)s->du * (se * se - te * te);
    }
    if (s->has_site) {
        float ku = s->site_ku[own];
        const float *ax = s->site_axis + own * n;
        float sn = dot_n(sv, ax, n), tn = dot_n(t, ax, n);
        dE = dE + ku * (sn * sn - tn * tn);
    }
    int acc;
    if (s->barker) {
        float e = n == 1 && cx->ising_exp ? memo_exp(cx->memo, cx->ising_exp, beta * dE) : expf(beta * dE);
        acc = (u0 * (1.0f + e)) < 1.0f;
        if (margin) {
            double xb = (double)beta * dE;       /* ln u0 + ln(1 + e^x), the Barker rule in log form */
            *margin = log((double)u0) + (xb > 0 ? xb + log1p(exp(-xb)) : log1p(exp(xb)));
        }
    } else {
        float x = (-beta) * dE;
        float p = n == 1 && cx->ising_exp ? memo_exp(cx->memo, cx->ising_exp, x) : expf(x);
        acc = u0 < p;
        if (margin) *margin = (double)beta * (double)dE + log((double)u0);
    }
    for (int k = 0; k < n; ++k) dst[k * nsite + own] = acc ? t[k] : sv[k];
    return acc;
}

/* ------------------------------------------------------------------------------------ public entries */

/* Sequential chain (classes colour-major then sublattice, raster order inside a class) for the replicas
 * listed in `reps`. spins: [nrep_stored][n][nsite] fp32 where row j holds replica reps[j]. The Philox
 * counter of replica r at site i is r * nsite + i (64-bit). Runs sweeps sweep0 .. sweep0 + nsweeps - 1.
 * class_lo/class_hi restrict a sweep to classes [lo, hi) of the sweep order (for per-class checks; pass
 * 0 and -1 for all). margins (optional): [nrep_stored][nsite] receives the last update's margin.
 * Returns the number of accepted updates. */
int64_t oracle_chain(const oracle_sys *s, float *spins, const int64_t *reps, int64_t nrep, const float *beta,
                     int64_t base_seed, int64_t sweep0, int64_t nsweeps, const oracle_prop *prop, exp_fn fn,
                     int class_lo, int class_hi, double *margins, int nthreads) {
    int64_t nsite = nsite_of(s);
    int nc = n_colours(s);
    sub_plan *plans = (sub_plan *)calloc(s->b, sizeof(sub_plan));
    for (int a = 0; a < s->b; ++a) plan_sublattice(s, a, &plans[a]);
    /* classes in sweep order */
    int ncls = 0;
    int *cls = (int *)malloc(sizeof(int) * 2 * nc * s->b);
    int64_t pl = (int64_t)s->L[2] * s->L[1] * s->L[0];
    for (int c = 0; c < nc; ++c)
        for (int a = 0; a < s->b; ++a) {
            int any = 0;
            for (int z = 0; z < s->L[2] && !any; ++z)
                for (int y = 0; y < s->L[1] && !any; ++y)
                    for (int x = 0; x < s->L[0] && !any; ++x)
                        if (s->mask[a * pl + ((int64_t)z * s->L[1] + y) * s->L[0] + x] &&
                            colour_of(s, x, y, z, a) == c)
                            any = 1;
            if (any) {
                cls[2 * ncls] = c;
                cls[2 * ncls + 1] = a;
                ++ncls;
            }
        }
    if (class_hi < 0 || class_hi > ncls) class_hi = ncls;
    int64_t total = 0;
    (void)nthreads;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total) num_threads(nthreads > 0 ? nthreads : 1)
    for (int64_t j = 0; j < nrep; ++j) {
        exp_memo *memo = (exp_memo *)calloc(1, sizeof(exp_memo));
        ctx_t cx = {s, plans, prop, fn, memo};
        float *S = spins + j * (int64_t)s->n * nsite;
        int64_t rep = reps[j];
        for (int64_t sw = sweep0; sw < sweep0 + nsweeps; ++sw) {
            int64_t seed = sweep_seed(base_seed, sw);
            for (int ci = class_lo; ci < class_hi; ++ci) {
                int c = cls[2 * ci], a = cls[2 * ci + 1];
                for (int z = 0; z < s->L[2]; ++z)
                    for (int y = 0; y < s->L[1]; ++y)
                        for (int x = 0; x < s->L[0]; ++x) {
                            int64_t own = a * pl + ((int64_t)z * s->L[1] + y) * s->L[0] + x;
                            if (!s->mask[own] || colour_of(s, x, y, z, a) != c) continue;
                            double mg;
                            total += update_site(&cx, S, S, nsite, rep, seed, rep * nsite, beta[rep], a, x, y, z,
                                                 margins ? &mg : NULL);
                            if (margins) margins[j * nsite + own] = mg;
                        }
            }
        }
        free(memo);
    }
    free(cls);
    free(plans);
    return total;
}

/* Synchronous control: every active site of every sublattice from the old state (SCK/sampler.py:271-279). */
void oracle_sync(const oracle_sys *s, float *spins, const int64_t *reps, int64_t nrep, const float *beta,
                 int64_t base_seed, int64_t sweep0, int64_t nsweeps, const oracle_prop *prop, exp_fn fn) {
    int64_t nsite = nsite_of(s);
    sub_plan *plans = (sub_plan *)calloc(s->b, sizeof(sub_plan));
    for (int a = 0; a < s->b; ++a) plan_sublattice(s, a, &plans[a]);
    int64_t pl = (int64_t)s->L[2] * s->L[1] * s->L[0];
    float *old = (float *)malloc(sizeof(float) * s->n * nsite);
    exp_memo *memo = (exp_memo *)calloc(1, sizeof(exp_memo));
    ctx_t cx = {s, plans, prop, fn, memo};
    for (int64_t j = 0; j < nrep; ++j) {
        float *S = spins + j * (int64_t)s->n * nsite;
        int64_t rep = reps[j];
        for (int64_t sw = sweep0; sw < sweep0 + nsweeps; ++sw) {
            int64_t seed = sweep_seed(base_seed, sw);
            memcpy(old, S, sizeof(float) * s->n * nsite);
            for (int a = 0; a < s->b; ++a)
                for (int z = 0; z < s->L[2]; ++z)
                    for (int y = 0; y < s->L[1]; ++y)
                        for (int x = 0; x < s->L[0]; ++x) {
                            int64_t own = a * pl + ((int64_t)z * s->L[1] + y) * s->L[0] + x;
                            if (!s->mask[own]) continue;
                            update_site(&cx, old, S, nsite, rep, seed, rep * nsite, beta[rep], a, x, y, z, NULL);
                        }
        }
    }
    free(memo);
    free(old);
    free(plans);
}

/* ------------------------------------------------------------------------------------------ bitwise */
/* Sequential bitwise chain on unpacked spins S[B][nsite] int8 (+-1, 0 masked) for all B replicas.
 * independent = 0: uniform of (word w, site i) is output w % 4 of rand4x(seed, (w / 4) * nsite + i),
 *   flip iff u < ptab_f[w][s][ze][k] (SCK/bitwise.py:150-155, 236-245).
 * independent = 1: 32-bit U of replica 32w + j has bit 31 - l equal to bit j of word R_l, R_{4g+q} output q
 *   of randint4x(key_g, w * nsite + i), key_g = (seed + (g + 1) * 104729) mod 2147483629; flip iff
 *   U < ptab_i[w][s][ze][k] as integers (SCK/bitwise.py:159-184, 236-245).
 * Z1 = Z + 1 is the table's inner size; field selects the s plane (0 otherwise). J-carrying bonds only. */
void oracle_bitwise(const oracle_sys *s, int8_t *S, int64_t B, const float *ptab_f, const int64_t *ptab_i, int Z1,
                    int field, int independent, int64_t base_seed, int64_t sweep0, int64_t nsweeps, int nthreads,
                    int64_t word0) {
    int64_t nsite = nsite_of(s);
    int64_t pl = (int64_t)s->L[2] * s->L[1] * s->L[0];
    int nc = n_colours(s);
    int64_t W = (B + 31) / 32;
    int64_t NT = 2LL * Z1 * Z1;
    (void)nthreads;
    for (int64_t sw = sweep0; sw < sweep0 + nsweeps; ++sw) {
        int64_t seed = sweep_seed(base_seed, sw);
        for (int c = 0; c < nc; ++c)
            for (int a = 0; a < s->b; ++a)
                for (int z = 0; z < s->L[2]; ++z)
                    for (int y = 0; y < s->L[1]; ++y)
                        for (int x = 0; x < s->L[0]; ++x) {
                            int64_t own = a * pl + ((int64_t)z * s->L[1] + y) * s->L[0] + x;
                            if (!s->mask[own] || colour_of(s, x, y, z, a) != c) continue;
                            int64_t nb[MAXBOND];
                            int ze = 0;
                            for (int i = 0; i < s->nbond; ++i) {
                                if (s->bond[6 * i] != a || s->bondJ[i] == 0.0) continue;
                                int64_t q = neighbour(s, i, x, y, z);
                                if (q >= 0 && s->mask[q]) nb[ze++] = q;
                            }
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
                            for (int64_t w = 0; w < W; ++w) {
                                uint32_t U[32];
                                float uf = 0.0f;
                                if (independent) {
                                    memset(U, 0, sizeof(U));
                                    for (int g = 0; g < 8; ++g) {
                                        int64_t kg = (seed + (int64_t)(g + 1) * 104729LL) % 2147483629LL;
                                        uint32_t rw[4];
                                        philox_raw4(kg, (uint64_t)(uint32_t)((word0 + w) * nsite + own), rw, 10);
                                        for (int q = 0; q < 4; ++q)
                                            for (int j = 0; j < 32; ++j)
                                                U[j] |= ((rw[q] >> j) & 1u) << (31 - (4 * g + q));
                                    }
                                } else {
                                    uint32_t r4[4];
                                    philox_raw4(seed, (uint64_t)(uint32_t)(((word0 + w) / 4) * nsite + own), r4, 10);
                                    uf = philox_u01(r4[(word0 + w) % 4]);
                                }
                                for (int j = 0; j < 32; ++j) {
                                    int64_t r = 32 * w + j;
                                    if (r >= B) break;
                                    int si = S[r * nsite + own];
                                    int tot = 0;
                                    for (int t = 0; t < ze; ++t) tot += S[r * nsite + nb[t]];
                                    int kk = (ze - si * tot) / 2;
                                    int sidx = field ? (si > 0) : 0;
                                    int64_t ti = w * NT + (int64_t)sidx * Z1 * Z1 + (int64_t)ze * Z1 + kk;
                                    int flip = independent ? ((int64_t)U[j] < ptab_i[ti]) : (uf < ptab_f[ti]);
                                    if (flip) S[r * nsite + own] = (int8_t)(-si);
                                }
                            }
                        }
    }
}

/* ---------------------------------------------------------------------------------------- observables */
/* Energy (float64, SCK/sampler.py:326-341 plus the uniaxial terms) and per-sublattice magnetization
 * [B][b][n] of spins [B][n][nsite] given as double. */
void oracle_observe(const oracle_sys *s, const double *spins, int64_t B, double *E, double *M) {
    int64_t nsite = nsite_of(s);
    int64_t pl = (int64_t)s->L[2] * s->L[1] * s->L[0];
    int n = s->n;
    for (int64_t r = 0; r < B; ++r) {
        const double *S = spins + r * n * nsite;
        double e = 0.0;
        for (int a = 0; a < s->b; ++a) {
            for (int k = 0; k < n; ++k) M[(r * s->b + a) * n + k] = 0.0;
            for (int z = 0; z < s->L[2]; ++z)
                for (int y = 0; y < s->L[1]; ++y)
                    for (int x = 0; x < s->L[0]; ++x) {
                        int64_t own = a * pl + ((int64_t)z * s->L[1] + y) * s->L[0] + x;
                        double sv[3] = {0, 0, 0};
                        for (int k = 0; k < n; ++k) {
                            sv[k] = S[k * nsite + own];
                            M[(r * s->b + a) * n + k] += sv[k];
                        }
                        for (int i = 0; i < s->nbond; ++i) {
                            if (s->bond[6 * i] != a) continue;
                            int64_t q = neighbour(s, i, x, y, z);
                            if (q < 0) continue;
                            const double *Ex = s->extra + 9 * i;
                            for (int k = 0; k < n; ++k)
                                for (int l = 0; l < n; ++l) {
                                    double Mkl = (k == l ? s->bondJ[i] : 0.0) + (s->has_extra[i] ? Ex[3 * k + l] : 0.0);
                                    e -= 0.5 * sv[k] * Mkl * S[l * nsite + q];
                                }
                        }
                        if (s->has_h)
                            for (int k = 0; k < n; ++k) e -= s->h[k] * sv[k];
                        if (s->has_cubic) {
                            double xx = sv[0], yy = sv[1], zz = sv[2];
                            e += (double)s->k1[own] * ((xx * yy) * (xx * yy) + (xx * zz) * (xx * zz) + (yy * zz) * (yy * zz)) +
                                 s->k2 * (xx * yy * zz) * (xx * yy * zz);
                        }
                        if (s->has_uniaxial) {
                            double se = 0.0;
                            for (int k = 0; k < n; ++k) se += sv[k] * s->du_axis[k];
                            e -= s->du * se * se;
                        }
                        if (s->has_site) {
                            double sn = 0.0;
                            for (int k = 0; k < n; ++k) sn += sv[k] * (double)s->site_axis[own * n + k];
                            e -= (double)s->site_ku[own] * sn * sn;
                        }
                    }
        }
        E[r] = e;
    }
}

/* Independent admissibility check of the colouring on the box: 1 when no bond joins equal colours. */
int oracle_colouring_ok(const oracle_sys *s) {
    int64_t pl = (int64_t)s->L[2] * s->L[1] * s->L[0];
    for (int i = 0; i < s->nbond; ++i) {
        int a = s->bond[6 * i], a2 = s->bond[6 * i + 1];
        for (int z = 0; z < s->L[2]; ++z)
            for (int y = 0; y < s->L[1]; ++y)
                for (int x = 0; x < s->L[0]; ++x) {
                    int64_t q = neighbour(s, i, x, y, z);
                    if (q < 0) continue;
                    int64_t rem = q - a2 * pl;
                    int qx = (int)(rem % s->L[0]), qy = (int)((rem / s->L[0]) % s->L[1]), qz = (int)(rem / ((int64_t)s->L[0] * s->L[1]));
                    if (colour_of(s, x, y, z, a) == colour_of(s, qx, qy, qz, a2)) return 0;
                }
    }
    return 1;
}

/* Raw Philox outputs for known-answer tests: out[4*j + q] for counters (c0[j], c1[j]) and key (k0, k1). */
void oracle_philox(uint32_t k0, uint32_t k1, const uint32_t *c0, const uint32_t *c1, const uint32_t *c2,
                   const uint32_t *c3, int64_t n, int rounds, uint32_t *out) {
    for (int64_t j = 0; j < n; ++j) {
        uint32_t c[4] = {c0[j], c1[j], c2[j], c3[j]};
        philox4x32(c, k0, k1, rounds);
        memcpy(out + 4 * j, c, 16);
    }
}

float oracle_u01(uint32_t x) { return philox_u01(x); }
int64_t oracle_sweep_seed(int64_t base, int64_t sweep) { return sweep_seed(base, sweep); }
