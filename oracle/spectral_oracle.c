/*
 * spectral_oracle.c -- CPU restatement of spectral-core (TEST INFRASTRUCTURE).
 *
 * This is the parity oracle and the timed CPU baseline; the product never
 * links or calls it.  Each function cites the reference file:line it
 * restates.  Build flags (oracle/Makefile): -O2 -ffp-contract=off, no
 * fast-math, so float arithmetic rounds exactly as the reference's Rust.
 *
 * Vector helpers restate glam 0.29's Vec3 (the reference's only math
 * dependency, not vendored): dot = (x*x' + y*y') + z*z', length =
 * sqrt(dot), normalize = v * (1/length), cross = (y z' - y' z, z x' - z' x,
 * x y' - x' y).  Parity against glam itself is unpinned at the bit level.
 */
#include "spectral_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* glam Vec3 restatement                                                     */
/* ------------------------------------------------------------------------- */
typedef struct { float x, y, z; } v3;

static inline v3 V(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 vl(const float *p) { return V(p[0], p[1], p[2]); }
static inline void vs(float *p, v3 a) { p[0] = a.x; p[1] = a.y; p[2] = a.z; }
static inline v3 add(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mul(v3 a, float s) { return V(a.x * s, a.y * s, a.z * s); }
static inline v3 divs(v3 a, float s) { return V(a.x / s, a.y / s, a.z / s); }
static inline v3 neg(v3 a) { return V(-a.x, -a.y, -a.z); }
static inline float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline float length(v3 a) { return sqrtf(dot(a, a)); }
static inline v3 normalize(v3 a) { return mul(a, 1.0f / length(a)); }
static inline v3 cross(v3 a, v3 b) {
    return V(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}

/* ------------------------------------------------------------------------- */
/* rng.rs:7-40                                                               */
/* ------------------------------------------------------------------------- */
uint32_t or_pcg32(uint32_t input) {
    /* rng.rs:7-11 */
    uint32_t state = input * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28) + 4u)) ^ state) * 277803737u;
    return (word >> 22) ^ word;
}

uint32_t or_mix2(uint32_t a, uint32_t b) { return or_pcg32(a ^ or_pcg32(b)); } /* rng.rs:38-40 */

uint32_t or_rng_new(uint32_t key, uint32_t seed) { return or_pcg32(key ^ or_pcg32(seed)); } /* rng.rs:19-22 */

static inline uint32_t next_u32(uint32_t *st) { /* rng.rs:25-28 */
    *st += 1u;
    return or_pcg32(*st);
}

static inline float next_f32(uint32_t *st) { /* rng.rs:31-34 */
    uint32_t bits = next_u32(st) >> 8;
    return (float)bits * (1.0f / (float)(1u << 24));
}

void or_rng_draws_u32(uint32_t key, uint32_t seed, uint32_t n, uint32_t *out) {
    uint32_t st = or_rng_new(key, seed);
    for (uint32_t i = 0; i < n; ++i) out[i] = next_u32(&st);
}

void or_rng_draws_f32(uint32_t key, uint32_t seed, uint32_t n, float *out) {
    uint32_t st = or_rng_new(key, seed);
    for (uint32_t i = 0; i < n; ++i) out[i] = next_f32(&st);
}

/* ------------------------------------------------------------------------- */
/* spectrum.rs:3-32                                                          */
/* ------------------------------------------------------------------------- */
#define LAMBDA_MIN 380.0f
#define LAMBDA_MAX 730.0f
#define LAMBDA_RANGE (LAMBDA_MAX - LAMBDA_MIN)
#define HERO_N 4

static inline float rem_euclid(float a, float b) {
    /* Rust f32::rem_euclid: r = a % b; if r < 0 { r + |b| } else { r } */
    float r = fmodf(a, b);
    return r < 0.0f ? r + fabsf(b) : r;
}

void or_hero_lambdas(float u, float out[4]) { /* spectrum.rs:20-32 */
    const float step = LAMBDA_RANGE / (float)HERO_N;
    for (int j = 0; j < HERO_N; ++j) {
        float off = rem_euclid(u * LAMBDA_RANGE + (float)j * step, LAMBDA_RANGE);
        out[j] = LAMBDA_MIN + off;
    }
}

/* ------------------------------------------------------------------------- */
/* sellmeier.rs:13-44 (f64 inside, f32 out)                                  */
/* ------------------------------------------------------------------------- */
float or_glass_n(uint32_t glass, float lambda_nm) {
    if (glass == 2u) { /* Water: Cauchy, sellmeier.rs:29-34 */
        double l_um = (double)lambda_nm / 1000.0;
        return (float)(1.3238 + 0.00314 / (l_um * l_um));
    }
    static const double B[2][3] = {{1.03961212, 0.231792344, 1.01046945},
                                   {1.73759695, 0.313747346, 1.89878101}};
    static const double C[2][3] = {{0.00600069867, 0.0200179144, 103.560653},
                                   {0.013188707, 0.0623068142, 155.23629}};
    const int g = glass == 1u ? 1 : 0;
    double l_um = (double)lambda_nm / 1000.0;
    double l2 = l_um * l_um;
    double s = 0.0;
    for (int k = 0; k < 3; ++k) s = s + B[g][k] * l2 / (l2 - C[g][k]);
    double n2 = 1.0 + s;
    return (float)sqrt(n2);
}

/* ------------------------------------------------------------------------- */
/* cie.rs:24-106                                                             */
/* ------------------------------------------------------------------------- */
#define TBL_N 71
static float g_cmf[TBL_N][3];
static float g_d65[TBL_N];
static float g_a[TBL_N];
static float g_start = 380.0f, g_step = 5.0f;
static int g_loaded = 0;

static int parse_csv(const char *path, int cols, float *dst, float *start, float *step) {
    /* Table::parse, cie.rs:24-45 */
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    char line[512];
    int n = 0;
    float first = NAN, last = NAN, st = NAN;
    while (fgets(line, sizeof line, f)) {
        char *p = line;
        while (*p == ' ' || *p == '\t') ++p;
        if (*p == '\n' || *p == '\r' || *p == 0) continue;
        if (n >= TBL_N) { fclose(f); return -2; }
        char *e;
        float nm = strtof(p, &e);
        p = e;
        for (int c = 0; c < cols; ++c) {
            while (*p == ',' || *p == ' ') ++p;
            dst[n * cols + c] = strtof(p, &e);
            p = e;
        }
        if (isnan(first)) first = nm;
        else if (isnan(st)) st = nm - last;
        last = nm;
        ++n;
    }
    fclose(f);
    if (n != TBL_N) return -3;
    if (start) *start = first;
    if (step) *step = st;
    return 0;
}

int or_load_tables(const char *dir) {
    char p[1024];
    snprintf(p, sizeof p, "%s/cie_1931_2deg.csv", dir);
    if (parse_csv(p, 3, &g_cmf[0][0], &g_start, &g_step)) return -1;
    snprintf(p, sizeof p, "%s/illuminant_d65.csv", dir);
    if (parse_csv(p, 1, g_d65, NULL, NULL)) return -2;
    snprintf(p, sizeof p, "%s/illuminant_a.csv", dir);
    if (parse_csv(p, 1, g_a, NULL, NULL)) return -3;
    g_loaded = 1;
    return 0;
}

static inline float clampf(float x, float lo, float hi) { /* Rust f32::clamp */
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}

/* Table::sample, cie.rs:47-61, for an n-column row table. */
static inline void table_sample(const float *vals, int cols, float nm, float *out) {
    float f = clampf((nm - g_start) / g_step, 0.0f, (float)(TBL_N - 1));
    int i = (int)floorf(f);
    float frac = f - (float)i;
    if (i + 1 >= TBL_N) {
        for (int c = 0; c < cols; ++c) out[c] = vals[(TBL_N - 1) * cols + c];
        return;
    }
    for (int c = 0; c < cols; ++c) {
        float a = vals[i * cols + c], b = vals[(i + 1) * cols + c];
        out[c] = a + (b - a) * frac;
    }
}

void or_cmf(float nm, float out[3]) { table_sample(&g_cmf[0][0], 3, nm, out); } /* cie.rs:79-82 */

float or_illuminant(uint32_t ill, float nm) { /* cie.rs:84-89; 2 = equal energy */
    float v = 1.0f;
    if (ill == 0u) table_sample(g_d65, 1, nm, &v);
    else if (ill == 1u) table_sample(g_a, 1, nm, &v);
    return v;
}

void or_integrate_illuminant(uint32_t ill, float out[3]) { /* cie.rs:92-106 */
    const float step = 5.0f;
    float x = 0, y = 0, z = 0, k = 0;
    float nm = LAMBDA_MIN;
    while (nm <= LAMBDA_MAX) {
        float s = or_illuminant(ill, nm);
        float c[3];
        or_cmf(nm, c);
        x += s * c[0] * step;
        y += s * c[1] * step;
        z += s * c[2] * step;
        k += s * c[1] * step;
        nm += step;
    }
    out[0] = x / k;
    out[1] = y / k;
    out[2] = z / k;
}

/* Photon power weight for a white beam of illuminant `ill` (extension; the
 * reference's photons carry power 1, which is ill = 2): S(lambda) / mean S
 * over the 71 table rows, the mean taken in double. */
float or_illuminant_scale(uint32_t ill) {
    if (ill != 0u && ill != 1u) return 1.0f;
    const float *t = ill == 0u ? g_d65 : g_a;
    double s = 0.0;
    for (int i = 0; i < TBL_N; ++i) s += (double)t[i];
    return (float)(1.0 / (s / (double)TBL_N));
}

void or_xyz_to_linear_srgb(const float xyz[3], float out[3]) { /* cie.rs:121-128 */
    float x = xyz[0], y = xyz[1], z = xyz[2];
    out[0] = 3.2406f * x - 1.5372f * y - 0.4986f * z;
    out[1] = -0.9689f * x + 1.8758f * y + 0.0415f * z;
    out[2] = 0.0557f * x - 0.2040f * y + 1.0570f * z;
}

void or_tonemap_to_u8(const float linear[3], float exposure, uint8_t out[3]) { /* cie.rs:131-144 */
    for (int c = 0; c < 3; ++c) {
        float v = fmaxf(linear[c] * exposure, 0.0f);
        float mapped = v / (1.0f + v);
        float gamma = mapped <= 0.0031308f ? 12.92f * mapped
                                           : 1.055f * powf(mapped, 1.0f / 2.4f) - 0.055f;
        out[c] = (uint8_t)(clampf(gamma, 0.0f, 1.0f) * 255.0f + 0.5f);
    }
}

/* ------------------------------------------------------------------------- */
/* geom.rs:35-175                                                            */
/* ------------------------------------------------------------------------- */
typedef struct {
    float t;
    v3 point;
    v3 normal;
    int front_face;
} hit_t;

static int sphere_intersect(v3 c, float radius, v3 o, v3 d, float t_min, float t_max, hit_t *h) {
    /* geom.rs:35-57 */
    v3 oc = sub(o, c);
    float a = dot(d, d);
    float half_b = dot(oc, d);
    float cc = dot(oc, oc) - radius * radius;
    float disc = half_b * half_b - a * cc;
    if (disc < 0.0f) return 0;
    float sq = sqrtf(disc);
    float t = (-half_b - sq) / a;
    if (t < t_min || t > t_max) {
        t = (-half_b + sq) / a;
        if (t < t_min || t > t_max) return 0;
    }
    v3 point = add(o, mul(d, t));
    v3 outward = divs(sub(point, c), radius);
    int front = dot(outward, d) < 0.0f;
    h->t = t;
    h->point = point;
    h->normal = front ? outward : neg(outward);
    h->front_face = front;
    return 1;
}

static int solid_intersect(const or_plane *pl, uint32_t n, v3 o, v3 d, float t_min, float t_max,
                           hit_t *h) {
    /* geom.rs:128-175 */
    float t_enter = -INFINITY, t_exit = INFINITY;
    v3 enter_n = V(0, 0, 0), exit_n = V(0, 0, 0);
    for (uint32_t i = 0; i < n; ++i) {
        v3 pn = vl(pl[i].normal);
        float denom = dot(pn, d);
        float dist = dot(pn, o) + pl[i].d;
        if (fabsf(denom) < 1e-8f) {
            if (dist > 0.0f) return 0;
            continue;
        }
        float t = -dist / denom;
        if (denom < 0.0f) {
            if (t > t_enter) { t_enter = t; enter_n = pn; }
        } else {
            if (t < t_exit) { t_exit = t; exit_n = pn; }
        }
    }
    if (t_enter > t_exit) return 0;
    float t_hit;
    v3 hn;
    if (t_enter > t_min) { t_hit = t_enter; hn = enter_n; }
    else if (t_exit > t_min) { t_hit = t_exit; hn = exit_n; }
    else return 0;
    if (t_hit > t_max) return 0;
    int front = dot(hn, d) < 0.0f;
    h->t = t_hit;
    h->point = add(o, mul(d, t_hit));
    h->normal = front ? hn : neg(hn);
    h->front_face = front;
    return 1;
}

static void hit_store(const hit_t *h, float out[8]) {
    out[0] = h->t;
    vs(out + 1, h->point);
    vs(out + 4, h->normal);
    out[7] = (float)h->front_face;
}

static hit_t hit_load(const float in[8]) {
    hit_t h;
    h.t = in[0];
    h.point = vl(in + 1);
    h.normal = vl(in + 4);
    h.front_face = in[7] != 0.0f;
    return h;
}

int or_sphere_intersect(const float center[3], float radius, const float o[3], const float d[3],
                        float t_min, float t_max, float hit_out[8]) {
    hit_t h;
    if (!sphere_intersect(vl(center), radius, vl(o), vl(d), t_min, t_max, &h)) return 0;
    hit_store(&h, hit_out);
    return 1;
}

int or_solid_intersect(const or_plane *planes, uint32_t n, const float o[3], const float d[3],
                       float t_min, float t_max, float hit_out[8]) {
    hit_t h;
    if (!solid_intersect(planes, n, vl(o), vl(d), t_min, t_max, &h)) return 0;
    hit_store(&h, hit_out);
    return 1;
}

/* Scene::intersect, scene.rs:64-78; returns the material index or -1. */
static int scene_intersect(const or_scene *s, v3 o, v3 d, hit_t *best) {
    int mat = -1;
    float closest = INFINITY;
    for (uint32_t i = 0; i < s->n_prims; ++i) {
        const or_prim *p = &s->prims[i];
        hit_t h;
        int ok;
        if (p->kind == 0u) ok = sphere_intersect(vl(p->center), p->radius, o, d, 1e-3f, closest, &h);
        else if (p->kind == 1u)
            ok = solid_intersect(s->planes + p->plane_start, p->plane_count, o, d, 1e-3f, closest, &h);
        else ok = 0;
        if (ok) {
            closest = h.t;
            *best = h;
            mat = (int)p->material;
        }
    }
    return mat;
}

int or_scene_intersect(const or_scene *s, const float o[3], const float d[3], float hit_out[8],
                       uint32_t *mat_out) {
    hit_t h;
    int m = scene_intersect(s, vl(o), vl(d), &h);
    if (m < 0) return 0;
    hit_store(&h, hit_out);
    if (mat_out) *mat_out = (uint32_t)m;
    return 1;
}

/* ------------------------------------------------------------------------- */
/* optics.rs:8-49                                                            */
/* ------------------------------------------------------------------------- */
static inline v3 reflect(v3 d, v3 n) { return sub(d, mul(n, 2.0f * dot(d, n))); } /* optics.rs:8-10 */

static int refract(v3 d, v3 n, float n1, float n2, v3 *out) { /* optics.rs:21-35 */
    float eta = n1 / n2;
    v3 nn = n;
    float cos_i = -dot(d, n);
    if (cos_i < 0.0f) {
        nn = neg(n);
        cos_i = -cos_i;
    }
    float k = 1.0f - eta * eta * (1.0f - cos_i * cos_i);
    if (k < 0.0f) return 0;
    float cos_t = sqrtf(k);
    *out = normalize(add(mul(d, eta), mul(nn, eta * cos_i - cos_t)));
    return 1;
}

float or_fresnel_reflectance(float cos_i, float n1, float n2) { /* optics.rs:39-49 */
    cos_i = clampf(cos_i, 0.0f, 1.0f);
    float r = n1 / n2;
    float sin_t2 = r * r * (1.0f - cos_i * cos_i);
    if (sin_t2 >= 1.0f) return 1.0f;
    float cos_t = sqrtf(1.0f - sin_t2);
    float qs = (n1 * cos_i - n2 * cos_t) / (n1 * cos_i + n2 * cos_t);
    float qp = (n1 * cos_t - n2 * cos_i) / (n1 * cos_t + n2 * cos_i);
    float rs = qs * qs;
    float rp = qp * qp;
    return 0.5f * (rs + rp);
}

void or_reflect(const float d[3], const float n[3], float out[3]) { vs(out, reflect(vl(d), vl(n))); }

int or_refract(const float d[3], const float n[3], float n1, float n2, float out[3]) {
    v3 r;
    if (!refract(vl(d), vl(n), n1, n2, &r)) return 0;
    vs(out, r);
    return 1;
}

/* ------------------------------------------------------------------------- */
/* material.rs:30-82                                                         */
/* ------------------------------------------------------------------------- */
static v3 cosine_hemisphere(v3 normal, uint32_t *rng) { /* material.rs:70-82 */
    float u1 = next_f32(rng);
    float u2 = next_f32(rng);
    float r = sqrtf(u1);
    float phi = 6.2831855f * u2; /* 2.0 * PI folded in f32 */
    float x = r * cosf(phi);
    float y = r * sinf(phi);
    float z = sqrtf(fmaxf(1.0f - u1, 0.0f));
    v3 a = fabsf(normal.x) > 0.9f ? V(0, 1, 0) : V(1, 0, 0);
    v3 t = normalize(cross(normal, a));
    v3 b = cross(normal, t);
    return normalize(add(add(mul(t, x), mul(b, y)), mul(normal, z)));
}

/* Returns 1 when a scatter happened (always, for the two materials);
 * *tir is set when refract() found total internal reflection. */
static int scatter(const or_material *m, v3 wo, const hit_t *h, float n_hero, uint32_t *rng,
                   v3 *dir, float *weight, int *tir) {
    *tir = 0;
    if (m->kind == 0u) {
        *dir = cosine_hemisphere(h->normal, rng);
        *weight = m->reflectance;
        return 1;
    }
    float cos_i = fabsf(-dot(wo, h->normal));
    float n1 = h->front_face ? 1.0f : n_hero;
    float n2 = h->front_face ? n_hero : 1.0f;
    float r = or_fresnel_reflectance(cos_i, n1, n2);
    *weight = 1.0f;
    if (next_f32(rng) < r) {
        *dir = reflect(wo, h->normal);
    } else if (!refract(wo, h->normal, n1, n2, dir)) {
        *dir = reflect(wo, h->normal);
        *tir = 1;
    }
    return 1;
}

int or_material_scatter(const or_material *m, const float wo[3], const float hit[8], float n_hero,
                        uint32_t *rng_state, float dir_out[3], float *weight_out) {
    hit_t h = hit_load(hit);
    v3 d;
    int tir;
    int ok = scatter(m, vl(wo), &h, n_hero, rng_state, &d, weight_out, &tir);
    vs(dir_out, d);
    return ok;
}

/* n_hero for a material at lambda (tracer.rs:76-79; dispersion-off uses
 * n(550 nm), as the GPU mirror's `spectral == 0`). */
static inline float material_n(const or_material *m, float lambda, uint32_t dispersion) {
    if (m->kind != 1u) return 1.0f;
    return or_glass_n(m->glass, dispersion ? lambda : 550.0f);
}

/* ------------------------------------------------------------------------- */
/* camera.rs:15-69                                                           */
/* ------------------------------------------------------------------------- */
void or_camera_look_at(const float origin[3], const float target[3], const float up[3],
                       float vfov_deg, float aspect, or_camera *out) {
    const float RADS = 3.14159265358979323846f / 180.0f; /* f32::to_radians */
    v3 o = vl(origin), tg = vl(target), u0 = vl(up);
    float theta = vfov_deg * RADS;
    float hh = tanf(theta / 2.0f);
    float viewport_h = 2.0f * hh;
    float viewport_w = aspect * viewport_h;
    v3 w = normalize(sub(o, tg));
    v3 u = normalize(cross(u0, w));
    v3 v = cross(w, u);
    v3 horizontal = mul(u, viewport_w);
    v3 vertical = mul(v, viewport_h);
    v3 ll = sub(sub(sub(o, divs(horizontal, 2.0f)), divs(vertical, 2.0f)), w);
    vs(out->origin, o);
    vs(out->lower_left, ll);
    vs(out->horizontal, horizontal);
    vs(out->vertical, vertical);
}

static inline v3 primary_ray(const or_camera *c, float s, float t) { /* camera.rs:41-45 */
    v3 d = sub(add(add(vl(c->lower_left), mul(vl(c->horizontal), s)), mul(vl(c->vertical), t)),
               vl(c->origin));
    return normalize(d);
}

void or_camera_primary_ray(const or_camera *c, float s, float t, float dir_out[3]) {
    vs(dir_out, primary_ray(c, s, t));
}

/* Camera::project, camera.rs:50-69, with the per-camera terms hoisted. */
typedef struct {
    v3 o, u, v, w;
    float vw, vh;
} proj_t;

static proj_t proj_make(const or_camera *c) {
    proj_t p;
    v3 hz = vl(c->horizontal), vt = vl(c->vertical);
    p.o = vl(c->origin);
    p.u = normalize(hz);
    p.v = normalize(vt);
    p.vw = length(hz);
    p.vh = length(vt);
    p.w = sub(sub(sub(p.o, mul(hz, 0.5f)), mul(vt, 0.5f)), vl(c->lower_left));
    return p;
}

static inline int project(const proj_t *P, v3 p, float *s, float *t, float *depth) {
    v3 dir = sub(p, P->o);
    float c = dot(dir, P->w);
    if (c >= -1e-6f) return 0;
    float ss = 0.5f + dot(dir, P->u) / (-c * P->vw);
    float tt = 0.5f + dot(dir, P->v) / (-c * P->vh);
    if (ss >= 0.0f && ss < 1.0f && tt >= 0.0f && tt < 1.0f) {
        *s = ss;
        *t = tt;
        *depth = -c;
        return 1;
    }
    return 0;
}

int or_camera_project(const or_camera *c, const float p[3], float out[3]) {
    proj_t P = proj_make(c);
    return project(&P, vl(p), &out[0], &out[1], &out[2]);
}

/* ------------------------------------------------------------------------- */
/* scene.rs:44-50 (+ sun-disk extension)                                     */
/* ------------------------------------------------------------------------- */
static inline float background_radiance(const or_scene_env *e, v3 dir) {
    float sky = (e->has_horizon && dir.y > 0.0f) ? e->horizon : 0.0f;
    float r = e->background + sky;
    if (e->sun_radiance > 0.0f && dot(dir, vl(e->sun_dir)) >= e->sun_cos) r = r + e->sun_radiance;
    return r;
}

float or_background_radiance(const or_scene_env *env, const float dir[3]) {
    return background_radiance(env, vl(dir));
}

/* ------------------------------------------------------------------------- */
/* tracer.rs:44-141                                                          */
/* ------------------------------------------------------------------------- */
static void camera_sample(const or_scene *s, const or_camera *cam, uint32_t w, uint32_t h,
                          uint32_t px, uint32_t py, uint32_t sample_idx, uint32_t seed,
                          const or_opts *o, float xyz[3], uint64_t *segs) {
    /* CpuTracer::sample_pixel, tracer.rs:44-111 */
    uint32_t pixel = py * w + px;
    uint32_t key = or_mix2(pixel, sample_idx);
    uint32_t rng = or_rng_new(key, seed);
    float s_ = ((float)px + next_f32(&rng)) / (float)w;
    float t_ = 1.0f - ((float)py + next_f32(&rng)) / (float)h;
    float lambda[HERO_N];
    or_hero_lambdas(next_f32(&rng), lambda);
    const float pdf = 1.0f / LAMBDA_RANGE;
    float rad[HERO_N] = {0, 0, 0, 0};
    float tp[HERO_N] = {1, 1, 1, 1};
    uint32_t valid = HERO_N;
    v3 ro = vl(cam->origin);
    v3 rd = primary_ray(cam, s_, t_);
    const uint32_t max_depth = o->max_depth;
    for (uint32_t b = 0; b < max_depth; ++b) {
        if (o->rr_depth && b >= o->rr_depth) {
            /* Russian roulette (extension): survive with q = clamp(max tp). */
            float q = tp[0];
            for (int k = 1; k < HERO_N; ++k) q = fmaxf(q, tp[k]);
            q = clampf(q, 0.05f, 0.95f);
            if (next_f32(&rng) >= q) break;
            for (int k = 0; k < HERO_N; ++k) tp[k] = tp[k] / q;
        }
        if (segs) ++*segs;
        hit_t hit;
        int mi = scene_intersect(s, ro, rd, &hit);
        if (mi < 0) {
            float bg = background_radiance(&s->env, rd);
            for (int k = 0; k < HERO_N; ++k) {
                float sp = or_illuminant(o->illuminant, lambda[k]);
                rad[k] += tp[k] * bg * sp;
            }
            break;
        }
        const or_material *m = &s->mats[mi];
        float n_hero = material_n(m, lambda[0], o->dispersion);
        v3 nd;
        float wgt;
        int tir;
        scatter(m, rd, &hit, n_hero, &rng, &nd, &wgt, &tir);
        for (int k = 0; k < HERO_N; ++k) tp[k] *= wgt;
        if (m->kind == 1u && valid > 1) {
            tp[1] = tp[2] = tp[3] = 0.0f;
            valid = 1;
        }
        ro = hit.point;
        rd = nd;
    }
    xyz[0] = xyz[1] = xyz[2] = 0.0f;
    for (int k = 0; k < HERO_N; ++k) {
        float c[3];
        or_cmf(lambda[k], c);
        float wk = rad[k] / pdf / (float)valid;
        xyz[0] += c[0] * wk;
        xyz[1] += c[1] * wk;
        xyz[2] += c[2] * wk;
    }
}

void or_camera_sample(const or_scene *s, const or_camera *cam, uint32_t w, uint32_t h,
                      uint32_t px, uint32_t py, uint32_t sample_idx, uint32_t seed,
                      const or_opts *o, float out[3]) {
    camera_sample(s, cam, w, h, px, py, sample_idx, seed, o, out, NULL);
}

typedef struct {
    const or_scene *s;
    const or_camera *cam;
    uint32_t w, h, spp, base, seed;
    const or_opts *o;
    float *sum;
    uint32_t p0, p1;
    uint64_t segs;
} cam_job;

static void *cam_worker(void *arg) {
    /* Renderer::accumulate, tracer.rs:115-141: in-order spp sum per pixel. */
    cam_job *j = (cam_job *)arg;
    for (uint32_t i = j->p0; i < j->p1; ++i) {
        uint32_t px = i % j->w, py = i / j->w;
        float acc[3] = {0, 0, 0};
        for (uint32_t si = 0; si < j->spp; ++si) {
            float c[3];
            camera_sample(j->s, j->cam, j->w, j->h, px, py, j->base + si, j->seed, j->o, c, &j->segs);
            acc[0] += c[0];
            acc[1] += c[1];
            acc[2] += c[2];
        }
        j->sum[3 * i + 0] += acc[0];
        j->sum[3 * i + 1] += acc[1];
        j->sum[3 * i + 2] += acc[2];
    }
    return NULL;
}

void or_camera_accumulate(const or_scene *s, const or_camera *cam, uint32_t w, uint32_t h,
                          uint32_t spp, uint32_t sample_base, uint32_t seed, const or_opts *o,
                          float *sum, uint32_t threads, or_stats *stats) {
    uint32_t npx = w * h;
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    cam_job jobs[256];
    pthread_t tid[256];
    for (uint32_t t = 0; t < threads; ++t) {
        cam_job *j = &jobs[t];
        j->s = s; j->cam = cam; j->w = w; j->h = h; j->spp = spp; j->base = sample_base;
        j->seed = seed; j->o = o; j->sum = sum; j->segs = 0;
        j->p0 = (uint32_t)((uint64_t)npx * t / threads);
        j->p1 = (uint32_t)((uint64_t)npx * (t + 1) / threads);
    }
    if (threads == 1) cam_worker(&jobs[0]);
    else {
        for (uint32_t t = 0; t < threads; ++t) pthread_create(&tid[t], NULL, cam_worker, &jobs[t]);
        for (uint32_t t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
    }
    if (stats) {
        for (uint32_t t = 0; t < threads; ++t) stats->segments += jobs[t].segs;
        stats->paths += (uint64_t)npx * spp;
    }
}

/* ------------------------------------------------------------------------- */
/* lighttrace.rs:42-166                                                      */
/* ------------------------------------------------------------------------- */
typedef struct {
    v3 corner, u, v, normal;
    uint32_t w, h;
} screen_t;

static screen_t screen_make(const or_screen *s) { /* Screen::new, lighttrace.rs:27-34 */
    screen_t S;
    S.corner = vl(s->corner);
    S.u = vl(s->u);
    S.v = vl(s->v);
    S.normal = normalize(cross(S.u, S.v));
    S.w = s->width;
    S.h = s->height;
    return S;
}

static int screen_coords(const screen_t *S, v3 o, v3 d, float *s, float *r, float *t) {
    /* lighttrace.rs:42-60 */
    float denom = dot(d, S->normal);
    if (fabsf(denom) < 1e-8f) return 0;
    float tt = dot(sub(S->corner, o), S->normal) / denom;
    if (tt <= 1e-4f) return 0;
    v3 p = add(o, mul(d, tt));
    v3 rel = sub(p, S->corner);
    float ss = dot(rel, S->u) / dot(S->u, S->u);
    float rr = dot(rel, S->v) / dot(S->v, S->v);
    if (ss >= 0.0f && ss < 1.0f && rr >= 0.0f && rr < 1.0f) {
        *s = ss;
        *r = rr;
        *t = tt;
        return 1;
    }
    return 0;
}

/* The photon spectra of the hero-comb option (hero_lanes = 4; an extension,
 * no reference counterpart): a photon carries the stratified comb of
 * spectrum.rs:20-32 drawn from its wavelength draw u, lane k weighing 1/4,
 * until it scatters off a dielectric, where the companions collapse onto
 * the hero lane (as a camera path's do, tracer.rs:86-91).  Its deposits are
 * spec * (material weights * geometry), spec = sum_k cmf(lambda_k) * s_k *
 * 0.25 (k in order) before the collapse and cmf(lambda_0) * s_0 after, s_k
 * the beam's SPD weight (1 for equal energy). */
static void comb_spec(float u, uint32_t ill, float ill_scale, float spec[3]) {
    float lam[4];
    or_hero_lambdas(u, lam);
    spec[0] = spec[1] = spec[2] = 0.0f;
    for (int k = 0; k < 4; ++k) {
        float c[3];
        or_cmf(lam[k], c);
        float s = ill <= 1u ? or_illuminant(ill, lam[k]) * ill_scale : 1.0f;
        for (int i = 0; i < 3; ++i) spec[i] += c[i] * s * 0.25f;
    }
}

static void hero_spec(float lambda0, uint32_t ill, float ill_scale, float spec[3]) {
    float c[3];
    or_cmf(lambda0, c);
    float s = ill <= 1u ? or_illuminant(ill, lambda0) * ill_scale : 1.0f;
    for (int i = 0; i < 3; ++i) spec[i] = c[i] * s;
}

static void screen_splat(const screen_t *S, float *sum, float s, float r, const float c[3],
                         float power) {
    /* Screen::texel_bilinear + splat, lighttrace.rs:65-91 (c = cmf(lambda), or
     * the photon's comb spectrum) */
    float fx = clampf(s * (float)S->w - 0.5f, 0.0f, (float)S->w - 1.0f);
    float fy = clampf((1.0f - r) * (float)S->h - 0.5f, 0.0f, (float)S->h - 1.0f);
    uint32_t x0 = (uint32_t)floorf(fx), y0 = (uint32_t)floorf(fy);
    uint32_t x1 = x0 + 1 < S->w - 1 ? x0 + 1 : S->w - 1;
    uint32_t y1 = y0 + 1 < S->h - 1 ? y0 + 1 : S->h - 1;
    float dx = fx - (float)x0, dy = fy - (float)y0;
    uint32_t xs[4] = {x0, x1, x0, x1}, ys[4] = {y0, y0, y1, y1};
    float ws[4] = {(1.0f - dx) * (1.0f - dy), dx * (1.0f - dy), (1.0f - dx) * dy, dx * dy};
    for (int k = 0; k < 4; ++k) {
        uint32_t i = ys[k] * S->w + xs[k];
        sum[3 * i + 0] += c[0] * power * ws[k];
        sum[3 * i + 1] += c[1] * power * ws[k];
        sum[3 * i + 2] += c[2] * power * ws[k];
    }
}

/* trace_photon, lighttrace.rs:125-153.  Returns 1 if it landed.  comb (may
 * be NULL): set while the photon carries the hero comb, cleared at its
 * first dielectric scatter. */
static int trace_photon(const or_scene *sc, const screen_t *S, v3 o, v3 d, float lambda,
                        float power0, uint32_t *rng, const or_opts *op, float *s, float *r,
                        float *power_out, uint64_t *segs, int *comb) {
    float power = power0;
    for (uint32_t b = 0; b < op->max_depth; ++b) {
        if (segs) ++*segs;
        hit_t hit;
        int mi = scene_intersect(sc, o, d, &hit);
        float ss, rr, st;
        int on_screen = screen_coords(S, o, d, &ss, &rr, &st);
        float scene_t = mi >= 0 ? hit.t : INFINITY;
        float screen_t = on_screen ? st : INFINITY;
        if (screen_t < scene_t) {
            *s = ss;
            *r = rr;
            *power_out = power;
            return 1;
        }
        if (mi < 0) return 0;
        const or_material *m = &sc->mats[mi];
        float n_hero = material_n(m, lambda, op->dispersion);
        v3 nd;
        float wgt;
        int tir;
        scatter(m, d, &hit, n_hero, rng, &nd, &wgt, &tir);
        power *= wgt;
        if (comb && m->kind == 1u) *comb = 0;
        o = hit.point;
        d = nd;
    }
    return 0;
}

/* trace_photon at a caller-chosen wavelength (lighttrace.rs:125-153). */
int or_lighttrace_photon(const or_scene *sc, const or_screen *scr, const float o[3], const float d[3],
                         float lambda, uint32_t *rng, const or_opts *op, float out[3]) {
    screen_t S = screen_make(scr);
    return trace_photon(sc, &S, vl(o), vl(d), lambda, 1.0f, rng, op, &out[0], &out[1], &out[2], NULL, NULL);
}

typedef struct {
    const or_scene *sc;
    const screen_t *S;
    const or_beam *beam;
    uint32_t i0, i1, seed;
    const or_opts *o;
    float *sum;
    float ill_scale;
    uint64_t segs, splats;
} lt_job;

static void *lt_worker(void *arg) {
    /* lighttrace::trace, lighttrace.rs:157-166 */
    lt_job *j = (lt_job *)arg;
    v3 bc = vl(j->beam->corner), bu = vl(j->beam->u), bv = vl(j->beam->v), bd = vl(j->beam->dir);
    for (uint32_t i = j->i0; i < j->i1; ++i) {
        uint32_t rng = or_rng_new(i, j->seed);
        float a = next_f32(&rng);
        float b = next_f32(&rng);
        v3 origin = add(add(bc, mul(bu, a)), mul(bv, b));
        float u = next_f32(&rng);
        float lambda = LAMBDA_MIN + LAMBDA_RANGE * u;
        const int hero = j->o->hero_lanes == 4u;
        float power0 = hero ? 1.0f
                            : (j->o->illuminant <= 1u ? or_illuminant(j->o->illuminant, lambda) * j->ill_scale
                                                      : 1.0f);
        float s, r, p;
        int comb = hero;
        if (trace_photon(j->sc, j->S, origin, bd, lambda, power0, &rng, j->o, &s, &r, &p, &j->segs,
                         hero ? &comb : NULL)) {
            float c[3];
            if (!hero) or_cmf(lambda, c);
            else if (comb) comb_spec(u, j->o->illuminant, j->ill_scale, c);
            else hero_spec(lambda, j->o->illuminant, j->ill_scale, c);
            screen_splat(j->S, j->sum, s, r, c, p);
            ++j->splats;
        }
    }
    return NULL;
}

/* Threads > 1 split the photon range into per-thread films reduced in
 * thread order (the all-core baseline variant; sums then differ from the
 * single-threaded order by float rounding only). */
void or_lighttrace(const or_scene *s, const or_screen *scr, const or_beam *beam, uint32_t n,
                   uint32_t photon_base, uint32_t seed, const or_opts *o, float *sum,
                   uint32_t threads, or_stats *stats) {
    screen_t S = screen_make(scr);
    size_t film = (size_t)3 * scr->width * scr->height;
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    lt_job jobs[256];
    pthread_t tid[256];
    float ill_scale = or_illuminant_scale(o->illuminant);
    for (uint32_t t = 0; t < threads; ++t) {
        lt_job *j = &jobs[t];
        j->sc = s; j->S = &S; j->beam = beam; j->seed = seed; j->o = o;
        j->i0 = photon_base + (uint32_t)((uint64_t)n * t / threads);
        j->i1 = photon_base + (uint32_t)((uint64_t)n * (t + 1) / threads);
        j->sum = threads == 1 ? sum : (float *)calloc(film, sizeof(float));
        j->ill_scale = ill_scale;
        j->segs = j->splats = 0;
    }
    if (threads == 1) lt_worker(&jobs[0]);
    else {
        for (uint32_t t = 0; t < threads; ++t) pthread_create(&tid[t], NULL, lt_worker, &jobs[t]);
        for (uint32_t t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
        for (uint32_t t = 0; t < threads; ++t) {
            for (size_t k = 0; k < film; ++k) sum[k] += jobs[t].sum[k];
            free(jobs[t].sum);
        }
    }
    if (stats) {
        for (uint32_t t = 0; t < threads; ++t) {
            stats->segments += jobs[t].segs;
            stats->splats += jobs[t].splats;
        }
        stats->paths += n;
    }
}

/* ------------------------------------------------------------------------- */
/* volume.rs:22-329                                                          */
/* ------------------------------------------------------------------------- */
float or_phase_hg(float g, float cos_theta) { /* volume.rs:88-92 */
    float g2 = g * g;
    float denom = powf(fmaxf(1.0f + g2 - 2.0f * g * cos_theta, 1e-6f), 1.5f);
    return (1.0f - g2) / (12.566371f * denom); /* 4.0 * PI folded in f32 */
}

#define R_BEAM 0.03f

static void splat_transverse(float *img, uint32_t width, uint32_t height, float s, float t,
                             const float color[3], float radius_px) {
    /* volume.rs:178-218 */
    float cx = s * (float)width - 0.5f;
    float cy = (1.0f - t) * (float)height - 0.5f;
    float r = fmaxf(radius_px, 0.5f);
    long x0 = (long)fmaxf(floorf(cx - r), 0.0f);
    long x1 = (long)ceilf(cx + r);
    if (x1 < 0) x1 = 0;
    if (x1 > (long)width - 1) x1 = (long)width - 1;
    long y0 = (long)fmaxf(floorf(cy - r), 0.0f);
    long y1 = (long)ceilf(cy + r);
    if (y1 < 0) y1 = 0;
    if (y1 > (long)height - 1) y1 = (long)height - 1;
    float wsum = 0.0f;
    for (long py = y0; py <= y1; ++py)
        for (long px = x0; px <= x1; ++px) {
            float dx = (float)px - cx, dy = (float)py - cy;
            wsum += fmaxf(1.0f - sqrtf(dx * dx + dy * dy) / r, 0.0f);
        }
    if (wsum <= 0.0f) {
        long px = (long)roundf(cx), py = (long)roundf(cy);
        if (px < 0) px = 0;
        if (px > (long)width - 1) px = (long)width - 1;
        if (py < 0) py = 0;
        if (py > (long)height - 1) py = (long)height - 1;
        size_t idx = (size_t)py * width + (size_t)px;
        img[3 * idx + 0] += color[0];
        img[3 * idx + 1] += color[1];
        img[3 * idx + 2] += color[2];
        return;
    }
    float inv = 1.0f / wsum;
    for (long py = y0; py <= y1; ++py)
        for (long px = x0; px <= x1; ++px) {
            float dx = (float)px - cx, dy = (float)py - cy;
            float w = fmaxf(1.0f - sqrtf(dx * dx + dy * dy) / r, 0.0f) * inv;
            if (w > 0.0f) {
                size_t idx = (size_t)py * width + (size_t)px;
                img[3 * idx + 0] += color[0] * w;
                img[3 * idx + 1] += color[1] * w;
                img[3 * idx + 2] += color[2] * w;
            }
        }
}

typedef struct {
    const or_scene *sc;
    const or_camera *cam;
    const or_beam *beam;
    uint32_t w, h, i0, i1, seed, est;
    const or_vol_params *vp;
    const float *zbuf;
    const or_opts *o;
    float *img;
    float ill_scale;
    uint64_t segs, splats;
    /* beam estimator, test hook: when box_group > 0, the largest pixel area
     * of the bounding box of the footprints of box_group consecutive steps
     * of one segment (steps k0 .. k0+box_group-1, those on the film and in
     * front of the z-buffer), footprints clamped to the film as
     * splat_transverse clamps them (volume.rs:178-218). */
    uint32_t box_group;
    uint64_t max_box;
} vol_job;

/* The footprint splat_transverse walks (volume.rs:180-190) joined into a
 * group's box, and the box's area folded into the largest seen. */
static void box_add(vol_job *j, int box[4], float s, float t, float radius_px) {
    float cx = s * (float)j->w - 0.5f, cy = (1.0f - t) * (float)j->h - 0.5f;
    float r = fmaxf(radius_px, 0.5f);
    int x0 = (int)fmaxf(floorf(cx - r), 0.0f), y0 = (int)fmaxf(floorf(cy - r), 0.0f);
    int x1 = (int)ceilf(cx + r), y1 = (int)ceilf(cy + r);
    x1 = x1 < 0 ? 0 : (x1 > (int)j->w - 1 ? (int)j->w - 1 : x1);
    y1 = y1 < 0 ? 0 : (y1 > (int)j->h - 1 ? (int)j->h - 1 : y1);
    if (x0 < box[0]) box[0] = x0;
    if (x1 > box[1]) box[1] = x1;
    if (y0 < box[2]) box[2] = y0;
    if (y1 > box[3]) box[3] = y1;
}

static void box_close(vol_job *j, int box[4]) {
    if (box[1] >= 0 && box[1] >= box[0] && box[3] >= box[2]) {
        uint64_t a = (uint64_t)(box[1] - box[0] + 1) * (uint64_t)(box[3] - box[2] + 1);
        if (a > j->max_box) j->max_box = a;
    }
    box[0] = INT32_MAX; box[1] = -1; box[2] = INT32_MAX; box[3] = -1;
}

static inline uint32_t pix_index(float s, float t, uint32_t width, uint32_t height) {
    /* volume.rs:134-135: float -> usize truncation, then min(dim-1) */
    uint32_t px = (uint32_t)(s * (float)width);
    uint32_t py = (uint32_t)((1.0f - t) * (float)height);
    if (px > width - 1) px = width - 1;
    if (py > height - 1) py = height - 1;
    return py * width + px;
}

static void *vol_worker(void *arg) {
    vol_job *j = (vol_job *)arg;
    const or_scene *sc = j->sc;
    const or_vol_params *vp = j->vp;
    proj_t P = proj_make(j->cam);
    v3 cam_o = vl(j->cam->origin);
    v3 bc = vl(j->beam->corner), bu = vl(j->beam->u), bv = vl(j->beam->v), bd = vl(j->beam->dir);
    const float INV_4PI = 1.0f / 12.566371f; /* volume.rs:36 */
    const uint32_t SEG = 4;                  /* SEG_SAMPLES, volume.rs:35,116 */
    for (uint32_t i = j->i0; i < j->i1; ++i) {
        uint32_t rng = or_rng_new(i, j->seed);
        float a = next_f32(&rng);
        float b = next_f32(&rng);
        v3 ro = add(add(bc, mul(bu, a)), mul(bv, b));
        float u = next_f32(&rng);
        float lambda = LAMBDA_MIN + LAMBDA_RANGE * u;
        float cmf[3];
        v3 rd = bd;
        float power;
        int comb = j->o->hero_lanes == 4u;  /* the hero comb's lanes still live */
        if (comb) {
            comb_spec(u, j->o->illuminant, j->ill_scale, cmf);  /* spectra in cmf, power = material weights */
            power = 1.0f;
        } else {
            or_cmf(lambda, cmf);
            power = j->o->illuminant <= 1u ? or_illuminant(j->o->illuminant, lambda) * j->ill_scale : 1.0f;
        }
        for (uint32_t bnc = 0; bnc < j->o->max_depth; ++bnc) {
            ++j->segs;
            hit_t hit;
            int mi = scene_intersect(sc, ro, rd, &hit);
            float seg_len = mi >= 0 ? hit.t : vp->max_dist;
            if (j->est == 0u) {
                /* point estimator, volume.rs:130-150 */
                for (uint32_t k = 0; k < SEG; ++k) {
                    float dist = next_f32(&rng) * seg_len;
                    v3 p = add(ro, mul(rd, dist));
                    float s, t, dep;
                    if (!project(&P, p, &s, &t, &dep)) continue;
                    uint32_t idx = pix_index(s, t, j->w, j->h);
                    v3 to = sub(cam_o, p);
                    float d = fmaxf(length(to), 1e-3f);
                    float zb = j->zbuf ? j->zbuf[idx] : INFINITY;
                    if (d < zb) {
                        float cos_theta = dot(rd, divs(to, d));
                        float phase = or_phase_hg(vp->g, cos_theta);
                        float trans = expf(-vp->sigma_t * d);
                        float contrib = power * vp->sigma_s * phase * trans / (d * d) * seg_len / (float)SEG;
                        j->img[3 * idx + 0] += cmf[0] * contrib;
                        j->img[3 * idx + 1] += cmf[1] * contrib;
                        j->img[3 * idx + 2] += cmf[2] * contrib;
                        ++j->splats;
                    }
                }
            } else if (j->est == 2u) {
                /* legacy isotropic estimator, volume.rs:51-64 */
                for (uint32_t k = 0; k < SEG; ++k) {
                    float dist = next_f32(&rng) * seg_len;
                    v3 p = add(ro, mul(rd, dist));
                    float s, t, dep;
                    if (!project(&P, p, &s, &t, &dep)) continue;
                    uint32_t idx = pix_index(s, t, j->w, j->h);
                    float contrib = power * vp->sigma_s * seg_len / (float)SEG * INV_4PI;
                    j->img[3 * idx + 0] += cmf[0] * contrib;
                    j->img[3 * idx + 1] += cmf[1] * contrib;
                    j->img[3 * idx + 2] += cmf[2] * contrib;
                    ++j->splats;
                }
            } else {
                /* beam estimator, volume.rs:263-311 */
                v3 p_end = add(ro, mul(rd, seg_len));
                float s0, t0, d0, s1, t1, d1;
                uint32_t m;
                if (project(&P, ro, &s0, &t0, &d0) && project(&P, p_end, &s1, &t1, &d1)) {
                    float dx = (s1 - s0) * (float)j->w;
                    float dy = (t1 - t0) * (float)j->h;
                    float mf = ceilf(sqrtf(dx * dx + dy * dy) * 2.0f);
                    uint64_t mm = (uint64_t)mf + 1u;
                    m = (uint32_t)(mm < 1 ? 1 : (mm > 4096 ? 4096 : mm));
                } else {
                    uint64_t mm = (uint64_t)ceilf(seg_len * 80.0f);
                    m = (uint32_t)(mm < 1 ? 1 : (mm > 4096 ? 4096 : mm));
                }
                float ds = seg_len / (float)m;
                int box[4] = {INT32_MAX, -1, INT32_MAX, -1};
                for (uint32_t k = 0; k < m; ++k) {
                    if (j->box_group && k > 0 && k % j->box_group == 0u) box_close(j, box);
                    float dist = ((float)k + 0.5f) * ds;
                    v3 p = add(ro, mul(rd, dist));
                    float s, t, dep;
                    if (!project(&P, p, &s, &t, &dep)) continue;
                    uint32_t idx = pix_index(s, t, j->w, j->h);
                    v3 to = sub(cam_o, p);
                    float d = fmaxf(length(to), 1e-3f);
                    float zb = j->zbuf ? j->zbuf[idx] : INFINITY;
                    if (!(d < zb)) continue;
                    v3 view = divs(to, d);
                    float cos_theta = dot(rd, view);
                    float phase = or_phase_hg(vp->g, cos_theta);
                    float trans = expf(-vp->sigma_t * d);
                    float weight = power * vp->sigma_s * phase * trans / (d * d) * ds;
                    float color[3] = {cmf[0] * weight, cmf[1] * weight, cmf[2] * weight};
                    v3 perp = cross(view, V(0, 1, 0));
                    if (dot(perp, perp) < 1e-8f) perp = cross(view, V(1, 0, 0));
                    perp = mul(normalize(perp), R_BEAM);
                    float s2, t2, d2, radius_px;
                    if (project(&P, add(p, perp), &s2, &t2, &d2)) {
                        float dpx = (s2 - s) * (float)j->w;
                        float dpy = (t2 - t) * (float)j->h;
                        radius_px = clampf(sqrtf(dpx * dpx + dpy * dpy), 1.5f, 24.0f);
                    } else {
                        radius_px = 1.0f;
                    }
                    splat_transverse(j->img, j->w, j->h, s, t, color, radius_px);
                    ++j->splats;
                    if (j->box_group) box_add(j, box, s, t, radius_px);
                }
                if (j->box_group) box_close(j, box);
            }
            if (mi < 0) break;
            const or_material *mt = &sc->mats[mi];
            float n_hero = material_n(mt, lambda, j->o->dispersion);
            v3 nd;
            float wgt;
            int tir;
            scatter(mt, rd, &hit, n_hero, &rng, &nd, &wgt, &tir);
            power *= wgt;
            if (comb && mt->kind == 1u) {  /* first dielectric: the comb collapses to the hero lane */
                hero_spec(lambda, j->o->illuminant, j->ill_scale, cmf);
                comb = 0;
            }
            ro = hit.point;
            rd = nd;
        }
    }
    return NULL;
}

void or_volume(const or_scene *s, const or_camera *cam, const or_beam *beam, uint32_t w,
               uint32_t h, uint32_t n, uint32_t photon_base, const or_vol_params *vp,
               const float *zbuf, uint32_t seed, uint32_t estimator, const or_opts *o,
               float *img, uint32_t threads, or_stats *stats) {
    size_t film = (size_t)3 * w * h;
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    vol_job jobs[256];
    pthread_t tid[256];
    float ill_scale = or_illuminant_scale(o->illuminant);
    for (uint32_t t = 0; t < threads; ++t) {
        vol_job *j = &jobs[t];
        j->sc = s; j->cam = cam; j->beam = beam; j->w = w; j->h = h; j->seed = seed;
        j->est = estimator; j->vp = vp; j->zbuf = zbuf; j->o = o; j->ill_scale = ill_scale;
        j->i0 = photon_base + (uint32_t)((uint64_t)n * t / threads);
        j->i1 = photon_base + (uint32_t)((uint64_t)n * (t + 1) / threads);
        j->img = threads == 1 ? img : (float *)calloc(film, sizeof(float));
        j->segs = j->splats = 0;
        j->box_group = 0;
        j->max_box = 0;
    }
    if (threads == 1) vol_worker(&jobs[0]);
    else {
        for (uint32_t t = 0; t < threads; ++t) pthread_create(&tid[t], NULL, vol_worker, &jobs[t]);
        for (uint32_t t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
        for (uint32_t t = 0; t < threads; ++t) {
            for (size_t k = 0; k < film; ++k) img[k] += jobs[t].img[k];
            free(jobs[t].img);
        }
    }
    if (stats) {
        for (uint32_t t = 0; t < threads; ++t) {
            stats->segments += jobs[t].segs;
            stats->splats += jobs[t].splats;
        }
        stats->paths += n;
    }
}

/* Test hook: the beam estimator (volume.rs:233-329) over photons
 * photon_base .. photon_base+n-1 on one thread, into img, returning the
 * largest bounding-box area (pixels) of `group` consecutive steps of one
 * segment -- the box a wave's LDS tile must hold when it takes a segment's
 * steps `group` at a time. */
uint64_t or_volume_beam_max_box(const or_scene *s, const or_camera *cam, const or_beam *beam, uint32_t w,
                                uint32_t h, uint32_t n, uint32_t photon_base, const or_vol_params *vp,
                                const float *zbuf, uint32_t seed, const or_opts *o, uint32_t group, float *img) {
    vol_job j;
    memset(&j, 0, sizeof j);
    j.sc = s; j.cam = cam; j.beam = beam; j.w = w; j.h = h; j.seed = seed;
    j.est = 1; j.vp = vp; j.zbuf = zbuf; j.o = o; j.ill_scale = or_illuminant_scale(o->illuminant);
    j.i0 = photon_base;
    j.i1 = photon_base + n;
    j.img = img;
    j.box_group = group ? group : 1u;
    vol_worker(&j);
    return j.max_box;
}

/* ------------------------------------------------------------------------- */
/* Per-photon states: cpu_simulate_photon_beam, vol_photons.rs:932-991       */
/* ------------------------------------------------------------------------- */
void or_photon_record_states(const or_scene *s, const or_beam *beam, const or_screen *scr,
                             uint32_t photon_idx, uint32_t seed, float max_dist, uint32_t mode,
                             const or_opts *o, or_photon_record *rec) {
    memset(rec, 0, sizeof *rec);
    screen_t S;
    if (mode == 2u) S = screen_make(scr);
    uint32_t rng = or_rng_new(photon_idx, seed);
    float a = next_f32(&rng);
    float b = next_f32(&rng);
    v3 ro = add(add(vl(beam->corner), mul(vl(beam->u), a)), mul(vl(beam->v), b));
    float lambda = LAMBDA_MIN + LAMBDA_RANGE * next_f32(&rng);
    v3 rd = vl(beam->dir);
    rec->lambda = lambda;
    vs(rec->states[0][0], ro);
    vs(rec->states[0][1], rd);
    rec->num_states = 1;
    for (uint32_t bnc = 0; bnc < o->max_depth; ++bnc) {
        hit_t hit;
        int mi = scene_intersect(s, ro, rd, &hit);
        if (mode == 2u) {
            float ss, rr, st;
            int on = screen_coords(&S, ro, rd, &ss, &rr, &st);
            float scene_t = mi >= 0 ? hit.t : INFINITY;
            if (on && st < scene_t) break;
        } else if (mode == 0u) {
            float seg_len = mi >= 0 ? hit.t : max_dist;
            for (int k = 0; k < 4; ++k) (void)(next_f32(&rng) * seg_len);
        }
        if (mi < 0) break;
        const or_material *m = &s->mats[mi];
        float n_hero = material_n(m, lambda, o->dispersion);
        v3 nd;
        float wgt;
        int tir;
        scatter(m, rd, &hit, n_hero, &rng, &nd, &wgt, &tir);
        /* tir_count counts bounces where refract() has no solution, whether
         * or not the Fresnel draw chose reflection (vol_photons.rs:971-976). */
        if (m->kind == 1u) {
            float n1 = hit.front_face ? 1.0f : n_hero, n2 = hit.front_face ? n_hero : 1.0f;
            v3 dummy;
            if (!refract(rd, hit.normal, n1, n2, &dummy)) rec->tir_count++;
        }
        ro = hit.point;
        rd = nd;
        if (rec->num_states < OR_MAX_PAIRS) {
            vs(rec->states[rec->num_states][0], ro);
            vs(rec->states[rec->num_states][1], rd);
        }
        rec->num_states++;
    }
}

/* ------------------------------------------------------------------------- */
/* Surface rim + depth (prism_dsotm.rs:59-74; z-buffer as prism_on_plane.rs:  */
/* 125-154 records the nearest-solid distance)                               */
/* ------------------------------------------------------------------------- */
void or_surface_rim(const or_scene *s, const or_camera *cam, uint32_t w, uint32_t h,
                    float *rim_xyz, float *depth) {
    for (uint32_t py = 0; py < h; ++py)
        for (uint32_t px = 0; px < w; ++px) {
            float ss = ((float)px + 0.5f) / (float)w;
            float tt = 1.0f - ((float)py + 0.5f) / (float)h;
            v3 rd = primary_ray(cam, ss, tt);
            size_t idx = (size_t)py * w + px;
            hit_t hit;
            int mi = scene_intersect(s, vl(cam->origin), rd, &hit);
            float x = 0, y = 0, z = 0, dep = INFINITY;
            if (mi >= 0) {
                float c = clampf(fabsf(dot(hit.normal, rd)), 0.0f, 1.0f);
                float rim = (1.0f - c) * (1.0f - c);
                x = 0.015f + rim * 0.45f;
                y = 0.02f + rim * 0.5f;
                z = 0.04f + rim * 0.6f;
                dep = hit.t;
            }
            if (rim_xyz) {
                rim_xyz[3 * idx + 0] = x;
                rim_xyz[3 * idx + 1] = y;
                rim_xyz[3 * idx + 2] = z;
            }
            if (depth) depth[idx] = dep;
        }
}

/* ------------------------------------------------------------------------- */
/* Viewer display (spectral-viewer main.rs).  The viewer's accum is u32      */
/* fixed point divided by SCALE; the films here are FP32 sums, so SCALE = 1  */
/* and k = (1 / SCALE) * inv_n * TARGET_N = inv_n * target_n.               */
/* ------------------------------------------------------------------------- */
static float display_k(uint32_t total_photons, float target_n) {
    float inv_n = 1.0f / (float)(total_photons > 1u ? total_photons : 1u); /* main.rs:144 */
    return inv_n * target_n;                                                  /* main.rs:145 */
}

/* fs_main, main.rs:131-174: Gaussian gather, composite, exposure, XYZ ->
 * linear sRGB (main.rs:119-124), tonemap (main.rs:126-130: Reinhard then
 * max(.,0) then pow 1/2.2); the Rgba8Unorm target stores
 * round(255 * clamp(v, 0, 1)). */
void or_display_resolve(const float *vol, const float *surf, uint32_t w, uint32_t h, uint32_t total_photons,
                        float target_n, float exposure, float kradius, uint8_t *rgba) {
    const float k = display_k(total_photons, target_n);
    const int radius = (int)kradius;                       /* main.rs:146 */
    const float sigma = fmaxf(kradius * 0.5f, 0.5f);       /* main.rs:147 */
    const float inv2s2 = 1.0f / (2.0f * sigma * sigma);    /* main.rs:148 */
    for (uint32_t py = 0; py < h; ++py)
        for (uint32_t px = 0; px < w; ++px) {
            float acc[3] = {0.0f, 0.0f, 0.0f}, wsum = 0.0f;
            for (int dy = -radius; dy <= radius; ++dy)     /* main.rs:151-165 */
                for (int dx = -radius; dx <= radius; ++dx) {
                    int sx = (int)px + dx, sy = (int)py + dy;
                    if (sx < 0 || sx >= (int)w || sy < 0 || sy >= (int)h) continue;
                    size_t s = (size_t)sy * w + (size_t)sx;
                    float g = expf(-(float)(dx * dx + dy * dy) * inv2s2);
                    for (int c = 0; c < 3; ++c) acc[c] = acc[c] + vol[3 * s + c] * g;
                    wsum = wsum + g;
                }
            size_t i = (size_t)py * w + px;
            float f = k / fmaxf(wsum, 1e-6f);              /* main.rs:168 */
            float xyz[3], lin[3];
            for (int c = 0; c < 3; ++c) {
                xyz[c] = acc[c] * f;
                if (surf) xyz[c] = xyz[c] + surf[3 * i + c];
                xyz[c] = xyz[c] * exposure;                /* main.rs:171 */
            }
            or_xyz_to_linear_srgb(xyz, lin);
            for (int c = 0; c < 3; ++c) {
                float t = fmaxf(lin[c] / (1.0f + lin[c]), 0.0f);
                float v = powf(t, 1.0f / 2.2f);
                rgba[4 * i + c] = (uint8_t)(clampf(v, 0.0f, 1.0f) * 255.0f + 0.5f);
            }
            rgba[4 * i + 3] = 255;
        }
}

static int cmp_float(const void *a, const void *b) {
    float x = *(const float *)a, y = *(const float *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* compute_exposure_p997, main.rs:635-643 (and prism_on_plane.rs:172-176 with
 * k = 1): the composited luminance's quantile, floored at 1e-9. */
float or_display_quantile(const float *vol, const float *surf, uint32_t w, uint32_t h, uint32_t total_photons,
                          float target_n, float quantile) {
    const size_t n = (size_t)w * h;
    const float k = display_k(total_photons, target_n);
    float *lums = (float *)malloc(n * sizeof(float));
    for (size_t i = 0; i < n; ++i) {
        lums[i] = vol[3 * i + 1] * k;
        if (surf) lums[i] = lums[i] + surf[3 * i + 1];
    }
    qsort(lums, n, sizeof(float), cmp_float);
    size_t idx = (size_t)((float)n * quantile);
    if (idx > n - 1) idx = n - 1;
    float p = lums[idx];
    free(lums);
    return fmaxf(p, 1e-9f);
}

/* ------------------------------------------------------------------------- */
/* The floor-caustic product image, prism_on_plane.rs:16-185                 */
/* ------------------------------------------------------------------------- */
/* box_blur, prism_on_plane.rs:40-74: horizontal then vertical, clamped. */
void or_box_blur(const float *in, float *out, uint32_t w, uint32_t h, int radius) {
    const float norm = 1.0f / (float)(2 * radius + 1);
    float *tmp = (float *)malloc((size_t)3 * w * h * sizeof(float));
    for (int pass = 0; pass < 2; ++pass) {
        const float *src = pass == 0 ? in : tmp;
        float *dst = pass == 0 ? tmp : out;
        for (int y = 0; y < (int)h; ++y)
            for (int x = 0; x < (int)w; ++x) {
                float acc[3] = {0.0f, 0.0f, 0.0f};
                for (int d = -radius; d <= radius; ++d) {
                    int xx = pass == 0 ? x + d : x, yy = pass == 0 ? y : y + d;
                    xx = xx < 0 ? 0 : (xx > (int)w - 1 ? (int)w - 1 : xx);
                    yy = yy < 0 ? 0 : (yy > (int)h - 1 ? (int)h - 1 : yy);
                    const float *p = src + 3 * ((size_t)yy * w + (size_t)xx);
                    for (int c = 0; c < 3; ++c) acc[c] += p[c];
                }
                for (int c = 0; c < 3; ++c) dst[3 * ((size_t)y * w + (size_t)x) + c] = acc[c] * norm;
            }
    }
    free(tmp);
}

/* Screen::sample, lighttrace.rs:98-113 (texel_bilinear as in screen_splat). */
static v3 screen_sample(const screen_t *S, const float *sum, float s, float r) {
    float fx = clampf(s * (float)S->w - 0.5f, 0.0f, (float)S->w - 1.0f);
    float fy = clampf((1.0f - r) * (float)S->h - 0.5f, 0.0f, (float)S->h - 1.0f);
    uint32_t x0 = (uint32_t)floorf(fx), y0 = (uint32_t)floorf(fy);
    uint32_t x1 = x0 + 1 < S->w - 1 ? x0 + 1 : S->w - 1;
    uint32_t y1 = y0 + 1 < S->h - 1 ? y0 + 1 : S->h - 1;
    float dx = fx - (float)x0, dy = fy - (float)y0;
    const float *a = sum + 3 * ((size_t)y0 * S->w + x0), *b = sum + 3 * ((size_t)y0 * S->w + x1);
    const float *c = sum + 3 * ((size_t)y1 * S->w + x0), *d = sum + 3 * ((size_t)y1 * S->w + x1);
    float o[3];
    for (int k = 0; k < 3; ++k)
        o[k] = a[k] * (1.0f - dx) * (1.0f - dy) + b[k] * dx * (1.0f - dy) + c[k] * (1.0f - dx) * dy +
               d[k] * dx * dy;
    return V(o[0], o[1], o[2]);
}

/* shade_floor, prism_on_plane.rs:16-38 */
static v3 shade_floor(const screen_t *S, const float *sum, float gain, const float *ambient,
                      const float *background, v3 o, v3 d, float *t) {
    float s, r;
    if (screen_coords(S, o, d, &s, &r, t)) {
        v3 c = screen_sample(S, sum, s, r);
        return V(ambient[0] + c.x * gain, ambient[1] + c.y * gain, ambient[2] + c.z * gain);
    }
    *t = INFINITY;
    return vl(background);
}

/* The surface pass + z-buffer, prism_on_plane.rs:125-154. */
void or_floor_surface(const or_scene *s, const or_camera *cam, const or_screen *floor, const float *floor_sum,
                      float gain, const float ambient[3], const float background[3], const float tint[3],
                      uint32_t w, uint32_t h, float *surface, float *depth) {
    screen_t S = screen_make(floor);
    for (uint32_t py = 0; py < h; ++py)
        for (uint32_t px = 0; px < w; ++px) {
            float ss = ((float)px + 0.5f) / (float)w;
            float tt = 1.0f - ((float)py + 0.5f) / (float)h;
            v3 o = vl(cam->origin), rd = primary_ray(cam, ss, tt);
            hit_t hit;
            int mi = scene_intersect(s, o, rd, &hit);
            float t_prism = mi >= 0 ? hit.t : INFINITY;
            float floor_t;
            v3 fc = shade_floor(&S, floor_sum, gain, ambient, background, o, rd, &floor_t);
            size_t idx = (size_t)py * w + px;
            v3 out;
            float dep;
            if (t_prism < floor_t) {
                float c = clampf(fabsf(dot(hit.normal, rd)), 0.0f, 1.0f);
                float q = 1.0f - c;
                float rim = q * (q * q); /* powi(3) */
                out = V(fc.x * 0.7f + tint[0] + rim * 0.6f, fc.y * 0.7f + tint[1] + rim * 0.6f,
                        fc.z * 0.7f + tint[2] + rim * 0.7f);
                dep = t_prism;
            } else {
                out = fc;
                dep = floor_t;
            }
            vs(surface + 3 * idx, out);
            if (depth) depth[idx] = dep;
        }
}

/* img = surface + vol, then tonemap_to_u8(xyz_to_linear_srgb(img), exposure):
 * prism_on_plane.rs:166-182. */
void or_composite_srgb8(const float *surface, const float *vol, uint32_t w, uint32_t h, float exposure,
                        uint8_t *rgb) {
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        float img[3], lin[3];
        for (int c = 0; c < 3; ++c) img[c] = surface[3 * i + c] + vol[3 * i + c];
        or_xyz_to_linear_srgb(img, lin);
        or_tonemap_to_u8(lin, exposure, rgb + 3 * i);
    }
}

uint32_t or_struct_sizes(uint32_t which) {
    switch (which) {
    case 0: return sizeof(or_prim);
    case 1: return sizeof(or_plane);
    case 2: return sizeof(or_material);
    case 3: return sizeof(or_scene_env);
    case 4: return sizeof(or_camera);
    case 5: return sizeof(or_beam);
    case 6: return sizeof(or_screen);
    case 7: return sizeof(or_opts);
    case 8: return sizeof(or_vol_params);
    case 9: return sizeof(or_photon_record);
    default: return 0;
    }
}
