// spx_device.hpp -- device-side physics for the gfx950 light-transport kernels.
//
// Every function restates a spectral-core function (file:line in its
// comment) with the reference's evaluation order.  The library is compiled
// with -ffp-contract=off and IEEE-rounded fp32 division and square root,
// so +,-,*,/,sqrt round exactly as the reference's Rust; Glass::n is
// evaluated in fp64 as the reference does, which keeps the TIR / Fresnel
// branch decisions identical to the CPU at the violet edge.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "spectral_hip.h"

namespace spx {

// ---------------------------------------------------------------------------
// Vec3 with glam 0.29 semantics: dot = (x*x'+y*y')+z*z', length = sqrt(dot),
// normalize = v * (1/length), cross = (y z' - y' z, z x' - z' x, x y' - x' y).
// ---------------------------------------------------------------------------
struct f3 {
    float x, y, z;
};

__host__ __device__ inline f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__host__ __device__ inline f3 ld3(const float *p) { return f3{p[0], p[1], p[2]}; }
__host__ __device__ inline f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__host__ __device__ inline f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
__host__ __device__ inline f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
__host__ __device__ inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline float length(f3 a) { return sqrtf(dot(a, a)); }
__host__ __device__ inline f3 normalize(f3 a) { return a * (1.0f / length(a)); }
__host__ __device__ inline f3 cross(f3 a, f3 b) {
    return f3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}

// ---------------------------------------------------------------------------
// Counter-based PCG hash RNG, rng.rs:7-40.
// ---------------------------------------------------------------------------
__host__ __device__ inline uint32_t pcg32(uint32_t input) {
    uint32_t state = input * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28) + 4u)) ^ state) * 277803737u;
    return (word >> 22) ^ word;
}
__host__ __device__ inline uint32_t mix2(uint32_t a, uint32_t b) { return pcg32(a ^ pcg32(b)); }
__host__ __device__ inline uint32_t rng_new(uint32_t key, uint32_t seed) {
    return pcg32(key ^ pcg32(seed));
}
__host__ __device__ inline uint32_t next_u32(uint32_t &st) {
    st += 1u;
    return pcg32(st);
}
// (u >> 8) * 2^-24: exact in fp32 (rng.rs:31-34).
__host__ __device__ inline float next_f32(uint32_t &st) {
    return (float)(next_u32(st) >> 8) * (1.0f / 16777216.0f);
}

// ---------------------------------------------------------------------------
// spectrum.rs:3-32
// ---------------------------------------------------------------------------
constexpr float kLambdaMin = 380.0f;
constexpr float kLambdaMax = 730.0f;
constexpr float kLambdaRange = kLambdaMax - kLambdaMin;
constexpr int kHeroN = 4;
constexpr float kHeroStep = kLambdaRange / (float)kHeroN;

// from_hero_u's rem_euclid(u * 350 + j * 87.5, 350) for u in [0, 1) and j in
// 0..3: the argument lies in [0, 612.5], below twice the modulus, where fmod
// is one subtraction and that subtraction is exact (Sterbenz: b <= a < 2b),
// so the result equals the reference's bit for bit without fmodf's loop.
__device__ inline float hero_lambda(float u, int j) {
    float a = u * kLambdaRange + (float)j * kHeroStep;
    return kLambdaMin + (a >= kLambdaRange ? a - kLambdaRange : a);
}

// ---------------------------------------------------------------------------
// Glass::n, sellmeier.rs:28-44 -- fp64 inside, fp32 out.
// ---------------------------------------------------------------------------
// No glass: scene creation admits glass ids 0..2 only.
constexpr uint32_t kNoGlass = 0xFFFFFFFFu;

__device__ inline float glass_n(uint32_t glass, float lambda_nm) {
    double l_um = (double)lambda_nm / 1000.0;
    if (glass == SPX_GLASS_WATER) return (float)(1.3238 + 0.00314 / (l_um * l_um));
    double l2 = l_um * l_um;
    double b0, b1, b2, c0, c1, c2;
    if (glass == SPX_GLASS_SF11) {
        b0 = 1.73759695; b1 = 0.313747346; b2 = 1.89878101;
        c0 = 0.013188707; c1 = 0.0623068142; c2 = 155.23629;
    } else {
        b0 = 1.03961212; b1 = 0.231792344; b2 = 1.01046945;
        c0 = 0.00600069867; c1 = 0.0200179144; c2 = 103.560653;
    }
    double s = 0.0;
    s = s + b0 * l2 / (l2 - c0);
    s = s + b1 * l2 / (l2 - c1);
    s = s + b2 * l2 / (l2 - c2);
    return (float)sqrt(1.0 + s);
}

// ---------------------------------------------------------------------------
// CIE tables (cie.rs:47-89).  Row r of the LDS table holds
// {xbar, ybar, zbar, S_illuminant} at 380 + 5 r nm, r = 0..70.
// ---------------------------------------------------------------------------
constexpr int kTblN = 71;

struct TblPos {
    int i;
    float frac;
    bool last;
};
// x / 5 correctly rounded, from one multiply and two FMAs: q = x * RN(1/5),
// r = x - q * 5 exactly, q + r * RN(1/5).  Equal to the IEEE quotient for
// every float x in [0, 350] (checked exhaustively, tests/test_div5.py);
// tbl_pos clamps its quotient to [0, 70], so outside that range a last-bit
// difference cannot reach the result either.
__device__ inline float div5(float x) {
    const float q = x * 0.2f;
    const float r = fmaf(-q, 5.0f, x);
    return fmaf(r, 0.2f, q);
}

__device__ inline TblPos tbl_pos(float nm) {
    float f = div5(nm - 380.0f);
    f = f < 0.0f ? 0.0f : f;
    f = f > (float)(kTblN - 1) ? (float)(kTblN - 1) : f;
    int i = (int)floorf(f);
    return TblPos{i, f - (float)i, i + 1 >= kTblN};
}
// Table::sample on the three CMF columns, without a branch for the last row
// (730 nm): there the position's fraction is exactly 0 (tbl_pos clamps to
// 70) and the second row read is the last row itself, so the blend
// a + (a - a) * 0 is a, bit for bit, for every entry but -0, which the
// tables do not hold (gen_tables.py checks).
__device__ inline f3 cmf_at(const float4 *tbl, TblPos p) {
    const float4 a = tbl[p.i];
    const float4 b = tbl[p.last ? p.i : p.i + 1];
    return f3{a.x + (b.x - a.x) * p.frac, a.y + (b.y - a.y) * p.frac, a.z + (b.z - a.z) * p.frac};
}
__device__ inline f3 cmf_lookup(const float4 *tbl, float nm) { return cmf_at(tbl, tbl_pos(nm)); }
// Table::sample on the illuminant column (1.0 for equal energy, a job-wide
// branch); the last row as in cmf_at.
__device__ inline float ill_at(const float4 *tbl, TblPos p, bool equal_energy) {
    if (equal_energy) return 1.0f;
    const float a = tbl[p.i].w;
    const float b = tbl[p.last ? p.i : p.i + 1].w;
    return a + (b - a) * p.frac;
}
__device__ inline float ill_lookup(const float4 *tbl, float nm, bool equal_energy) {
    if (equal_energy) return 1.0f;
    return ill_at(tbl, tbl_pos(nm), false);
}

// ---------------------------------------------------------------------------
// Scene (scene.rs, geom.rs) staged in LDS.
// ---------------------------------------------------------------------------
struct SceneData {
    uint32_t n_prims, n_planes, n_mats, pad;
    spx_env env;
    spx_prim prims[SPX_MAX_PRIMS];
    spx_plane planes[SPX_MAX_PLANES];
    spx_material mats[SPX_MAX_MATERIALS];
};

struct Hit {
    float t;
    f3 point;
    f3 normal;
    bool front_face;
};

// Sphere::intersect, geom.rs:35-57: the hit distance (the point and normal
// are formed by scene_hit for the nearest primitive only).  Both roots are
// computed and the near one taken when it lies in range, instead of
// branching (a miss, the far root): the same roots, tests and result.
__device__ inline bool sphere_t(f3 c, float radius, f3 o, f3 d, float t_min, float t_max, float &t_out) {
    f3 oc = o - c;
    float a = dot(d, d);
    float half_b = dot(oc, d);
    float cc = dot(oc, oc) - radius * radius;
    float disc = half_b * half_b - a * cc;
    const float sq = sqrtf(disc);
    const float t0 = (-half_b - sq) / a;
    const float t1 = (-half_b + sq) / a;
    const bool in0 = !(t0 < t_min || t0 > t_max);
    const bool in1 = !(t1 < t_min || t1 > t_max);
    t_out = in0 ? t0 : t1;
    return !(disc < 0.0f) & (in0 | in1);
}

// ConvexSolid::intersect (slab method), geom.rs:128-175: the hit distance
// and the index of the plane it lies on (n: none, whose normal is zero, as
// the reference's initial entering / exiting normal).  Every lane walks all
// n planes: a ray parallel to a plane it lies outside of (the reference's
// early `return None`) is flagged and rejected after the loop, and each
// plane's update is a select on its index, so the loop has no per-lane exit
// and no per-lane branches; the distances, comparisons and result are the
// reference's.
__device__ inline bool solid_t(const spx_plane *pl, uint32_t n, f3 o, f3 d, float t_min, float t_max,
                               float &t_out, uint32_t &plane_out) {
    float t_enter = -INFINITY, t_exit = INFINITY;
    uint32_t ie = n, ix = n;
    bool outside = false;
    for (uint32_t i = 0; i < n; ++i) {
        const f3 pn = ld3(pl[i].normal);
        const float denom = dot(pn, d);
        const float dist = dot(pn, o) + pl[i].d;
        const bool parallel = fabsf(denom) < 1e-8f;
        outside = outside || (parallel && dist > 0.0f);
        const float t = -dist / denom;
        const bool enter = !parallel && denom < 0.0f && t > t_enter;
        const bool leave = !parallel && !(denom < 0.0f) && t < t_exit;
        t_enter = enter ? t : t_enter;
        ie = enter ? i : ie;
        t_exit = leave ? t : t_exit;
        ix = leave ? i : ix;
    }
    if (outside || t_enter > t_exit) return false;
    float th;
    uint32_t ih;
    if (t_enter > t_min) {
        th = t_enter;
        ih = ie;
    } else if (t_exit > t_min) {
        th = t_exit;
        ih = ix;
    } else {
        return false;
    }
    if (th > t_max) return false;
    t_out = th;
    plane_out = ih;
    return true;
}

// Scene::intersect, scene.rs:64-78.  Returns the material index or -1.  The
// nearest primitive's hit point and normal (Sphere / ConvexSolid::intersect's
// HitRecord, geom.rs:50-56, 165-174) are formed once, after the loop, from its
// distance.
__device__ inline int scene_hit(const SceneData &s, f3 o, f3 d, Hit &best) {
    float closest = INFINITY;
    int bp = -1;
    uint32_t bplane = 0;
    for (uint32_t i = 0; i < s.n_prims; ++i) {
        const spx_prim &p = s.prims[i];
        float t;
        uint32_t plane = 0;
        bool ok;
        if (p.kind == SPX_PRIM_SPHERE) ok = sphere_t(ld3(p.center), p.radius, o, d, 1e-3f, closest, t);
        else ok = solid_t(s.planes + p.plane_start, p.plane_count, o, d, 1e-3f, closest, t, plane);
        if (ok) {
            closest = t;
            bp = (int)i;
            bplane = plane;
        }
    }
    if (bp < 0) return -1;
    const spx_prim &p = s.prims[bp];
    const f3 point = o + d * closest;
    f3 outward;
    if (p.kind == SPX_PRIM_SPHERE) outward = (point - ld3(p.center)) / p.radius;
    else outward = bplane < p.plane_count ? ld3(s.planes[p.plane_start + bplane].normal) : mk(0.f, 0.f, 0.f);
    const bool front = dot(outward, d) < 0.0f;
    best.t = closest;
    best.point = point;
    best.normal = front ? outward : -outward;
    best.front_face = front;
    return (int)p.material;
}

// background_radiance, scene.rs:44-50, plus the sun-disk extension.
__device__ inline float background_radiance(const spx_env &e, f3 dir) {
    float sky = (e.has_horizon && dir.y > 0.0f) ? e.horizon : 0.0f;
    float r = e.background + sky;
    if (e.sun_radiance > 0.0f && dot(dir, ld3(e.sun_dir)) >= e.sun_cos) r = r + e.sun_radiance;
    return r;
}

// ---------------------------------------------------------------------------
// optics.rs:8-49
// ---------------------------------------------------------------------------
__device__ inline f3 reflect(f3 d, f3 n) { return d - n * (2.0f * dot(d, n)); }

// refract and fresnel_reflectance take their branches as selects (the
// total-internal-reflection outcomes are computed and discarded), so a
// wave's lanes stay together; the same operations and results (refract
// writes `out` either way: a NaN direction where it returns false).
__device__ inline bool refract(f3 d, f3 n, float n1, float n2, f3 &out) {
    const float eta = n1 / n2;
    const float c0 = -dot(d, n);
    const bool flip = c0 < 0.0f;
    const f3 nn = flip ? -n : n;
    const float cos_i = flip ? -c0 : c0;
    const float k = 1.0f - eta * eta * (1.0f - cos_i * cos_i);
    const float cos_t = sqrtf(k);
    out = normalize(d * eta + nn * (eta * cos_i - cos_t));
    return !(k < 0.0f);
}

__device__ inline float fresnel_reflectance(float cos_i, float n1, float n2) {
    cos_i = cos_i < 0.0f ? 0.0f : (cos_i > 1.0f ? 1.0f : cos_i);
    const float r = n1 / n2;
    const float sin_t2 = r * r * (1.0f - cos_i * cos_i);
    const float cos_t = sqrtf(1.0f - sin_t2);
    const float qs = (n1 * cos_i - n2 * cos_t) / (n1 * cos_i + n2 * cos_t);
    const float qp = (n1 * cos_t - n2 * cos_i) / (n1 * cos_t + n2 * cos_i);
    return sin_t2 >= 1.0f ? 1.0f : 0.5f * (qs * qs + qp * qp);
}

// ---------------------------------------------------------------------------
// material.rs:30-82
// ---------------------------------------------------------------------------
__device__ inline f3 cosine_hemisphere(f3 normal, uint32_t &rng) {
    float u1 = next_f32(rng);
    float u2 = next_f32(rng);
    float r = sqrtf(u1);
    float phi = 6.2831855f * u2;  // (2.0 * PI) folded in fp32, then * u2
    float x = r * cosf(phi);
    float y = r * sinf(phi);
    float z = sqrtf(fmaxf(1.0f - u1, 0.0f));
    f3 a = fabsf(normal.x) > 0.9f ? mk(0.f, 1.f, 0.f) : mk(1.f, 0.f, 0.f);
    f3 t = normalize(cross(normal, a));
    f3 b = cross(normal, t);
    return normalize(t * x + b * y + normal * z);
}

// Material::scatter.  Returns the weight; *tir reports a refract() failure.
__device__ inline float scatter(const spx_material &m, f3 wo, const Hit &h, float n_hero,
                                uint32_t &rng, f3 &dir, bool &tir) {
    tir = false;
    if (m.kind == SPX_MAT_LAMBERTIAN) {
        dir = cosine_hemisphere(h.normal, rng);
        return m.reflectance;
    }
    float cos_i = fabsf(-dot(wo, h.normal));
    float n1 = h.front_face ? 1.0f : n_hero;
    float n2 = h.front_face ? n_hero : 1.0f;
    float r = fresnel_reflectance(cos_i, n1, n2);
    // reflect when the draw says so or refraction fails (material.rs:66-78),
    // as selects: both directions are formed, one is kept
    const bool reflected = next_f32(rng) < r;
    const f3 refl = reflect(wo, h.normal);
    f3 refr;
    const bool refracted = refract(wo, h.normal, n1, n2, refr);
    dir = reflected || !refracted ? refl : refr;
    tir = !reflected && !refracted;
    return 1.0f;
}

// n_hero: tracer.rs:76-79; dispersion off = n(550 nm) (trace.wgsl:480).
__device__ inline float material_n(const spx_material &m, float lambda, uint32_t dispersion) {
    if (m.kind != SPX_MAT_DIELECTRIC) return 1.0f;
    return glass_n(m.glass, dispersion ? lambda : 550.0f);
}

// ---------------------------------------------------------------------------
// camera.rs:41-69.  Proj holds Camera::project's per-camera terms
// (camera_proj_params, vol_photons.rs:100-115), computed on the host with
// the same expressions.
// ---------------------------------------------------------------------------
struct Proj {
    float o[3], u[3], v[3], w[3];
    float vw, vh;
    float pad[2];
};

__device__ inline f3 primary_ray(const spx_camera &c, float s, float t) {
    f3 d = ld3(c.lower_left) + ld3(c.horizontal) * s + ld3(c.vertical) * t - ld3(c.origin);
    return normalize(d);
}

__device__ inline bool project(const Proj &P, f3 p, float &s, float &t) {
    f3 dir = p - ld3(P.o);
    float c = dot(dir, ld3(P.w));
    if (c >= -1e-6f) return false;
    float ss = 0.5f + dot(dir, ld3(P.u)) / (-c * P.vw);
    float tt = 0.5f + dot(dir, ld3(P.v)) / (-c * P.vh);
    if (ss >= 0.0f && ss < 1.0f && tt >= 0.0f && tt < 1.0f) {
        s = ss;
        t = tt;
        return true;
    }
    return false;
}

// Nearest pixel, volume.rs:134-135.
__device__ inline void pix_xy(float s, float t, uint32_t w, uint32_t h, uint32_t &px, uint32_t &py) {
    px = (uint32_t)(s * (float)w);
    py = (uint32_t)((1.0f - t) * (float)h);
    px = px > w - 1 ? w - 1 : px;
    py = py > h - 1 ? h - 1 : py;
}
__device__ inline uint32_t pix_index(float s, float t, uint32_t w, uint32_t h) {
    uint32_t px, py;
    pix_xy(s, t, w, h, px, py);
    return py * w + px;
}

// phase_hg, volume.rs:88-92
__device__ inline float phase_hg(float g, float cos_theta) {
    float g2 = g * g;
    float denom = powf(fmaxf(1.0f + g2 - 2.0f * g * cos_theta, 1e-6f), 1.5f);
    return (1.0f - g2) / (12.566371f * denom);  // (4.0 * PI) folded in fp32
}

// ---------------------------------------------------------------------------
// Film splat: hardware fp32 atomic add (global_atomic_add_f32, no return).
// ---------------------------------------------------------------------------
// One film deposit a segment may make: pixel (px, py) = index idx, value v;
// ok = false when there is none.
struct Dep {
    uint32_t px, py, idx;
    f3 v;
    bool ok;
};

// Adds v to film pixel idx.
__device__ inline void film_add(float *film, uint32_t idx, f3 v) {
    unsafeAtomicAdd(film + 3 * idx + 0, v.x);
    unsafeAtomicAdd(film + 3 * idx + 1, v.y);
    unsafeAtomicAdd(film + 3 * idx + 2, v.z);
}

}  // namespace spx
