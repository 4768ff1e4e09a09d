// spx_paths.hpp -- per-path / per-photon transport routines shared by the
// plain, wavefront and batched kernels.  Each restates one reference loop.
#pragma once

#include "spx_internal.hpp"

namespace spx {

// Cooperative copy of the scene and the 71-row CIE table into LDS.
__device__ inline void stage_scene(SceneData &dst, const SceneData *src, float4 *tbl,
                                   const float4 *tsrc) {
    const uint32_t *s = reinterpret_cast<const uint32_t *>(src);
    uint32_t *d = reinterpret_cast<uint32_t *>(&dst);
    constexpr uint32_t words = sizeof(SceneData) / 4;
    for (uint32_t k = threadIdx.x; k < words; k += blockDim.x) d[k] = s[k];
    for (uint32_t k = threadIdx.x; k < (uint32_t)kTblN; k += blockDim.x) tbl[k] = tsrc[k];
}

// ---------------------------------------------------------------------------
// Camera path: CpuTracer::sample_pixel, tracer.rs:44-111 (+ optional
// Russian roulette before bounce b >= rr_depth: survive with
// q = clamp(max throughput, 0.05, 0.95), throughput /= q).
// ---------------------------------------------------------------------------
// A path's radiance is added once, at its escape (tracer.rs:63-70), as
// tp[k] * bg * S(lambda_k) onto zero; the state keeps bg (zero until the
// path escapes) and camera_xyz forms each lane's radiance with the same
// operations, so the CMF and illuminant lookups share one table position.
struct CamState {
    f3 ro, rd;
    uint32_t rng;
    float lam0;     // hero wavelength (lane 0)
    float u_hero;   // the comb draw; lane k is hero_lambda(u_hero, k)
    float tp[4];
    float esc;      // background radiance at the escape (0 = not escaped)
    uint32_t valid;
};

__device__ inline void camera_begin(const CamParams &P, uint32_t px, uint32_t py, uint32_t sidx,
                                    CamState &st) {
    uint32_t pixel = py * P.w + px;
    st.rng = rng_new(mix2(pixel, sidx), P.seed);
    float s = ((float)px + next_f32(st.rng)) / (float)P.w;
    float t = 1.0f - ((float)py + next_f32(st.rng)) / (float)P.h;
    st.u_hero = next_f32(st.rng);
    st.lam0 = hero_lambda(st.u_hero, 0);
#pragma unroll
    for (int k = 0; k < kHeroN; ++k) st.tp[k] = 1.0f;
    st.esc = 0.0f;
    st.valid = kHeroN;
    st.ro = ld3(P.cam.origin);
    st.rd = primary_ray(P.cam, s, t);
}

// One bounce-loop iteration b.  Returns false when the path has ended.
__device__ inline bool camera_bounce(const SceneData &S, const float4 *tbl, const CamParams &P,
                                     uint32_t b, CamState &st) {
    if (P.rr_depth && b >= P.rr_depth) {
        float q = fmaxf(fmaxf(fmaxf(st.tp[0], st.tp[1]), st.tp[2]), st.tp[3]);
        q = q < 0.05f ? 0.05f : (q > 0.95f ? 0.95f : q);
        if (next_f32(st.rng) >= q) return false;
#pragma unroll
        for (int k = 0; k < kHeroN; ++k) st.tp[k] = st.tp[k] / q;
    }
    Hit h;
    int mi = scene_hit(S, st.ro, st.rd, h);
    if (mi < 0) {
        st.esc = background_radiance(S.env, st.rd);
        return false;
    }
    const spx_material &m = S.mats[mi];
    float nh = material_n(m, st.lam0, P.dispersion);
    f3 nd;
    bool tir;
    float w = scatter(m, st.rd, h, nh, st.rng, nd, tir);
#pragma unroll
    for (int k = 0; k < kHeroN; ++k) st.tp[k] *= w;
    if (m.kind == SPX_MAT_DIELECTRIC && st.valid > 1) {
        st.tp[1] = st.tp[2] = st.tp[3] = 0.0f;
        st.valid = 1;
    }
    st.ro = h.point;
    st.rd = nd;
    return true;
}

// XYZ of a finished path, tracer.rs:102-110.  Lane k's radiance is
// 0 + tp[k] * bg * S(lambda_k) (tracer.rs:66-68; exactly tp[k] * bg * S, as
// the product is never -0), and dividing by valid_lanes, 1 or 4, is exact
// scaling by 1 or 0.25.
__device__ inline f3 camera_xyz(const float4 *tbl, const CamState &st, bool equal_energy) {
    const float pdf = 1.0f / kLambdaRange;
    const float inv_valid = st.valid == 1 ? 1.0f : 0.25f;
    f3 xyz = mk(0.f, 0.f, 0.f);
#pragma unroll
    for (int k = 0; k < kHeroN; ++k) {
        const TblPos p = tbl_pos(hero_lambda(st.u_hero, k));
        const f3 c = cmf_at(tbl, p);
        const float rad = st.tp[k] * st.esc * ill_at(tbl, p, equal_energy);
        const float wk = rad / pdf * inv_valid;
        xyz.x += c.x * wk;
        xyz.y += c.y * wk;
        xyz.z += c.z * wk;
    }
    return xyz;
}

__device__ inline f3 camera_path(const SceneData &S, const float4 *tbl, const CamParams &P,
                                 uint32_t px, uint32_t py, uint32_t sidx) {
    CamState st;
    camera_begin(P, px, py, sidx, st);
    for (uint32_t b = 0; b < P.max_depth; ++b)
        if (!camera_bounce(S, tbl, P, b, st)) break;
    return camera_xyz(tbl, st, P.illuminant == SPX_ILLUM_E);
}

// ---------------------------------------------------------------------------
// Photon emission: lighttrace.rs:159-161 / volume.rs:119-124.
// ---------------------------------------------------------------------------
struct PhotonState {
    f3 ro, rd;
    uint32_t rng;
    float lambda;
    float power;
    float n;           // Glass::n of glass n_glass at this photon's wavelength
    uint32_t n_glass;  // kNoGlass until the first dielectric hit
    f3 cmf;            // the CIE matching functions at this photon's wavelength (the hero
                       // comb: the photon's spectrum, see photon_comb)
    uint32_t comb;     // the hero comb's companion lanes are still live
    float u;           // the wavelength draw (lambda = 380 + 350 u)
};

// The photon spectra of the hero-comb option (P.hero; an extension with no
// reference counterpart): the photon carries the stratified comb of
// spectrum.rs:20-32 drawn from its wavelength draw, lane k weighing 1/4,
// until it scatters off a dielectric, where the companions collapse onto
// the hero lane (tracer.rs:86-91).  Deposits are spectrum * (material
// weights * geometry): the spectrum is sum_k cmf(lambda_k) * s_k * 0.25 in
// lane order before the collapse and cmf(lambda_0) * s_0 after, s_k the
// beam's SPD weight (1 for equal energy), and the power starts at 1.
__device__ inline f3 photon_comb(const PhotonParams &P, const float4 *tbl, float u) {
    f3 spec = mk(0.f, 0.f, 0.f);
#pragma unroll
    for (int k = 0; k < kHeroN; ++k) {
        const float lam = hero_lambda(u, k);
        const f3 c = cmf_lookup(tbl, lam);
        const float s = P.illuminant == SPX_ILLUM_E ? 1.0f : ill_lookup(tbl, lam, false) * P.ill_scale;
        spec.x += c.x * s * 0.25f;
        spec.y += c.y * s * 0.25f;
        spec.z += c.z * s * 0.25f;
    }
    return spec;
}

__device__ inline f3 photon_hero(const PhotonParams &P, const float4 *tbl, float lam) {
    const f3 c = cmf_lookup(tbl, lam);
    const float s = P.illuminant == SPX_ILLUM_E ? 1.0f : ill_lookup(tbl, lam, false) * P.ill_scale;
    return mk(c.x * s, c.y * s, c.z * s);
}

// A photon's start in two parts: photon_start holds its costly terms (the
// RNG key and first three draws, the wavelength's table lookups and, for
// the scene's first glass g0, Glass::n -- the value photon_n would compute
// at the photon's first hit of g0), so that a wave can compute the starts
// of its next photons together with every lane busy; photon_resume builds
// the state from them.  photon_begin is the two in a row.
struct PhotonStart {
    uint32_t rng;  // after the three draws
    float a, b, u;
    float power, n;
    f3 cmf;
};
constexpr uint32_t kPhotonStartWords = 9;

__device__ inline PhotonStart photon_start(const PhotonParams &P, const float4 *tbl, uint32_t i, uint32_t g0) {
    PhotonStart s;
    s.rng = rng_new(i, P.seed);
    s.a = next_f32(s.rng);
    s.b = next_f32(s.rng);
    s.u = next_f32(s.rng);
    const float lambda = kLambdaMin + kLambdaRange * s.u;  // = the comb's hero lane, hero_lambda(u, 0)
    if (P.hero) {
        s.power = 1.0f;
        s.cmf = photon_comb(P, tbl, s.u);
    } else {
        s.power = P.illuminant == SPX_ILLUM_E ? 1.0f : ill_lookup(tbl, lambda, false) * P.ill_scale;
        s.cmf = cmf_lookup(tbl, lambda);  // the wavelength is fixed: looked up once
    }
    s.n = g0 != kNoGlass ? glass_n(g0, P.dispersion ? lambda : 550.0f) : 1.0f;
    return s;
}

__device__ inline void photon_resume(const PhotonParams &P, const PhotonStart &s, uint32_t g0, PhotonState &st) {
    st.rng = s.rng;
    st.ro = ld3(P.beam.corner) + ld3(P.beam.u) * s.a + ld3(P.beam.v) * s.b;
    st.u = s.u;
    st.lambda = kLambdaMin + kLambdaRange * s.u;
    st.rd = ld3(P.beam.dir);
    st.n_glass = g0;
    st.n = s.n;
    st.comb = P.hero;
    st.power = s.power;
    st.cmf = s.cmf;
}

__device__ inline void photon_begin(const PhotonParams &P, const float4 *tbl, uint32_t i,
                                    PhotonState &st) {
    photon_resume(P, photon_start(P, tbl, i, kNoGlass), kNoGlass, st);
}

// Glass of the scene's first dielectric material, or kNoGlass.
__device__ inline uint32_t scene_first_glass(const SceneData &S) {
    for (uint32_t k = 0; k < S.n_mats; ++k)
        if (S.mats[k].kind == SPX_MAT_DIELECTRIC) return S.mats[k].glass;
    return kNoGlass;
}

// A deposit at film pixel (px, py) = index idx; sinks keyed by film tile
// overload it to use the pixel directly.
template <class Sink>
__device__ inline void deposit_xy(Sink &sink, uint32_t px, uint32_t py, uint32_t idx, f3 v) {
    sink.add(idx, v);
}

// Hooks a deposit sink may take: a new photon starts, the photon scattered
// off a material of `kind`.  Sinks that track photons overload these.
template <class Sink>
__device__ inline void sink_photon_begin(Sink &) {}
template <class Sink>
__device__ inline void sink_scattered(Sink &, uint32_t) {}

// material_n at the photon's wavelength.  The wavelength is fixed for the
// photon's life, so Glass::n (fp64 Sellmeier: the costliest step of a
// dielectric hit) is computed at its first hit of a glass and reused, the
// same value, at later hits of that glass.
__device__ inline float photon_n(const spx_material &m, const PhotonParams &P, PhotonState &st) {
    if (m.kind != SPX_MAT_DIELECTRIC) return 1.0f;
    if (st.n_glass != m.glass) {
        st.n = material_n(m, st.lambda, P.dispersion);
        st.n_glass = m.glass;
    }
    return st.n;
}

// Scatter at a hit (the tail of every photon loop).  Returns tir.
__device__ inline bool photon_scatter(const SceneData &S, const float4 *tbl, const PhotonParams &P, int mi, const Hit &h,
                                      PhotonState &st) {
    const spx_material &m = S.mats[mi];
    float nh = photon_n(m, P, st);
    f3 nd;
    bool tir;
    float w = scatter(m, st.rd, h, nh, st.rng, nd, tir);
    st.power *= w;
    if (st.comb && m.kind == SPX_MAT_DIELECTRIC) {  // the comb collapses onto the hero lane
        st.cmf = photon_hero(P, tbl, st.lambda);
        st.comb = 0;
    }
    st.ro = h.point;
    st.rd = nd;
    return tir;
}

// Screen::screen_coords, lighttrace.rs:42-60 (normal, |u|^2, |v|^2 hoisted).
__device__ inline bool screen_coords(const PhotonParams &P, f3 o, f3 d, float &s, float &r, float &t) {
    f3 n = ld3(P.scr_n);
    float denom = dot(d, n);
    if (fabsf(denom) < 1e-8f) return false;
    float tt = dot(ld3(P.scr_corner) - o, n) / denom;
    if (tt <= 1e-4f) return false;
    f3 p = o + d * tt;
    f3 rel = p - ld3(P.scr_corner);
    float ss = dot(rel, ld3(P.scr_u)) / P.scr_uu;
    float rr = dot(rel, ld3(P.scr_v)) / P.scr_vv;
    if (ss >= 0.0f && ss < 1.0f && rr >= 0.0f && rr < 1.0f) {
        s = ss;
        r = rr;
        t = tt;
        return true;
    }
    return false;
}

// Screen::splat's bilinear footprint (texel_bilinear, lighttrace.rs:65-91):
// texel k = 0..3 is (x0,y0), (x1,y0), (x0,y1), (x1,y1), weighted
// (1-dx)(1-dy), dx(1-dy), (1-dx)dy, dx dy, of the colour c * power.
struct ScreenTexels {
    uint32_t x0, y0, x1, y1;
    float w[4];
    f3 cp;
};

__device__ inline ScreenTexels screen_texels(const PhotonParams &P, float s, float r, f3 c, float power) {
    ScreenTexels T;
    float W = (float)P.scr_w, H = (float)P.scr_h;
    float fx = s * W - 0.5f;
    fx = fx < 0.0f ? 0.0f : (fx > W - 1.0f ? W - 1.0f : fx);
    float fy = (1.0f - r) * H - 0.5f;
    fy = fy < 0.0f ? 0.0f : (fy > H - 1.0f ? H - 1.0f : fy);
    T.x0 = (uint32_t)floorf(fx);
    T.y0 = (uint32_t)floorf(fy);
    T.x1 = T.x0 + 1 < P.scr_w - 1 ? T.x0 + 1 : P.scr_w - 1;
    T.y1 = T.y0 + 1 < P.scr_h - 1 ? T.y0 + 1 : P.scr_h - 1;
    float dx = fx - (float)T.x0, dy = fy - (float)T.y0;
    T.w[0] = (1.0f - dx) * (1.0f - dy);
    T.w[1] = dx * (1.0f - dy);
    T.w[2] = (1.0f - dx) * dy;
    T.w[3] = dx * dy;
    T.cp = c * power;
    return T;
}

__device__ inline Dep screen_texel(const PhotonParams &P, const ScreenTexels &T, int k) {
    Dep d;
    d.px = (k & 1) ? T.x1 : T.x0;
    d.py = (k & 2) ? T.y1 : T.y0;
    d.idx = d.py * P.scr_w + d.px;
    d.v = T.cp * T.w[k];
    d.ok = true;
    return d;
}

// Screen::splat (texel_bilinear + 4-texel deposit), lighttrace.rs:65-91.
template <class Sink>
__device__ inline void screen_splat(const PhotonParams &P, float s, float r, f3 c, float power, Sink &sink) {
    const ScreenTexels T = screen_texels(P, s, r, c, power);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const Dep d = screen_texel(P, T, k);
        deposit_xy(sink, d.px, d.py, d.idx, d.v);
    }
}

// Screen::sample, lighttrace.rs:98-113: the bilinear read of texel_bilinear's
// four texels, each term sum * wx * wy in the reference's order.
__device__ inline f3 screen_sample(const PhotonParams &P, const float *sum, float s, float r) {
    const float W = (float)P.scr_w, H = (float)P.scr_h;
    float fx = s * W - 0.5f;
    fx = fx < 0.0f ? 0.0f : (fx > W - 1.0f ? W - 1.0f : fx);
    float fy = (1.0f - r) * H - 0.5f;
    fy = fy < 0.0f ? 0.0f : (fy > H - 1.0f ? H - 1.0f : fy);
    const uint32_t x0 = (uint32_t)floorf(fx), y0 = (uint32_t)floorf(fy);
    const uint32_t x1 = x0 + 1 < P.scr_w - 1 ? x0 + 1 : P.scr_w - 1;
    const uint32_t y1 = y0 + 1 < P.scr_h - 1 ? y0 + 1 : P.scr_h - 1;
    const float dx = fx - (float)x0, dy = fy - (float)y0;
    const float *a = sum + 3 * ((size_t)y0 * P.scr_w + x0), *b = sum + 3 * ((size_t)y0 * P.scr_w + x1);
    const float *c = sum + 3 * ((size_t)y1 * P.scr_w + x0), *d = sum + 3 * ((size_t)y1 * P.scr_w + x1);
    float o[3];
    for (int k = 0; k < 3; ++k)
        o[k] = a[k] * (1.0f - dx) * (1.0f - dy) + b[k] * dx * (1.0f - dy) + c[k] * (1.0f - dx) * dy +
               d[k] * dx * dy;
    return mk(o[0], o[1], o[2]);
}

// One lighttrace bounce (trace_photon body, lighttrace.rs:133-151).
// Returns 0 = continue, 1 = landed (splatted), 2 = lost.
template <class Sink>
__device__ inline int lighttrace_bounce(const SceneData &S, const float4 *tbl, const PhotonParams &P,
                                        PhotonState &st, Sink &sink) {
    Hit h;
    int mi = scene_hit(S, st.ro, st.rd, h);
    float ss, rr, stt;
    bool on = screen_coords(P, st.ro, st.rd, ss, rr, stt);
    float scene_t = mi >= 0 ? h.t : INFINITY;
    float screen_t = on ? stt : INFINITY;
    if (screen_t < scene_t) {
        screen_splat(P, ss, rr, st.cmf, st.power, sink);
        return 1;
    }
    if (mi < 0) return 2;
    photon_scatter(S, tbl, P, mi, h, st);
    sink_scattered(sink, S.mats[mi].kind);
    return 0;
}

// A job-wide value known to be the same on every lane, moved to a scalar
// register so that branches on it are scalar branches: the compiler cannot
// prove a value read from the job table uniform, and would otherwise
// if-convert such a branch and evaluate both sides on every lane.
__device__ inline float uniform_f32(float v) {
    return __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(v)));
}
__device__ inline const float *uniform_ptr(const float *p) {
    const uint64_t a = (uint64_t)(uintptr_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
    return (const float *)(uintptr_t)(((uint64_t)hi << 32) | lo);
}

// phase_hg (volume.rs:88-92) with x^1.5 = x * sqrt(x) and a hardware
// reciprocal: within a few ulp.
__device__ inline float phase_hg_approx(float g, float cos_theta) {
    float g2 = g * g;
    float x = fmaxf(1.0f + g2 - 2.0f * g * cos_theta, 1e-6f);
    return (1.0f - g2) * __builtin_amdgcn_rcpf(12.566371f * (x * __builtin_amdgcn_sqrtf(x)));
}

// Point / legacy estimator deposits for one segment (volume.rs:130-150,
// 51-64).  ISO: g == 0, where phase_hg(0, c) = 1 / (4 pi * 1^1.5) for every
// c, bit for bit.  Without a z-buffer every depth passes (+INF).
// One of a segment's four samples: draws its distance, then projects and
// weighs it (cam_o = the camera origin, zbuf = the z-buffer or null).
template <bool LEGACY, bool ISO>
__device__ inline Dep point_sample(const PhotonParams &P, const PhotonState &st, f3 cmf, float seg_len,
                                   uint32_t &rng, f3 cam_o, const float *zbuf) {
    Dep dp;
    float dist = next_f32(rng) * seg_len;
    f3 p = st.ro + st.rd * dist;
    // Camera::project (project(), camera.rs:41-69) with its tests folded into
    // dp.ok instead of returning early: a sample off the film still forms its
    // (unused) pixel and weight, so a segment's four samples run without
    // per-lane branches.  The same operations decide the pixel and the tests.
    const f3 dir = p - ld3(P.proj.o);
    const float c = dot(dir, ld3(P.proj.w));
    const float ss = 0.5f + dot(dir, ld3(P.proj.u)) / (-c * P.proj.vw);
    const float tt = 0.5f + dot(dir, ld3(P.proj.v)) / (-c * P.proj.vh);
    bool ok = !(c >= -1e-6f) & (ss >= 0.0f) & (ss < 1.0f) & (tt >= 0.0f) & (tt < 1.0f);
    pix_xy(ok ? ss : 0.0f, ok ? tt : 0.0f, P.w, P.h, dp.px, dp.py);  // in range either way
    dp.idx = dp.py * P.w + dp.px;
    if (LEGACY) {
        const float inv4pi = 1.0f / 12.566371f;
        float contrib = st.power * P.vp.sigma_s * seg_len / 4.0f * inv4pi;
        dp.v = cmf * contrib;
    } else {
        f3 to = cam_o - p;
        const float d2 = dot(to, to);
        // the z-buffer test (whether the deposit happens) on the IEEE distance
        if (zbuf) ok = ok & (fmaxf(sqrtf(d2), 1e-3f) < zbuf[dp.idx]);
        // the weight (how much it deposits) from the hardware square root,
        // reciprocal and exp2, each within a few ulp
        const float d = fmaxf(__builtin_amdgcn_sqrtf(d2), 1e-3f);
        const float inv_d = __builtin_amdgcn_rcpf(d);
        float phase = ISO ? 1.0f / 12.566371f : phase_hg_approx(P.vp.g, dot(st.rd, to * inv_d));
        float trans = __expf(-P.vp.sigma_t * d);
        float contrib = st.power * P.vp.sigma_s * phase * trans * (inv_d * inv_d) * seg_len * 0.25f;
        dp.v = cmf * contrib;
    }
    dp.ok = ok;
    return dp;
}

template <bool LEGACY, bool ISO, class Sink>
__device__ inline void point_deposits(const PhotonParams &P, const PhotonState &st, f3 cmf, float seg_len,
                                      uint32_t &rng, Sink &sink) {
    const f3 cam_o = ld3(P.proj.o);
    const float *zbuf = uniform_ptr(P.zbuf);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const Dep d = point_sample<LEGACY, ISO>(P, st, cmf, seg_len, rng, cam_o, zbuf);
        if (d.ok) deposit_xy(sink, d.px, d.py, d.idx, d.v);
    }
}

template <bool LEGACY, class Sink>
__device__ inline void volume_point_segment(const PhotonParams &P, const float4 *tbl,
                                            const PhotonState &st, f3 cmf, float seg_len,
                                            uint32_t &rng, Sink &sink) {
    if (LEGACY) point_deposits<true, true>(P, st, cmf, seg_len, rng, sink);
    else if (uniform_f32(P.vp.g) == 0.0f) point_deposits<false, true>(P, st, cmf, seg_len, rng, sink);
    else point_deposits<false, false>(P, st, cmf, seg_len, rng, sink);
}

// Beam estimator deposits for one segment (volume.rs:263-311, 178-218).
//
// A texel with dx^2 + dy^2 >= r^2 has weight max(1 - sqrt(d2) / r, 0) = 0
// exactly (sqrt and the division round monotonically and r / r = 1), adds
// +0 to the weight sum and is not deposited, so each row walks only the
// span where d2 < r^2 can hold, |px - cx| < h = sqrt(r^2 - dy^2): the
// integers in [cx - h', cx + h'] with h' = h + 2^-5 pixel.  The margin
// covers every rounding on the way: r^2 - dy^2 (r <= 24) is off by at most
// 2^-14, which moves its square root by at most 2^-7 near zero; the
// hardware square root adds an ulp; cx -+ h' rounds by at most 2^-8 for
// films under 2^16 pixels (2^-7 under 2^17).  An exact test,
// fma(-r, r, d2) >= 0, skips the candidates outside: the same sums and
// deposits, in the same row-major order, as walking the whole box.
// Coordinates fit in 32 bits (the box is clamped to the film).
__device__ inline void tent_row_span(float cx, float r, float dy, int x0, int x1, int &xs, int &xe) {
    const float hw = __builtin_amdgcn_sqrtf(fmaxf(fmaf(-dy, dy, r * r), 0.0f)) + 0.03125f;
    xs = max(x0, (int)ceilf(cx - hw));
    xe = min(x1, (int)floorf(cx + hw));
}

// A texel's tent weight max(1 - sqrt(d2) / r, 0) (volume.rs:198-205).  FAST:
// the hardware square root (v_sqrt_f32, 1 ulp) times the splat's reciprocal
// radius instead of the IEEE-rounded square root and division (17 + 11
// instructions): each weight moves by about an ulp and the splat's weights
// still sum to one, since the weight sum and the deposits use the same
// values.  The texel set itself (fma(-r, r, d2) < 0) is exact either way.
template <bool FAST>
__device__ inline float tent_weight(float d2, float r, float inv_r) {
    if (FAST) return fmaxf(1.0f - __builtin_amdgcn_sqrtf(d2) * inv_r, 0.0f);
    return fmaxf(1.0f - sqrtf(d2) / r, 0.0f);
}

template <class Sink, bool FAST = false>
__device__ inline void splat_transverse(const PhotonParams &P, float s, float t, f3 color,
                                        float radius_px, Sink &sink) {
    const int W = (int)P.w, H = (int)P.h;
    const float cx = s * (float)W - 0.5f;
    const float cy = (1.0f - t) * (float)H - 0.5f;
    const float r = fmaxf(radius_px, 0.5f);
    const float inv_r = FAST ? 1.0f / r : 0.0f;
    const int x0 = (int)fmaxf(floorf(cx - r), 0.0f);
    int x1 = (int)ceilf(cx + r);
    x1 = x1 < 0 ? 0 : (x1 > W - 1 ? W - 1 : x1);
    const int y0 = (int)fmaxf(floorf(cy - r), 0.0f);
    int y1 = (int)ceilf(cy + r);
    y1 = y1 < 0 ? 0 : (y1 > H - 1 ? H - 1 : y1);
    float wsum = 0.0f;
    for (int py = y0; py <= y1; ++py) {
        const float dy = (float)py - cy;
        int xs, xe;
        tent_row_span(cx, r, dy, x0, x1, xs, xe);
        for (int px = xs; px <= xe; ++px) {
            const float dx = (float)px - cx;
            const float d2 = dx * dx + dy * dy;
            if (fmaf(-r, r, d2) >= 0.0f) continue;
            wsum += tent_weight<FAST>(d2, r, inv_r);
        }
    }
    if (wsum <= 0.0f) {
        int px = (int)roundf(cx), py = (int)roundf(cy);
        px = px < 0 ? 0 : (px > W - 1 ? W - 1 : px);
        py = py < 0 ? 0 : (py > H - 1 ? H - 1 : py);
        deposit_xy(sink, (uint32_t)px, (uint32_t)py, (uint32_t)(py * W + px), color);
        return;
    }
    const float inv = 1.0f / wsum;
    // FAST (the wave-shared LDS tile): each lane starts its deposit rows at
    // row (lane mod rows) and wraps around, so lanes whose tents sit half a
    // pixel apart add to different rows in one LDS instruction instead of
    // to the same texels; the same deposits, in another order.
    const int nrows = y1 - y0 + 1;
    int py = y0 + (FAST && nrows > 1 ? (int)(__lane_id() % (uint32_t)nrows) : 0);
    for (int i = 0; i < nrows; ++i, py = py == y1 ? y0 : py + 1) {
        const float dy = (float)py - cy;
        int xs, xe;
        tent_row_span(cx, r, dy, x0, x1, xs, xe);
        for (int px = xs; px <= xe; ++px) {
            const float dx = (float)px - cx;
            const float d2 = dx * dx + dy * dy;
            if (fmaf(-r, r, d2) >= 0.0f) continue;
            const float w = tent_weight<FAST>(d2, r, inv_r) * inv;
            if (w > 0.0f) deposit_xy(sink, (uint32_t)px, (uint32_t)py, (uint32_t)(py * W + px), color * w);
        }
    }
}

// Beam estimator, split so that a segment's steps can also be spread over
// a wave (the wavefront beam engine): the step count of the segment from ro
// along rd of length seg_len (volume.rs:263-276), and step k's tent splat
// (volume.rs:277-311).
__device__ inline uint32_t beam_step_count(const PhotonParams &P, f3 ro, f3 rd, float seg_len) {
    f3 p_end = ro + rd * seg_len;
    float s0, t0, s1, t1;
    if (project(P.proj, ro, s0, t0) && project(P.proj, p_end, s1, t1)) {
        float dx = (s1 - s0) * (float)P.w;
        float dy = (t1 - t0) * (float)P.h;
        float mf = ceilf(sqrtf(dx * dx + dy * dy) * 2.0f);
        uint64_t mm = (uint64_t)mf + 1u;
        return (uint32_t)(mm < 1 ? 1 : (mm > 4096 ? 4096 : mm));
    }
    uint64_t mm = (uint64_t)ceilf(seg_len * 80.0f);
    return (uint32_t)(mm < 1 ? 1 : (mm > 4096 ? 4096 : mm));
}

// Step k of a segment: its pixel-space centre, tent radius and colour
// (volume.rs:277-305); false when the point is off the film or behind the
// z-buffer.
struct BeamSplat {
    float s, t, radius;
    float weight;  // the step's scalar weight; color = cmf * weight
    f3 color;
};

// The camera projection with hardware reciprocals (v_rcp_f32, 1 ulp) in
// place of the two IEEE divisions: for quantities that only scale a splat
// (its radius), not for where a splat lands.
__device__ inline bool project_approx(const Proj &P, f3 p, float &s, float &t) {
    f3 dir = p - ld3(P.o);
    float c = dot(dir, ld3(P.w));
    if (c >= -1e-6f) return false;
    float ss = 0.5f + dot(dir, ld3(P.u)) * __builtin_amdgcn_rcpf(-c * P.vw);
    float tt = 0.5f + dot(dir, ld3(P.v)) * __builtin_amdgcn_rcpf(-c * P.vh);
    if (ss >= 0.0f && ss < 1.0f && tt >= 0.0f && tt < 1.0f) {
        s = ss;
        t = tt;
        return true;
    }
    return false;
}

// FAST: the step's weight and radius -- the values that scale its splat --
// from hardware square roots, reciprocals and exp2 (each within a few ulp);
// its centre (s, t), pixel and camera distance (the z-buffer test) stay
// IEEE-rounded, so every splat lands where the reference puts it.
template <bool FAST = false>
__device__ inline bool beam_step_prep(const PhotonParams &P, f3 ro, f3 rd, float power, f3 cmf, float ds, uint32_t k,
                                      BeamSplat &o) {
    const f3 cam_o = ld3(P.proj.o);
    float dist = ((float)k + 0.5f) * ds;
    f3 p = ro + rd * dist;
    float s, t;
    if (!project(P.proj, p, s, t)) return false;
    uint32_t idx = pix_index(s, t, P.w, P.h);
    f3 to = cam_o - p;
    float d = fmaxf(length(to), 1e-3f);
    float zb = P.zbuf ? P.zbuf[idx] : INFINITY;
    if (!(d < zb)) return false;
    f3 view;
    float weight;
    if (FAST) {
        const float inv_d = __builtin_amdgcn_rcpf(d);
        view = to * inv_d;
        const float phase = phase_hg_approx(P.vp.g, dot(rd, view));
        const float trans = __expf(-P.vp.sigma_t * d);
        weight = power * P.vp.sigma_s * phase * trans * (inv_d * inv_d) * ds;
    } else {
        view = to / d;
        float cos_theta = dot(rd, view);
        float phase = phase_hg(P.vp.g, cos_theta);
        float trans = expf(-P.vp.sigma_t * d);
        weight = power * P.vp.sigma_s * phase * trans / (d * d) * ds;
    }
    f3 perp = cross(view, mk(0.f, 1.f, 0.f));
    if (dot(perp, perp) < 1e-8f) perp = cross(view, mk(1.f, 0.f, 0.f));
    perp = FAST ? perp * (__builtin_amdgcn_rsqf(dot(perp, perp)) * 0.03f)
                : normalize(perp) * 0.03f;  // R_BEAM, volume.rs:172
    float s2, t2, radius_px;
    if (FAST ? project_approx(P.proj, p + perp, s2, t2) : project(P.proj, p + perp, s2, t2)) {
        float dpx = (s2 - s) * (float)P.w;
        float dpy = (t2 - t) * (float)P.h;
        radius_px = FAST ? __builtin_amdgcn_sqrtf(dpx * dpx + dpy * dpy) : sqrtf(dpx * dpx + dpy * dpy);
        radius_px = radius_px < 1.5f ? 1.5f : (radius_px > 24.0f ? 24.0f : radius_px);
    } else {
        radius_px = 1.0f;
    }
    o.s = s;
    o.t = t;
    o.radius = radius_px;
    o.weight = weight;
    o.color = cmf * weight;
    return true;
}

template <class Sink>
__device__ inline void beam_step(const PhotonParams &P, f3 ro, f3 rd, float power, f3 cmf, float ds, uint32_t k,
                                 Sink &sink) {
    BeamSplat b;
    if (beam_step_prep(P, ro, rd, power, cmf, ds, k, b)) splat_transverse(P, b.s, b.t, b.color, b.radius, sink);
}

template <class Sink>
__device__ inline void volume_beam_segment(const PhotonParams &P, const PhotonState &st, f3 cmf,
                                           float seg_len, Sink &sink) {
    const uint32_t m = beam_step_count(P, st.ro, st.rd, seg_len);
    const float ds = seg_len / (float)m;
    for (uint32_t k = 0; k < m; ++k) beam_step(P, st.ro, st.rd, st.power, cmf, ds, k, sink);
}

// One volume bounce: segment deposits, then scatter.  Returns false when
// the photon has escaped (its last segment is still deposited).
template <uint32_t EST, class Sink>
__device__ inline bool volume_bounce(const SceneData &S, const float4 *tbl, const PhotonParams &P,
                                     PhotonState &st, Sink &sink) {
    Hit h;
    int mi = scene_hit(S, st.ro, st.rd, h);
    float seg_len = mi >= 0 ? h.t : P.vp.max_dist;
    const f3 cmf = st.cmf;
    if (EST == SPX_EST_POINT) volume_point_segment<false>(P, tbl, st, cmf, seg_len, st.rng, sink);
    else if (EST == SPX_EST_LEGACY) volume_point_segment<true>(P, tbl, st, cmf, seg_len, st.rng, sink);
    else volume_beam_segment(P, st, cmf, seg_len, sink);
    if (mi < 0) return false;
    photon_scatter(S, tbl, P, mi, h, st);
    sink_scattered(sink, S.mats[mi].kind);
    return true;
}

// The film copy this block deposits into.
__device__ inline float *deposit_film(const PhotonParams &P) {
    return P.replicas > 1 ? P.film + (size_t)(blockIdx.x % P.replicas) * P.rep_stride : P.film;
}

// Sinks ---------------------------------------------------------------------
struct GlobalSink {
    float *film;
    __device__ inline void add(uint32_t idx, f3 v) { film_add(film, idx, v); }
};

struct NullSink {
    __device__ inline void add(uint32_t, f3) {}
};

}  // namespace spx
