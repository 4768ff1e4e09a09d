// k_plain.hip -- the plain (one-job-per-launch) kernels and utility kernels.
//
// Plain camera: one thread per pixel summing its spp samples in sample order
// (the reference's per-pixel sum, tracer.rs:120-134, and the GPU mirror's
// thread-per-pixel launch, trace.wgsl:526-537).  Plain photons: one thread
// per photon (volume.wgsl:542-647) with fp32 film atomics.
#include "spx_paths.hpp"

namespace spx {

__global__ __launch_bounds__(256) void k_camera_plain(CamParams P) {
    __shared__ SceneData S;
    __shared__ float4 tbl[kTblN];
    stage_scene(S, P.scene, tbl, P.tbl);
    __syncthreads();
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.w * P.h) return;
    uint32_t px = i % P.w, py = i / P.w;
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    for (uint32_t k = 0; k < P.spp; ++k) {
        f3 c = camera_path(S, tbl, P, px, py, P.sample_base + k);
        ax += c.x;
        ay += c.y;
        az += c.z;
    }
    // AccumBuffer::sum += batch (tracer.rs:135-139): one owner per pixel.
    P.film[3 * i + 0] += ax;
    P.film[3 * i + 1] += ay;
    P.film[3 * i + 2] += az;
}

hipError_t launch_camera_plain(const CamParams &p, hipStream_t s) {
    uint32_t n = p.w * p.h;
    hipLaunchKernelGGL(k_camera_plain, dim3((n + 255) / 256), dim3(256), 0, s, p);
    return hipGetLastError();
}

template <uint32_t EST>
__global__ __launch_bounds__(256) void k_volume_plain(PhotonParams P) {
    __shared__ SceneData S;
    __shared__ float4 tbl[kTblN];
    stage_scene(S, P.scene, tbl, P.tbl);
    __syncthreads();
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    PhotonState st;
    photon_begin(P, tbl, P.photon_base + i, st);
    GlobalSink sink{deposit_film(P)};
    for (uint32_t b = 0; b < P.max_depth; ++b)
        if (!volume_bounce<EST>(S, tbl, P, st, sink)) break;
}

__global__ __launch_bounds__(256) void k_lighttrace_plain(PhotonParams P) {
    __shared__ SceneData S;
    __shared__ float4 tbl[kTblN];
    stage_scene(S, P.scene, tbl, P.tbl);
    __syncthreads();
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    PhotonState st;
    photon_begin(P, tbl, P.photon_base + i, st);
    GlobalSink sink{deposit_film(P)};
    for (uint32_t b = 0; b < P.max_depth; ++b)
        if (lighttrace_bounce(S, tbl, P, st, sink) != 0) break;
}

hipError_t launch_photon_plain(const PhotonParams &p, int lighttrace, hipStream_t s) {
    dim3 grid((p.n + 255) / 256), block(256);
    if (lighttrace) hipLaunchKernelGGL(k_lighttrace_plain, grid, block, 0, s, p);
    else if (p.estimator == SPX_EST_POINT) hipLaunchKernelGGL(k_volume_plain<SPX_EST_POINT>, grid, block, 0, s, p);
    else if (p.estimator == SPX_EST_BEAM) hipLaunchKernelGGL(k_volume_plain<SPX_EST_BEAM>, grid, block, 0, s, p);
    else hipLaunchKernelGGL(k_volume_plain<SPX_EST_LEGACY>, grid, block, 0, s, p);
    return hipGetLastError();
}

// Per-photon ray states: the same emission / intersect / scatter code as the
// film kernels, recording (origin, dir) after every scatter
// (cpu_simulate_photon_beam, vol_photons.rs:932-991).
__global__ __launch_bounds__(256) void k_photon_records(PhotonParams P, uint32_t first, uint32_t count,
                                                        uint32_t mode, spx_photon_record *out) {
    __shared__ SceneData S;
    __shared__ float4 tbl[kTblN];
    stage_scene(S, P.scene, tbl, P.tbl);
    __syncthreads();
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= count) return;
    PhotonState st;
    photon_begin(P, tbl, first + j, st);
    spx_photon_record rec;
    rec.lambda = st.lambda;
    rec.tir_count = 0;
    rec.pad = 0;
    for (int a = 0; a < 9; ++a)
        for (int b = 0; b < 2; ++b)
            for (int c = 0; c < 4; ++c) rec.states[a][b][c] = 0.0f;
    rec.states[0][0][0] = st.ro.x; rec.states[0][0][1] = st.ro.y; rec.states[0][0][2] = st.ro.z;
    rec.states[0][1][0] = st.rd.x; rec.states[0][1][1] = st.rd.y; rec.states[0][1][2] = st.rd.z;
    uint32_t ns = 1;
    for (uint32_t b = 0; b < P.max_depth; ++b) {
        Hit h;
        int mi = scene_hit(S, st.ro, st.rd, h);
        if (mode == 2u) {
            float ss, rr, stt;
            bool on = screen_coords(P, st.ro, st.rd, ss, rr, stt);
            float scene_t = mi >= 0 ? h.t : INFINITY;
            if (on && stt < scene_t) break;
        } else if (mode == 0u) {
            float seg_len = mi >= 0 ? h.t : P.vp.max_dist;
            for (int k = 0; k < 4; ++k) (void)(next_f32(st.rng) * seg_len);
        }
        if (mi < 0) break;
        const spx_material &m = S.mats[mi];
        if (m.kind == SPX_MAT_DIELECTRIC) {
            float nh = material_n(m, st.lambda, P.dispersion);
            float n1 = h.front_face ? 1.0f : nh, n2 = h.front_face ? nh : 1.0f;
            f3 dummy;
            if (!refract(st.rd, h.normal, n1, n2, dummy)) rec.tir_count++;
        }
        photon_scatter(S, tbl, P, mi, h, st);
        if (ns < 9) {
            rec.states[ns][0][0] = st.ro.x; rec.states[ns][0][1] = st.ro.y; rec.states[ns][0][2] = st.ro.z;
            rec.states[ns][1][0] = st.rd.x; rec.states[ns][1][1] = st.rd.y; rec.states[ns][1][2] = st.rd.z;
        }
        ++ns;
    }
    rec.num_states = ns;
    out[j] = rec;
}

hipError_t launch_photon_records(const PhotonParams &p, uint32_t first, uint32_t count, uint32_t mode,
                                 spx_photon_record *out, hipStream_t s) {
    hipLaunchKernelGGL(k_photon_records, dim3((count + 255) / 256), dim3(256), 0, s, p, first, count,
                       mode, out);
    return hipGetLastError();
}

// Surface rim + z-buffer: prism_dsotm.rs:59-74 / vol_photons.rs:1517-1538.
__global__ __launch_bounds__(256) void k_surface_rim(const SceneData *scene, spx_camera cam, uint32_t w,
                                                     uint32_t h, float *rim, float *depth) {
    __shared__ SceneData S;
    const uint32_t *src = reinterpret_cast<const uint32_t *>(scene);
    uint32_t *dst = reinterpret_cast<uint32_t *>(&S);
    for (uint32_t k = threadIdx.x; k < sizeof(SceneData) / 4; k += blockDim.x) dst[k] = src[k];
    __syncthreads();
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w * h) return;
    uint32_t px = i % w, py = i / w;
    float s = ((float)px + 0.5f) / (float)w;
    float t = 1.0f - ((float)py + 0.5f) / (float)h;
    f3 rd = primary_ray(cam, s, t);
    Hit hit;
    int mi = scene_hit(S, ld3(cam.origin), rd, hit);
    float x = 0.f, y = 0.f, z = 0.f, dep = INFINITY;
    if (mi >= 0) {
        float c = fabsf(dot(hit.normal, rd));
        c = c < 0.0f ? 0.0f : (c > 1.0f ? 1.0f : c);
        float r = (1.0f - c) * (1.0f - c);
        x = 0.015f + r * 0.45f;
        y = 0.02f + r * 0.5f;
        z = 0.04f + r * 0.6f;
        dep = hit.t;
    }
    if (rim) {
        rim[3 * i + 0] = x;
        rim[3 * i + 1] = y;
        rim[3 * i + 2] = z;
    }
    if (depth) depth[i] = dep;
}

hipError_t launch_surface_rim(const SceneData *scene, const spx_camera &cam, uint32_t w, uint32_t h,
                              float *rim, float *depth, hipStream_t s) {
    hipLaunchKernelGGL(k_surface_rim, dim3((w * h + 255) / 256), dim3(256), 0, s, scene, cam, w, h, rim,
                       depth);
    return hipGetLastError();
}

// xyz_to_linear_srgb + tonemap_to_u8, cie.rs:121-144.
__global__ void k_resolve_srgb8(const float *film, uint32_t n, float scale, float exposure, uint8_t *rgb) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float x = film[3 * i] * scale, y = film[3 * i + 1] * scale, z = film[3 * i + 2] * scale;
    float lin[3] = {3.2406f * x - 1.5372f * y - 0.4986f * z, -0.9689f * x + 1.8758f * y + 0.0415f * z,
                    0.0557f * x - 0.2040f * y + 1.0570f * z};
    for (int c = 0; c < 3; ++c) {
        float v = fmaxf(lin[c] * exposure, 0.0f);
        float mapped = v / (1.0f + v);
        float g = mapped <= 0.0031308f ? 12.92f * mapped : 1.055f * powf(mapped, 1.0f / 2.4f) - 0.055f;
        g = g < 0.0f ? 0.0f : (g > 1.0f ? 1.0f : g);
        rgb[3 * i + c] = (uint8_t)(g * 255.0f + 0.5f);
    }
}

hipError_t launch_resolve_srgb8(const float *film, uint32_t w, uint32_t h, float scale, float exposure,
                                uint8_t *rgb, hipStream_t s) {
    uint32_t n = w * h;
    hipLaunchKernelGGL(k_resolve_srgb8, dim3((n + 255) / 256), dim3(256), 0, s, film, n, scale, exposure, rgb);
    return hipGetLastError();
}

// The viewer's blit pass on an FP32 film (spectral-viewer main.rs:131-174):
// Gaussian-weighted mean over the (2r+1)^2 neighbourhood inside the film,
// times k = target_n / max(total_photons, 1); plus the surface layer; times
// exposure; XYZ -> linear sRGB (the shader's matrix, cie.rs:121-128);
// Reinhard, clamp below at 0, gamma 1/2.2; the Rgba8Unorm store rounds
// 255 * clamp(v, 0, 1) to the nearest byte.  One thread per pixel.
__global__ __launch_bounds__(256) void k_display_resolve(const float *vol, const float *surf, uint32_t w,
                                                         uint32_t h, float k, int radius, float inv2s2,
                                                         float exposure, uint8_t *rgba) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w * h) return;
    const int px = (int)(i % w), py = (int)(i / w);
    float ax = 0.0f, ay = 0.0f, az = 0.0f, wsum = 0.0f;
    for (int dy = -radius; dy <= radius; ++dy)
        for (int dx = -radius; dx <= radius; ++dx) {
            const int sx = px + dx, sy = py + dy;
            if (sx < 0 || sx >= (int)w || sy < 0 || sy >= (int)h) continue;
            const size_t s = (size_t)sy * w + (size_t)sx;
            const float g = expf(-(float)(dx * dx + dy * dy) * inv2s2);
            ax += vol[3 * s + 0] * g;
            ay += vol[3 * s + 1] * g;
            az += vol[3 * s + 2] * g;
            wsum += g;
        }
    const float f = k / fmaxf(wsum, 1e-6f);
    float x = ax * f, y = ay * f, z = az * f;
    if (surf) {
        x += surf[3 * i + 0];
        y += surf[3 * i + 1];
        z += surf[3 * i + 2];
    }
    x *= exposure;
    y *= exposure;
    z *= exposure;
    const float lin[3] = {3.2406f * x - 1.5372f * y - 0.4986f * z, -0.9689f * x + 1.8758f * y + 0.0415f * z,
                          0.0557f * x - 0.2040f * y + 1.0570f * z};
    uchar4 o;
    uint8_t *oc = &o.x;
    for (int c = 0; c < 3; ++c) {
        const float t = fmaxf(lin[c] / (1.0f + lin[c]), 0.0f);
        const float v = powf(t, 1.0f / 2.2f);
        oc[c] = (uint8_t)((v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v)) * 255.0f + 0.5f);
    }
    o.w = 255;
    reinterpret_cast<uchar4 *>(rgba)[i] = o;
}

hipError_t launch_display_resolve(const float *vol, const float *surf, uint32_t w, uint32_t h, float k,
                                  int radius, float inv2s2, float exposure, uint8_t *rgba, hipStream_t s) {
    const uint32_t n = w * h;
    hipLaunchKernelGGL(k_display_resolve, dim3((n + 255) / 256), dim3(256), 0, s, vol, surf, w, h, k, radius,
                       inv2s2, exposure, rgba);
    return hipGetLastError();
}

// Radix-select pass over the composited luminance L_i = vol_Y[i] * k
// (+ surf_Y[i]): a 256-bin histogram of the key digit at `shift` among the
// keys whose bits under `mask` equal `prefix`.  Keys order like the floats
// (sign bit flipped for non-negative values, all bits for negative ones).
__device__ inline uint32_t float_order_key(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__global__ __launch_bounds__(256) void k_lum_digit_hist(const float *vol, const float *surf, uint32_t n, float k,
                                                        uint32_t prefix, uint32_t mask, uint32_t shift,
                                                        uint32_t *hist) {
    __shared__ uint32_t sh[256];
    sh[threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float L = vol[3 * (size_t)i + 1] * k;
        if (surf) L = L + surf[3 * (size_t)i + 1];
        const uint32_t key = float_order_key(L);
        if ((key & mask) == prefix) atomicAdd(&sh[(key >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (sh[threadIdx.x]) atomicAdd(&hist[threadIdx.x], sh[threadIdx.x]);
}

hipError_t launch_lum_digit_hist(const float *vol, const float *surf, uint32_t n, float k, uint32_t prefix,
                                 uint32_t mask, uint32_t shift, uint32_t *hist, hipStream_t s) {
    uint32_t blocks = (n + 255) / 256;
    blocks = blocks < 1 ? 1 : (blocks > 1024 ? 1024 : blocks);
    hipLaunchKernelGGL(k_lum_digit_hist, dim3(blocks), dim3(256), 0, s, vol, surf, n, k, prefix, mask, shift,
                       hist);
    return hipGetLastError();
}

// The same radix select with its decisions kept on the device, in a
// file-scope scratch array (no allocation per call): four 256-bin
// histograms, then {prefix, mask, rank low word, rank high word}.  Every
// call on a device shares that device's copy; calls that overlap on one
// device race.  Host code reaches the array only through the address
// hipGetSymbolAddress returns for the current device.
constexpr uint32_t kQuantileHist = 4 * 256;
constexpr uint32_t kQuantileWords = kQuantileHist + 4;
static __device__ __align__(256) uint32_t scratch[kQuantileWords];

int quantile_scratch_address(uint32_t **out) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    *out = static_cast<uint32_t *>(p);
    return 0;
}

__global__ __launch_bounds__(256) void k_quantile_init(uint32_t *q, unsigned long long rank) {
    for (uint32_t i = threadIdx.x; i < kQuantileHist; i += blockDim.x) q[i] = 0u;
    if (threadIdx.x == 0) {
        q[kQuantileHist + 0] = 0u;  // prefix
        q[kQuantileHist + 1] = 0u;  // mask
        q[kQuantileHist + 2] = (uint32_t)rank;
        q[kQuantileHist + 3] = (uint32_t)(rank >> 32);
    }
}

// One pass: the histogram of the digit at `shift` among keys matching the
// prefix selected so far.
__global__ __launch_bounds__(256) void k_lum_digit_hist_dev(const float *vol, const float *surf, uint32_t n,
                                                            float k, uint32_t shift, uint32_t *q, uint32_t pass) {
    __shared__ uint32_t sh[256];
    sh[threadIdx.x] = 0;
    const uint32_t prefix = __builtin_amdgcn_readfirstlane(q[kQuantileHist + 0]);
    const uint32_t mask = __builtin_amdgcn_readfirstlane(q[kQuantileHist + 1]);
    __syncthreads();
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float L = vol[3 * (size_t)i + 1] * k;
        if (surf) L = L + surf[3 * (size_t)i + 1];
        const uint32_t key = float_order_key(L);
        if ((key & mask) == prefix) atomicAdd(&sh[(key >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (sh[threadIdx.x]) atomicAdd(&q[pass * 256 + threadIdx.x], sh[threadIdx.x]);
}

// The digit holding the rank-th key of the pass's histogram: one lane walks
// the 256 bins (as the host loop of the allocating form does).
__global__ void k_quantile_select(uint32_t *q, uint32_t pass, uint32_t shift) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const uint32_t *h = q + pass * 256;
    unsigned long long rank = (unsigned long long)q[kQuantileHist + 2] |
                              ((unsigned long long)q[kQuantileHist + 3] << 32);
    uint32_t b = 0;
    unsigned long long below = 0;
    while (b < 255u && rank >= below + h[b]) below += h[b++];
    rank -= below;
    q[kQuantileHist + 0] |= b << shift;
    q[kQuantileHist + 1] |= 0xFFu << shift;
    q[kQuantileHist + 2] = (uint32_t)rank;
    q[kQuantileHist + 3] = (uint32_t)(rank >> 32);
}

// The four passes of the select, stream-ordered, with no host round trip;
// the selected key ends in q[kQuantileHist] (the prefix word).
hipError_t launch_quantile_select_dev(const float *vol, const float *surf, uint32_t n, float k,
                                      unsigned long long rank, uint32_t *q, hipStream_t s) {
    hipLaunchKernelGGL(k_quantile_init, dim3(1), dim3(256), 0, s, q, rank);
    hipError_t e = hipGetLastError();
    uint32_t blocks = (n + 255) / 256;
    blocks = blocks < 1 ? 1 : (blocks > 1024 ? 1024 : blocks);
    for (uint32_t pass = 0; pass < 4 && e == hipSuccess; ++pass) {
        const uint32_t shift = 24u - 8u * pass;
        hipLaunchKernelGGL(k_lum_digit_hist_dev, dim3(blocks), dim3(256), 0, s, vol, surf, n, k, shift, q, pass);
        if ((e = hipGetLastError()) != hipSuccess) break;
        hipLaunchKernelGGL(k_quantile_select, dim3(1), dim3(64), 0, s, q, pass, shift);
        e = hipGetLastError();
    }
    return e;
}

uint32_t quantile_scratch_key_offset() { return kQuantileHist; }

// box_blur, prism_on_plane.rs:40-74: one pass of the separable box filter
// (dx = 1: along rows, else along columns), clamp-to-edge, taps summed in
// order then scaled by 1/(2r+1).
__global__ __launch_bounds__(256) void k_box_blur_pass(const float *in, float *out, uint32_t w, uint32_t h,
                                                       int radius, int along_x) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w * h) return;
    const int x = (int)(i % w), y = (int)(i / w);
    const float norm = 1.0f / (float)(2 * radius + 1);
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
    for (int d = -radius; d <= radius; ++d) {
        int xx = along_x ? x + d : x, yy = along_x ? y : y + d;
        xx = xx < 0 ? 0 : (xx > (int)w - 1 ? (int)w - 1 : xx);
        yy = yy < 0 ? 0 : (yy > (int)h - 1 ? (int)h - 1 : yy);
        const float *p = in + 3 * ((size_t)yy * w + (size_t)xx);
        a0 += p[0];
        a1 += p[1];
        a2 += p[2];
    }
    out[3 * (size_t)i + 0] = a0 * norm;
    out[3 * (size_t)i + 1] = a1 * norm;
    out[3 * (size_t)i + 2] = a2 * norm;
}

hipError_t launch_box_blur(const float *in, float *tmp, float *out, uint32_t w, uint32_t h, uint32_t radius,
                           hipStream_t s) {
    const uint32_t n = w * h;
    hipLaunchKernelGGL(k_box_blur_pass, dim3((n + 255) / 256), dim3(256), 0, s, in, tmp, w, h, (int)radius, 1);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_box_blur_pass, dim3((n + 255) / 256), dim3(256), 0, s, tmp, out, w, h, (int)radius, 0);
    return hipGetLastError();
}

// prism_on_plane's surface pass: shade_floor (prism_on_plane.rs:16-38) and
// the per-pixel composite of prism and floor (prism_on_plane.rs:125-154).
__device__ inline f3 shade_floor(const PhotonParams &F, const float *floor_sum, const FloorShade &sh, f3 o, f3 d,
                                 float &t) {
    float s, r;
    if (screen_coords(F, o, d, s, r, t)) {
        const f3 c = screen_sample(F, floor_sum, s, r);
        return mk(sh.ambient[0] + c.x * sh.gain, sh.ambient[1] + c.y * sh.gain, sh.ambient[2] + c.z * sh.gain);
    }
    t = INFINITY;
    return mk(sh.background[0], sh.background[1], sh.background[2]);
}

__global__ __launch_bounds__(256) void k_floor_surface(const SceneData *scene, spx_camera cam, PhotonParams F,
                                                       const float *floor_sum, FloorShade sh, uint32_t w,
                                                       uint32_t h, float *surface, float *depth) {
    __shared__ SceneData S;
    const uint32_t *src = reinterpret_cast<const uint32_t *>(scene);
    uint32_t *dst = reinterpret_cast<uint32_t *>(&S);
    for (uint32_t k = threadIdx.x; k < sizeof(SceneData) / 4; k += blockDim.x) dst[k] = src[k];
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w * h) return;
    const uint32_t px = i % w, py = i / w;
    const float s = ((float)px + 0.5f) / (float)w;
    const float t = 1.0f - ((float)py + 0.5f) / (float)h;
    const f3 o = ld3(cam.origin), rd = primary_ray(cam, s, t);
    Hit hit;
    const int mi = scene_hit(S, o, rd, hit);
    const float t_prism = mi >= 0 ? hit.t : INFINITY;
    float floor_t;
    const f3 fc = shade_floor(F, floor_sum, sh, o, rd, floor_t);
    f3 out;
    float dep;
    if (t_prism < floor_t) {
        float c = fabsf(dot(hit.normal, rd));
        c = c < 0.0f ? 0.0f : (c > 1.0f ? 1.0f : c);
        const float q = 1.0f - c;
        const float rim = q * (q * q);  // powi(3)
        out = mk(fc.x * 0.7f + sh.tint[0] + rim * 0.6f, fc.y * 0.7f + sh.tint[1] + rim * 0.6f,
                 fc.z * 0.7f + sh.tint[2] + rim * 0.7f);
        dep = t_prism;
    } else {
        out = fc;
        dep = floor_t;
    }
    surface[3 * (size_t)i + 0] = out.x;
    surface[3 * (size_t)i + 1] = out.y;
    surface[3 * (size_t)i + 2] = out.z;
    if (depth) depth[i] = dep;
}

hipError_t launch_floor_surface(const SceneData *scene, const spx_camera &cam, const PhotonParams &F,
                                const float *floor_sum, const FloorShade &sh, uint32_t w, uint32_t h,
                                float *surface, float *depth, hipStream_t s) {
    hipLaunchKernelGGL(k_floor_surface, dim3((w * h + 255) / 256), dim3(256), 0, s, scene, cam, F, floor_sum, sh,
                       w, h, surface, depth);
    return hipGetLastError();
}

// tonemap_to_u8(xyz_to_linear_srgb(surface + vol), exposure): prism_on_plane.rs:166-182.
__global__ void k_composite_srgb8(const float *surf, const float *vol, uint32_t n, float exposure, uint8_t *rgb) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float x = surf[3 * i] + vol[3 * i], y = surf[3 * i + 1] + vol[3 * i + 1],
                z = surf[3 * i + 2] + vol[3 * i + 2];
    const float lin[3] = {3.2406f * x - 1.5372f * y - 0.4986f * z, -0.9689f * x + 1.8758f * y + 0.0415f * z,
                          0.0557f * x - 0.2040f * y + 1.0570f * z};
    for (int c = 0; c < 3; ++c) {
        const float v = fmaxf(lin[c] * exposure, 0.0f);
        const float mapped = v / (1.0f + v);
        float g = mapped <= 0.0031308f ? 12.92f * mapped : 1.055f * powf(mapped, 1.0f / 2.4f) - 0.055f;
        g = g < 0.0f ? 0.0f : (g > 1.0f ? 1.0f : g);
        rgb[3 * i + c] = (uint8_t)(g * 255.0f + 0.5f);
    }
}

hipError_t launch_composite_srgb8(const float *surf, const float *vol, uint32_t w, uint32_t h, float exposure,
                                  uint8_t *rgb, hipStream_t s) {
    const uint32_t n = w * h;
    hipLaunchKernelGGL(k_composite_srgb8, dim3((n + 255) / 256), dim3(256), 0, s, surf, vol, n, exposure, rgb);
    return hipGetLastError();
}

// Known-answer kernels ---------------------------------------------------------
__global__ void k_rng_draws(uint32_t key, uint32_t seed, uint32_t n, uint32_t *out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    uint32_t st = rng_new(key, seed);
    for (uint32_t i = 0; i < n; ++i) out[i] = next_u32(st);
}

hipError_t launch_rng_draws(uint32_t key, uint32_t seed, uint32_t n, uint32_t *out, hipStream_t s) {
    hipLaunchKernelGGL(k_rng_draws, dim3(1), dim3(64), 0, s, key, seed, n, out);
    return hipGetLastError();
}

__global__ void k_glass_n(uint32_t glass, const float *lam, uint32_t n, float *out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = glass_n(glass, lam[i]);
}

hipError_t launch_glass_n(uint32_t glass, const float *lam, uint32_t n, float *out, hipStream_t s) {
    hipLaunchKernelGGL(k_glass_n, dim3((n + 255) / 256), dim3(256), 0, s, glass, lam, n, out);
    return hipGetLastError();
}

__global__ void k_fill(float *dst, float v, size_t n) {
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = v;
}

hipError_t launch_fill(float *dst, float v, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill, dim3((uint32_t)blocks), dim3(256), 0, s, dst, v, n);
    return hipGetLastError();
}

__global__ void k_copy(float4 *dst, const float4 *src, size_t n4) {
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) dst[i] = src[i];
}

hipError_t launch_copy(float *dst, const float *src, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if ((n & 3) || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15)) return hipErrorInvalidValue;
    size_t n4 = n / 4, blocks = (n4 + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_copy, dim3((uint32_t)blocks), dim3(256), 0, s, (float4 *)dst, (const float4 *)src, n4);
    return hipGetLastError();
}

}  // namespace spx
