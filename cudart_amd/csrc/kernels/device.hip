// Device memory, streams and the renderer entry points (ParallelRenderer,
// src/include/renderer.h:102-126, src/renderer/parallel_renderer.cu).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../host/common.hpp"
#include "../host/scene.hpp"
#include "kernel_api.hpp"
#include "renderer.hpp"

namespace prt {

int hip_fail(hipError_t e, const char* what) {
    return fail(PRT_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr)                                      \
    do {                                                   \
        hipError_t e_ = (expr);                            \
        if (e_ != hipSuccess) return hip_fail(e_, #expr);  \
    } while (0)

void release_device_scene(DeviceScene& d) {
    if (d.device < 0) return;
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(d.device);
    void* ptrs[] = {d.nodes, d.tris, d.shade, d.mats, d.lights, d.texels, d.texinfo, d.lin};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    (void)hipSetDevice(prev);
    d = DeviceScene();
}

namespace {
template <class T>
hipError_t upload(void** dst, const std::vector<T>& v, hipStream_t stream) {
    size_t bytes = std::max<size_t>(v.size() * sizeof(T), 16);
    hipError_t e = hipMalloc(dst, bytes);
    if (e != hipSuccess) return e;
    if (!v.empty()) e = copy_wait(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, stream);
    return e;
}
}  // namespace

// Uploads the scene to `device` once (copyMemoryToCuda, parallel_renderer.cu:89-126)
// and refreshes the material table when it changed.
int device_scene(prt_scene* s, int device, SceneArgs& out, hipStream_t stream) {
    std::lock_guard<std::mutex> g(s->dev_mutex);
    DeviceScene* d = nullptr;
    for (auto& x : s->devices)
        if (x.device == device) d = &x;
    if (!d) {
        DeviceScene n;
        n.device = device;
        hipError_t e = hipSuccess;
        if (e == hipSuccess) e = upload(&n.nodes, s->bvh.nodes, stream);
        if (e == hipSuccess) e = upload(&n.tris, s->tris, stream);
        if (e == hipSuccess) e = upload(&n.shade, s->shade, stream);
        if (e == hipSuccess) e = upload(&n.mats, s->mats, stream);
        if (e == hipSuccess) e = upload(&n.lights, s->lights, stream);
        if (e == hipSuccess) e = upload(&n.texels, s->texels, stream);
        if (e == hipSuccess) e = upload(&n.texinfo, s->texinfo, stream);
        if (e == hipSuccess) e = hipMalloc(&n.lin, 256 * sizeof(float));
        if (e == hipSuccess) e = copy_wait(n.lin, s->lin, 256 * sizeof(float), hipMemcpyHostToDevice, stream);
        n.mats_version = s->mats_version;
        if (e != hipSuccess) {
            release_device_scene(n);
            return hip_fail(e, "uploading scene");
        }
        s->devices.push_back(n);
        d = &s->devices.back();
    } else if (d->mats_version != s->mats_version) {
        HIP_TRY(copy_wait(d->mats, s->mats.data(), s->mats.size() * sizeof(float), hipMemcpyHostToDevice, stream));
        d->mats_version = s->mats_version;
    }
    out.nodes = static_cast<const float4*>(d->nodes);
    out.tris = static_cast<const float4*>(d->tris);
    out.shade = static_cast<const float4*>(d->shade);
    out.mats = static_cast<const float4*>(d->mats);
    out.lights = static_cast<const uint32_t*>(d->lights);
    out.texels = static_cast<const uint8_t*>(d->texels);
    out.texinfo = static_cast<const uint4*>(d->texinfo);
    out.lin = static_cast<const float*>(d->lin);
    out.lights_area = s->lights_area;
    out.inv_lights_area = 1.f / s->lights_area;
    out.num_lights = uint32_t(s->lights.size());
    out.num_records = s->num_triangles;
    return PRT_OK;
}

CamArgs cam_args(const prt_camera& c) {
    CamArgs a;
    std::memcpy(a.eye, c.eye, sizeof a.eye);
    std::memcpy(a.dir, c.dir, sizeof a.dir);
    std::memcpy(a.up, c.up, sizeof a.up);
    std::memcpy(a.right, c.right, sizeof a.right);
    a.fov = c.fov;
    a.aspect = c.aspect;
    a.aperture_width = c.aperture_width;
    a.focus_distance = c.focus_distance;
    a.xpixels = c.xpixels;
    a.ypixels = c.ypixels;
    return a;
}

// gammaCorrectPixel (src/renderer/renderer.cu:495-504) as a table over the 8-bit input.
void gamma_table(uint8_t lut[256]) {
    float invGamma = 1.f / 2.2f;
    for (int b = 0; b < 256; ++b) {
        float f = std::pow(float(b) * 1.f / 255.f, invGamma);
        lut[b] = uint8_t(f * 255.f);
    }
}

int ensure_spill(prt_renderer* r, uint32_t threads) {
    // Entries the deepest walk of this BVH can hold: a node visit pushes at most its
    // three farther children, once per level 0..max_depth.
    const uint32_t need_entries = 3 * (r->scene->bvh.max_depth + 1);
    if (need_entries <= uint32_t(kStackLds)) return PRT_OK;
    size_t bytes = size_t(kStackMax - kStackLds) * threads * sizeof(int);
    if (r->spill && r->spill_threads >= threads) return PRT_OK;
    if (r->spill) (void)hipFree(r->spill);
    r->spill = nullptr;
    HIP_TRY(hipMalloc(reinterpret_cast<void**>(&r->spill), bytes));
    r->spill_threads = threads;
    return PRT_OK;
}

int set_device(int device) {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) return fail(PRT_ERR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= count) return fail(PRT_ERR_INVALID_ARG, "device ordinal out of range");
    HIP_TRY(hipSetDevice(device));
    return PRT_OK;
}

}  // namespace prt

using namespace prt;

extern "C" int prt_device_count(int* count) {
    clear_error();
    if (!count) return fail(PRT_ERR_INVALID_ARG, "null argument");
    hipError_t e = hipGetDeviceCount(count);
    if (e != hipSuccess) {
        *count = 0;
        if (e == hipErrorNoDevice) return PRT_OK;
        return hip_fail(e, "hipGetDeviceCount");
    }
    return PRT_OK;
}

extern "C" void prt_render_config_default(prt_render_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof *c);
    c->max_depth = 6;
    c->rr_start = 3;
    c->seed = 0x5EED;
    c->view_id = 0;
    c->device = 0;
    c->integrator = PRT_INTEGRATOR_MEGAKERNEL;
    c->spp_per_launch = 0;
    c->count_rays = 0;
}

extern "C" int prt_renderer_create(prt_scene* scene, const prt_camera* cam, uint32_t w, uint32_t h, uint32_t spp,
                                   const prt_render_config* cfg, prt_renderer** out) {
    clear_error();
    if (!scene || !out) return fail(PRT_ERR_INVALID_ARG, "null argument");
    if (w == 0 || h == 0 || w > 65536 || h > 65536) return fail(PRT_ERR_INVALID_ARG, "image size out of range");
    // Pixel indices and counts are 32-bit throughout (kernels, tonemap, planner).
    if (uint64_t(w) * h >= (uint64_t(1) << 32)) return fail(PRT_ERR_LIMIT, "image of 2^32 pixels or more");
    if (scene->lights.empty())
        return fail(PRT_ERR_NO_EMITTERS, "scene has no EMISSIVE triangles (the reference divides by a zero light area)");
    prt_render_config c;
    if (cfg) c = *cfg;
    else prt_render_config_default(&c);
    if (c.max_depth == 0 || c.max_depth > 64) return fail(PRT_ERR_INVALID_ARG, "max_depth must be 1..64");
    if (c.integrator != PRT_INTEGRATOR_MEGAKERNEL && c.integrator != PRT_INTEGRATOR_WAVEFRONT)
        return fail(PRT_ERR_INVALID_ARG, "unknown integrator");
    const prt_camera& cm = cam ? *cam : scene->camera;
    if (cm.xpixels != int32_t(w) || cm.ypixels != int32_t(h))
        return fail(PRT_ERR_INVALID_ARG, "camera xpixels/ypixels differ from the image size");
    int rc = set_device(c.device);
    if (rc != PRT_OK) return rc;
    PRT_TRY_BEGIN
    // A failure part-way releases whatever was set up (stream, buffers).
    std::unique_ptr<prt_renderer, int (*)(prt_renderer*)> r(new prt_renderer, prt_renderer_destroy);
    r->scene = scene;
    r->cam = cm;
    r->w = w;
    r->h = h;
    r->spp = spp;
    r->cfg = c;
    HIP_TRY(hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking));
    rc = device_scene(scene, c.device, r->sc, r->stream);
    if (rc != PRT_OK) return rc;
    HIP_TRY(hipEventCreate(&r->ev0));
    HIP_TRY(hipEventCreate(&r->ev1));
    const size_t px = size_t(w) * h;
    HIP_TRY(hipMalloc(reinterpret_cast<void**>(&r->accum), px * sizeof(float4)));
    HIP_TRY(hipMemsetAsync(r->accum, 0, px * sizeof(float4), r->stream));  // F9: the reference never clears it
    HIP_TRY(hipMalloc(reinterpret_cast<void**>(&r->counters), kCntNum * sizeof(unsigned long long)));
    HIP_TRY(hipMemsetAsync(r->counters, 0, kCntNum * sizeof(unsigned long long), r->stream));
    uint8_t lut[256];
    gamma_table(lut);
    HIP_TRY(hipMalloc(reinterpret_cast<void**>(&r->gamma_lut), 256));
    HIP_TRY(copy_wait(r->gamma_lut, lut, 256, hipMemcpyHostToDevice, r->stream));
    if (c.keep_samples && spp > 0)
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&r->samples), px * spp * sizeof(float4)));
    r->tiles_x = (w + 7) / 8;
    r->tiles = r->tiles_x * ((h + 7) / 8);
    rc = ensure_spill(r.get(), ((r->tiles + 7) / 8) * 8 * kBlock);
    if (rc != PRT_OK) return rc;
    HIP_TRY(hipStreamSynchronize(r->stream));
    *out = r.release();
    return PRT_OK;
    PRT_TRY_END
}

extern "C" int prt_render_spp(prt_renderer* r, uint32_t n) {
    clear_error();
    if (!r) return fail(PRT_ERR_INVALID_ARG, "null argument");
    if (n == 0) return PRT_OK;
    int rc = set_device(r->cfg.device);
    if (rc != PRT_OK) return rc;
    rc = device_scene(r->scene, r->cfg.device, r->sc, r->stream);
    if (rc != PRT_OK) return rc;
    if (r->samples) {
        if (r->samples_done == 0) {
            r->samples_ok = true;
            r->samples_version = r->scene->mats_version;
        } else if (r->samples_version != r->scene->mats_version) {
            r->samples_ok = false;  // this accumulation mixes material settings
        }
    }
    if (r->cfg.integrator == PRT_INTEGRATOR_WAVEFRONT) return render_wavefront(r, n);
    MegaArgs a;
    a.sc = r->sc;
    a.cam = cam_args(r->cam);
    a.path = PathArgs{r->cfg.seed, r->cfg.view_id, r->cfg.max_depth, r->cfg.rr_start};
    a.width = r->w;
    a.height = r->h;
    a.tiles_x = r->tiles_x;
    a.tiles = r->tiles;
    a.accum = r->accum;
    a.samples = r->samples;
    a.samples_cap = r->samples ? r->spp : 0;
    a.counters = r->counters;
    a.spill = r->spill;
    a.spill_stride = r->spill_threads;
    const uint32_t per = r->cfg.spp_per_launch ? r->cfg.spp_per_launch : 1u;
    HIP_TRY(hipEventRecord(r->ev0, r->stream));
    for (uint32_t done = 0; done < n;) {
        uint32_t k = std::min(per, n - done);
        a.s0 = r->samples_done + done;
        a.nspp = k;
        HIP_TRY(launch_megakernel(a, r->cfg.count_rays != 0, r->stream));
        done += k;
    }
    HIP_TRY(hipEventRecord(r->ev1, r->stream));
    HIP_TRY(hipEventSynchronize(r->ev1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, r->ev0, r->ev1));
    r->last_ms = ms;
    r->samples_done += n;
    r->samples_total += uint64_t(n) * r->w * r->h;
    return PRT_OK;
}

extern "C" int prt_renderer_samples_rendered(const prt_renderer* r, uint32_t* n) {
    clear_error();
    if (!r || !n) return fail(PRT_ERR_INVALID_ARG, "null argument");
    *n = r->samples_done;
    return PRT_OK;
}

extern "C" int prt_renderer_reset(prt_renderer* r) {
    clear_error();
    if (!r) return fail(PRT_ERR_INVALID_ARG, "null argument");
    int rc = set_device(r->cfg.device);
    if (rc != PRT_OK) return rc;
    HIP_TRY(hipMemsetAsync(r->accum, 0, size_t(r->w) * r->h * sizeof(float4), r->stream));
    HIP_TRY(hipStreamSynchronize(r->stream));
    r->samples_done = 0;
    r->samples_ok = false;
    return PRT_OK;
}

static int mean_to(prt_renderer* r, float* dst_device) {
    if (r->samples_done == 0) return fail(PRT_ERR_INVALID_ARG, "no samples rendered yet");
    HIP_TRY(launch_mean(r->accum, r->w * r->h, r->samples_done, dst_device, r->stream));
    HIP_TRY(hipStreamSynchronize(r->stream));
    return PRT_OK;
}

extern "C" int prt_get_radiance(prt_renderer* r, float* rgb) {
    clear_error();
    if (!r || !rgb) return fail(PRT_ERR_INVALID_ARG, "null argument");
    int rc = set_device(r->cfg.device);
    if (rc != PRT_OK) return rc;
    const size_t bytes = size_t(r->w) * r->h * 3 * sizeof(float);
    float* tmp = nullptr;
    HIP_TRY(hipMalloc(reinterpret_cast<void**>(&tmp), bytes));
    rc = mean_to(r, tmp);
    if (rc == PRT_OK) {
        hipError_t e = copy_wait(rgb, tmp, bytes, hipMemcpyDeviceToHost, r->stream);
        if (e != hipSuccess) rc = hip_fail(e, "copying radiance");
    }
    (void)hipFree(tmp);
    return rc;
}

extern "C" int prt_get_radiance_device(prt_renderer* r, void* device_rgb) {
    clear_error();
    if (!r || !device_rgb) return fail(PRT_ERR_INVALID_ARG, "null argument");
    int rc = set_device(r->cfg.device);
    if (rc != PRT_OK) return rc;
    return mean_to(r, static_cast<float*>(device_rgb));
}

extern "C" int prt_copy_image_bytes(prt_renderer* r, uint8_t* rgba) {
    clear_error();
    if (!r || !rgba) return fail(PRT_ERR_INVALID_ARG, "null argument");
    if (r->samples_done == 0) return fail(PRT_ERR_INVALID_ARG, "no samples rendered yet");
    int rc = set_device(r->cfg.device);
    if (rc != PRT_OK) return rc;
    const size_t n = size_t(r->w) * r->h;
    uint8_t* tmp = nullptr;
    HIP_TRY(hipMalloc(reinterpret_cast<void**>(&tmp), n * 4));
    hipError_t e = launch_tonemap(r->accum, uint32_t(n), 0.f, r->samples_done, r->gamma_lut, tmp, r->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
    if (e == hipSuccess) e = copy_wait(rgba, tmp, n * 4, hipMemcpyDeviceToHost, r->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return hip_fail(e, "tonemap");
    return PRT_OK;
}

extern "C" int prt_first_hit(prt_renderer* r, uint32_t sample, int32_t* tri, int32_t* mat, float* t) {
    clear_error();
    if (!r) return fail(PRT_ERR_INVALID_ARG, "null argument");
    int rc = set_device(r->cfg.device);
    if (rc != PRT_OK) return rc;
    const size_t n = size_t(r->w) * r->h;
    char* buf = nullptr;
    HIP_TRY(hipMalloc(reinterpret_cast<void**>(&buf), n * 12));
    FirstHitArgs a;
    a.sc = r->sc;
    a.cam = cam_args(r->cam);
    a.path = PathArgs{r->cfg.seed, r->cfg.view_id, r->cfg.max_depth, r->cfg.rr_start};
    a.width = r->w;
    a.height = r->h;
    a.sample = sample;
    a.tiles_x = r->tiles_x;
    a.tiles = r->tiles;
    a.tri = reinterpret_cast<int32_t*>(buf);
    a.mat = reinterpret_cast<int32_t*>(buf + n * 4);
    a.t = reinterpret_cast<float*>(buf + n * 8);
    a.spill = r->spill;
    a.spill_stride = r->spill_threads;
    hipError_t e = launch_first_hit(a, r->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
    if (e == hipSuccess && tri) e = copy_wait(tri, a.tri, n * 4, hipMemcpyDeviceToHost, r->stream);
    if (e == hipSuccess && mat) e = copy_wait(mat, a.mat, n * 4, hipMemcpyDeviceToHost, r->stream);
    if (e == hipSuccess && t) e = copy_wait(t, a.t, n * 4, hipMemcpyDeviceToHost, r->stream);
    (void)hipFree(buf);
    if (e != hipSuccess) return hip_fail(e, "first hit");
    return PRT_OK;
}

extern "C" int prt_get_stats(prt_renderer* r, prt_stats* st) {
    clear_error();
    if (!r || !st) return fail(PRT_ERR_INVALID_ARG, "null argument");
    int rc = set_device(r->cfg.device);
    if (rc != PRT_OK) return rc;
    unsigned long long c[kCntNum];
    HIP_TRY(copy_wait(c, r->counters, sizeof c, hipMemcpyDeviceToHost, r->stream));
    st->primary_rays = c[kCntPrimary];
    st->extension_rays = c[kCntExtension];
    st->shadow_rays = c[kCntShadow];
    st->nodes_visited = c[kCntNodes];
    st->tris_tested = c[kCntTris];
    st->hits = c[kCntHits];
    st->samples = r->samples_total;
    st->last_render_ms = r->last_ms;
    st->anyhit_nodes = c[kCntAnyNodes];
    st->anyhit_tris = c[kCntAnyTris];
    for (int k = 0; k < 3; ++k) {
        st->kernel_ms[k] = r->kernel_ms[k];
        st->kernel_launches[k] = r->kernel_launches[k];
    }
    return PRT_OK;
}

extern "C" int prt_reset_stats(prt_renderer* r) {
    clear_error();
    if (!r) return fail(PRT_ERR_INVALID_ARG, "null argument");
    int rc = set_device(r->cfg.device);
    if (rc != PRT_OK) return rc;
    HIP_TRY(fill_wait(r->counters, 0, kCntNum * sizeof(unsigned long long), r->stream));
    r->samples_total = 0;
    return PRT_OK;
}

extern "C" int prt_renderer_destroy(prt_renderer* r) {
    if (!r) return PRT_OK;
    (void)hipSetDevice(r->cfg.device);
    release_wavefront(r);
    if (r->stream) (void)hipStreamSynchronize(r->stream);
    void* ptrs[] = {r->accum, r->counters, r->gamma_lut, r->spill, r->samples};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (r->ev0) (void)hipEventDestroy(r->ev0);
    if (r->ev1) (void)hipEventDestroy(r->ev1);
    if (r->stream) (void)hipStreamDestroy(r->stream);
    delete r;
    return PRT_OK;
}
