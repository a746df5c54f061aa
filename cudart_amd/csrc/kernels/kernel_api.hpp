// Plain-data interface between the host library and the HIP kernels.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <hip/hip_runtime.h>

namespace prt {

constexpr int kBlock = 64;       // one wave64 per workgroup: an 8x8 pixel tile
constexpr int kStackLds = 32;    // traversal stack entries per lane in LDS (megakernel)
#ifndef PRT_QUEUE_STACK_LDS
#define PRT_QUEUE_STACK_LDS 16
#endif
constexpr int kStackLdsQueue = PRT_QUEUE_STACK_LDS;  // LDS entries of the queue-based traversal kernels
constexpr int kStackMax = 128;   // total entries (the rest in the global spill area)

// Device view of a built scene. Layouts:
//   nodes   4 x float4 per four-wide node (64 B, quantized child boxes)
//   tris    3 x float4 per record: {v1.xyz, id}{e1.xyz, material}{e2.xyz, 0} (48 B)
//   shade   4 x float4 per record: normals, uvs, area (64 B)
//   mats    4 x float4 per material: {kd, bsdf}{ka*3*A_L, tex}{ks, ns}{ni, dcoef, 0, 0}
struct SceneArgs {
    const float4* nodes;
    const float4* tris;
    const float4* shade;
    const float4* mats;
    const uint32_t* lights;   // record index of each emissive triangle, by original id
    const uint8_t* texels;    // RGBA8, all textures back to back
    const uint4* texinfo;     // per texture: first texel, width, height, 0
    const float* lin;         // 256-entry RGBA8 -> linear table
    float lights_area;
    float inv_lights_area;    // 1.f / lights_area, as samplePixel computes it
    uint32_t num_lights;
    uint32_t num_records;
};

struct CamArgs {
    float eye[3], dir[3], up[3], right[3];
    float fov, aspect, aperture_width, focus_distance;
    int32_t xpixels, ypixels;
};

struct PathArgs {
    uint32_t seed, view;
    uint32_t max_depth, rr_start;
};

// Counter slots (unsigned long long each).
enum CounterSlot {
    kCntPrimary = 0,
    kCntExtension,
    kCntShadow,
    kCntNodes,
    kCntTris,
    kCntHits,
    kCntSamples,
    kCntAnyNodes,  // wavefront shadow kernel's share of kCntNodes
    kCntAnyTris,   // and of kCntTris
    kCntNum
};

struct MegaArgs {
    SceneArgs sc;
    CamArgs cam;
    PathArgs path;
    uint32_t width, height;
    uint32_t s0, nspp;
    uint32_t tiles_x, tiles;  // 8x8 tiles
    float4* accum;            // width*height, .xyz = running sum
    float4* samples;          // nullptr, or each sample's radiance at [sample * width * height + pixel]
    uint32_t samples_cap;     // samples stored per pixel
    unsigned long long* counters;  // kCntNum or nullptr
    int* spill;               // (kStackMax - kStackLds) * spill_stride ints
    uint32_t spill_stride;
};

struct FirstHitArgs {
    SceneArgs sc;
    CamArgs cam;
    PathArgs path;
    uint32_t width, height, sample;
    uint32_t tiles_x, tiles;
    int32_t* tri;
    int32_t* mat;
    float* t;
    int* spill;
    uint32_t spill_stride;
};

// Launchers (render.hip). Return hipError_t.
hipError_t launch_megakernel(const MegaArgs& a, bool count, hipStream_t s);
hipError_t launch_first_hit(const FirstHitArgs& a, hipStream_t s);
hipError_t launch_tonemap(const float4* accum, uint32_t n, float inv_unused, uint32_t samples, const uint8_t* gamma_lut,
                          uint8_t* out, hipStream_t s);
hipError_t launch_mean(const float4* accum, uint32_t n, uint32_t samples, float* out_rgb, hipStream_t s);

// Workgroup order: consecutive workgroups land on different XCDs (round robin),
// so tile k is given to workgroup (k % per) * 8 + k / per, making each XCD's L2
// see a contiguous band of tiles.
__host__ __device__ inline uint32_t xcd_tile(uint32_t wg, uint32_t tiles) {
    const uint32_t per = (tiles + 7) / 8;
    uint32_t xcd = wg % 8, slot = wg / 8;
    uint32_t t = xcd * per + slot;
    return t;
}

}  // namespace prt
