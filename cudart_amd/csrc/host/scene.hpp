// Built scene: everything the kernels read, packed on the host; device copies
// are made per device on first use.
#pragma once

#include <cstdint>
#include <mutex>
#include <vector>

#include "bvh.hpp"
#include "prt.h"

namespace prt {

// Device-resident copies of one scene on one device (pointers are device pointers).
struct DeviceScene {
    int device = -1;
    void* nodes = nullptr;
    void* tris = nullptr;
    void* shade = nullptr;
    void* mats = nullptr;
    void* lights = nullptr;
    void* texels = nullptr;
    void* texinfo = nullptr;
    void* lin = nullptr;
    uint64_t mats_version = 0;
};

void release_device_scene(DeviceScene& d);  // device.hip

}  // namespace prt

struct prt_scene {
    uint32_t num_triangles = 0;
    uint32_t num_materials = 0;
    prt::BvhLayout bvh;
    std::vector<float> tris;   // 12 floats per record
    std::vector<float> shade;  // 16 floats per record
    std::vector<float> mats;   // 16 floats per material
    std::vector<prt_material> materials;  // MTL-level values
    std::vector<uint32_t> lights;         // record indices, by original id
    float lights_area = 0.f;
    std::vector<uint8_t> texels;          // RGBA8, all textures
    std::vector<uint32_t> texinfo;        // 4 per texture: first texel, w, h, 0
    float lin[256];                       // RGBA8 -> linear
    prt_camera camera{};
    uint64_t mats_version = 1;
    std::mutex dev_mutex;
    std::vector<prt::DeviceScene> devices;
};

namespace prt {
void pack_material(const prt_material& m, float lights_area, float out[16]);
}
