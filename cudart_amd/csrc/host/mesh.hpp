// Owned host scene description (what the OBJ loader and the procedural
// generators produce) and the loaders around it.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "prt.h"

struct prt_mesh {
    std::vector<float> positions;     // n*9
    std::vector<float> normals;       // n*9
    std::vector<float> uvs;           // n*6
    std::vector<uint32_t> material_ids;
    std::vector<prt_material> materials;
    std::vector<std::string> material_names;
    std::vector<std::vector<uint8_t>> texture_data;
    std::vector<prt_texture> textures;
    std::vector<std::string> mesh_names;  // one per objl mesh, in file order
    std::vector<uint32_t> mesh_first_tri; // first triangle of each mesh
    prt_camera camera{};
    bool has_camera = false;

    uint32_t num_triangles() const { return uint32_t(material_ids.size()); }
    void add_triangle(const float p[9], const float nrm[9], const float uv[6], uint32_t mat) {
        positions.insert(positions.end(), p, p + 9);
        normals.insert(normals.end(), nrm, nrm + 9);
        uvs.insert(uvs.end(), uv, uv + 6);
        material_ids.push_back(mat);
    }
};

namespace prt {

// objl::Material() defaults (src/include/obj_load.h:151-161): Ni = 0 (F18, kept),
// diffuse = 0, type LAMBERT (= DIFFUSE), colours 0.
prt_material default_material();

int load_obj(const std::string& path, prt_mesh& out);
int camera_from_json(const std::string& path, uint32_t w, uint32_t h, prt_camera& out);
void camera_look_at(const float eye[3], const float target[3], const float up[3], float vfov_deg, uint32_t w,
                    uint32_t h, prt_camera& out);
int generate_scene(const std::string& kind, const std::string& params, uint32_t seed, prt_mesh& out);

}  // namespace prt
