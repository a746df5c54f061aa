/*
 * prt.h — C ABI of the MI355X path tracer (libprt.so).
 *
 * This is the drop-in boundary for CudaRT's per-pixel path-tracing hot path.
 * Every entry point replaces one reference interface, cited as file:line in the
 * CudaRT tree (src/...). Plain pointers and sizes only; no C++ or torch types.
 *
 * Conventions
 *   - Every function returns an int status: PRT_OK (0) or a negative PRT_ERR_*.
 *     The reference calls exit(1) on any CUDA error (src/include/cuda_error_check.h:9-21)
 *     and throws from loaders (src/include/loaders.h:19-23); here nothing exits or
 *     throws across the ABI. prt_last_error() returns this thread's last message.
 *   - Handles are opaque. The library owns all device memory. Output buffers are
 *     caller-owned host memory unless a function says otherwise.
 *   - One renderer handle is used by one host thread at a time; each handle owns
 *     one HIP stream. Calls are synchronous.
 */
#ifndef PRT_H
#define PRT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PRT_ABI_VERSION 1

/* Status codes. */
#define PRT_OK 0
#define PRT_ERR_INVALID_ARG (-1)
#define PRT_ERR_OUT_OF_MEMORY (-2)
#define PRT_ERR_HIP (-3)
#define PRT_ERR_NO_EMITTERS (-4) /* reference divides by a zero light area (src/renderer/renderer.cu:138) */
#define PRT_ERR_UNSUPPORTED (-5)
#define PRT_ERR_IO (-6)
#define PRT_ERR_PARSE (-7)
#define PRT_ERR_NO_DEVICE (-8)
#define PRT_ERR_LIMIT (-9) /* a size beyond a documented maximum */

/* Surface types: refl_t, src/include/material.h:8-15 (same numeric values). */
#define PRT_DIFFUSE 0
#define PRT_SPECULAR 1
#define PRT_DIFFSPEC 2
#define PRT_MICROFACET 3
#define PRT_REFRACTIVE 4
#define PRT_EMISSIVE 5

#define PRT_MAX_MATERIALS 128 /* MAX_MATERIALS, src/include/renderer.h:22 */
#define PRT_NO_TEXTURE (-1)   /* NO_TEXTURE, src/include/material.h:6 */

/* Integrators. */
#define PRT_INTEGRATOR_MEGAKERNEL 0 /* one thread per pixel-sample, whole path in one kernel */
#define PRT_INTEGRATOR_WAVEFRONT 1  /* extend/shade/shadow/compact queues in one HIP graph */

/* Pinhole / thin-lens camera. Same fields and meaning as Camera
 * (src/include/camera.h:9-22); built by prt_camera_from_json the way
 * CameraJsonLoader::getCamera does (src/context/json_loader.cpp:26-48). */
typedef struct prt_camera {
    float eye[3];
    float dir[3];   /* unit view direction */
    float up[3];    /* unit up */
    float right[3]; /* unit, normalize(cross(dir, up)) */
    float fov;      /* tan(vertical field of view / 2) */
    float aspect;   /* width / height */
    float aperture_width; /* 0 = pinhole */
    float focus_distance;
    int32_t xpixels;
    int32_t ypixels;
} prt_camera;

/* One material as the MTL file states it (objl::Material, src/include/obj_load.h:148-200,
 * converted by materialFromMtl, src/geometry/scene.cpp:175-190). ka is the MTL Ka: the
 * library applies LIGHTS_GAIN (x3, scene.cpp:19) and the total light area
 * (scene.cpp:161-164) itself. */
typedef struct prt_material {
    int32_t bsdf;       /* PRT_DIFFUSE .. PRT_EMISSIVE */
    float kd[3];
    float ka[3];
    float ks[3];
    float ns;
    float ni;           /* index of refraction for PRT_REFRACTIVE */
    float diffuse_coef; /* probability of the diffuse lobe for PRT_DIFFSPEC */
    int32_t tex_kd;     /* index into prt_scene_desc.textures or PRT_NO_TEXTURE */
} prt_material;

/* An RGBA8 texture, row-major, rows bottom-to-top in v. */
typedef struct prt_texture {
    uint32_t width;
    uint32_t height;
    const uint8_t *rgba; /* width*height*4 bytes */
} prt_texture;

/* Flat scene description: triangle soup in original (file) order.
 * Mirrors what Scene::loadTriangles produces (src/geometry/scene.cpp:74-138). */
typedef struct prt_scene_desc {
    uint32_t num_triangles;
    const float *positions;      /* [n][3 vertices][xyz] */
    const float *normals;        /* [n][3 vertices][xyz] shading normals */
    const float *uvs;            /* [n][3 vertices][uv], may be NULL (all zero) */
    const uint32_t *material_ids;/* [n] */
    uint32_t num_materials;
    const prt_material *materials;
    uint32_t num_textures;
    const prt_texture *textures;
    prt_camera camera;
} prt_scene_desc;

/* Renderer settings. prt_render_config_default fills the reference defaults. */
typedef struct prt_render_config {
    uint32_t max_depth;     /* path vertices; c_maxBounces = 6 (src/renderer/renderer.cu:25,33) */
    uint32_t rr_start;      /* first bounce with Russian roulette; 3 (renderer.cu:253) */
    uint32_t seed;          /* Philox key word 0 */
    uint32_t view_id;       /* Philox key word 1 */
    int32_t device;         /* HIP device ordinal; the reference picks device 1 when >1 (main.cpp:160-162) */
    uint32_t integrator;    /* PRT_INTEGRATOR_* */
    uint32_t spp_per_launch;/* megakernel: samples per pixel per launch (0 = choose) */
    uint32_t count_rays;    /* 1 = keep per-kind ray / node / triangle counters */
    uint32_t wavefront_group;   /* wavefront: side g (1, 2, 4 or 8) of the pixel block whose
                                   64 / g^2 consecutive samples make one wave's 64 camera paths
                                   (0 = the smallest g that divides the tile) */
    uint32_t wavefront_queue;   /* wavefront: path slots in flight (0 = choose) */
    uint32_t profile_kernels;   /* wavefront: 1 = launch each chunk's kernels without a graph and
                                   time every extend / shade / shadow launch with events */
    uint32_t wavefront_blocks_per_cu; /* wavefront: resident traversal workgroups per compute unit
                                         (0 = as many as fit) */
    uint32_t keep_samples;  /* 1 = also store each pixel-sample's radiance (16 B per pixel-sample, for
                               the first `spp` samples), so the adjoint replays paths without
                               tracing them first */
} prt_render_config;

/* Counters accumulated since the renderer was created (or reset). */
typedef struct prt_stats {
    uint64_t primary_rays;
    uint64_t extension_rays;
    uint64_t shadow_rays;
    uint64_t nodes_visited;   /* four-wide nodes fetched (64 B each) */
    uint64_t tris_tested;     /* 48-byte intersection records fetched */
    uint64_t hits;            /* closest hits shaded (64-byte shading records fetched) */
    uint64_t samples;         /* pixel-samples completed */
    double last_render_ms;    /* device time of the last prt_render_spp / batch call */
    uint64_t anyhit_nodes;    /* wavefront: the part of nodes_visited fetched by shadow rays */
    uint64_t anyhit_tris;     /* wavefront: the part of tris_tested fetched by shadow rays */
    double kernel_ms[3];      /* profile_kernels: device time of the extend, shade and shadow
                                 launches of the last wavefront render */
    uint32_t kernel_launches[3];
} prt_stats;

/* BVH statistics of a built scene. */
typedef struct prt_bvh_info {
    uint32_t num_nodes;        /* four-wide nodes (64 B each: quantized child boxes) */
    uint32_t num_leaves;
    uint32_t max_depth;
    uint32_t max_leaf_size;
    uint64_t node_bytes;
    uint64_t triangle_bytes;   /* intersection records */
    uint64_t shading_bytes;    /* shading records */
    float sah_cost;
    uint32_t node_read_bytes;  /* bytes one node visit reads (the four boxes and child words) */
} prt_bvh_info;

typedef struct prt_mesh prt_mesh;         /* owned host scene description (loader / generator output) */
typedef struct prt_scene prt_scene;       /* built scene: BVH + lights + materials, uploaded on demand */
typedef struct prt_renderer prt_renderer; /* one image: accumulation buffer + stream */

/* ---- errors / version ---------------------------------------------------- */
const char *prt_last_error(void);
int prt_abi_version(void);
int prt_device_count(int *count); /* cudaGetDeviceCount, src/main.cpp:155 */

/* ---- scene input (src/context, src/geometry) ------------------------------ */

/* Loads an OBJ (+ its mtllib) the way objl::Loader::LoadFile does
 * (src/include/obj_load.h:441-723, LoadMaterials :1016-1186) and converts it the
 * way Scene::loadTriangles does (src/geometry/scene.cpp:74-138). One material per
 * newmtl name (fixes the value-dedupe of scene.cpp:82-97,192-203). */
int prt_mesh_load_obj(const char *obj_path, prt_mesh **out);

/* Procedural scenes from a seed (every draw from Philox keyed by seed and a stream id).
 * kind:
 *   "spheres"          displaced level-`level` icospheres (instances=200, radius=1,
 *                      disp=0.05, spread=10, materials=32, ground=100, emission=15):
 *                      the 1M-triangle scene;
 *   "cornell-spheres"  Cornell-like box of displaced UV spheres (tris=100000,
 *                      materials=16, emission=15);
 *   "mixed"            the "spheres" scene sized to about `tris` triangles (tris=100000);
 *   "terrain-spheres"  fBm height field (grid=1415, extent=200) plus icospheres
 *                      (instances=2000) with RGBA8 textures (textures=16, texsize=1024,
 *                      materials=64).
 * width/height set the aspect of the stored camera. params: key=value pairs
 * separated by ';' overriding the defaults above. */
int prt_mesh_generate(const char *kind, const char *params, uint32_t seed, prt_mesh **out);

/* Fills desc with pointers into the mesh (valid until prt_mesh_destroy). */
int prt_mesh_get_desc(const prt_mesh *mesh, prt_scene_desc *desc);
/* Replaces material i of the mesh (e.g. to force the ceiling quad of cornell.obj to
 * be an emitter: its MTL has no `type`, so it loads as diffuse). */
int prt_mesh_set_material(prt_mesh *mesh, uint32_t index, const prt_material *m);
/* Copies the name of material i (the MTL newmtl name) into buf. */
int prt_mesh_material_name(const prt_mesh *mesh, uint32_t index, char *buf, size_t cap);
/* Sets the camera stored with the mesh. */
int prt_mesh_set_camera(prt_mesh *mesh, const prt_camera *cam);
int prt_mesh_destroy(prt_mesh *mesh);

/* Camera from a Maya JSON file: CameraJsonLoader::getCamera (src/context/json_loader.cpp:26-48). */
int prt_camera_from_json(const char *json_path, uint32_t width, uint32_t height, prt_camera *out);
/* Pinhole camera looking from eye at target (vertical fov in degrees). */
int prt_camera_look_at(const float eye[3], const float target[3], const float up[3], float vfov_deg,
                       uint32_t width, uint32_t height, prt_camera *out);

/* .env reader: EnvLoader (src/context/env_loader.cpp:14-37). Writes the three
 * paths (MESHES_PATH, CAMERA_PATH, TEXTURES_PATH) into caller buffers. */
int prt_env_load(const char *env_path, char *meshes, char *cameras, char *textures, size_t cap);

/* ---- scene build ----------------------------------------------------------- */

/* Copies desc, builds the light list (Scene::constructLightList, src/geometry/scene.cpp:148-165),
 * the binned-SAH BVH in the four-wide layout (replaces constructBVH, src/geometry/bvh.cpp:65-323)
 * and packs materials (ParallelRenderer::createMaterialsData, src/renderer/parallel_renderer.cu:252-279).
 * Host-only work: no GPU is touched until a renderer is created. */
int prt_scene_create(const prt_scene_desc *desc, prt_scene **out);
int prt_scene_bvh_info(const prt_scene *scene, prt_bvh_info *info);
/* Copies the built layout out, for host-side validation:
 * nodes: num_nodes*16 floats (64 B each: origin.xyz, then a word of per-axis exponent
 * bytes e+127, four child words (>= 0 inner node, 0x7FFFFFFF empty slot, negative = leaf
 * 0x80000000 | (count-1) << 28 | first record), then six words of 8-bit planes
 * q_lo.x q_hi.x q_lo.y q_hi.y q_lo.z q_hi.z, byte c for child c, each plane
 * fmaf(q, 2^e, origin), and two zero words); tris: n*12 floats (48 B each);
 * tri_order: n uint32 (original triangle id of each record). Any pointer may be NULL. */
int prt_scene_export_bvh(const prt_scene *scene, float *nodes, float *tris, uint32_t *tri_order);
int prt_scene_num_lights(const prt_scene *scene, uint32_t *num_lights, float *lights_area);
int prt_scene_destroy(prt_scene *scene);

/* ---- renderer: ParallelRenderer (src/include/renderer.h:102-126) ---------- */

void prt_render_config_default(prt_render_config *cfg);

/* ParallelRenderer ctor (src/renderer/parallel_renderer.cu:37-77): allocates the
 * accumulation buffer and uploads the scene to cfg->device. cam may be NULL to use
 * the scene's camera. width/height need not be multiples of 16 (fixes parallel_renderer.cu:42,290). */
int prt_renderer_create(prt_scene *scene, const prt_camera *cam, uint32_t width, uint32_t height,
                        uint32_t spp, const prt_render_config *cfg, prt_renderer **out);

/* Renders n more samples per pixel: renderOneSamplePerPixel x n
 * (src/renderer/parallel_renderer.cu:288-301, terminal_launcher.cpp:13-15). */
int prt_render_spp(prt_renderer *r, uint32_t n);
int prt_renderer_samples_rendered(const prt_renderer *r, uint32_t *n); /* getSamplesRendered, renderer.h:91 */
/* Clears the accumulation buffer and the sample counter. */
int prt_renderer_reset(prt_renderer *r);

/* Mean radiance accum/n, fp32 RGB, W*H*3 floats, row-major from the top row
 * (p_imgBuffer[idx]/(sampleNumber+1), parallel_renderer.cu:340-341). */
int prt_get_radiance(prt_renderer *r, float *rgb);
/* Same into device memory owned by the caller (e.g. a torch tensor on the renderer's device). */
int prt_get_radiance_device(prt_renderer *r, void *device_rgb);

/* 8-bit tonemap then gamma: float3ToUchar4 (renderer.h:149-156) followed by
 * gammaCorrectPixel (renderer.cu:495-504), as copyImageBytes does (parallel_renderer.cu:303-310).
 * W*H*4 bytes RGBA. Negative and NaN radiance map to 0. */
int prt_copy_image_bytes(prt_renderer *r, uint8_t *rgba);

/* Primary-ray closest hit of every pixel for one sample index: camera ray
 * (Camera::computeCameraRay, src/geometry/camera.cu:32-57) then intersectBVH
 * (renderer.cu:283-330). Writes original triangle id and material id (-1 on a miss)
 * and the hit distance (FLT_MAX on a miss). Any output pointer may be NULL. */
int prt_first_hit(prt_renderer *r, uint32_t sample, int32_t *tri, int32_t *mat, float *t);

int prt_get_stats(prt_renderer *r, prt_stats *stats);
int prt_reset_stats(prt_renderer *r);
int prt_renderer_destroy(prt_renderer *r);

/* ---- batched wavefront (no reference counterpart) ------------------------- */

/* One view of one scene in a batch. */
typedef struct prt_job {
    prt_scene *scene;
    prt_camera camera;      /* xpixels/ypixels give the resolution */
    uint32_t spp;
    uint32_t view_id;       /* Philox key word 1 for this view */
    float *out_rgb;         /* host, xpixels*ypixels*3 mean radiance; may be NULL */
} prt_job;

/* Tunables picked by prt_autotune. Zero fields mean "choose". */
typedef struct prt_tune {
    uint32_t queue_slots;   /* path slots per chunk (the batch runs in ceil(paths / slots) chunks) */
    uint32_t group_side;    /* camera-path grouping of wf_generate: side g (1, 2, 4, 8) of the pixel
                               block of one wave's 64 paths (as prt_render_config.wavefront_group) */
    uint32_t block_size;    /* threads per workgroup of the traversal kernels: only 64 (one wave)
                               is built; 0 or 64 are accepted, anything else is PRT_ERR_INVALID_ARG */
    uint32_t blocks_per_cu; /* persistent traversal workgroups per compute unit */
} prt_tune;

/* Renders every job in one persistent wavefront launch sequence (a captured HIP
 * graph replayed until all work drains). Per-job images equal the megakernel's for
 * the same seed/view_id/depth. stats may be NULL. */
int prt_render_batch(prt_job *jobs, uint32_t num_jobs, const prt_render_config *cfg,
                     const prt_tune *tune, prt_stats *stats);

/* Times candidate tunings on a trial of the given jobs (each view at its own width
 * and samples per pixel over 8 rows spread across its field, cut into as many
 * chunks as the full run would take) and writes the fastest into best: traversal
 * occupancy (blocks_per_cu 20..32), then the camera-path grouping (g = 1, 2, 4, 8),
 * then the chunk count (the fewest chunks that fit, twice and four times as many),
 * each searched with the others held at their best so far. best_ms = the trial's time. */
int prt_autotune(prt_job *jobs, uint32_t num_jobs, const prt_render_config *cfg, prt_tune *best,
                 double *best_ms);

/* ---- path-replay adjoint (no reference counterpart) ----------------------- */

/* Gradients of the image loss sum_px,c (I - target)^2 / (W*H*3) with respect to every
 * material's diffuse albedo kd (RGB) and emission scale e (MTL Ka = e*(1,1,1)).
 * The adjoint replays each path with the same Philox stream as the forward render
 * (path-replay backpropagation), so nothing per path is stored.
 * target_rgb: host W*H*3; image_rgb: host W*H*3 mean radiance of the forward pass
 * at the same seed and spp (NULL = render it here). d_albedo: num_materials*3,
 * d_emission: num_materials (host, overwritten). loss may be NULL. */
int prt_adjoint(prt_renderer *r, const float *target_rgb, const float *image_rgb, uint32_t spp,
                float *d_albedo, float *d_emission, float *loss);

/* The same for an arbitrary upstream gradient: grad_rgb = dLoss/dI (host, W*H*3) of the
 * mean-radiance image of `spp` samples at the renderer's seed; the backward pass of an
 * autograd function. hess = d2Loss/dI2 of one image element, taken as constant and
 * diagonal (2/(W*H*3) for the mean squared error; 0 for a first-order gradient): it
 * weights the Russian-roulette decisions, whose survival probability depends on the
 * albedos and which change the loss by a finite step. d_albedo: M*3, d_emission: M
 * (overwritten). Supports max_depth <= 16. */
int prt_adjoint_grad(prt_renderer *r, const float *grad_rgb, float hess, uint32_t spp, float *d_albedo,
                     float *d_emission);

/* prt_adjoint_grad for callers that cannot allocate (code under HIP graph capture,
 * tight optimisation or serving loops, harnesses that forbid hipMalloc). No
 * counterpart in the reference. It calls no hipMalloc / hipFree: grad_rgb_device is
 * the caller's device memory on the renderer's device (W*H*3 floats, dLoss/dI), the
 * per-workgroup partial gradients live in a static 16 MiB device array (2^21
 * doubles) that each GPU holds until the process ends, and it never writes a
 * per-sample radiance buffer: without radiance kept by the forward pass
 * (keep_samples) each sample's primal path is traced inside the replay. Limit:
 * ceil8(ceil(W/8) * ceil(H/8)) * 4M + 4M <= 2^21 (e.g. 1920x1080 with 16 materials);
 * larger calls fail with PRT_ERR_LIMIT. The gradients equal prt_adjoint_grad's for
 * the same inputs. Calls that overlap on one GPU (several host threads) share its
 * array and race; serialise them. d_albedo: M*3, d_emission: M (host, overwritten). */
int prt_adjoint_grad_static(prt_renderer *r, const float *grad_rgb_device, float hess, uint32_t spp,
                            float *d_albedo, float *d_emission);

/* Changes material parameters of a built scene in place (re-uploaded on next render). */
int prt_scene_set_material_params(prt_scene *scene, uint32_t index, const float kd[3], float emission);

/* ---- output files ----------------------------------------------------------- */

/* RGBA8 PNG, as Launcher::saveToImage writes it (src/context/launcher.cpp:13-21). */
int prt_write_png(const char *path, uint32_t width, uint32_t height, const uint8_t *rgba);
/* Float RGB portable float map (radiance, for image diffs). */
int prt_write_pfm(const char *path, uint32_t width, uint32_t height, const float *rgb);

#ifdef __cplusplus
}
#endif

#endif /* PRT_H */
