/*
 * ORACLE — test infrastructure only.
 *
 * C entry points of liboracle.so: a CPU restatement of CudaRT's per-pixel path
 * tracer (src/renderer/renderer.cu, src/geometry/{camera.cu,geometry.cu,bvh.cpp,scene.cpp},
 * src/renderer/sequential_renderer.cu). It is the checker for the HIP path and the
 * CPU baseline that is timed beside it. Only the test suite, the smoke check and the
 * benchmark's CPU-baseline leg load it; libprt.so never links or calls it.
 *
 * Scene input uses the struct layouts of include/prt.h (data only).
 */
#ifndef ORACLE_API_H
#define ORACLE_API_H

#include <stdint.h>

#include "prt.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_scene oracle_scene;

typedef struct oracle_cfg {
    uint32_t max_depth; /* c_maxBounces (renderer.cu:25,33) */
    uint32_t rr_start;  /* renderer.cu:253 */
    uint32_t seed;      /* Philox key word 0 */
    uint32_t view_id;   /* Philox key word 1 */
} oracle_cfg;

/* counters[0..2] = primary, extension, shadow rays. */
int oracle_scene_create(const prt_scene_desc *desc, oracle_scene **out);
int oracle_scene_destroy(oracle_scene *s);
int oracle_scene_info(const oracle_scene *s, uint32_t *num_nodes, uint32_t *max_depth,
                      uint32_t *num_lights, float *lights_area);
/* Exports the reference-layout BVH (LinearBVHNode, bvh.h:18-27): per node
 * lo[3], hi[3] as 6 floats; offset, ntris, axis as 3 ints; tri_order = original ids in leaf order. */
int oracle_scene_export_bvh(const oracle_scene *s, float *bounds, int32_t *meta, uint32_t *tri_order);

/* Accumulates samples [s0, s0+n) of every pixel into accum (W*H*3, in/out), adding
 * samples in increasing index order. threads <= 0 means all OpenMP threads. */
int oracle_render_f32(const oracle_scene *s, const prt_camera *cam, uint32_t w, uint32_t h,
                      uint32_t s0, uint32_t n, const oracle_cfg *cfg, float *accum,
                      uint64_t counters[3], int threads);
int oracle_render_f64(const oracle_scene *s, const prt_camera *cam, uint32_t w, uint32_t h,
                      uint32_t s0, uint32_t n, const oracle_cfg *cfg, double *accum,
                      uint64_t counters[3], int threads);
/* Renders only the listed pixels (index y*w+x), n samples from s0, into rgb (count*3). */
int oracle_render_pixels_f32(const oracle_scene *s, const prt_camera *cam, uint32_t w, uint32_t h,
                             const uint32_t *pixels, uint32_t count, uint32_t s0, uint32_t n,
                             const oracle_cfg *cfg, float *rgb, uint64_t counters[3], int threads);

int oracle_render_pixels_f64(const oracle_scene *s, const prt_camera *cam, uint32_t w, uint32_t h,
                             const uint32_t *pixels, uint32_t count, uint32_t s0, uint32_t n,
                             const oracle_cfg *cfg, double *rgb, uint64_t counters[3], int threads);

/* Primary-ray closest hit per pixel for one sample: original triangle id, material id, t. */
int oracle_first_hit(const oracle_scene *s, const prt_camera *cam, uint32_t w, uint32_t h,
                     uint32_t sample, const oracle_cfg *cfg, int32_t *tri, int32_t *mat, float *t,
                     int brute_force, int threads);
/* Closest hit of arbitrary rays (tMin = 1e-5, tMax = FLT_MAX). */
int oracle_trace(const oracle_scene *s, uint32_t n, const float *orig, const float *dir, int32_t *tri,
                 float *t, float *u, float *v, int brute_force, int threads);

/* Single-function entry points used by known-answer tests. */
int oracle_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
int oracle_sincosf(uint32_t n, const float *x, float *s, float *c);
int oracle_triangle_intersect(const float v1[3], const float e1[3], const float e2[3], const float o[3],
                              const float d[3], float *t, float *u, float *v);
int oracle_box_hit(const float lo[3], const float hi[3], const float o[3], const float d[3]);
int oracle_fresnel(const float wo[3], const float n[3], float ior, float *re, float *tr, float tdir[3]);
int oracle_diffuse_dir(const float n[3], float u1, float u2, float dir[3], float *cos_weight);
int oracle_camera_ray(const prt_camera *cam, int32_t x, int32_t y, float u0, float u1, float u2, float u3,
                      float o[3], float d[3]);
int oracle_tonemap(uint32_t n, const float *rgb, uint8_t *rgba);
int oracle_texture_table(float out[256]);
/* Bilinear texture lookup with wrap (the kernels' tex_sample) of n (u, v) pairs. */
int oracle_texture_sample(const prt_texture *tex, uint32_t n, const double *uv, double *rgb, int fp64);

/* Path-replay adjoint: gradients of sum (I - T)^2 / (W*H*3) with respect to every
 * material's kd (RGB) and emission scale, accumulated in fp64. The replayed paths use
 * fp32 arithmetic (fp64 = 0, the same paths the HIP kernels trace) or fp64 (fp64 = 1).
 * image = mean radiance of the forward pass over the same spp and seed.
 * hess = d2Loss/dI2 per element weighting the roulette decisions (2/(W*H*3) for this
 * loss; 0 for the first-order estimate of the gradient of the loss of E[I]).
 * pathwise = 0: derivative of the expected loss (the HIP kernels' estimator);
 * pathwise = 1: derivative of the loss for these fixed random numbers, with every
 * roulette decision held (1/p differentiated at every step) -- what a finite
 * difference with a step too small to flip a decision measures.
 * d_albedo: M*3, d_emission: M (overwritten). */
int oracle_adjoint(const oracle_scene *s, const prt_camera *cam, uint32_t w, uint32_t h, uint32_t spp,
                   const oracle_cfg *cfg, const float *image, const float *target, double *d_albedo,
                   double *d_emission, double hess, int pathwise, int fp64, int threads);
/* Replaces kd and emission scale of material i (ka = e * (1,1,1) in MTL units). */
int oracle_scene_set_material_params(oracle_scene *s, uint32_t index, const float kd[3], float emission);

#ifdef __cplusplus
}
#endif

#endif
