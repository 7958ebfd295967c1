/*
 * pm_scene.h -- synthetic-workload producers of libpm_gather (gfx950).
 *
 * These entry points generate the inputs of the gather on the device: the
 * query points of a camera pass over a small triangle scene, and a surrogate
 * photon array.  They stand in for the reference's Embree ray casting and
 * photon tracing (include/scene.h:307-348, include/integrator.h:393-469) for
 * benchmark and test workloads; they use a counter-based hash RNG, not the
 * reference's PCG32 stream, so they reproduce its distributions, not its
 * exact samples.
 */
#ifndef PM_SCENE_H
#define PM_SCENE_H

#include <stdint.h>

#include "pm_gather.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One scene triangle (80 bytes): vertices, per-vertex shading normals
 * (interpolated without renormalising, include/triangle.h:64-71), the
 * material index and a light flag (1 = emitting face, include/scene.h:41-51). */
typedef struct pmx_triangle {
  float v[3][3];
  float n[3][3];
  int32_t material;
  int32_t is_light;
} pmx_triangle;

/* Pinhole camera of include/camera.h:21-47: `position` is the sensor centre,
 * the pinhole sits at position + focal * forward. */
typedef struct pmx_camera {
  float position[3];
  float forward[3];
  float right[3];
  float up[3];
  float focal;
} pmx_camera;

/* Generates width*height*spp gather queries, pixel-major then sample.
 * depth 0: the primary diffuse hit (include/integrator.h:305-308);
 * depth 1: the final-gathering secondary hit of a cosine-sampled bounce from
 * the primary hit (include/integrator.h:215-244).  Camera samples that miss,
 * or hit a light at the primary vertex, and bounces that escape, are redrawn.
 * tris and cam are host memory; out (m pm_query records) and fg_mat_opt
 * (m int32: material of the primary hit, may be NULL) are device memory.
 * Runs on the null stream and synchronises. */
int pmx_fg_queries(pm_context* ctx, const pmx_triangle* tris, int n_tri,
                   const pmx_camera* cam, int width, int height, int spp,
                   int depth, uint64_t seed, pm_query* out, int32_t* fg_mat_opt);

/* Photon tracing of include/integrator.h:393-469 over the triangle scene:
 * each path leaves a light triangle (uniform choice, uniform point, cosine
 * direction; initial flux Le * n_lights * area * pi), stores a photon
 * (flux, position, -direction) at every Lambert hit, continues with Russian
 * roulette from its second bounce (p = min(max flux component, 1)) and
 * bounces as its material samples it (Lambert: cosine; mirror; glass with
 * Schlick Fresnel, include/material.h:73-245), with the light-leak cosine
 * term of include/integrator.h:25-46.  Paths are numbered from 0 with a
 * counter-based RNG and their photons concatenated in (path, bounce) order
 * until n_store photons are stored; *n_emit receives the number of paths
 * emitted (the estimate's divisor).  Triangles with is_light = 1 are the
 * lights, all with radiance light_le.  mats (n_mat records) and mat_ior
 * (n_mat floats, may be NULL: 1.5) are host memory; the SoA outputs (3 *
 * n_store floats each) are device memory.  Runs on the null stream and
 * synchronises. */
int pmx_trace_photons(pm_context* ctx, const pmx_triangle* tris, int n_tri,
                      const pm_material* mats, const float* mat_ior, int n_mat,
                      const float* light_le, int64_t n_store, int max_depth,
                      uint64_t seed, float* pos_soa, float* wi_soa, float* flux_soa,
                      int64_t* n_emit);

/* Surrogate photon array: n photons uniform by area over the triangles,
 * wi cosine-distributed about the face normal, flux = power * albedo of the
 * face's material * (0.5 + 0.5 u).  tris, mat_albedo (n_mat*3 floats) and
 * power (3 floats) are host memory; the three SoA outputs (3n floats each)
 * are device memory.  Runs on the null stream and synchronises. */
int pmx_surrogate_photons(pm_context* ctx, const pmx_triangle* tris, int n_tri,
                          const float* mat_albedo, int n_mat, const float* power,
                          int64_t n, uint64_t seed, float* pos_soa, float* wi_soa,
                          float* flux_soa);

#ifdef __cplusplus
}
#endif

#endif /* PM_SCENE_H */
