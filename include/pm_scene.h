/*
 * pm_scene.h -- synthetic-workload producers of libpm_gather (gfx950).
 *
 * These entry points build a triangle scene with a BVH on the device and
 * generate the inputs of the gather from it: traced photon maps and the
 * query points of a camera pass, plus a surrogate photon array.  They stand
 * in for the reference's Embree ray casting and photon tracing
 * (include/scene.h:275-348, include/integrator.h:393-563); they use a
 * counter-based hash RNG, not the reference's PCG32 stream, so they
 * reproduce its distributions, not its exact samples.
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

/* ---- scene: triangles, materials, lights and a BVH ----------------------- */

/* A triangle scene on the device with its bounding-volume hierarchy.
 * Replaces Scene::build / Scene::intersect (include/scene.h:275-348): the
 * reference hands its triangles to Embree; here the library builds a binned
 * surface-area-heuristic BVH on the host and traverses it in HIP kernels. */
typedef struct pmx_scene pmx_scene;

/* Copies the scene to the device and builds its BVH.  tris (n_tri records),
 * mats (n_mat records, rho and type: include/material.h Lambert / Mirror /
 * Glass), mat_ior (n_mat floats, may be NULL: 1.5 for every glass material)
 * are host memory.  Emission (include/light.h:17-45), host memory too:
 *   mat_le (n_mat x 3 floats, may be NULL): a triangle whose material has a
 *     positive emission component is a light of that radiance, as the
 *     reference makes an AreaLight of every face with a positive MTL Ke
 *     (include/scene.h:41-52);
 *   light_le (3 floats, may be NULL): the radiance of every other triangle
 *     with is_light = 1.
 * At least one of them must be given.  Triangle materials outside [0, n_mat)
 * read as material 0.  Synchronous. */
int pmx_scene_create(pm_context* ctx, const pmx_triangle* tris, int n_tri,
                     const pm_material* mats, const float* mat_ior, int n_mat,
                     const float* light_le, const float* mat_le, pmx_scene** out_scene);
int pmx_scene_destroy(pmx_scene* scene);

typedef struct pmx_scene_info {
  int32_t n_tri;
  int32_t n_lights;
  int32_t bvh_nodes;   /* internal nodes */
  int32_t bvh_leaves;
  int32_t bvh_depth;   /* deepest internal node, root = 0 */
  int32_t reserved;
  double sah_cost;     /* expected node + triangle tests per ray */
  double light_area;   /* summed area of the light triangles */
} pmx_scene_info;

int pmx_scene_get_info(const pmx_scene* scene, pmx_scene_info* out);

#define PMX_HIT_BVH 0          /* traverse the BVH */
#define PMX_HIT_BRUTE_FORCE 1  /* test every triangle (same results, for checks) */

/* Closest hits of n rays (Scene::intersect, include/scene.h:307-348): for
 * ray i with origin org_soa[{0,n,2n}+i] and direction dir_soa[...], the
 * triangle with the smallest t in (1e-5, inf) of a two-sided
 * Moller-Trumbore test, ties to the lower triangle index; out_tri = -1 and
 * out_t = inf on a miss.  out_u / out_v (may be NULL) are the barycentrics
 * of vertices 1 and 2.  mem PM_MEM_DEVICE: every array is device memory and
 * the call is asynchronous on the stream; PM_MEM_HOST: host arrays, staged
 * through device memory, synchronous. */
int pmx_scene_intersect(const pmx_scene* scene, int64_t n, const float* org_soa,
                        const float* dir_soa, int method, int32_t* out_tri, float* out_t,
                        float* out_u, float* out_v, int mem, void* stream);

/* ---- workload producers over a scene -------------------------------------- */

/* Generates width*height*spp gather queries, pixel-major then sample.
 * depth 0: the primary diffuse hit (include/integrator.h:305-308);
 * depth 1: the final-gathering secondary hit of a cosine-sampled bounce from
 * the primary hit (include/integrator.h:215-244).  Camera samples that miss,
 * or hit a light or a non-Lambert surface at the primary vertex, and bounces
 * that escape, are redrawn; after 64 draws without a query point the record
 * is a placeholder with material -1 (its estimate is zero).  Final-gathering
 * rays are followed through up to 16 specular bounces, each sampled as the
 * material does for camera paths.  cam is host memory; out (m pm_query records)
 * and fg_mat_opt (m int32: material of the primary hit, may be NULL) are
 * device memory.  Runs on the null stream and synchronises. */
int pmx_fg_queries(const pmx_scene* scene, const pmx_camera* cam, int width, int height,
                   int spp, int depth, uint64_t seed, pm_query* out, int32_t* fg_mat_opt);

#define PMX_PHOTONS_GLOBAL 0    /* a photon at every Lambert hit */
#define PMX_PHOTONS_CAUSTICS 1  /* only at the first Lambert hit after specular bounces */

/* Photon tracing of include/integrator.h:393-563.  Each path leaves a light
 * triangle (uniform choice, uniform point, cosine direction; initial flux
 * Le * n_lights * area * pi), continues with Russian roulette from its
 * second bounce (p = min(max flux component, 1)) and bounces as its
 * material samples it (Lambert: cosine; mirror; glass with Schlick Fresnel,
 * include/material.h:73-245), with the light-leak cosine term of
 * include/integrator.h:25-46.  GLOBAL stores a photon (flux, position,
 * -direction) at every Lambert hit (:431-438); CAUSTICS ends a path at its
 * first Lambert hit and stores a photon there only if every earlier hit was
 * specular (:496-512).  Paths are numbered from 0 with a counter-based RNG
 * (the kind is part of the stream) and their photons concatenated in
 * (path, bounce) order until n_store photons are stored; *n_emit receives
 * the number of paths emitted (the estimate's divisor).  Fails with
 * PM_ERR_INVALID_ARG when 2^26 paths store less than 1e-4 of n_store (for
 * example caustics in a scene without specular surfaces).  The SoA outputs
 * (3 * n_store floats each) are device memory.  Runs on the null stream and
 * synchronises. */
int pmx_trace_photons(const pmx_scene* scene, int kind, int64_t n_store, int max_depth,
                      uint64_t seed, float* pos_soa, float* wi_soa, float* flux_soa,
                      int64_t* n_emit);

/* The reference's own count semantics (PhotonMapping::build,
 * include/integrator.h:393-563): exactly n_paths paths are emitted and every
 * photon they store goes into a new photon map of the scene's context, in
 * (path, bounce) order; *n_stored receives the number stored (possibly 0).
 * The map is not built yet (pm_photonmap_build).  Synchronous. */
int pmx_trace_photon_map(const pmx_scene* scene, int kind, int64_t n_paths, int max_depth,
                         uint64_t seed, pm_photonmap** out_map, int64_t* n_stored);

/* ---- rendering --------------------------------------------------------------- */

/* Parameters of a render (include/integrator.h:49-130, 290-392;
 * examples/main.cpp:22-70). */
typedef struct pmx_render_params {
  int32_t width, height, spp;
  int32_t max_depth;              /* eye-path depth limit (maxDepth) */
  int32_t final_gathering_depth;  /* photon mapping: diffuse hits at a smaller depth get
                                     direct light, caustics and a final-gathering ray */
  int32_t k;                      /* nEstimationGlobal, used for both maps
                                     (include/integrator.h:139, 161) */
  double n_emit_global;           /* paths emitted for the global map (its divisor) */
  double n_emit_caustics;         /* paths emitted for the caustics map */
  uint64_t seed;
  int32_t batch_pixels;           /* pixels per wavefront batch; 0: library default */
  int32_t out_mem;                /* PM_MEM_DEVICE (0) or PM_MEM_HOST: where out_mean and
                                     out_var_opt live */
} pmx_render_params;

typedef struct pmx_render_stats {
  int64_t samples;
  int64_t dropped_samples;        /* NaN, negative or beyond 2^30 (not accumulated) */
  int64_t rays;                   /* closest-hit queries: camera, bounce and shadow rays */
  int64_t gathers_global;         /* photon-density estimates on each map */
  int64_t gathers_caustics;
  int32_t batches;
  int32_t iterations;             /* wavefront shading rounds, summed over batches */
  double ms;                      /* wall time of the call */
} pmx_render_stats;

/* PhotonMapping::integrate for every pixel sample
 * (include/integrator.h:290-377), the reference's render loop
 * (examples/main.cpp:33-70): sample (i, j, s) shoots the camera ray of
 * sensor point u = (2 (j + r0) - width) / height, v = (2 (i + r1) - height) /
 * height; a sample whose radiance is NaN or negative is dropped, and every
 * pixel is divided by spp.  Wavefront: eye paths are shaded in rounds
 * (light hits, next-event estimation with shadow rays, specular splits at
 * depth < 3, final-gathering rays), the photon-density estimates of each
 * round run as pm_gather_radiance calls on global_map (the radiance at
 * diffuse hits at depth >= final_gathering_depth and at final-gathering
 * hits) and caustics_map (may be NULL when final_gathering_depth is 0), and
 * sample radiances accumulate in 32.32 fixed point, so the image does not
 * depend on scheduling.  Both maps must be built.  out_mean (height x width
 * x 3 floats, row i = sensor v) and out_var_opt (the variance of each pixel
 * mean, same shape, may be NULL) are device or host memory as
 * params->out_mem says.  Synchronous. */
int pmx_render_photon_mapping(const pmx_scene* scene, const pmx_camera* cam,
                              pm_photonmap* global_map, pm_photonmap* caustics_map,
                              const pmx_render_params* params, float* out_mean,
                              float* out_var_opt, pmx_render_stats* stats_opt);

/* PathTracing::integrate (include/integrator.h:49-101) for every pixel
 * sample, the same camera, accumulation and outputs as above (k, maps and
 * final_gathering_depth unused).  Synchronous. */
int pmx_render_path_tracing(const pmx_scene* scene, const pmx_camera* cam,
                            const pmx_render_params* params, float* out_mean,
                            float* out_var_opt, pmx_render_stats* stats_opt);

/* The reference's per-ray Integrator::integrate(ray, scene, sampler)
 * (include/integrator.h:11-18, 60-105, 566-569) for n rays.  The radiance
 * arriving along ray i (rays: 6 floats per ray, origin then direction) is one
 * sample of PathTracing::integrate or PhotonMapping::integrate, drawn from the
 * random stream (params->seed, i).  params gives max_depth and, for photon
 * mapping, final_gathering_depth, k and n_emit_global / n_emit_caustics
 * (width, height and spp are ignored; batch_pixels bounds the rays per
 * wavefront batch as it bounds pixels).  A sample whose radiance is NaN,
 * negative or beyond 2^30 gives 0: the reference's render loop skips such a
 * sample and still divides by the sample count (examples/main.cpp:49-56,
 * 67).  rays and out_rgb (3n floats) are host or device memory as
 * params->out_mem says.  The maps must be built (caustics_map may be NULL
 * when final_gathering_depth is 0).  Synchronous. */
#define PMX_INTEGRATOR_PATH_TRACING 0
#define PMX_INTEGRATOR_PHOTON_MAPPING 1
int pmx_integrate_rays(const pmx_scene* scene, int integrator, pm_photonmap* global_map,
                       pm_photonmap* caustics_map, const pmx_render_params* params, int64_t n,
                       const float* rays, float* out_rgb);

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
