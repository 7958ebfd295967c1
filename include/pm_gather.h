/*
 * pm_gather.h -- C ABI of the MI355X (gfx950) photon-gather library.
 *
 * This is the drop-in boundary for the kNN photon gather of a Jensen-style
 * photon mapper.  Every entry point is plain C: opaque handles, plain
 * pointers and sizes, an int status.  Streams are hipStream_t passed as
 * void* (NULL = the null stream).  Reference citations are relative to the
 * yumcyaWiz/photon_mapping checkout.
 *
 * Semantics kept from the reference:
 *   - neighbour sets are the k smallest (FP32 d^2, photon index) keys, with
 *     d^2 = ((0 + dx*dx) + dy*dy) + dz*dz evaluated without FMA
 *     (include/photon_map.h:33-41, 107-147);
 *   - neighbour lists come out farthest first, i.e. in descending
 *     (d^2, index) order, and r^2 is the k-th smallest d^2
 *     (include/photon_map.h:166-183);
 *   - k > N returns all N photons; an empty map returns an empty set with
 *     r^2 = 0 (include/photon_map.h:112, 173-174);
 *   - photon indices are positions in the array handed to the map
 *     (include/photon_map.h:195-200).
 *
 * Domain: photon and query coordinates must be finite and within +-2^60
 * (every squared distance is then finite in FP32).  pm_photonmap_build
 * rejects a map outside it with PM_ERR_INVALID_ARG.  A query outside it (a
 * NaN or infinite coordinate, or one beyond +-2^60) is left out of the
 * gather and gets r^2 = NaN, radiance NaN, key = all ones, a kNN row of -1
 * and no adjoint contribution (the reference's radiance is NaN there too,
 * and its render loop drops such a sample, examples/main.cpp:49-56); every
 * other query of the call is unaffected.
 */
#ifndef PM_GATHER_H
#define PM_GATHER_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define PM_OK 0
#define PM_ERR_INVALID_ARG 1 /* bad size, null pointer, bad photons, k < 1 */
#define PM_ERR_HIP 2         /* a HIP runtime call failed                  */
#define PM_ERR_OOM 3         /* device allocation failed                   */
#define PM_ERR_NOT_BUILT 4   /* query on a map whose build() has not run   */
#define PM_ERR_UNSUPPORTED 5 /* request outside a documented limit         */

/* ---- memory spaces of pointer arguments --------------------------------- */
#define PM_MEM_DEVICE 0 /* pointers are device memory; call is asynchronous  */
#define PM_MEM_HOST 1   /* pointers are host memory; call stages and syncs   */

/* ---- material types ------------------------------------------------------ */
/* include/material.h:73-245; only Lambert is non-zero inside the estimate.   */
#define PM_MATERIAL_LAMBERT 0
#define PM_MATERIAL_MIRROR 1
#define PM_MATERIAL_GLASS 2

typedef struct pm_context pm_context;
typedef struct pm_photonmap pm_photonmap;

/* One gather query (40 bytes): the surface point of a diffuse hit.
 * pos = SurfaceInfo::position, ns = SurfaceInfo::shadingNormal (not
 * necessarily unit length, include/triangle.h:64-71), wo = the outgoing
 * direction (-ray.direction), material = index into the material table.
 * Replaces the (wo, IntersectInfo) pair of
 * PhotonMapping::computeRadianceWithPhotonMap, include/integrator.h:133-134. */
typedef struct pm_query {
  float pos[3];
  float ns[3];
  float wo[3];
  int32_t material;
} pm_query;

/* One material (16 bytes): Lambert albedo rho (include/material.h:73-105). */
typedef struct pm_material {
  float rho[3];
  int32_t type;
} pm_material;

/* Per-call statistics of the last gather/knn/adjoint on a context.  Kernel
 * times are HIP-event times on the call's stream (filled only when the
 * PM_OPT_TIMING option is on, which synchronises the stream). */
typedef struct pm_stats {
  double sort_ms;      /* query Hilbert keys + anchors + radix sort       */
  double gather_ms;    /* the persistent gather kernel                    */
  double post_ms;      /* knn row sort / adjoint flush                    */
  double build_ms;     /* last photon-map build (incl. anchors)           */
  int64_t groups;      /* query groups processed by the last launch       */
  int64_t fallbacks;   /* groups that re-ran with an unbounded radius     */
  int64_t refinements; /* histogram refinement passes in the last launch  */
  int64_t candidates1; /* photons swept for the run-centre radius bounds  */
  int64_t candidates2; /* photons of the per-lane candidate lists         */
  int64_t cut_lists;   /* candidate lists coarsened to fit (x1000: whole) */
  int64_t runs;        /* runs of nearby queries (a group holds >= 1)     */
  /* wave cycles per phase of a query group, summed over groups: query and
   * radius bound, run-centre bound, candidate lists, histogram pass,
   * thresholds/refinement, selection pass, stash/scatter/outputs, queue */
  int64_t phase_cycles[8];
  int64_t candidates3; /* photons of the selection-pass lists, after they
                        * shrink to the lane windows' sphere              */
  int64_t split_runs;  /* runs halved because their sphere needed more
                        * subtrees than a candidate list holds            */
  int64_t narrowed_runs; /* one-query runs whose loose radius bound was
                          * narrowed over leaf lists first                */
} pm_stats;

/* Options for pm_context_set_option. */
#define PM_OPT_TIMING 1         /* 0/1: record HIP events (syncs stream)    */
#define PM_OPT_WAVES_PER_CU 2   /* persistent waves per CU (default 16)     */
#define PM_OPT_ANCHOR_STRIDE 3  /* photons per radius anchor, 1..4096 (4)   */
#define PM_OPT_LEAF_SIZE 4      /* default kd-tree leaf size (default 8)    */
#define PM_OPT_XCD_AWARE 5      /* 0/1: per-XCD work queues (default 1)     */
#define PM_OPT_COUNT_STATS 6    /* 0/1: count fallbacks/refinements         */
#define PM_OPT_RUN_SPLIT 7      /* query-run split factor x1000 (def. 900)  */
#define PM_OPT_TRACE_RECORDS 8  /* photon tracing: most photons past a path's
                                 * cache kept from the count pass, 0..2^30
                                 * (default 262144); when more are made, or
                                 * with 0, the paths that overflowed their
                                 * cache are traced again                   */

/* ---- errors / version ----------------------------------------------------- */
/* Thread-local text of the last failing call ("" when none). */
const char* pm_last_error(void);
int pm_version(void);

/* ---- context --------------------------------------------------------------- */
int pm_context_create(int device, pm_context** out_ctx);
int pm_context_destroy(pm_context* ctx);
int pm_context_set_option(pm_context* ctx, int option, int64_t value);
int pm_context_get_stats(pm_context* ctx, pm_stats* out);
int pm_context_synchronize(pm_context* ctx, void* stream);

/* ---- photon map ---------------------------------------------------------------
 * Replaces PhotonMap::setPhotons + the KdTree it owns
 * (include/photon_map.h:186-206).  The map copies the photons; positions
 * must be finite.  pos_soa / wi_soa / flux_soa are each 3*n floats laid out
 * as [x0..x_{n-1}, y0..y_{n-1}, z0..z_{n-1}] (flux: r, g, b).
 * Photon.throughput is the flux, Photon.wi the incident direction
 * (include/photon_map.h:10-22). */
int pm_photonmap_create(pm_context* ctx, int64_t n, const float* pos_soa,
                        const float* wi_soa, const float* flux_soa, int mem,
                        pm_photonmap** out_map);
/* Same, from the reference's 36-byte Photon records
 * {throughput[3], position[3], wi[3]} (include/photon_map.h:10-13). */
int pm_photonmap_create_aos(pm_context* ctx, int64_t n, const float* photons,
                            int mem, pm_photonmap** out_map);
int pm_photonmap_destroy(pm_photonmap* map);
/* Replaces PhotonMap::build / KdTree::buildTree (include/photon_map.h:157-164,
 * 202-206): GPU Hilbert-curve sort, balanced tree over leaf buckets of leaf_size
 * photons (0 = context default; 4..256). */
int pm_photonmap_build(pm_photonmap* map, int leaf_size, void* stream);
/* Replaces the flux of every photon (3*n SoA, original photon order)
 * without rebuilding the tree. */
int pm_photonmap_update_flux(pm_photonmap* map, const float* flux_soa, int mem,
                             void* stream);
/* PhotonMap::getNPhotons (include/photon_map.h:194). */
int64_t pm_photonmap_size(const pm_photonmap* map);
/* PhotonMap::getIthPhoton (include/photon_map.h:195): writes the 9 floats
 * {throughput, position, wi} of photon i to host memory. */
int pm_photonmap_get_photon(pm_photonmap* map, int64_t i, float* out9);

/* ---- kNN ---------------------------------------------------------------------
 * Batched PhotonMap::queryKNearestPhotons (include/photon_map.h:208-211).
 * q_soa: 3*m floats [x.., y.., z..].  out_idx: m rows of k int32, each row
 * farthest first; slots past min(k, N) hold -1.  out_r2: m floats.
 * Any k (rows of more than 8192 are sorted through a device scratch of
 * m * 8 * (the power of two >= k) bytes). */
int pm_knn(pm_photonmap* map, int64_t m, const float* q_soa, int k,
           int32_t* out_idx, float* out_r2, int mem, void* stream);

/* ---- radiance gather ----------------------------------------------------------
 * Batched PhotonMapping::computeRadianceWithPhotonMap
 * (include/integrator.h:133-152) and computeCausticsWithPhotonMap (:155-175,
 * same formula on another map): for each query
 *   L = [sum_{i in kNN} f(wo, wi_i) * flux_i] / (n_emit * PI * r^2)
 * with f = rho/PI when dot(wo, ns) >= 0 and dot(wi_i, ns) >= 0, else 0.
 * out_rgb: m*3 floats; out_r2: m floats (may be NULL); out_key (may be NULL):
 * m uint64 neighbour-set keys (bits(r^2) << 32 | index of the k-th
 * neighbour), which encode the neighbour set exactly and are what
 * pm_gather_adjoint reuses. */
int pm_gather_radiance(pm_photonmap* map, int64_t m, const pm_query* q,
                       const pm_material* mats, int n_mats, int k,
                       double n_emit, float* out_rgb, float* out_r2,
                       uint64_t* out_key, int mem, void* stream);

/* ---- allocation-free radiance gather (no reference counterpart) ----------------
 * For callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  Same
 * arguments, semantics and results (bit for bit) as pm_gather_radiance with
 * PM_MEM_DEVICE: every pointer is device memory and the call is
 * asynchronous on the stream.  The call allocates nothing and does not
 * synchronise: its scratch (descriptor, query sort, nearest anchors,
 * counters, and radius bounds when needed) lives in a fixed device array of
 * the library, PM_NOALLOC_SCRATCH_BYTES on each GPU it runs on, set up at
 * pm_context_create and held until the process ends.  Limits:
 *   - m <= PM_NOALLOC_MAX_QUERIES per call (else PM_ERR_UNSUPPORTED);
 *   - the map's radius bounds for min(k, N) are read from its cache when an
 *     earlier gather at that k computed them; otherwise they are computed
 *     into the scratch array on every call, which needs at most
 *     PM_NOALLOC_MAX_ANCHORS anchors (maps of up to 1,048,576 photons at the
 *     default anchor stride of 4), else PM_ERR_UNSUPPORTED.
 * All calls on one GPU share its scratch array, which keeps the last call's
 * contents: calls that overlap on one GPU (other streams, other threads,
 * other contexts) race, and must be ordered by the caller. */
#define PM_NOALLOC_MAX_QUERIES (1 << 20)
#define PM_NOALLOC_MAX_ANCHORS (1 << 18)
#define PM_NOALLOC_SCRATCH_BYTES (25 << 20)
int pm_gather_radiance_noalloc(pm_photonmap* map, int64_t m, const pm_query* q,
                               const pm_material* mats, int n_mats, int k, double n_emit,
                               float* out_rgb, float* out_r2, uint64_t* out_key,
                               void* stream);

/* ---- adjoint (no reference counterpart) ----------------------------------------
 * Reverse-mode derivative of pm_gather_radiance for dl_drgb (m*3):
 *   dl_dflux[c*N + i]  += sum_q s_qi (rho_c/PI) dl_drgb[q,c] / (n_emit PI r_q^2)
 *   dl_dalbedo[3*m + c] += sum_{q: mat=m} dl_drgb[q,c] sum_i s_qi flux_ic
 *                                          / (PI n_emit PI r_q^2)
 * key == NULL: the neighbour sets are recomputed; otherwise the keys from
 * the forward pass are reused.  dl_dflux is 3*N SoA in original photon order,
 * dl_dalbedo is n_mats*3; both are accumulated into (caller zeroes them).
 * fg_mat (may be NULL): per-query index of the primary-hit material of a
 * final-gathering bounce; when given, the forward value is rho_fg * L, so
 * the incoming gradient is scaled by rho_fg and dl_dalbedo[fg_mat] receives
 * dl_drgb * L (include/integrator.h:227-243).
 * With keys and PM_MEM_DEVICE, a call on the same query pointer, m, map and
 * stream as the context's last single-instance gather (no other gather,
 * kNN or map build of the context in between) takes that call's query order
 * instead of sorting the queries again.  The results do not depend on the
 * order; only the speed does. */
int pm_gather_adjoint(pm_photonmap* map, int64_t m, const pm_query* q,
                      const pm_material* mats, int n_mats, int k,
                      double n_emit, const float* dl_drgb, const uint64_t* key,
                      const int32_t* fg_mat, float* dl_dflux,
                      float* dl_dalbedo, int mem, void* stream);

/* ---- batched persistent gather (no reference counterpart) -----------------------
 * Many independent gathers (photon map, k, n_emit, query list) in one
 * persistent launch fed by an atomic work queue.  All pointers are device
 * memory.  Results equal pm_gather_radiance on each instance bit for bit. */
typedef struct pm_instance {
  pm_photonmap* map;
  int64_t m;
  const pm_query* q;
  const pm_material* mats;
  int32_t n_mats;
  int32_t k;
  double n_emit;
  float* out_rgb;
  float* out_r2;    /* may be NULL */
  uint64_t* out_key; /* may be NULL */
} pm_instance;
int pm_gather_batched(pm_context* ctx, int n_inst, const pm_instance* inst,
                      void* stream);

/* ---- prepared batched gather (no reference counterpart) ---------------------------
 * A batched gather prepared once and launched many times, e.g. once per frame
 * inside a HIP graph capture.  pm_plan_create validates the instances,
 * computes the radius anchors they need, keeps its own copies of them and of
 * the instance descriptors in device memory, and reserves every workspace of
 * the launch (it synchronises the stream).  pm_plan_launch then only enqueues
 * kernels on the stream -- no host synchronisation, allocation or memset --
 * reading the queries and writing the outputs at the addresses given at
 * creation; each launch equals pm_gather_batched on those instances bit for
 * bit.  The maps must stay built while the plan lives (their flux may change
 * through pm_photonmap_update_flux); launches of one plan must not overlap. */
typedef struct pm_plan pm_plan;
int pm_plan_create(pm_context* ctx, int n_inst, const pm_instance* inst,
                   void* stream, pm_plan** out_plan);
int pm_plan_launch(pm_plan* plan, void* stream);
int pm_plan_destroy(pm_plan* plan);

/* ---- autotuner (no reference counterpart) ----------------------------------------
 * Times pm_gather_radiance on the caller's workload (device pointers) under a
 * small grid of settings, coordinate by coordinate: leaf size {8, 16, 32}
 * (the map is rebuilt), query-run split factor {0.5 .. 1.25} and persistent
 * waves per CU {8, 12, 16, 20}, each timed as the median of five calls, the
 * value in force timed again in each sweep beside the candidates.  The fastest
 * setting stays in force (map leaf size, context options); results do not
 * depend on it.  Synchronises. */
typedef struct pm_tune_report {
  int32_t leaf_size;    /* chosen leaf size                                  */
  int32_t waves_per_cu; /* chosen persistent waves per CU                    */
  double run_split;     /* chosen query-run split factor                     */
  double best_ms;       /* query sort + gather time of the chosen setting    */
  double initial_ms;    /* the same with the settings found on entry         */
  int32_t n_tried;      /* settings timed                                    */
  int32_t pad_;
} pm_tune_report;
int pm_autotune(pm_photonmap* map, int64_t m, const pm_query* q,
                const pm_material* mats, int n_mats, int k, double n_emit,
                void* stream, pm_tune_report* out);

#ifdef __cplusplus
}
#endif

#endif /* PM_GATHER_H */
