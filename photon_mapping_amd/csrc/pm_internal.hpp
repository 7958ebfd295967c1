// Internal declarations shared by the HIP translation units of libpm_gather.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <deque>
#include <string>
#include <vector>

#include "pm_gather.h"

namespace pm {

// ---- tuning constants ------------------------------------------------------
constexpr int kWave = 64;           // CDNA wavefront width
#ifndef PM_BINS
#define PM_BINS 28
#endif
constexpr int kBins = PM_BINS;      // selection histogram bins per query
#ifndef PM_STASH_CAP
#define PM_STASH_CAP 11
#endif
constexpr int kStashCap = PM_STASH_CAP;  // per-query candidates kept past the cut
#ifndef PM_GRID_BITS
#define PM_GRID_BITS 10
#endif
constexpr int kGridBits = PM_GRID_BITS;  // space-filling-curve grid bits per axis
// bits of a curve code: the top 32 bits of the 3 * kGridBits-bit index
constexpr int kCodeBits = 3 * kGridBits;
// Bit 31 of a query's curve key marks a query outside the supported domain
// (kInvalidKeyBit), so curve codes must stay below it.
static_assert(kGridBits >= 6 && kGridBits <= 10, "curve grid bits");
constexpr uint32_t kInvalidKeyBit = 1u << 31;
// Query and photon coordinates must be finite and within +-kCoordLimit
// (2^60): every squared distance then stays finite in FP32.
constexpr float kCoordLimit = 1152921504606846976.0f;
#ifndef PM_QUERY_SORT_LO_BIT
#define PM_QUERY_SORT_LO_BIT 0
#endif
constexpr int kQuerySortLoBit = PM_QUERY_SORT_LO_BIT;  // query keys sort from this bit up
#ifndef PM_ANCHOR_WINDOW
#define PM_ANCHOR_WINDOW 4
#endif
#ifndef PM_ANCHOR_STRIDE
#define PM_ANCHOR_STRIDE 4
#endif
constexpr int kAnchorWindow = PM_ANCHOR_WINDOW;  // anchors tried on each side of a query's
constexpr int kAnchorStride = PM_ANCHOR_STRIDE;  // default photons per radius anchor
constexpr int kAnchorTabBits = 16;  // code prefix bits of the anchor lookup table
constexpr int kRowSortLds = 8192;   // kNN row sort: longest row sorted in LDS (64 KiB)
constexpr int kXcds = 8;            // MI355X accelerator complex dies

// Layout of the per-launch counter block: kXcds work-queue counters, then
// statistics (only written when statistics are requested).
constexpr int kStatBase = 2 * kXcds;
constexpr int kStatFallback = kStatBase + 0;   // groups re-run unbounded
constexpr int kStatRefine = kStatBase + 1;     // histogram refinement rounds
constexpr int kStatCand1 = kStatBase + 2;      // centre-bound candidate photons
constexpr int kStatCand2 = kStatBase + 3;      // per-lane pass candidate photons
constexpr int kStatOverflow = kStatBase + 4;   // candidate lists cut short
constexpr int kStatRuns = kStatBase + 5;       // runs of nearby lanes
constexpr int kStatCand3 = kStatBase + 6;      // selection-pass candidate photons
constexpr int kStatSplit = kStatBase + 7;      // runs split for an overflowing list
constexpr int kStatNarrowed = kStatBase + 8;   // one-lane runs whose sphere was narrowed
static_assert(kStatNarrowed < 32, "statistics end below the phase counters");
// Wave cycles per phase of a group (uint64 each, from word kPhaseBase).
constexpr int kPhaseBase = 32;
constexpr int kPhases = 8;
constexpr int kCounterWords = kPhaseBase + 2 * kPhases;

// ---- error plumbing ----------------------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define PM_HIP(call)                                   \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return pm::hip_fail(e_, #call); \
  } while (0)

// ---- device buffer -----------------------------------------------------------
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  ~DevBuf();
  int reserve(size_t need);  // grows (never shrinks); contents not kept
  void release();
  template <typename T>
  T* as() const { return static_cast<T*>(ptr); }
};

// Device scratch of one gather call: query keys and permutation (double
// buffered for the radix sort), sort histograms and scan partials, the
// nearest anchor of every query, and the work-queue / statistics
// counters.
struct Workspace {
  DevBuf keys, vals, keys2, vals2, counts, partials, q_anchor, counters, tiles;
};

// One tile of a segmented radix sort: elements [base, base + n) of segment
// seg, whose digit-d count lives at cbase + d * stride of the count table.
constexpr int kSortTileElems = 4096;
struct SortTile {
  int64_t base;
  uint32_t n;
  uint32_t cbase;
  uint32_t stride;
  uint32_t seg;
};

// ---- gather instance descriptor (device-resident, read with scalar loads) ----
// One per photon-map x query-batch pairing.  All pointers are device memory.
struct alignas(16) GatherDesc {
  // photon map, sorted by Hilbert code ("slot" order)
  const float4* pos4;      // slot -> (x, y, z, original index as bits)
  const float4* wif4;      // slot -> (wi.x, wi.y, wi.z, flux.r)
  const float2* flux2;     // slot -> (flux.g, flux.b)
  const float4* boxes;     // heap node -> [2n] = min, [2n+1] = max
  const float4* anchors;   // anchor -> (centroid xyz, k-radius bound)
  const uint32_t* anchor_code;  // anchor -> Hilbert code of its first slot
  const uint32_t* anchor_tab;   // code prefix -> first anchor whose code is >= it
  const int32_t* inv_perm; // original index -> slot
  const float* pos_orig;   // 3*n positions in original photon order (SoA)
  // queries
  const pm_query* q;       // RADIANCE / ADJOINT: local query records
  const float* q_soa;      // KNN: 3*m positions
  const pm_material* mats;
  const uint32_t* q_anchor;  // local sorted position -> nearest anchor (slot order)
  // outputs
  float* out_rgb;
  float* out_r2;
  uint64_t* out_key;
  int32_t* out_slots;      // KNN: m rows of k photon indices (unsorted)
  float* anchor_out;       // ANCHOR: written radius bound per anchor
  // adjoint
  const float* dl_drgb;
  const uint64_t* key_in;
  const int32_t* fg_mat;
  float* dl_dflux;         // ADJOINT: 3*n gradients in slot order (accumulated)
  float* dl_dalbedo;
  // sizes / parameters
  int64_t m;               // queries of this instance
  int64_t q_base;          // first sorted position of this instance
  int64_t group_base;      // first global group of this instance
  int64_t n_groups;
  float qlo[3];            // curve quantisation origin
  float qscale[3];         // curve quantisation scale (cells per unit)
  int32_t n;               // photons
  int32_t leaf_size;
  int32_t n_leaves;
  int32_t first_leaf;      // heap index of leaf 0 (power of two)
  int32_t n_anchors;
  int32_t anchor_stride;
  int32_t n_mats;
  int32_t k;               // requested k
  int32_t k_eff;           // min(k, n)
  float n_emit;
  int32_t all_members;     // n <= k: every photon is a neighbour
  int32_t pad_;
};

enum GatherMode : int {
  kModeRadiance = 0,
  kModeKnn = 1,
  kModeAnchor = 2,    // anchor radius bounds (anchor_bound_kernel)
  kModeAdjoint = 3,
};

struct GatherLaunch {
  const GatherDesc* descs;     // device array
  const int64_t* group_prefix; // device, n_inst + 1 entries
  const uint32_t* perm;        // sorted position -> global query index
  const int64_t* q_prefix;     // device, n_inst + 1 (global query offsets)
  uint32_t* counters;          // 2 * kXcds work counters (+ stats)
  int64_t total_groups;
  int32_t n_inst;
  int32_t xcd_aware;
  int32_t count_stats;
  int32_t mode;
  float run_split;             // run cut threshold / smallest lane radius bound
  int32_t pad_;
};

// ---- host-side photon map --------------------------------------------------------
struct Context;

struct PhotonMapImpl {
  Context* ctx = nullptr;
  int64_t n = 0;
  bool built = false;
  int leaf_size = 8;
  int n_leaves = 0;
  int first_leaf = 1;
  int anchor_stride = kAnchorStride;
  int n_anchors = 0;
  // radius bounds of the anchors, one set per k (least recently used
  // replaced); each holds the centroids with w = the k-NN radius bound.  A
  // set pinned by the batch being prepared (pin == that batch's epoch) is
  // never replaced: the cache grows past kAnchorSets instead (a deque, so
  // the sets never move).
  struct AnchorSet {
    int k = -1;
    uint64_t stamp = 0;
    uint64_t pin = 0;
    DevBuf buf;
  };
  static constexpr int kAnchorSets = 4;
  std::deque<AnchorSet> aset = std::deque<AnchorSet>(kAnchorSets);
  uint64_t anchor_clock = 0;
  float qlo[3] = {0, 0, 0};
  float qscale[3] = {1, 1, 1};
  // original-order inputs (SoA, 3n each)
  DevBuf pos_soa, wi_soa, flux_soa;
  // built structure
  DevBuf pos4, wif4, flux2, boxes, anchors, anchor_code, anchor_tab, inv_perm, codes;
  DevBuf desc;             // one GatherDesc for single-instance calls
};

struct Context {
  int device = 0;
  int n_cu = 256;
  int waves_per_cu = 16;
  int anchor_stride = kAnchorStride;
  int leaf_size = 8;
  int xcd_aware = 1;
  float run_split = 0.9f;
  int64_t trace_records = int64_t{1} << 18;  // photon tracing's overflow records
  uint64_t batch_epoch = 0;  // pins the anchor sets of the batch being prepared
  int timing = 0;
  int count_stats = 0;
  pm_stats stats{};
  // workspace
  Workspace ws;
  DevBuf ws_slots;
  DevBuf ws_stage_in, ws_stage_out, ws_stage_aux;
  DevBuf ws_tmp;
  DevBuf ws_grad;  // adjoint: per-photon gradients in slot order (3 n floats)
  // The query order of this context's last single-instance gather on device
  // queries (its permutation, in ws.vals or ws.vals2, lasts until the next
  // sort or map build of the context, which clear this record).  An adjoint
  // with the forward keys on the same queries, map and stream takes it
  // instead of sorting the queries again; its results do not depend on the
  // order, only its speed does.
  struct QueryOrder {
    const void* q = nullptr;
    int64_t m = 0;
    const void* map = nullptr;
    hipStream_t s = nullptr;
    const uint32_t* perm = nullptr;
  };
  QueryOrder last_order;
  // descriptor staging: pinned host arena mirrored into a device arena;
  // `stage_done` is recorded after each upload so the host side is only
  // rewritten once the previous copy has completed.
  void* stage_host = nullptr;
  DevBuf stage_dev;
  size_t stage_cap = 0;
  size_t stage_off = 0;
  hipEvent_t stage_done = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
};

// Descriptor staging (host -> device) for one API call.
int stage_begin(Context* ctx, size_t max_bytes, hipStream_t s);
void* stage_put(Context* ctx, const void* src, size_t bytes);  // returns device ptr
int stage_commit(Context* ctx, hipStream_t s);

// ---- launchers (defined in the .hip files) -----------------------------------------
// Stable radix sort of (key, value) pairs on bits [lo_bit, bits) of the key
// (8-bit digits from lo_bit up; bits below lo_bit keep their input order).
// Returns the index (0 or 1) of the buffer pair that holds the result.
int radix_sort_u32(Workspace* ws, uint32_t* keys, uint32_t* vals,
                   uint32_t* keys_alt, uint32_t* vals_alt, int64_t n, int bits,
                   hipStream_t s, int* which, int lo_bit = 0);
// Reserves the sort histograms and scan partials for n keys.
int sort_reserve(Workspace* ws, int64_t n);
// In-place exclusive prefix sum of n uint32 words (plain kernels, no
// allocation); partial holds scan_partial_words(n) words.
int exclusive_scan_u32(uint32_t* partial, uint32_t* data, int64_t n, hipStream_t s);
int64_t scan_partial_words(int64_t n);
// uint32 words of the sort histograms and of the scan partials for n keys.
int64_t sort_count_words(int64_t n);
int64_t sort_partial_words(int64_t n);
// Segmented sort of n keys in n_seg segments: the tiles it may use, and the
// words of its histograms and scan partials.
int64_t sort_seg_tiles(int64_t n, int n_seg);
int64_t sort_seg_count_words(int64_t n, int n_seg);
int64_t sort_seg_partial_words(int64_t n, int n_seg);
// Stable sort of every segment of a tile table (sort_tiles_kernel) on bits
// [lo_bit, bits): each segment's pairs stay in its own range.
int radix_sort_seg_u32_raw(uint32_t* counts, uint32_t* partials, const SortTile* tiles,
                           int64_t n_tiles, uint32_t* keys, uint32_t* vals, uint32_t* keys_alt,
                           uint32_t* vals_alt, int64_t n, int bits, hipStream_t s, int* which,
                           int lo_bit);
// The same sorts over caller-provided histogram and partial storage (at
// least the words above): they allocate nothing.
int radix_sort_u32_raw(uint32_t* counts, uint32_t* partials, uint32_t* keys, uint32_t* vals,
                       uint32_t* keys_alt, uint32_t* vals_alt, int64_t n, int bits,
                       hipStream_t s, int* which, int lo_bit = 0);

int build_photon_map(PhotonMapImpl* map, hipStream_t s);
int update_flux(PhotonMapImpl* map, const float* flux_dev, hipStream_t s);
// Anchor radius bounds of the map for k_eff (computed on first use, then
// cached); *out receives the anchor array the gather must read.
int ensure_anchors(PhotonMapImpl* map, int k_eff, hipStream_t s, const float4** out,
                   uint64_t pin_epoch = 0);
// The cached anchor set of k_eff (marked most recently used), or nullptr.
const float4* cached_anchors(PhotonMapImpl* map, int k_eff);
// Descriptor of the anchor-bound pass of the map for k_eff writing to out
// (4 floats per anchor; the centroids must already be there).
GatherDesc anchor_desc(const PhotonMapImpl* map, int k_eff, float* out);
// Launches the anchor-bound pass of a descriptor already in device memory
// (dev_gp: its group prefix {0, groups}).
int launch_anchor_bounds(Context* ctx, const GatherDesc* dev_desc, const int64_t* dev_gp,
                         int64_t n_anchors, uint32_t* counters, hipStream_t s);
// Stores d at dst and the prefixes {0, q_total, 0, g_total} at prefix (4
// int64) with a one-thread kernel: no host staging.
int launch_store_desc(const GatherDesc& d, GatherDesc* dst, int64_t* prefix, int64_t q_total,
                      int64_t g_total, hipStream_t s);
void fill_desc_map_part(const PhotonMapImpl* map, GatherDesc* d);

// Hilbert keys for all queries of all instances, stable sort, and (when
// need_anchors) the nearest radius anchor of every sorted query
// (ws->q_anchor, global sorted order; instance i starts at q_prefix[i] there,
// as in query order).
// Reserves every buffer prepare_queries and the gather launch use.
int prepare_reserve(Workspace* ws, int n_inst, int64_t total_q, int64_t total_groups);
int prepare_queries(Workspace* ws, int n_inst, const GatherDesc* dev_descs,
                    const int64_t* dev_q_prefix, const int64_t* dev_g_prefix,
                    int64_t total_q, int64_t total_groups, hipStream_t s,
                    const uint32_t** perm_out, bool need_anchors = true);
// Device storage of prepare_queries: keys / keys2 hold total_q uint32 keys,
// vals / vals2 / q_anchor total_q words; counts and partials the
// sort_count_words / sort_partial_words of total_q for one instance, the
// sort_seg_count_words / sort_seg_partial_words for several, which also use
// sort_seg_tiles tiles.
struct QueryScratch {
  void* keys;
  void* keys2;
  uint32_t* vals;
  uint32_t* vals2;
  uint32_t* q_anchor;
  uint32_t* counts;
  uint32_t* partials;
  SortTile* tiles;  // sort_seg_tiles entries when n_inst > 1, else unused
};
// prepare_queries over caller-provided storage: allocates nothing.
int prepare_queries_raw(const QueryScratch& qs, int n_inst, const GatherDesc* dev_descs,
                        const int64_t* dev_q_prefix, int64_t total_q, hipStream_t s,
                        const uint32_t** perm_out, bool need_anchors);

int launch_gather(Context* ctx, const GatherLaunch& L, hipStream_t s);
// Sorts m kNN rows of k indices farthest first (rows longer than
// kRowSortLds keys sort in `scratch`, which it grows).
int launch_knn_rowsort(const GatherDesc* dev_desc, int64_t m, int k, int32_t* out_idx,
                       DevBuf* scratch, hipStream_t s);
int launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t s);
int launch_fill_f32(float* p, float v, int64_t n, hipStream_t s);
// out[c n + i] += acc[c n + inv_perm[i]], c = 0, 1, 2 (slot-order gradients
// into original photon order).
int launch_slot_grad_add(const float* acc, const int32_t* inv_perm, int64_t n, float* out,
                         hipStream_t s);

}  // namespace pm
