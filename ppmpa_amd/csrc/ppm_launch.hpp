// ppm_launch.hpp — host-side launchers of the gfx950 kernels (ppm_kernels.hip),
// called by the C ABI in ppm_api.cpp.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ppm_layout.hpp"

namespace ppm {

// Device-side scene (pointers into device memory + camera + bounds).
struct SceneBuf {
  DevShape* shapes = nullptr;
  DevMaterial* mats = nullptr;  // n_mats + 1 (M_AIR last)
  DevLight* lights = nullptr;
  DevShape* lin = nullptr;      // shapes tested one by one (== shapes when there is no BVH)
  Node* lin_box = nullptr;      // padded box of each lin shape; lo.w = 1 marks a plane
  Node* bvh = nullptr;          // 2^(bvh_depth+1) - 1 nodes, or null
  DevShape* bprims = nullptr;   // n_bprims shapes in leaf order
  uint32_t n_lin = 0, n_bprims = 0, bvh_depth = 0;
  uint32_t bvh_leaf = 1;        // primitives per BVH leaf
  uint32_t n_shapes = 0, n_mats = 0, n_lights = 0;
  uint32_t n_emit_total = 0;  // photons per pass over all lights
  DevCamera cam{};
  float4 bmin{}, bmax{};
};

// Dynamic LDS a BVH kernel's launch gives a scene: the copies of a small tree
// (all nodes and the primitives as compact spheres), 0 for a large tree or
// none.  The scene view's dyn_lds flag comes from the same test.
constexpr uint32_t BVH_DYN_LDS_BYTES = 255u * 32u + 512u * (16u + 4u);
__host__ __device__ inline uint32_t bvh_dyn_lds(const SceneBuf& b) {
  const bool small = b.bvh != nullptr && b.n_bprims > 0 && b.n_bprims <= BVH_SMALL_PRIMS &&
                     ((2u << b.bvh_depth) - 1u) <= 255u;
  return small ? BVH_DYN_LDS_BYTES : 0u;
}
// The instantiation of a BVH kernel for scene buffers b: 0 no BVH, 1 a large
// tree (read from global memory only), 2 a small tree with its LDS copies.
inline int bvh_kind(const SceneBuf& b) { return b.bvh ? (bvh_dyn_lds(b) != 0u ? 2 : 1) : 0; }

// Photons in regions: region r holds count[r] records starting at
// rec + r * region_cap; offset[r] is the exclusive prefix of the counts, so
// the set reads densely as i -> region r = last with offset[r] <= i.
// A traced set has one region per tracing wave; an uploaded set has one.
struct PhotonBuf {
  PhotonRec* rec = nullptr;
  double* rec64 = nullptr;   // double positions (uploaded sets only)
  uint32_t* count = nullptr;
  uint32_t* offset = nullptr;
  uint32_t n_regions = 0;    // regions in use
  uint32_t max_regions = 0;  // capacity of count/offset
  uint32_t region_cap = 0;
  uint64_t cap = 0;          // records allocated
};

// One-sweep LSD radix sort of a pass's keys (grid buckets, query Morton
// codes), 8-bit digits.  One histogram kernel counts every digit of every
// key; then one kernel per digit, whose blocks take tiles of OS_TILE keys in
// ticket order, rank each tile stably in LDS, find the earlier tiles' keys
// of each digit by decoupled look-back over their published counts, and
// write the tile out one digit run at a time.  ctl: the digit counts
// [digit][256], then one ticket per digit pass on its own 128-byte line
// (cleared by the kernel that writes the keys); look: [digit][tile][256]
// look-back words (cleared by the histogram kernel).
constexpr uint32_t OS_THREADS = 256;
constexpr uint32_t OS_KEYS_PER_THREAD = 16;
constexpr uint32_t OS_TILE = OS_THREADS * OS_KEYS_PER_THREAD;
constexpr uint32_t OS_MAX_DIGITS = 4;
constexpr uint32_t OS_TICKETS = OS_MAX_DIGITS * 256;  // ticket of digit pass p at OS_TICKETS + 32 p
constexpr uint32_t OS_CTL_WORDS = OS_TICKETS + 32 * OS_MAX_DIGITS;
constexpr uint32_t OS_BLOCKS_PER_CU = 2;  // persistent scatter blocks
constexpr uint32_t OS_HIST_BLOCKS = 64;   // of 1024 threads
struct OneSweep {
  uint32_t* ctl = nullptr;
  uint32_t* look = nullptr;
  uint32_t tiles = 0;   // tiles of the capacity
  uint32_t blocks = 0;  // persistent scatter blocks
  uint64_t cap = 0;     // keys at most (< 2^31 - 1)
};

// A photon map in one of two forms.  Tree (ppm_map_build: any radius per
// call): photons in leaf order under a heap-ordered box tree.  Grid (the
// passes of a session or batch, whose queries share the pass radius): photons
// sorted by hashed cell, cells[b] = [start, end) of bucket b.
struct MapBuf {
  uint64_t cap = 0;
  PhotonRec* sorted = nullptr;  // photons in tree / bucket order
  double* sorted64 = nullptr;   // optional double positions in that order (3 per photon)
  uint32_t* src_idx = nullptr;  // that order -> index in the input set
  Node* nodes = nullptr;        // tree: 2 * lpad_max - 1
  uint32_t lpad_max = 0;
  uint2* cells = nullptr;       // grid: cells_max buckets
  uint32_t cells_max = 0;       // a power of two
  uint32_t* keys[2] = {nullptr, nullptr};
  uint32_t* vals[2] = {nullptr, nullptr};
  uint32_t* ghist = nullptr;    // 256 * sort_blocks
  uint32_t* totals = nullptr;   // 256
  uint32_t sort_blocks = 0;
  OneSweep os;                  // grid: the per-pass bucket sort
  int morton_bits = 10;         // per axis
};

// Eye-pass buffers.  Queries: slot t in [0, q_primary) is pixel t's primary
// query (8x8-tile order); the k-th deeper query of pixel t (k < sec_max)
// sits at q_primary * (1 + k) + t, so round k of a tile reads 64 consecutive
// slots; further ones go to an overflow region of ovf_cap slots taken with a
// global counter.  sec_count[t] = deeper queries of pixel t kept in place.
// The gather sorts the in-place queries along a Morton curve of their
// positions (qkeys/qvals, per-tile counts tile_count -> tile_off) so that
// each wave gathers 64 spatially adjacent queries; res[slot] receives each
// query's weighted radiance.
struct EyeBuf {
  Query* q = nullptr;
  float4* qpos = nullptr;  // q[slot].a again, densely: the sort keys' only input
  uint32_t q_primary = 0;
  uint32_t sec_max = 0;
  uint32_t* sec_count = nullptr;
  uint32_t ovf_cap = 0;
  float4* pix = nullptr;  // per-pixel float radiance of the pass (row-major)
  uint32_t n_pix = 0;
  uint32_t n_tiles = 0;
  uint32_t tiles_cap = 0;  // the buffers hold the tiles of any tile shape
  uint32_t* tile_count = nullptr;   // queries per tile (of the last pass until k_tile_count)
  uint32_t* tile_off = nullptr;
  uint32_t* tile_order = nullptr;   // the eye kernel's tile order (k_tile_order)
  uint32_t* qkeys[2] = {nullptr, nullptr};
  uint32_t* qvals[2] = {nullptr, nullptr};
  OneSweep qos;         // the query sort of sessions
  uint32_t q_fine = 0;  // 1: fine query sort keys (QUERY_FINE_*)
  float4* res = nullptr;
  // pending refraction nodes of the eye trees: wave w of the eye kernel
  // keeps a pool of 64 * estack_depth entries, entry e being the three
  // float4 estack[(w * 64 * estack_depth + e) * 3 + c] (origin + node meta,
  // direction + weight r, weight g b + pixel of the tile)
  float4* estack = nullptr;
  uint32_t estack_lanes = 0;  // 64 * the most eye waves a launch may use
  uint32_t estack_depth = 0;  // pool entries per lane of a wave
};

// Eye waves per CU at most (the launch configurations a batch may tune to).
constexpr uint32_t EYE_WAVES_PER_CU_MAX = 32;
// Grid sizes of a session's persistent kernels, per CU: eye waves, shade
// blocks of 4 waves, photon-tracing waves (= photon regions).  Each kernel
// leaves slots free beside it for the other branch of its pass and the
// other pipeline's pass: filling the GPU with every kernel (16 eye waves,
// 8 shade blocks, 16 tracing waves) serialised them.  Caustic workload
// 0.855 -> 0.730 ms per pass with these, mesh 95.6 -> 87.6 ms.
constexpr uint32_t EYE_WAVES_PER_CU = 6;
constexpr uint32_t SHADE_BLOCKS_PER_CU = 5;
constexpr uint32_t TRACE_WAVES_PER_CU = 8;

// Next pass of the state (the first one: first_pass); a state that runs
// every stride-th pass advances by stride.
hipError_t launch_pass_begin(hipStream_t s, PassState* st, uint32_t first_pass, double r2_first, double alpha,
                             double power, uint32_t stride = 1);
hipError_t launch_pass_end(hipStream_t s, PassState* st);

// Number of tracing waves (= photon regions) used for n photons.
uint32_t trace_waves(uint64_t n_photon, int n_cu);

hipError_t launch_trace_photons(hipStream_t s, const SceneBuf& sc, PassState* st, PhotonBuf& pb, uint64_t seed,
                                int use_classic, int max_depth);
hipError_t launch_compact(hipStream_t s, const PhotonBuf& pb, const PassState* st, PhotonRec* dense);

hipError_t launch_map_build(hipStream_t s, MapBuf& m, PassState* st, const PhotonBuf& in, float4 bmin, float4 bmax);

// The pass's photon grid (cells of side >= 2.1 r over the scene bounds,
// hashed into m.cells): setup, bucket keys, sort, clear, cell ranges, the
// photon records in bucket order.  Buckets per pass: the next power of two
// >= the stored photons, at least 1024, at most m.cells_max <= 2^24.
constexpr uint32_t GRID_KEY_BITS = 24;  // bucket keys < 2^24: three 8-bit sort passes
constexpr uint32_t GRID_MAX_BUCKETS = 1u << GRID_KEY_BITS;

// Query sort keys (gather order): a Morton code of the query position in the
// scene box with the 2-bit level of points outside the box above it.  Coarse
// keys, 7 bits per axis (7.8 mm cells in a 1 m box), 23 bits: three 8-bit
// passes.  Eye buffers of at least QUERY_FINE_PIXELS pixels hold so many
// queries per coarse cell that the gather loses its coherence (mesh workload,
// 1920x1080: 43.8 against 41.0 ms per pass), and sort on fine keys, 10 bits
// per axis, 32 bits: four passes.  Keys start in buffer (bits / 8) & 1 so that
// the sorted order ends in qvals[0].  A query's result does not depend on the
// order, so the image does not either.
constexpr uint32_t QUERY_MORTON_MAX = 127;
constexpr uint32_t QUERY_LEVEL_SHIFT = 21;
constexpr uint32_t QUERY_KEY_BITS = 24;
constexpr uint32_t QUERY_KEYS_FIRST = (QUERY_KEY_BITS / 8u) & 1u;
constexpr uint32_t QUERY_FINE_MORTON_MAX = 1023;
constexpr uint32_t QUERY_FINE_LEVEL_SHIFT = 30;
constexpr uint32_t QUERY_FINE_KEY_BITS = 32;
constexpr uint32_t QUERY_FINE_PIXELS = 1u << 20;

// Blocks of 256 threads for a grid-stride kernel over about `items`
// elements, four per thread.  The grid kernels are sized from the photons a
// pass is expected to store (twice those emitted) rather than from the
// capacity (emitted times the depth limit): in a batch the capacity-sized
// grids summed to over a million mostly idle blocks per round.
inline uint32_t grid_stride_blocks(uint64_t items) {
  const uint64_t b = (items + 1023) / 1024;
  return (uint32_t)(b < 1 ? 1 : (b > 16384 ? 16384 : b));
}
// Blocks for the overflow queries (a grid-stride loop: there are rarely many).
constexpr uint32_t OVF_BLOCKS = 16;
inline uint64_t next_pow2_u64(uint64_t v) {
  uint64_t p = 1;
  while (p < v) p <<= 1;
  return p;
}

// expect: the photons the pass is expected to store, which sizes the grid
// kernels (0: the set's capacity).
hipError_t launch_grid_build(hipStream_t s, MapBuf& m, PassState* st, const PhotonBuf& in, float4 bmin, float4 bmax,
                             uint64_t expect = 0);

hipError_t launch_eye(hipStream_t s, const SceneBuf& sc, PassState* st, const EyeBuf& eb, uint64_t seed,
                      int use_classic, int max_depth, int n_cu);
// One pass of the classic ray tracer (trace_ray_classic) into eb.pix:
// k_eye_classic, the query sort, k_shade_classic (direct light + ambient),
// the overflow queries and the per-pixel sum.
hipError_t launch_classic(hipStream_t s, const SceneBuf& sc, PassState* st, const EyeBuf& eb, uint64_t seed,
                          float4 ambient, int n_cu);

// Query order of the gather: per-tile offsets, Morton keys of the eye
// queries, their sort (needs only the eye pass, not the photons).
hipError_t launch_query_prep(hipStream_t s, const SceneBuf& sc, const MapBuf& m, PassState* st, const EyeBuf& eb);

// Gather (over the pass's photon grid or a box-tree map) + direct light of
// one pass over the Morton-sorted queries, then the per-pixel sum (after
// launch_query_prep, run here unless `prepared`).  sum/sumsq != null: the
// pixel's pass radiance is added into the FP64 sums; otherwise it is left in
// pix.  ev (optional, 2 events) brackets the k_shade launch alone.
hipError_t launch_gather(hipStream_t s, const SceneBuf& sc, const MapBuf& m, PassState* st, const EyeBuf& eb,
                         int pfilter, int use_classic, int n_cu, double* sum, double* sumsq,
                         hipEvent_t* ev = nullptr, bool prepared = false);

// Test hooks (radius query / gather on host-provided query points).
// mode 0: count, 1: list, 2: gather rgb
hipError_t launch_query(hipStream_t s, const MapBuf& m, PassState* st, const float4* qf, const double* q64,
                        const float4* qn, uint32_t nq, double r2, int refine64, int mode, int pfilter,
                        uint32_t* counts, const uint64_t* offsets, uint32_t* idx_out, float4* rgb_out, int n_cu);

// The same hooks over a pass's photon grid (launch_grid_build for a radius
// >= sqrt(r2)): index sets (mode 0 count, 1 list) through the shade
// kernel's cell walk, flags bit 0 set for a query beyond the grid's span;
// and the gather through the shade kernel itself with the direct light off
// (sc: a scene with no BVH whose bmin is the grid's; eb.q/eb.res the
// queries and results, order their order, st with r2f, scale, n_qsort set
// and the work counters zeroed).
hipError_t launch_grid_query(hipStream_t s, const MapBuf& m, const PassState* st, float4 bmin, const float4* qf,
                             const double* q64, uint32_t nq, double r2, int refine64, int mode, uint32_t* counts,
                             const uint64_t* offsets, uint32_t* idx_out, uint32_t* flags);
hipError_t launch_grid_shade(hipStream_t s, const SceneBuf& sc, const MapBuf& m, PassState* st, const EyeBuf& eb,
                             const uint32_t* order, int pfilter, int n_cu);

// The gather without allocation: at most NOALLOC_MAX_QUERIES queries per
// call (= PPM_NOALLOC_MAX_QUERIES of ppm.h), in a fixed device area of at
// most NOALLOC_SCRATCH_MAX_BYTES per GPU, for devices 0..NOALLOC_MAX_DEVICES-1.
// noalloc_init(device), with that device current, takes the area's address
// (once per device); launch_gather_noalloc then runs ppm_gather's kernels on
// a tree map (grid = false: launch_query mode 2, results * k) or a grid map
// (launch_grid_shade over grid_scene, results as the shade scaled them),
// copying q_xyz/q_nrm in and rgb_out (nq x 3 doubles) out on stream s;
// *overflow_out receives the grid shade's flags once s has run.
constexpr uint32_t NOALLOC_MAX_QUERIES = 65536;
constexpr size_t NOALLOC_SCRATCH_MAX_BYTES = 9u << 20;
constexpr uint32_t NOALLOC_MAX_DEVICES = 64;
hipError_t noalloc_init(int device);
hipError_t launch_gather_noalloc(hipStream_t s, int device, const MapBuf& m, const PassState* map_st, bool grid,
                                 const SceneBuf& grid_scene, const double* q_xyz, const double* q_nrm, uint32_t nq,
                                 double r2, double k, float scale, int refine64, int pfilter, int n_cu,
                                 double* rgb_out, uint32_t* overflow_out);

// Scene BVH over the n shapes shapes[ids[i]]: Morton keys of the centroids
// (cmin, cscale = 1023 / extent), a radix sort (count *np on the device),
// leaves of `leaf` shapes copied into bprims, levels folded up.
hipError_t launch_bvh_build(hipStream_t s, const DevShape* shapes, const uint32_t* ids, uint32_t n, float4 cmin,
                            float4 cscale, uint32_t depth, uint32_t leaf, uint32_t* keys[2], uint32_t* vals[2],
                            const uint32_t* np,
                            uint32_t* ghist, uint32_t* totals, uint32_t sort_blocks, Node* pbox, DevShape* bprims,
                            Node* nodes);

// One scene of a batch: everything its pass kernels read (device memory).
struct BatchItem {
  SceneBuf sc;
  PassState* st;
  PhotonBuf pb;
  MapBuf m;
  EyeBuf eb;
  double* sum;
  double* sumsq;
  uint64_t seed;
  double r2_first, alpha, power;
  float4 mscale;             // Morton scale of the scene bounds (QUERY_MORTON_MAX + 1 cells per axis)
  uint32_t first_pass;       // this pipeline's first pass
  uint32_t pass_stride;      // passes between two of this pipeline's (the batch's pipelines)
  int32_t uc, maxd, pfilter;
  uint32_t trace_per_wave;   // photons per tracing wave
  uint32_t eye_waves;        // persistent eye waves
};

// Launch kinds of a batched pass; tab[k * (n_all + 1) + s] is the first block
// of scene s in a kind-k launch (tab[k * (n_all + 1) + n_all] the total).
enum BatchKind : int {
  BK_TRACE = 0,  // photon tracing waves / 4
  BK_MORTON,     // photon grid keys
  BK_PHIST,      // photon sort histogram (blocks of 1024)
  BK_POS,        // photon sort digit passes (persistent blocks)
  BK_LEAVES,     // grid cell ranges
  BK_CLEAR,      // grid bucket clear
  BK_EYE,        // eye waves / 4
  BK_PIX,        // primary slots / 256 (tile counts, accumulation)
  BK_QKEYS,      // primary slots / 256 + the overflow blocks
  BK_QHIST,      // query sort histogram (blocks of 1024)
  BK_QOS,        // query sort digit passes (persistent blocks)
  BK_SHADE,      // shade blocks
  BK_OVF,        // overflow query blocks
  BK_KINDS
};

float4 morton_scale(float4 bmin, float4 bmax, uint32_t qmax);

// Adjoint of the pass just run on these buffers (eye queries, their results,
// the map), for gimg = dL/d(the pass's image) per pixel: adds dL/dc1 of the
// Simple materials into grad (3 per material).  Not graph-capturable (it
// clears its scratch with hipMemsetAsync).
hipError_t launch_adjoint(hipStream_t s, const SceneBuf& sc, const MapBuf& m, PassState* st, const PhotonBuf& pb,
                          const EyeBuf& eb, int pfilter, int n_cu, uint64_t seed, int use_classic, int max_depth,
                          const float4* gimg, float* adj_sorted, float* adj_slot, double* grad);

// One pass of the first n scenes of a batch (scenes sorted by pass count,
// so the scenes still running are a prefix); blocks[k] = tab[k*(n_all+1)+n].
hipError_t launch_batch_pass(hipStream_t s, hipStream_t s2, hipEvent_t fork, hipEvent_t join,
                             const BatchItem* items, const uint32_t* tab, const uint32_t* blocks, uint32_t n,
                             uint32_t n_all, bool bvh, uint32_t dyn);  // dyn: the BVH kernels' dynamic LDS

// Test hook: nearest hit of n rays (unit directions).
hipError_t launch_intersect(hipStream_t s, const SceneBuf& sc, const float4* o, const float4* d, uint32_t n,
                            float* t_out, int32_t* obj_out);

}  // namespace ppm
