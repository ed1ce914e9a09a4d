// ppm_layout.hpp — memory layout shared by the host code and the gfx950
// kernels: scene records, photon records, tree nodes, gather queries, pass
// state, and the constants of the reference.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ppm {

// Reference constants.
constexpr float NEARLY0_F = 1.0e-4f;                 // ray/mod.rs:16
constexpr float INF_F = __builtin_huge_valf();
constexpr float PI_F = 3.14159265358979323846f;
constexpr float TWO_PI_F = 6.28318530717958647692f;
constexpr float FOUR_PI_F = 12.5663706143591729539f;  // PI4, ray/mod.rs:18
constexpr float INV_PI_F = 0.318309886183790671538f;  // ONE_PI, surface.rs:13
constexpr float SR_HALF_F = 0.159154943091895335769f;  // 1/(2 pi), tracer.rs:25
constexpr float LIGHT_SLACK_F = 0.002f;                // tracer.rs:286
constexpr float PARA_DIV_F = 0.2f;                     // light.rs:162
constexpr float K_CONE_F = 1.1f;                       // tracer.rs:198
constexpr float FAC_K_F = (float)(1.0 - 2.0 / (3.0 * 1.1));  // tracer.rs:199
constexpr float GAUSS_ALPHA_F = 0.918f;                // tracer.rs:208
constexpr float GAUSS_BETA_F = 1.953f;                 // tracer.rs:209
constexpr float GAUSS_E_BETA_F = (float)(1.0 - 0.14184788965323);  // tracer.rs:210
constexpr float GAUSS_CORR_F = 0.5f;                   // tracer.rs:211

constexpr uint32_t PHILOX_M0 = 0xD2511F53u;
constexpr uint32_t PHILOX_M1 = 0xCD9E8D57u;
constexpr uint32_t PHILOX_W0 = 0x9E3779B9u;
constexpr uint32_t PHILOX_W1 = 0xBB67AE85u;

// Stream tags (counter word 2).
constexpr uint32_t TAG_PHOTON = 0x10000000u;
constexpr uint32_t TAG_EYE = 0x20000000u;  // | depth << 16 | refraction path bits

// Enumerations (mirror include/ppm.h).
enum : uint32_t { SHAPE_POINT = 0, SHAPE_PLAIN = 1, SHAPE_SPHERE = 2, SHAPE_POLYGON = 3, SHAPE_PARALLELOGRAM = 4 };
enum : uint32_t { SURF_NOTHING = 0, SURF_SIMPLE = 1, SURF_TS = 2, SURF_DISNEY = 3, SURF_BRADY = 4 };
enum : uint32_t { LIGHT_POINT = 0, LIGHT_PARALLELOGRAM = 1, LIGHT_SUN = 2 };
enum : int { FILTER_NONE = 0, FILTER_CONE = 1, FILTER_GAUSS = 2 };

constexpr int MAX_DEPTH_LIMIT = 16;  // deepest MAX_TRACE the kernels accept
constexpr int CLASSIC_DEPTH = 10;    // trace_ray_classic's own depth limit (tracer.rs:222)
#ifndef PPM_LEAF
#define PPM_LEAF 32
#endif
constexpr int LEAF = PPM_LEAF;       // photons per map leaf (<= one wavefront)

// Scene BVH: a heap-ordered complete binary tree over leaves of a few
// primitives in Morton order of their centroids: BVH_LEAF_SMALL for a tree
// of at most BVH_SMALL_PRIMS primitives (which then fits the kernels' LDS
// copy whole), one otherwise (the mesh of 327,680 triangles: 105 ms per pass
// with one, 111 with two, 118 with four, 139 with eight primitives per
// leaf).  Traversal defers at most one node per level, kept as one bit per
// level of a 32-bit trail, so BVH_MAX_DEPTH < 32.
constexpr uint32_t BVH_LEAF_SMALL = 4;
constexpr uint32_t BVH_SMALL_PRIMS = 512;
constexpr uint32_t BVH_MAX_DEPTH = 24;   // up to 16M leaves
constexpr uint32_t BVH_MIN_PRIMS = 64;   // automatic mode: fewer bounded shapes are tested one by one
constexpr float BVH_EMPTY = 3.0e38f;     // an empty node is the point box (BVH_EMPTY,)^3

// --- scene records (80 B) -------------------------------------------------
struct alignas(16) DevShape {
  uint32_t kind;
  uint32_t mat;
  float dist;    // Plain
  float radius;  // Sphere
  float4 p;      // position / centre / p0; w = the shape's index in the scene (bits)
  float4 n;      // unit normal
  float4 d1;
  float4 d2;
};

struct alignas(16) DevMaterial {
  uint32_t surface;
  float diff;         // diffuseness / scatterness
  float metal;
  float rough;
  float density_pow;  // 1/(10^(5(1-sqrt(rough)))+1), surface.rs:52,63
  float alpha;        // rough^4, surface.rs:64
  float ior_avg;      // (ior.r+ior.g+ior.b)/3, physics.rs:193
  float pad;
  float4 em;  // emittance
  float4 ior;
  float4 c1;  // reflectance / albedo_diff
  float4 c2;  // specular_refl / albedo_spec
  float4 q;   // survival probability of a Simple diffuse bounce (= c1 unless set)
};

struct alignas(16) DevLight {
  uint32_t kind;
  float flux;
  uint32_t first_photon;  // photons [first, first + count) of a pass come from here
  uint32_t n_photon;
  float4 color;
  float4 pos;    // w: the light's own shape among the shapes tested one by one (uint bits; ~0: none)
  float4 nvec;
  float4 d1;
  float4 d2;
  float4 dir;
  // a ParallelogramLight's 25 sample points (light.rs:164-170): sample s at
  // (u, v) = (0.1 + 0.2 (s / 5), 0.1 + 0.2 (s % 5)) is pos + u d1 + v d2,
  // formed as fma(v, d2, fma(u, d1, pos)) in float
  float4 tss[25];
};

struct alignas(16) DevCamera {
  int32_t xreso, yreso;
  int32_t progressive, antialias;
  int32_t blur, pfilter;
  // the eye's tiles: (1 << tile_shift) x (64 >> tile_shift) pixels, one per
  // wave, tiles_x of them per row (tile_shift 2..5: 4x16, 8x8, 16x4, 32x2)
  int32_t tiles_x, tile_shift;
  float4 eye, origin, esx, esy, eex, eey;
};

// --- photon record (32 B in HBM; 28 B of payload) ---------------------------
// pm = (x, y, z, meta), dm = (dx, dy, dz, w); meta bits 0-1 channel,
// 2-6 depth, 7-31 photon index; w = the photon's weight, the product of
// c1/q over its diffuse survivals (exactly 1 when q = c1).
struct alignas(16) PhotonRec {
  float4 pm;
  float4 dm;
};
// The photon index has 25 bits: photons per pass (and per uploaded set) are
// limited to PHOTON_INDEX_MAX, checked where sets are made.
constexpr uint64_t PHOTON_INDEX_MAX = 1ull << 25;
__host__ __device__ inline uint32_t pack_meta(uint32_t wl, uint32_t depth, uint32_t pid) {
  return (wl & 3u) | ((depth & 31u) << 2) | (pid << 7);
}

// The per-pass grid keeps its sorted copy in a gather form of the same 32 B:
// pm = (x, y, z, [channel 0]), dm = (w dx, w dy, w dz, [channel 1]), the
// flags 1.0f or 0.0f.  A gather then adds v * flag per channel (channel 2:
// 1 - both flags), exact for flags of 0 and 1, in place of unpacking the
// meta bits and selecting, and takes the weight with the direction.  The
// index and depth stay in the map's src_idx and the trace's own records.
// Gathers read four records per step from any start inside a cell's range:
// up to GRID_READ_SLACK records past the last photon, which the grid build
// keeps allocated and zeroed.
constexpr uint32_t GRID_READ_SLACK = 3;
__device__ inline PhotonRec grid_rec(const PhotonRec& r) {
  const uint32_t wl = __float_as_uint(r.pm.w) & 3u;
  const float w = r.dm.w;
  return PhotonRec{make_float4(r.pm.x, r.pm.y, r.pm.z, wl == 0u ? 1.0f : 0.0f),
                   make_float4(r.dm.x * w, r.dm.y * w, r.dm.z * w, wl == 1u ? 1.0f : 0.0f)};
}

// --- tree node (32 B) ------------------------------------------------------
struct alignas(16) Node {
  float4 lo;  // xyz min
  float4 hi;  // xyz max
};

// --- gather query (48 B) -----------------------------------------------------
// a = (x, y, z, pixel), b = (nx, ny, nz, material), w = (wr, wg, wb, 0)
struct alignas(16) Query {
  float4 a;
  float4 b;
  float4 w;
};
constexpr uint32_t NO_PIXEL = 0xFFFFFFFFu;
// EyeBuf::sec_count[t]: the pixel's deeper in-place queries, with this bit
// set when the pixel has its primary query in slot t.
constexpr uint32_t SEC_PRIMARY = 1u << 31;

// --- per-pass device state -----------------------------------------------------
struct alignas(128) PassState {
  double r2;          // r_i^2 of the current pass
  double alpha;
  double r2_first;    // r_1^2
  double power;       // per-photon power
  float r2f;          // (float) r2
  float scale;        // (float)(power / (pi r2))
  uint32_t pass;      // 1-based index of the current pass
  uint32_t n_stored;  // photons stored this pass
  uint32_t next_photon;
  uint32_t n_queries;  // secondary gather queries this pass
  uint32_t overflow;   // set when a capacity was exceeded
  uint32_t nonfinite;  // set when a stored photon coordinate is not finite
  uint32_t n_leaves;   // leaves of the current map
  uint32_t depth;      // tree depth (leaves = 1 << depth padded)
  uint32_t n_qsort;    // gather queries sorted this pass
  uint32_t pad0;
  uint64_t total_stored;
  uint64_t total_queries;
  // photon grid of the pass (k_grid_setup): cell side 1 / g_inv_c, query
  // reach before its per-point rounding margin, bucket mask, cell index
  // clamp [-1, g_lim]
  float g_inv_c;
  float g_reach;
  uint32_t g_mask;
  int32_t g_lim;
  // the two ends of the eye's tile order (costly tiles first; zeroed by
  // k_pass_begin)
  uint32_t tiles_front;
  uint32_t tiles_back;
  // work counters of the persistent kernels (zeroed by k_pass_begin):
  // work[k][c][0] hands out items c, c + 8, c + 16, ... of kernel k (0 the
  // query groups of k_shade, 1 the tiles of k_eye); each counter has its own
  // 128-byte line
  alignas(128) uint32_t work[2][8][32];
};
constexpr uint32_t WORK_SHADE = 0, WORK_EYE = 1;

}  // namespace ppm
