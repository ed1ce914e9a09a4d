// ppm_dev.hpp — device-side types and arithmetic of the PPM-PA hot path
// (float traversal/intersection/shading, gfx950).
//
// Every function restates the reference function it cites; the reference is
// FP64 Rust, this is FP32 HIP.  Where FP32 could produce a result that FP64
// produces only on a measure-zero set (NaN from 1/0 etc.), the FP64 outcome
// is taken and the guard is commented.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ppm_layout.hpp"

namespace ppm {

// ---------------------------------------------------------------------------
// 3-vectors (algebra.rs:50-181)

struct V3 {
  float x, y, z;
};

__device__ __forceinline__ V3 mk(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 mk(float4 a) { return V3{a.x, a.y, a.z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 operator*(float s, V3 a) { return V3{a.x * s, a.y * s, a.z * s}; }
// dot and cross with their fused multiply-adds written out, so that they
// round the same way wherever they are inlined (the intersection code that
// the BVH and the shape list share must give bit-identical distances).
__device__ __forceinline__ float dot(V3 a, V3 b) { return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)); }
// algebra.rs:174-180 (component order kept)
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
  return V3{fmaf(a.y, b.z, -(b.y * a.z)), fmaf(a.z, b.x, -(b.z * a.x)), fmaf(a.x, b.y, -(b.x * a.y))};
}
// algebra.rs:151-158: None for a zero vector.
__device__ __forceinline__ bool normalize(V3 a, V3& out) {
  // v_rsq_f32 (about 1 ulp) instead of sqrt and an IEEE division, each a
  // sequence of about a dozen instructions; normalisation runs for every
  // ray and every shadow sample
  const float d2 = dot(a, a);
  if (d2 == 0.0f) return false;
  const float inv = __builtin_amdgcn_rsqf(d2);
  out = a * inv;
  return true;
}
// c ? a : b per component: a ternary over whole V3 values can make the
// compiler pick between two temporaries by address, which puts them in
// scratch memory.
__device__ __forceinline__ V3 sel(bool c, V3 a, V3 b) { return mk(c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z); }
__device__ __forceinline__ V3 normalize_or(V3 a, V3 fallback) {
  V3 r = fallback;
  const bool ok = normalize(a, r);
  return sel(ok, r, fallback);
}

// RGB triple (physics.rs Color / optics.rs Radiance share this shape).
struct C3 {
  float r, g, b;
};
__device__ __forceinline__ C3 c3(float r, float g, float b) { return C3{r, g, b}; }
__device__ __forceinline__ C3 c3(float4 a) { return C3{a.x, a.y, a.z}; }
__device__ __forceinline__ C3 operator+(C3 a, C3 b) { return C3{a.r + b.r, a.g + b.g, a.b + b.b}; }
__device__ __forceinline__ C3 operator*(C3 a, C3 b) { return C3{a.r * b.r, a.g * b.g, a.b * b.b}; }
__device__ __forceinline__ C3 operator*(C3 a, float s) { return C3{a.r * s, a.g * s, a.b * s}; }
__device__ __forceinline__ bool nonzero(C3 a) { return a.r != 0.0f || a.g != 0.0f || a.b != 0.0f; }
__device__ __forceinline__ float chan(float4 c, uint32_t wl) {
  return wl == 0 ? c.x : (wl == 1 ? c.y : c.z);
}

// ---------------------------------------------------------------------------
// Counter-based RNG: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel
// random numbers: as easy as 1, 2, 3", SC'11).  The reference draws from the
// unseeded rand::thread_rng (physics.rs:324 and others), which nothing can
// reproduce; every stream here is keyed by (seed) and indexed by
// (element, pass, stream tag, block) so that the FP64 oracle draws the same
// numbers.

__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int i = 0; i < 10; ++i) {
    uint32_t lo0 = PHILOX_M0 * c.x;
    uint32_t hi0 = __umulhi(PHILOX_M0, c.x);
    uint32_t lo1 = PHILOX_M1 * c.z;
    uint32_t hi1 = __umulhi(PHILOX_M1, c.z);
    c = make_uint4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
    k0 += PHILOX_W0;
    k1 += PHILOX_W1;
  }
  return c;
}

// Uniform in [0, 1) with 24 random bits: exact in float and in double.
__device__ __forceinline__ float u01(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

// A sequential stream of uniforms: draw n comes from block n/4, lane n%4.
struct Rng {
  uint32_t k0, k1;     // key = seed
  uint32_t c0, c1, c2; // element, pass, tag
  uint32_t n;          // draws consumed
  uint4 blk;

  __device__ __forceinline__ void init(uint64_t seed, uint32_t elem, uint32_t pass, uint32_t tag) {
    k0 = (uint32_t)seed;
    k1 = (uint32_t)(seed >> 32);
    c0 = elem;
    c1 = pass;
    c2 = tag;
    n = 0;
  }
  __device__ __forceinline__ float next() {
    uint32_t j = n & 3u;
    if (j == 0) blk = philox4x32_10(make_uint4(c0, c1, c2, n >> 2), k0, k1);
    ++n;
    uint32_t v = j == 0 ? blk.x : (j == 1 ? blk.y : (j == 2 ? blk.z : blk.w));
    return u01(v);
  }
};

// ---------------------------------------------------------------------------
// Scene view passed by value to every kernel.

struct SceneView {
  const DevShape* shapes;   // in scene order
  const DevMaterial* mats;  // n_mats + 1 entries: the last one is M_AIR
  const DevLight* lights;
  const DevShape* lin;      // shapes tested one by one (all of them, or the unbounded ones beside a BVH)
  const Node* lin_box;      // padded boxes of lin; lo.w = 1: a plane (no box)
  const Node* bvh;          // BVH nodes (null: no BVH)
  const DevShape* bprims;   // BVH primitives in leaf order
  uint32_t n_lin;
  uint32_t n_bprims;
  uint32_t bvh_depth;       // leaves at level bvh_depth
  uint32_t bvh_leaf;        // primitives per leaf
  uint32_t n_shapes;
  uint32_t n_mats;          // index of M_AIR
  uint32_t n_lights;
  uint32_t n_emit;           // photons per pass over all lights
  uint32_t lds;              // in LDS (scene_to_lds): bit 0 lin and lin_box, 1 mats, 2 bvh, 3 compact bprims
  uint32_t dyn_lds;          // the launch gave BVH_DYN_LDS_BYTES of dynamic LDS for the BVH copies
  DevCamera cam;
  float4 bmin, bmax;        // Morton quantisation box
};

// ---------------------------------------------------------------------------
// LDS copies of the shapes tested one by one (with their boxes) and of the
// materials, for the kernels that walk rays: every ray tests each of those
// shapes, and from global memory each test waits for its shape's scalar
// loads.  scene_to_lds(), called by all threads of a block before any ray
// work, fills them and sets the view's lds bits; a view without those bits
// (any other kernel, or a scene too large for the copies) reads the global
// arrays.

constexpr uint32_t LIN_LDS = 32;  // shapes
constexpr uint32_t MAT_LDS = 48;  // materials (M_AIR included)
__shared__ DevShape g_lin_lds[LIN_LDS];
__shared__ Node g_linbox_lds[LIN_LDS];
__shared__ DevMaterial g_mat_lds[MAT_LDS];

__device__ __forceinline__ void copy16(float4* dst, const float4* src, uint32_t n) {
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

// A small BVH (kernels that walk it, BVH = true): every node of a tree of at
// most BVH_LDS_NODES nodes, and its primitives, up to BVH_LDS_PRIMS, as
// compact spheres (centre, radius^2; w = -1 for a primitive of another kind,
// which is then read from global memory) with their scene indices.
// They live in dynamic LDS, which a launch requests only for a scene whose
// tree is that small (bvh_dyn_lds on the host): a kernel walking a large
// tree (the mesh) then holds no 18 KB of unused LDS per block, which had
// cost it occupancy (mesh 105.8 -> 98.6 ms per pass without it).
constexpr uint32_t BVH_LDS_NODES = 255;  // leaf level <= 7: 128 leaves
constexpr uint32_t BVH_LDS_PRIMS = 512;
static_assert(BVH_LEAF_SMALL * ((BVH_LDS_NODES + 1u) / 2u) <= BVH_LDS_PRIMS, "every leaf slot of a small tree in LDS");
extern __shared__ float4 g_dyn_lds[];
__device__ __forceinline__ Node* bvh_lds_nodes() { return reinterpret_cast<Node*>(g_dyn_lds); }
__device__ __forceinline__ float4* bvh_lds_sph() { return g_dyn_lds + 2u * BVH_LDS_NODES; }
__device__ __forceinline__ int32_t* bvh_lds_idx() {
  return reinterpret_cast<int32_t*>(g_dyn_lds + 2u * BVH_LDS_NODES + BVH_LDS_PRIMS);
}

template <int BV = 0, bool BVH = (BV != 0)>
__device__ __forceinline__ void scene_to_lds(SceneView& sc) {
  const bool lin = sc.n_lin <= LIN_LDS && sc.lin_box != nullptr;
  const bool mat = sc.n_mats + 1u <= MAT_LDS;
  if (lin) {
    copy16((float4*)g_lin_lds, (const float4*)sc.lin, sc.n_lin * (uint32_t)(sizeof(DevShape) / 16));
    copy16((float4*)g_linbox_lds, (const float4*)sc.lin_box, sc.n_lin * (uint32_t)(sizeof(Node) / 16));
  }
  if (mat) copy16((float4*)g_mat_lds, (const float4*)sc.mats, (sc.n_mats + 1u) * (uint32_t)(sizeof(DevMaterial) / 16));
  bool nodes = false, prims = false;
  if (BV == 2 && sc.dyn_lds && sc.n_bprims > 0) {
    const uint32_t n_nodes = (2u << sc.bvh_depth) - 1u;
    nodes = n_nodes <= BVH_LDS_NODES;
    prims = sc.n_bprims <= BVH_LDS_PRIMS;
    if (nodes) copy16((float4*)bvh_lds_nodes(), (const float4*)sc.bvh, n_nodes * (uint32_t)(sizeof(Node) / 16));
    // every slot of the tree's leaves: the slots past the last primitive
    // hold w = -2 (no primitive), so a leaf's tests need no bound
    const uint32_t slots = nodes ? (BVH_LEAF_SMALL << sc.bvh_depth) : sc.n_bprims;
    if (prims)
      for (uint32_t j = threadIdx.x; j < slots; j += blockDim.x) {
        float4 c = make_float4(0.0f, 0.0f, 0.0f, -2.0f);
        int32_t idx = 0;
        if (j < sc.n_bprims) {
          const DevShape s = sc.bprims[j];
          c = s.kind == SHAPE_SPHERE ? make_float4(s.p.x, s.p.y, s.p.z, s.radius * s.radius)
                                     : make_float4(0.0f, 0.0f, 0.0f, -1.0f);
          idx = (int32_t)__float_as_uint(s.p.w);
        }
        bvh_lds_sph()[j] = c;
        bvh_lds_idx()[j] = idx;
      }
  }
  __syncthreads();
  sc.lds = (lin ? 1u : 0u) | (mat ? 2u : 0u) | (nodes ? 4u : 0u) | (prims ? 8u : 0u);
}

// ---------------------------------------------------------------------------
// Intersection (tracer.rs:306-359, geometry.rs:132-202).

struct Hit {
  float t;
  int32_t obj;
};

// Nearest-hit update (when ok): the earlier of two equal distances is the
// shape that comes first in the scene, as the stable sort of tracer.rs:335
// keeps it.  Written as selects: the tests of the walks' inner loops then
// issue no branches.
__device__ __forceinline__ void take(Hit& h, float t, int32_t idx, bool ok = true) {
  const bool better = ok & (t >= NEARLY0_F) & ((t < h.t) | ((t == h.t) & (idx < h.obj)));
  h.t = better ? t : h.t;
  h.obj = better ? idx : h.obj;
}

// distance_sphere (geometry.rs:179-193) of the sphere (c, radius^2 = rr).
// The reference offers both roots t0 -+ t2 (one when t2 = 0); of the two,
// the farther can only be taken when the nearer lies short of NEARLY0, so
// one take of that root gives what taking both in turn gives.
// Taken only when ok.
__device__ __forceinline__ void sphere_hit(V3 c, float rr, int32_t idx, V3 o, V3 d, Hit& h, bool ok = true) {
#pragma clang fp contract(off)
  V3 oc = c - o;
  float t0 = dot(oc, d);
  float t1 = rr - (dot(oc, oc) - t0 * t0);
  const float t2 = __builtin_amdgcn_sqrtf(fmaxf(t1, 0.0f));  // v_sqrt_f32, ~1 ulp
  const float tn = t0 - t2;
  take(h, tn >= NEARLY0_F ? tn : t0 + t2, idx, ok & (t1 > 0.0f));
}

// distance_plain (geometry.rs:170-177), distance_sphere (:179-193) and
// method_moller (:149-165) folded into a nearest-hit update.  Candidates with
// t < NEARLY0 are dropped (tracer.rs:324).
__device__ __forceinline__ void shape_hit(const DevShape& s, int32_t idx, V3 o, V3 d, Hit& h) {
#pragma clang fp contract(off)
  switch (s.kind) {
    case SHAPE_PLAIN: {
      V3 n = mk(s.n);
      float c = dot(n, d);
      if (c != 0.0f) take(h, (s.dist + dot(n, o)) * __builtin_amdgcn_rcpf(-c), idx);  // v_rcp_f32, ~1 ulp
      break;
    }
    case SHAPE_SPHERE:
      sphere_hit(mk(s.p), s.radius * s.radius, idx, o, d, h);
      break;
    case SHAPE_POLYGON:
    case SHAPE_PARALLELOGRAM: {
      // computed whatever det is, and taken only for det != 0
      V3 d1 = mk(s.d1), d2 = mk(s.d2);
      V3 re2 = cross(d, d2);
      float det = dot(re2, d1);
      V3 pp = o - mk(s.p);
      V3 te1 = cross(pp, d1);
      float inv = __builtin_amdgcn_rcpf(det);
      float u = dot(re2, pp) * inv;
      float v = dot(te1, d) * inv;
      float t = dot(te1, d2) * inv;
      float lim = s.kind == SHAPE_POLYGON ? 1.0f : 2.0f;
      const bool out = (u < 0.0f) | (u > 1.0f) | (v < 0.0f) | (v > 1.0f) | (u + v > lim);
      take(h, t, idx, (det != 0.0f) & !out);
      break;
    }
    default:
      break;  // Point: never hit (geometry.rs:134)
  }
}

// Read-only device data through the constant address space: loads at a
// wave-uniform index become scalar (s_load) loads shared by the 64 lanes.
// (The host pass of hipcc also parses device bodies; it has no such space.)
template <typename T>
__device__ __forceinline__ T ldc(const T* p, uint32_t i) {
#if defined(__HIP_DEVICE_COMPILE__)
  return ((const __attribute__((address_space(4))) T*)p)[i];
#else
  return p[i];
#endif
}

// Shape i of lin, its box, material m: from LDS when scene_to_lds put them
// there, else from global memory.
__device__ __forceinline__ DevShape lin_shape(const SceneView& sc, uint32_t i) {
  if (sc.lds & 1u) return g_lin_lds[i];
  return ldc(sc.lin, i);
}
__device__ __forceinline__ Node lin_box(const SceneView& sc, uint32_t i) {
  if (sc.lds & 1u) return g_linbox_lds[i];
  return ldc(sc.lin_box, i);
}
__device__ __forceinline__ DevMaterial material(const SceneView& sc, uint32_t m) {
  if (sc.lds & 2u) return g_mat_lds[m];
  return sc.mats[m];
}

// Slab test of a node box: the ray o + t d, given as inv = 1/d and noi =
// -o/d, enters the box at tn within [0, tmax].  Each plane distance is one
// FMA, b inv + noi, whose rounding amounts to about an ulp of the
// coordinates, far inside the boxes' padding.  An axis whose two plane
// distances are NaN (a ray in the plane of a flat box) does not constrain:
// fminf/fmaxf drop NaN operands.
// 1/d for the slab tests, each component of d first kept at least 1e-20
// from zero: with an infinite 1/d, b/d - o/d would be inf - inf = NaN and
// leave that axis unconstrained for the whole walk (such rays walked large
// parts of a 327k-triangle tree); a finite 1e20 puts the axis's planes
// at +-huge distances, as the exact difference did.  Primitive tests use d.
__device__ __forceinline__ V3 slab_inv(V3 d) {
  const float m = 1e-20f;
  return mk(1.0f / (fabsf(d.x) < m ? copysignf(m, d.x) : d.x), 1.0f / (fabsf(d.y) < m ? copysignf(m, d.y) : d.y),
            1.0f / (fabsf(d.z) < m ? copysignf(m, d.z) : d.z));
}

__device__ __forceinline__ bool slab(const Node& b, V3 inv, V3 noi, float tmax, float& tn) {
  const float x0 = fmaf(b.lo.x, inv.x, noi.x), x1 = fmaf(b.hi.x, inv.x, noi.x);
  const float y0 = fmaf(b.lo.y, inv.y, noi.y), y1 = fmaf(b.hi.y, inv.y, noi.y);
  const float z0 = fmaf(b.lo.z, inv.z, noi.z), z1 = fmaf(b.hi.z, inv.z, noi.z);
  const float t0 = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), 0.0f));
  const float t1 = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tmax));
  tn = t0;
  return t0 <= t1;
}

// BVH node i: from the LDS copy (NL: scene_to_lds put the nodes there), else
// from global memory.  The walks are instantiated for each source behind one
// uniform test, so a node load is an LDS or a global load, not a flat load
// through a selected pointer with its 64-bit address arithmetic.
template <bool NL>
__device__ __forceinline__ Node bvh_node(const SceneView& sc, uint32_t i) {
  if (NL) return bvh_lds_nodes()[i];
  return sc.bvh[i];
}
// The nearest-hit test of primitive j: from its LDS compact sphere when
// scene_to_lds put the primitives there, else from global memory.
template <bool CACHED>
__device__ __forceinline__ void bprim_hit(const SceneView& sc, uint32_t j, V3 o, V3 d, Hit& h) {
  if (CACHED && (sc.lds & 8u)) {
    const float4 c = bvh_lds_sph()[j];
    if (c.w >= 0.0f) {
      sphere_hit(mk(c.x, c.y, c.z), c.w, bvh_lds_idx()[j], o, d, h);
      return;
    }
  }
  const DevShape s = sc.bprims[j];
  shape_hit(s, (int32_t)__float_as_uint(s.p.w), o, d, h);
}

// The nearest-hit tests of small-tree leaf `leaf`, all from the LDS copies
// (its BVH_LEAF_SMALL slots, padded past the last primitive): the slots'
// loads are issued together and the sphere tests run without branches.
__device__ __forceinline__ void leaf_hit_lds(const SceneView& sc, uint32_t leaf, V3 o, V3 d, Hit& h) {
  const uint32_t base = leaf * BVH_LEAF_SMALL;
  float4 c[BVH_LEAF_SMALL];
  int32_t id[BVH_LEAF_SMALL];
  bool other = false;
#pragma unroll
  for (uint32_t k = 0; k < BVH_LEAF_SMALL; ++k) {
    c[k] = bvh_lds_sph()[base + k];
    id[k] = bvh_lds_idx()[base + k];
  }
#pragma unroll
  for (uint32_t k = 0; k < BVH_LEAF_SMALL; ++k) {
    sphere_hit(mk(c[k].x, c[k].y, c[k].z), c[k].w, id[k], o, d, h, c[k].w >= 0.0f);
    other |= c[k].w == -1.0f;
  }
  if (other)  // primitives of another kind, from global memory
    for (uint32_t k = 0; k < BVH_LEAF_SMALL; ++k)
      if (c[k].w == -1.0f) {
        const DevShape s = sc.bprims[base + k];
        shape_hit(s, (int32_t)__float_as_uint(s.p.w), o, d, h);
      }
}

// The BVH is a complete binary tree in heap order (children of n: 2n+1,
// 2n+2), and a depth-first walk only ever defers the sibling of the child it
// enters.  So instead of a stack, a bit trail (Laine, "Restart trail for
// stackless BVH traversal", HPG 2010) keeps bit k set while the sibling of
// the walk's level-k node is still to be visited; trail_pop takes the
// deepest one, the node a stack would pop, from the current node's
// ancestor at that level.  No LDS or scratch per lane.
__device__ __forceinline__ uint32_t heap_level(uint32_t n) { return 31u - __clz(n + 1u); }
__device__ __forceinline__ uint32_t trail_pop(uint32_t& trail, uint32_t node) {
  const uint32_t k = 31u - __clz(trail);
  trail &= ~(1u << k);
  const uint32_t a = ((node + 1u) >> (heap_level(node) - k)) - 1u;  // the level-k ancestor
  return ((a - 1u) ^ 1u) + 1u;                                       // its sibling
}

// Nearest hit over the BVH primitives: ordered depth-first traversal (the
// nearer child first, the farther one deferred on the bit trail).  A node is
// entered while its box starts no later than the best hit, so equal
// distances still reach the lower shape index.
template <bool CACHED, bool NL>
__device__ __forceinline__ void bvh_nearest_walk(const SceneView& sc, V3 o, V3 d, Hit& h) {
  const V3 inv = slab_inv(d);
  const V3 noi = mk(-o.x * inv.x, -o.y * inv.y, -o.z * inv.z);
  const uint32_t first_leaf = (1u << sc.bvh_depth) - 1u;
  uint32_t node = 0, trail = 0;
  float tn;
  if (!slab(bvh_node<NL>(sc, 0), inv, noi, h.t, tn)) return;
  while (true) {
    if (node >= first_leaf) {
      if (NL) {
        leaf_hit_lds(sc, node - first_leaf, o, d, h);
      } else {
        const uint32_t base = (node - first_leaf) * sc.bvh_leaf;
        const uint32_t end = min(base + sc.bvh_leaf, sc.n_bprims);
        for (uint32_t j = base; j < end; ++j) bprim_hit<CACHED>(sc, j, o, d, h);
      }
    } else {
      const uint32_t l = 2u * node + 1u;
      float tl, tr;
      const bool hl = slab(bvh_node<NL>(sc, l), inv, noi, h.t, tl);
      const bool hr = slab(bvh_node<NL>(sc, l + 1u), inv, noi, h.t, tr);
      if (hl && hr) {
        const bool left_first = tl <= tr;
        node = left_first ? l : l + 1u;
        trail |= 1u << heap_level(node);
        continue;
      }
      if (hl || hr) {
        node = hl ? l : l + 1u;
        continue;
      }
    }
    bool more = false;
    while (trail != 0u) {
      node = trail_pop(trail, node);
      if (slab(bvh_node<NL>(sc, node), inv, noi, h.t, tn)) {
        more = true;
        break;
      }
    }
    if (!more) break;
  }
}
template <bool CACHED>
__device__ __forceinline__ void bvh_nearest(const SceneView& sc, V3 o, V3 d, Hit& h) {
  if (sc.n_bprims == 0) return;  // a scene without a BVH in a batch that has one
  if (CACHED && (sc.lds & 12u) == 12u)  // nodes and primitives in LDS
    bvh_nearest_walk<CACHED, true>(sc, o, d, h);
  else
    bvh_nearest_walk<CACHED, false>(sc, o, d, h);
}

// calc_intersection (tracer.rs:306-335): the nearest hit over every shape.
// BVH = false: the scene has no BVH and `lin` holds all shapes.
template <int BV, bool BVH = (BV != 0)>
__device__ __forceinline__ Hit nearest(const SceneView& sc, V3 o, V3 d) {
  Hit h{INF_F, -1};
  for (uint32_t i = 0; i < sc.n_lin; ++i) {
    const DevShape s = lin_shape(sc, i);
    shape_hit(s, (int32_t)__float_as_uint(s.p.w), o, d, h);
  }
  if (BVH) bvh_nearest<BV == 2>(sc, o, d, h);
  return h;
}

struct Isect {
  V3 pos;
  V3 n;       // faces the incoming ray
  int32_t obj;
  uint32_t mat;
  bool inside_to_out;  // io == Out (the normal was flipped)
};

// calc_intersection tail (tracer.rs:336-348): position, normal, flip.
__device__ __forceinline__ bool finish_hit(const SceneView& sc, V3 o, V3 d, Hit h, Isect& is) {
#pragma clang fp contract(off)
  if (h.obj < 0) return false;
  const DevShape& s = sc.shapes[h.obj];
  V3 p = mk(o.x + d.x * h.t, o.y + d.y * h.t, o.z + d.z * h.t);  // Ray::target (geometry.rs:56-58)
  V3 n;
  if (s.kind == SHAPE_SPHERE) {
    if (!normalize(p - mk(s.p), n)) return false;  // get_normal None
  } else {
    n = mk(s.n);
  }
  bool flip = dot(n, d) > 0.0f;
  is.pos = p;
  is.n = sel(flip, -n, n);
  is.obj = h.obj;
  is.mat = s.mat;
  is.inside_to_out = flip;
  return true;
}

template <int BV, bool BVH = (BV != 0)>
__device__ __forceinline__ bool intersect(const SceneView& sc, V3 o, V3 d, Isect& is) {
  return finish_hit(sc, o, d, nearest<BV>(sc, o, d), is);
}

// ---------------------------------------------------------------------------
// Physics (physics.rs).

// relative_ior (physics.rs:200-205): ior2/ior1, 1 when ior1 == 0.
__device__ __forceinline__ float relative_ior(float i1, float i2) { return i1 == 0.0f ? 1.0f : i2 / i1; }

// specular_reflection (physics.rs:214-221).
__device__ __forceinline__ V3 specular_reflection(V3 n, V3 v, float& cos1) {
  float c = -dot(v, n);
  if (c < 0.0f) {
    cos1 = -c;
    return n;
  }
  cos1 = c;
  return normalize_or(v + (2.0f * c) * n, n);
}

// specular_refraction (physics.rs:270-285).  eta == 0 (ior 0 surfaces) is
// None here: in double g0 = cos1^2 - 1 < 0 except on a measure-zero set, but
// in float cos1 rounds to exactly 1 near normal incidence and 1/0 would give
// NaN.
__device__ __forceinline__ bool specular_refraction(V3 n, V3 v, float eta, V3& t, float& cos2) {
  cos2 = 0.0f;
  float c1 = -dot(v, n);
  if (c1 < 0.0f || eta == 0.0f) return false;
  float g0 = eta * eta + c1 * c1 - 1.0f;
  if (g0 < 0.0f) return false;
  float g = sqrtf(g0);
  V3 tv = (1.0f / eta) * (v + (c1 - g) * n);
  if (!normalize(tv, t)) return false;
  cos2 = g / eta;
  return true;
}

// (1 - c)^5 as in schlick (physics.rs:316-318) and reflection_index
// (surface.rs:512-515).
__device__ __forceinline__ float pow5(float x) {
  float x2 = x * x;
  return x2 * x2 * x;
}
__device__ __forceinline__ float schlick(float f0, float c) { return f0 + (1.0f - f0) * pow5(1.0f - c); }

// generate_random_dir_by_angle (algebra.rs:225-235) + diffuse_reflection
// (surface.rs:408-417): uniform on the sphere, flipped into n's hemisphere.
__device__ __forceinline__ V3 diffuse_reflection(V3 n, Rng& rng) {
  float phi = TWO_PI_F * rng.next();
  float xi = 2.0f * rng.next() - 1.0f;
  float xi2 = __builtin_amdgcn_sqrtf(fmaxf(1.0f - xi * xi, 0.0f));
  const float sp = __sinf(phi), cp = __cosf(phi);  // = __sincosf, without its pointer outputs
  V3 d = normalize_or(mk(xi2 * cp, xi, xi2 * sp), mk(0.0f, 1.0f, 0.0f));
  return sel(dot(n, d) > 0.0f, d, -d);
}

// reflection_glossy (physics.rs:232-261).  xi1 = xi0^(pw c0) is formed as
// exp(t) with 1 - xi1 = -expm1(t), t = pw c0 ln xi0, so that the tiny
// deviation of near-mirror surfaces (pw ~ 1e-5) keeps float precision.
__device__ __forceinline__ V3 reflection_glossy(V3 n, V3 r, float pw, Rng& rng) {
  V3 u;
  if (!normalize(cross(mk(0.00424f, 1.0f, 0.00764f), r), u))
    u = normalize_or(cross(mk(1.0f, 0.00424f, 0.00764f), r), mk(1.0f, 0.0f, 0.0f));
  V3 v = cross(u, r);
  float c0 = dot(n, r);
  float xi0 = rng.next();
  float xi2 = TWO_PI_F * rng.next();
  float e = pw * c0;
  float xi1, one_m;
  if (xi0 == 0.0f) {  // 0^e: 1 when e == 0, else 0
    xi1 = e == 0.0f ? 1.0f : (e > 0.0f ? 0.0f : INF_F);
    one_m = 1.0f - xi1;
  } else {
    float t = e * logf(xi0);
    xi1 = expf(t);
    one_m = -expm1f(t);
  }
  float s = sqrtf(fmaxf(one_m * (1.0f + xi1), 0.0f));
  const float sp = __sinf(xi2), cp = __cosf(xi2);
  float x = cp * s, y = xi1, z = sp * s;
  V3 wi = x * u + y * r + z * v;
  if (dot(n, wi) < 0.0f) wi = (-x) * u + y * r - z * v;
  return normalize_or(wi, mk(1.0f, 0.0f, 0.0f));
}

// ---------------------------------------------------------------------------
// Surfaces (surface.rs).

__device__ __forceinline__ bool is_black(float4 c) { return c.x == 0.0f && c.y == 0.0f && c.z == 0.0f; }
__device__ __forceinline__ bool is_white(float4 c) { return c.x == 1.0f && c.y == 1.0f && c.z == 1.0f; }

// Surface::reflect (surface.rs:68-100)
__device__ __forceinline__ bool surf_reflect(const DevMaterial& m, float cos) {
  if (m.surface == SURF_SIMPLE) return !(m.diff == 1.0f || (cos == 1.0f && is_black(m.c2)));
  if (m.surface == SURF_TS) return m.metal == 1.0f ? !is_black(m.c2) : true;
  return false;
}
// Surface::refract (surface.rs:102-133)
__device__ __forceinline__ bool surf_refract(const DevMaterial& m, float cos) {
  if (m.surface == SURF_SIMPLE) return !(cos == 0.0f && is_white(m.c2));
  if (m.surface == SURF_TS) return m.metal == 0.0f && m.diff < 1.0f && !is_black(m.c1);
  return false;
}
// Surface::store_photon (surface.rs:289-310)
__device__ __forceinline__ bool surf_store(const DevMaterial& m) {
  if (m.surface == SURF_SIMPLE) return m.diff > 0.0f;
  if (m.surface == SURF_TS) return m.metal != 1.0f && m.diff != 0.0f;
  return true;
}
// Surface::power_glossy (surface.rs:381-402)
__device__ __forceinline__ float power_glossy(const DevMaterial& m) {
  return (m.surface == SURF_SIMPLE || m.surface == SURF_TS) ? m.density_pow : 0.0f;
}

// The BSDF of surface.rs:135-206 is linear in (di, si, ti):
//   result = wd*di + ws*si + wt*ti
// Simple (surface.rs:150-153):
//   f = reflection_index(specular_refl, cos0)
//   wd = diffuseness * reflectance / pi
//   ws = (1 - diffuseness) * f
//   wt = (1 - diffuseness) * (1 - metalness) * (1 - f)
// TS (surface.rs:164-197):
//   f = reflection_index(albedo_spec, cos0); f2 = 1 - f
//   metalness == 0: wd = f2*albedo_diff*scatterness/pi,
//                   wt = f2*albedo_diff*(1 - scatterness); else wd = wt = 0
//   ws = f
struct BsdfW {
  C3 wd, ws, wt;
};
__device__ __forceinline__ C3 reflection_index(float4 col, float c) {
  float c2 = pow5(1.0f - c);
  return C3{col.x + (1.0f - col.x) * c2, col.y + (1.0f - col.y) * c2, col.z + (1.0f - col.z) * c2};
}
__device__ __forceinline__ BsdfW bsdf_weights(const DevMaterial& m, float cos0) {
  BsdfW w;
  C3 z{0.0f, 0.0f, 0.0f};
  w.wd = z; w.ws = z; w.wt = z;
  if (m.surface == SURF_SIMPLE) {
    C3 f = reflection_index(m.c2, cos0);
    float od = 1.0f - m.diff;
    w.wd = c3(m.c1) * (m.diff * INV_PI_F);
    w.ws = f * od;
    float k = od * (1.0f - m.metal);
    w.wt = C3{(1.0f - f.r) * k, (1.0f - f.g) * k, (1.0f - f.b) * k};
  } else if (m.surface == SURF_TS) {
    C3 f = reflection_index(m.c2, cos0);
    C3 f2{1.0f - f.r, 1.0f - f.g, 1.0f - f.b};
    if (m.metal == 0.0f) {
      C3 a = f2 * c3(m.c1);
      w.wd = a * (m.diff * INV_PI_F);
      w.wt = a * (1.0f - m.diff);
    }
    w.ws = f;
  }
  return w;
}

// ---------------------------------------------------------------------------
// Lights (light.rs).

// Σ over one light of light.get_radiance(dists) zipped with cos^2
// (tracer.rs:263-290, light.rs:93-150).  The reference prepends RADIANCE0 to
// the radiance list before the zip (light.rs:132), so sample i's cos pairs
// with sample i-1's radiance and the last radiance is dropped: the running
// `prev` carries that shift.
// illuminated (tracer.rs:280-290) as a shadow test: the sample is lit iff
// the ray p -> ld meets some shape and its nearest hit lies no more than
// LIGHT_SLACK (in squared distance) short of the light.  The squared distance
// |(p + t d) - p|^2 of a hit never decreases with t (every float operation in
// it is monotonic), so that is: some shape is hit, and no hit has
// sq_l - sq_o > LIGHT_SLACK (an occluder).  The occluders are looked for
// among the candidates of `cand` (shapes that can meet a segment from p to
// the light) and in the BVH, where a lane stops at its first one; only when
// none of those shapes is hit at all are the other shapes tested, for any
// hit.  The decision equals that of the nearest-hit test except for a
// nearest sphere hit whose normal cannot be formed (a sphere of radius ~0).
__device__ __forceinline__ bool occluder_hit(const Hit& h, V3 p, V3 d, float sq_l, bool& any) {
#pragma clang fp contract(off)
  if (h.obj < 0) return false;
  any = true;
  const V3 q = mk(p.x + d.x * h.t, p.y + d.y * h.t, p.z + d.z * h.t);  // as finish_hit forms it
  const V3 od = q - p;
  return sq_l - dot(od, od) > LIGHT_SLACK_F;
}
__device__ __forceinline__ bool occludes(const DevShape& s, V3 p, V3 d, float sq_l, bool& any) {
  Hit h{INF_F, -1};
  shape_hit(s, 0, p, d, h);
  return occluder_hit(h, p, d, sq_l, any);
}
// occludes() for BVH primitive j (LDS compact sphere when there is one)
template <bool CACHED>
__device__ __forceinline__ bool bprim_occludes(const SceneView& sc, uint32_t j, V3 p, V3 d, float sq_l, bool& any) {
  if (CACHED && (sc.lds & 8u)) {
    const float4 c = bvh_lds_sph()[j];
    if (c.w >= 0.0f) {
      Hit h{INF_F, -1};
      sphere_hit(mk(c.x, c.y, c.z), c.w, 0, p, d, h);
      return occluder_hit(h, p, d, sq_l, any);
    }
  }
  return occludes(sc.bprims[j], p, d, sq_l, any);
}
// bprim_occludes over small-tree leaf `leaf` from the LDS copies, as
// leaf_hit_lds; true at the first occluder, whose slot goes to occ.
__device__ __forceinline__ bool leaf_occludes_lds(const SceneView& sc, uint32_t leaf, V3 p, V3 d, float sq_l,
                                                  bool& any, uint32_t& occ) {
  const uint32_t base = leaf * BVH_LEAF_SMALL;
  float4 c[BVH_LEAF_SMALL];
#pragma unroll
  for (uint32_t k = 0; k < BVH_LEAF_SMALL; ++k) c[k] = bvh_lds_sph()[base + k];
#pragma unroll
  for (uint32_t k = 0; k < BVH_LEAF_SMALL; ++k) {
    Hit h{INF_F, -1};
    sphere_hit(mk(c[k].x, c[k].y, c[k].z), c[k].w, 0, p, d, h, c[k].w >= 0.0f);
    if (occluder_hit(h, p, d, sq_l, any) || (c[k].w == -1.0f && occludes(sc.bprims[base + k], p, d, sq_l, any))) {
      occ = base + k;
      return true;
    }
  }
  return false;
}

// Any-hit form of bvh_nearest for the shadow test: true at the first
// occluder, whose primitive slot goes to occ; sets any when some primitive
// is hit.  Once a hit is known only boxes entered before t_occ (beyond which
// no hit can be an occluder) are searched.
template <bool CACHED, bool NL>
__device__ __forceinline__ bool bvh_shadow_walk(const SceneView& sc, V3 o, V3 d, float sq_l, bool& any,
                                                uint32_t& occ) {
  const V3 inv = slab_inv(d);
  const V3 noi = mk(-o.x * inv.x, -o.y * inv.y, -o.z * inv.z);
  const float t_occ = sqrtf(fmaxf(sq_l - LIGHT_SLACK_F, 0.0f)) * 1.001f + 1e-4f;
  const uint32_t first_leaf = (1u << sc.bvh_depth) - 1u;
  uint32_t node = 0, trail = 0;
  float tn;
  if (!slab(bvh_node<NL>(sc, 0), inv, noi, any ? t_occ : INF_F, tn)) return false;
  while (true) {
    if (node >= first_leaf) {
      if (NL) {
        if (leaf_occludes_lds(sc, node - first_leaf, o, d, sq_l, any, occ)) return true;
      } else {
        const uint32_t base = (node - first_leaf) * sc.bvh_leaf;
        const uint32_t end = min(base + sc.bvh_leaf, sc.n_bprims);
        for (uint32_t j = base; j < end; ++j)
          if (bprim_occludes<CACHED>(sc, j, o, d, sq_l, any)) {
            occ = j;
            return true;
          }
      }
    } else {
      const uint32_t l = 2u * node + 1u;
      const float tmax = any ? t_occ : INF_F;
      float tl, tr;
      const bool hl = slab(bvh_node<NL>(sc, l), inv, noi, tmax, tl);
      const bool hr = slab(bvh_node<NL>(sc, l + 1u), inv, noi, tmax, tr);
      if (hl && hr) {
        const bool left_first = tl <= tr;
        node = left_first ? l : l + 1u;
        trail |= 1u << heap_level(node);
        continue;
      }
      if (hl || hr) {
        node = hl ? l : l + 1u;
        continue;
      }
    }
    bool more = false;
    while (trail != 0u) {
      node = trail_pop(trail, node);
      if (slab(bvh_node<NL>(sc, node), inv, noi, any ? t_occ : INF_F, tn)) {
        more = true;
        break;
      }
    }
    if (!more) break;
  }
  return false;
}
// occ: the primitive that occluded the lane's previous shadow ray from the
// same point (the samples of one light go to nearby points), tested before
// the walk; the walk's first occluder replaces it.  An occluder found either
// way makes the same decision: the ray is shadowed, and some primitive is
// hit.
template <bool CACHED>
__device__ __forceinline__ bool bvh_shadow(const SceneView& sc, V3 o, V3 d, float sq_l, bool& any, uint32_t& occ) {
  if (sc.n_bprims == 0) return false;
  if (occ < sc.n_bprims && bprim_occludes<CACHED>(sc, occ, o, d, sq_l, any)) return true;
  if (CACHED && (sc.lds & 12u) == 12u) return bvh_shadow_walk<CACHED, true>(sc, o, d, sq_l, any, occ);
  return bvh_shadow_walk<CACHED, false>(sc, o, d, sq_l, any, occ);
}

// Linear shapes that can meet a segment from p to a point of light L's
// parallelogram (bit i: shape i of lin; shapes from the 65th on always
// count): a plane unless p and the four corners lie strictly on one side, a
// bounded shape whose padded box meets the box of p and the corners.
constexpr uint32_t CAND_BITS = 64;
__device__ __forceinline__ uint64_t occluder_candidates(const SceneView& sc, const DevLight& L, V3 p) {
  const V3 a = mk(L.pos), b = a + mk(L.d1), c = a + mk(L.d2), e = b + mk(L.d2);
  V3 lo = mk(fminf(fminf(p.x, a.x), fminf(fminf(b.x, c.x), e.x)), fminf(fminf(p.y, a.y), fminf(fminf(b.y, c.y), e.y)),
             fminf(fminf(p.z, a.z), fminf(fminf(b.z, c.z), e.z)));
  V3 hi = mk(fmaxf(fmaxf(p.x, a.x), fmaxf(fmaxf(b.x, c.x), e.x)), fmaxf(fmaxf(p.y, a.y), fmaxf(fmaxf(b.y, c.y), e.y)),
             fmaxf(fmaxf(p.z, a.z), fmaxf(fmaxf(b.z, c.z), e.z)));
  const float m = fmaxf(fmaxf(fmaxf(fabsf(lo.x), fabsf(lo.y)), fmaxf(fabsf(lo.z), fabsf(hi.x))),
                        fmaxf(fabsf(hi.y), fabsf(hi.z)));
  const float pad = 1e-5f * (1.0f + m);
  lo = lo - mk(pad, pad, pad);
  hi = hi + mk(pad, pad, pad);
  uint64_t mask = 0;
  const uint32_t n = min(sc.n_lin, CAND_BITS);
  for (uint32_t i = 0; i < n; ++i) {
    const Node bx = lin_box(sc, i);
    bool may;
    if (bx.lo.w != 0.0f) {
      const DevShape s = lin_shape(sc, i);
      const V3 nn = mk(s.n);
      const float tol = 1e-5f * (1.0f + fabsf(s.dist) + m);
      const float e0 = dot(nn, p) + s.dist, e1 = dot(nn, a) + s.dist, e2 = dot(nn, b) + s.dist;
      const float e3 = dot(nn, c) + s.dist, e4 = dot(nn, e) + s.dist;
      const bool above = e0 > tol && e1 > tol && e2 > tol && e3 > tol && e4 > tol;
      const bool below = e0 < -tol && e1 < -tol && e2 < -tol && e3 < -tol && e4 < -tol;
      may = !(above || below);
    } else {
      may = !(bx.hi.x < lo.x || bx.lo.x > hi.x || bx.hi.y < lo.y || bx.lo.y > hi.y || bx.hi.z < lo.z ||
              bx.lo.z > hi.z);
    }
    if (may) mask |= 1ull << i;
  }
  return mask;
}

// The light's own shape `self` (bit of cand), whose plane has normal nq
// (= d1 x d2, |nq| = nqn): the sample point lies on it at (u, v) in
// [0.1, 0.9]^2, so a ray to it that is not grazing (at least ~6 degrees off
// the plane) and neither very short nor very long meets it at the sample's
// distance up to float error far below the 0.1 margin of (u, v) and the
// slack of the occluder test: hit, not an occluder.  Such samples skip its
// test and count as having met a shape; the others test it as any other
// candidate.
struct LightSelf {
  uint32_t self;
  V3 nq;
  float nqn;
};

// min over the box bx of dot(nv, x - q): above a small pad, the box lies wholly
// on the positive side of the plane through q with normal nv
__device__ __forceinline__ float box_beyond(const Node& bx, V3 nv, V3 q) {
  const float x = nv.x > 0.0f ? bx.lo.x : bx.hi.x;
  const float y = nv.y > 0.0f ? bx.lo.y : bx.hi.y;
  const float z = nv.z > 0.0f ? bx.lo.z : bx.hi.z;
  return dot(nv, mk(x - q.x, y - q.y, z - q.z));
}

// The pyramid with apex p over the samples' area of light L, the
// parallelogram of (u, v) in [0.1, 0.9]^2 that holds the TSS points (corner
// j: a + u_j e1 + v_j e2, counter-clockwise in (u, v)).  Face j: the sides
// j = 0..3 through p and corners j, j + 1, the base (j = 4) through a, each
// with its outward normal f through q and a pad on dot(f, x - q) that
// bounds its rounding: padk times |x||y| of the cross product x X y that
// forms f, so it holds also where a face is nearly degenerate.
struct PyrBase {
  V3 a, e1, e2, kc;
  float dmax, padk, l1, l2;
};
__device__ __forceinline__ PyrBase pyr_base(const DevLight& L, V3 p) {
  PyrBase b;
  b.a = mk(L.pos);
  b.e1 = mk(L.d1);
  b.e2 = mk(L.d2);
  b.kc = b.a + 0.5f * b.e1 + 0.5f * b.e2;
  float dmax = 0.0f;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const V3 r = b.a + (j == 1 || j == 2 ? 0.9f : 0.1f) * b.e1 + (j >= 2 ? 0.9f : 0.1f) * b.e2 - p;
    dmax = fmaxf(dmax, dot(r, r));
  }
  b.dmax = sqrtf(dmax);
  const float m =
      fmaxf(fmaxf(fabsf(p.x), fabsf(p.y)), fmaxf(fabsf(p.z), fmaxf(fmaxf(fabsf(b.a.x), fabsf(b.a.y)), fabsf(b.a.z))));
  b.l1 = sqrtf(dot(b.e1, b.e1));
  b.l2 = sqrtf(dot(b.e2, b.e2));
  b.padk = 1e-5f * (1.0f + m + b.dmax + b.l1 + b.l2);
  return b;
}
__device__ __forceinline__ void pyr_face(const PyrBase& b, V3 p, int j, V3& f, V3& q, float& pad) {
  if (j < 4) {
    const int j1 = (j + 1) & 3;
    const V3 ra = b.a + (j == 1 || j == 2 ? 0.9f : 0.1f) * b.e1 + (j >= 2 ? 0.9f : 0.1f) * b.e2 - p;
    const V3 rb = b.a + (j1 == 1 || j1 == 2 ? 0.9f : 0.1f) * b.e1 + (j1 >= 2 ? 0.9f : 0.1f) * b.e2 - p;
    f = cross(ra, rb);
    f = sel(dot(f, b.kc - p) > 0.0f, -1.0f * f, f);
    q = p;
    pad = b.padk * sqrtf(dot(ra, ra)) * sqrtf(dot(rb, rb));
  } else {
    f = cross(b.e1, b.e2);
    f = sel(dot(f, p - b.a) > 0.0f, -1.0f * f, f);
    q = b.a;
    pad = b.padk * b.l1 * b.l2;
  }
}

// The candidates (cand, a bit per shape of lin) that a shadow ray from p to a
// sample of L still has to test.  Both cuts are exact:
// - a parallelogram, polygon or plane whose plane holds p while the samples'
//   area lies strictly on one side of it: a ray from p to a sample x meets
//   that plane at |t| = |e(p)| |x - p| / |e(x) - e(p)| (e: signed distance to
//   the plane), at most (|e(p)| + err) dmax / (emin - |e(p)| - err) with dmax
//   the farthest corner of the area, emin the nearest corner's |e| and err a
//   bound on the float error of the shape's own test; kept below 1e-5, the
//   test's t is under NEARLY0 and it reports no hit, so the shape is dropped
//   for every sample (need_o, need_s);
// - a shape whose padded box lies beyond a face of the pyramid with apex p
//   over the samples' area (the TSS points lie in the parallelogram of
//   (u, v) in [0.1, 0.9]^2): an occluding hit lies on the segment from p to
//   the sample, inside that pyramid, so such a shape is met, if at all,
//   beyond the sample.  It never occludes and matters only through `any`,
//   which the samples that skip the light's own shape (LightSelf) already
//   hold: dropped for those samples only (need_s, which also drops self).
__device__ __forceinline__ void shadow_need(const SceneView& sc, const DevLight& L, V3 p, uint64_t cand, uint64_t wcand,
                                            uint32_t self, uint64_t& need_o, uint64_t& need_s) {
  need_o = cand;
  need_s = self < CAND_BITS ? cand & ~(1ull << self) : cand;
  // the wave's candidates other than the light's own shape (wave-uniform):
  // none (queries whose boxes to the light meet no other shape) leaves
  // nothing to drop
  const uint64_t others = self < CAND_BITS ? wcand & ~(1ull << self) : wcand;
  if (others == 0) return;
  const PyrBase pb = pyr_base(L, p);
  const V3 a = pb.a, e1 = pb.e1, e2 = pb.e2;
  const float dmax = pb.dmax;
  // pyramid faces one at a time (few registers held across the shape loop)
  uint64_t out = 0;
#pragma unroll 1
  for (int j = 0; j < 5; ++j) {
    V3 f, q;
    float pad;
    pyr_face(pb, p, j, f, q, pad);
    for (uint64_t w = others; w != 0; w &= w - 1) {
      const uint32_t i = (uint32_t)__builtin_ctzll(w);
      const Node bx = lin_box(sc, i);
      if (bx.lo.w == 0.0f && box_beyond(bx, f, q) > pad) out |= 1ull << i;
    }
  }
  need_s &= ~out;
  for (uint64_t w = others; w != 0; w &= w - 1) {
    const uint32_t i = (uint32_t)__builtin_ctzll(w);
    if (((cand >> i) & 1ull) == 0) continue;
    const DevShape s = lin_shape(sc, i);
    if (s.kind != SHAPE_PLAIN && s.kind != SHAPE_POLYGON && s.kind != SHAPE_PARALLELOGRAM) continue;
    V3 nn, q;
    float err;
    if (s.kind == SHAPE_PLAIN) {
      nn = mk(s.n);
      q = (-s.dist) * nn;
      err = 1e-6f * (fabsf(s.dist) + sqrtf(dot(p, p)));
    } else {
      const V3 d1 = mk(s.d1), d2 = mk(s.d2);
      const V3 c = cross(d1, d2);
      const float cl = sqrtf(dot(c, c));
      if (!(cl > 0.0f)) continue;
      nn = (1.0f / cl) * c;
      q = mk(s.p);
      const V3 pp = p - q;
      err = 1e-6f * sqrtf(dot(pp, pp)) * sqrtf(dot(d1, d1)) * sqrtf(dot(d2, d2)) / cl;
    }
    // distances of p and of the four corners to the shape's plane
    const float ep = dot(nn, p - q);
    const float ga = dot(nn, a - q), gu = dot(nn, e1), gv = dot(nn, e2);
    const float g0 = ga + 0.1f * gu + 0.1f * gv, g1 = ga + 0.9f * gu + 0.1f * gv;
    const float g2 = ga + 0.9f * gu + 0.9f * gv, g3 = ga + 0.1f * gu + 0.9f * gv;
    const float gmin = fminf(fminf(g0, g1), fminf(g2, g3)), gmax = fmaxf(fmaxf(g0, g1), fmaxf(g2, g3));
    const float emin = gmin > 0.0f ? gmin : (gmax < 0.0f ? -gmax : 0.0f);
    const float ap = fabsf(ep) + err + 1e-6f * (fabsf(ga) + fabsf(gu) + fabsf(gv));
    if (emin > 2.0f * ap && ap * dmax <= 1e-5f * (emin - ap)) {
      need_o &= ~(1ull << i);
      need_s &= ~(1ull << i);
    }
  }
}

template <int BV, bool BVH = (BV != 0)>
__device__ __forceinline__ bool sample_lit(const SceneView& sc, V3 p, V3 n, V3 ld, uint64_t cand, uint64_t need_o,
                                           uint64_t need_s, uint64_t wu_o, uint64_t wu_s, const LightSelf& ls,
                                           uint32_t& occ, float& cos0sq, float& sqd) {
  V3 ld3;
  if (!normalize(ld, ld3)) return false;
  float c0 = dot(n, ld3);
  if (c0 < 0.0f) return false;
  const float sq_l = dot(ld, ld);
  const bool self = ls.self < CAND_BITS && sq_l >= 1e-4f && sq_l <= 100.0f && fabsf(dot(ld3, ls.nq)) >= 0.1f * ls.nqn;
  bool any = self;
  const uint64_t need = self ? need_s : need_o;
  // the shapes some active lane of the wave needs (wave-uniform), each tested
  // by the lanes that need it
  const uint64_t wu = __ballot(!self) != 0 ? wu_o : wu_s;
  for (uint64_t m = wu; m != 0; m &= m - 1) {
    const uint32_t i = (uint32_t)__builtin_ctzll(m);
    if (((need >> i) & 1ull) == 0) continue;
    if (occludes(lin_shape(sc, i), p, ld3, sq_l, any)) return false;
  }
  for (uint32_t i = CAND_BITS; i < sc.n_lin; ++i)
    if (occludes(lin_shape(sc, i), p, ld3, sq_l, any)) return false;
  if (BVH && bvh_shadow<BV == 2>(sc, p, ld3, sq_l, any, occ)) return false;
  // no occluder; lit iff the ray meets something at all (calc_intersection
  // returns None otherwise, tracer.rs:284)
  for (uint32_t i = 0; i < min(sc.n_lin, CAND_BITS) && !any; ++i) {
    if ((cand >> i) & 1ull) continue;
    Hit h{INF_F, -1};
    shape_hit(lin_shape(sc, i), 0, p, ld3, h);
    any = h.obj >= 0;
  }
  if (!any) return false;
  cos0sq = c0 * c0;
  sqd = sq_l;
  return true;
}

__device__ __forceinline__ C3 light_radiance(const DevLight& L, float sqd) {
  if (L.kind == LIGHT_POINT) {
    float l0 = L.flux / (FOUR_PI_F * sqd);
    return c3(L.color) * l0;
  }
  if (L.kind == LIGHT_PARALLELOGRAM) {
    // per lit shadow sample: the reciprocal (v_rcp_f32, ~1 ulp) in place of
    // an IEEE division
    float l0 = (2.0f * L.flux * PARA_DIV_F * PARA_DIV_F * (1.0f / FOUR_PI_F)) * __builtin_amdgcn_rcpf(sqd);
    return c3(L.color) * l0;
  }
  return c3(L.color) * L.flux;  // SunLight
}

// Σ over the lights of the reference's direct term (tracer.rs:136-144).
// Only a ParallelogramLight contributes: a PointLight or SunLight yields at
// most one sample (light.rs:116-117,132-147), which the RADIANCE0 shift
// pairs with zero radiance, so their term is exactly 0 and no shadow ray is
// cast for them.
// Sample s of the TSS grid (light.rs:164-170), row-major in the first
// coordinate: (0.1 + 0.2 (s / 5), 0.1 + 0.2 (s % 5)) as the float literals
// 0.1, 0.3, 0.5, 0.7, 0.9; the light record holds the 25 points, formed once
// when the scene is made (a wave-uniform scalar load per sample).
__device__ __forceinline__ V3 tss_point(const SceneView& sc, uint32_t li, int s) {
  const float4 t = ldc(&sc.lights[li].tss[0], (uint32_t)s);
  return mk(t.x, t.y, t.z);
}

template <int BV, bool BVH = (BV != 0)>
__device__ C3 direct_light(const SceneView& sc, V3 p, V3 n) {
  C3 sum{0.0f, 0.0f, 0.0f};
  for (uint32_t li = 0; li < sc.n_lights; ++li) {
    const DevLight L = ldc(sc.lights, li);
    if (L.kind != LIGHT_PARALLELOGRAM) continue;
    const V3 ln = mk(L.nvec);
    const uint64_t cand = occluder_candidates(sc, L, p);
    LightSelf ls;
    ls.self = __float_as_uint(L.pos.w);
    ls.nq = cross(mk(L.d1), mk(L.d2));
    ls.nqn = sqrtf(dot(ls.nq, ls.nq));
    // the union of the candidates of the wave's active lanes (one ballot per
    // shape): the 25 samples then loop over these shapes only, instead of
    // every shape with a per-lane test (caustic shade 0.75 -> 0.64 ms)
    uint64_t wunion = 0;
    const uint32_t nb = min(sc.n_lin, CAND_BITS);
    for (uint32_t i = 0; i < nb; ++i)
      if (__ballot(((cand >> i) & 1ull) != 0)) wunion |= 1ull << i;
    uint64_t need_o, need_s;
    shadow_need(sc, L, p, cand, wunion, ls.self, need_o, need_s);
    uint64_t wu_o = 0, wu_s = 0;
    for (uint64_t w = wunion; w != 0; w &= w - 1) {
      const uint32_t i = (uint32_t)__builtin_ctzll(w);
      if (__ballot(((need_o >> i) & 1ull) != 0)) wu_o |= 1ull << i;
      if (__ballot(((need_s >> i) & 1ull) != 0)) wu_s |= 1ull << i;
    }
    C3 prev{0.0f, 0.0f, 0.0f};
    uint32_t occ = ~0u;  // the BVH primitive that shadowed the last sample
    for (int s = 0; s < 25; ++s) {
      const V3 ld = tss_point(sc, li, s) - p;
      if (!(dot(ln, ld) < 0.0f)) continue;  // light.rs:129-131
      float csq, sqd;
      if (!sample_lit<BV>(sc, p, n, ld, cand, need_o, need_s, wu_o, wu_s, ls, occ, csq, sqd)) continue;
      sum = sum + prev * csq;
      prev = light_radiance(L, sqd);
    }
  }
  return sum;
}

// ---------------------------------------------------------------------------
// Camera::generate_ray (camera.rs:58-75) for pixel (y, x); draws come from
// the pixel's root stream: blur offsets first, then the AA jitter.
__device__ __forceinline__ void camera_ray(const DevCamera& cam, int x, int y, Rng& rng, V3& o, V3& d) {
  V3 blur = mk(0.0f, 0.0f, 0.0f);
  if (cam.blur) {
    float r1 = rng.next() - 0.5f;
    float r2 = rng.next() - 0.5f;
    blur = r1 * mk(cam.eex) + r2 * mk(cam.eey);
  }
  float r3 = 0.0f, r4 = 0.0f;
  if (cam.progressive && cam.antialias) {
    r3 = rng.next() - 0.5f;
    r4 = rng.next() - 0.5f;
  }
  o = mk(cam.eye) + blur;
  V3 dir = mk(cam.origin) + ((float)x + r3) * mk(cam.esx) + ((float)y + r4) * mk(cam.esy) - blur;
  d = normalize_or(dir, mk(0.0f, 0.0f, 1.0f));
}

// ---------------------------------------------------------------------------
// Radius predicate of KdTree::within with squared_euclidean: the left fold
// ((0 + dx^2) + dy^2) + dz^2 over the axes, compared inclusively.  Evaluated
// with no contraction so that the host float brute force reproduces it bit
// for bit.

__device__ __forceinline__ float dist2_nc(float qx, float qy, float qz, float px, float py, float pz) {
#pragma clang fp contract(off)
  float dx = qx - px;
  float dy = qy - py;
  float dz = qz - pz;
  float s = dx * dx + dy * dy;
  return s + dz * dz;
}

// Squared distance from q to an axis-aligned box, with the same operation
// order: never larger than dist2_nc(q, p) for any p inside the box, because
// float subtraction, squaring and addition are monotonic.
__device__ __forceinline__ float box_dist2_nc(float qx, float qy, float qz, float4 lo, float4 hi) {
#pragma clang fp contract(off)
  float dx = fmaxf(fmaxf(lo.x - qx, qx - hi.x), 0.0f);
  float dy = fmaxf(fmaxf(lo.y - qy, qy - hi.y), 0.0f);
  float dz = fmaxf(fmaxf(lo.z - qz, qz - hi.z), 0.0f);
  float s = dx * dx + dy * dy;
  return s + dz * dz;
}

// filter_cone (tracer.rs:197-204) and filter_gauss (tracer.rs:206-216) as
// functions of t = d2 / r2 (both reference filters depend on d2 only through
// that ratio).  Hardware square root and exponential (a few ulp) and
// products by the constants' reciprocals: the weight is formed per photon
// in the gather, and these keep it within 1e-6 of the exact one.
__device__ __forceinline__ float filter_weight(int pf, float t) {
  if (pf == FILTER_CONE) {
    const float d = __builtin_amdgcn_sqrtf(t) * (1.0f / K_CONE_F);
    return d > 1.0f ? 0.0f : (1.0f - d) * (1.0f / FAC_K_F);
  }
  if (pf == FILTER_GAUSS) {
    const float e_r = 1.0f - __expf(-GAUSS_BETA_F * t * 0.5f);
    return e_r > GAUSS_E_BETA_F ? 0.0f : GAUSS_ALPHA_F * (1.0f - e_r * (1.0f / GAUSS_E_BETA_F)) + GAUSS_CORR_F;
  }
  return 1.0f;
}

}  // namespace ppm
