// ppm_kernels.hip — gfx950 kernels of one PPM-PA pass:
//   photon emission + bouncing   (light.rs:67-91, tracer.rs:31-125)
//   photon map build             (photonmap.rs:23-29 + crate kdtree)
//   eye-ray tree + direct light  (tracer.rs:129-177, 263-290, camera.rs:58-75)
//   radius query + gather        (tracer.rs:179-216, optics.rs:224-233)
//   FP64 per-pixel accumulation  (util/averager2.rb:49-62)
//
// Wavefront-64 design: one photon / pixel / query per lane; wave-wide ballots
// compact live photons and stored records; the map is a balanced binary tree
// over Morton-sorted 64-photon leaves, traversed wave-coherently so that one
// leaf is staged once into LDS (coalesced loads) and tested by all 64 lanes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "ppm_dev.hpp"
#include "ppm_launch.hpp"

namespace ppm {

// Block index and grid size as a kernel body sees them: a plain launch passes
// blockIdx.x / gridDim.x, a batched launch the block's place within its
// scene's share of the grid.
struct Blk {
  uint32_t b, n;
};

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }

__host__ __device__ inline SceneView make_view(const SceneBuf& b) {
  SceneView v;
  v.shapes = b.shapes;
  v.mats = b.mats;
  v.lights = b.lights;
  v.lin = b.lin;
  v.lin_box = b.lin_box;
  v.bvh = b.bvh;
  v.bprims = b.bprims;
  v.n_lin = b.n_lin;
  v.n_bprims = b.n_bprims;
  v.bvh_depth = b.bvh_depth;
  v.bvh_leaf = b.bvh_leaf;
  v.n_shapes = b.n_shapes;
  v.n_mats = b.n_mats;
  v.n_lights = b.n_lights;
  v.n_emit = b.n_emit_total;
  v.lds = 0;
  v.dyn_lds = bvh_dyn_lds(b) != 0u ? 1u : 0u;
  v.cam = b.cam;
  v.bmin = b.bmin;
  v.bmax = b.bmax;
  return v;
}

// ===========================================================================
// Pass state (util/iterator.rb:33-38): r_{p}^2 = r_{p-1}^2 (p - 1 + alpha) / p,
// in double on the device so that a captured pass loop needs no host input.

// A state that runs every stride-th pass (one of several pass pipelines)
// steps through the passes in between, so r2 of every pass is the same FP64
// product as in a single pipeline.
__device__ __forceinline__ void k_pass_begin_body(PassState* st, uint32_t first_pass, double r2_first, double alpha,
                                                  double power, uint32_t stride, Blk blk) {
  if (blk.b != 0) return;
  if (threadIdx.x < 16) st->work[threadIdx.x >> 3][threadIdx.x & 7u][0] = 0;
  if (threadIdx.x != 0) return;
  uint32_t pass = st->pass;
  double a = st->alpha, r2 = st->r2, pw = st->power;
  if (pass == 0) {
    pass = first_pass;
    a = alpha;
    pw = power;
    st->alpha = alpha;
    st->r2_first = r2_first;
    st->power = power;
    r2 = r2_first;
    for (uint32_t k = 1; k < first_pass; ++k) r2 *= ((double)k + alpha) / ((double)k + 1.0);
  } else {
    for (uint32_t j = 0; j < stride; ++j) {
      pass += 1;
      double k = (double)(pass - 1);
      r2 *= (k + a) / (k + 1.0);
    }
  }
  st->pass = pass;
  st->r2 = r2;
  st->r2f = (float)r2;
  st->scale = (float)(pw / (3.14159265358979323846 * r2));
  st->n_stored = 0;
  st->next_photon = 0;
  st->n_queries = 0;
  st->tiles_front = 0;
  st->tiles_back = 0;
}

__global__ void k_pass_begin(PassState* st, uint32_t first_pass, double r2_first, double alpha, double power,
                             uint32_t stride) {
  k_pass_begin_body(st, first_pass, r2_first, alpha, power, stride, Blk{blockIdx.x, gridDim.x});
}

__device__ __forceinline__ void k_pass_end_body(PassState* st, Blk blk) {
  if (threadIdx.x != 0 || blk.b != 0) return;
  st->total_stored += st->n_stored;
  st->total_queries += st->n_qsort + min(st->n_queries, 0xFFFFFFFFu);
}

__global__ void k_pass_end(PassState* st) {
  k_pass_end_body(st, Blk{blockIdx.x, gridDim.x});
}

hipError_t launch_pass_begin(hipStream_t s, PassState* st, uint32_t first_pass, double r2_first, double alpha,
                             double power, uint32_t stride) {
  hipLaunchKernelGGL(k_pass_begin, dim3(1), dim3(64), 0, s, st, first_pass, r2_first, alpha, power, stride);
  return hipGetLastError();
}
hipError_t launch_pass_end(hipStream_t s, PassState* st) {
  hipLaunchKernelGGL(k_pass_end, dim3(1), dim3(64), 0, s, st);
  return hipGetLastError();
}

// ===========================================================================
// Photon emission (light.rs:67-91) and one bounce (tracer.rs:31-125).

struct PhotonState {
  V3 o, d;
  uint32_t wl, med, depth, pid;
  float w;  // weight: the product of c1/q over the diffuse survivals
};

// Light::generate_photon for photon `pid` of the pass.
__device__ __forceinline__ void emit_photon(const SceneView& sc, uint32_t pid, Rng& rng, PhotonState& p) {
  uint32_t li = 0;
  for (uint32_t i = 0; i < sc.n_lights; ++i)
    if (pid >= ldc(sc.lights, i).first_photon) li = i;
  const DevLight& L = sc.lights[li];
  // select_wavelength / decide_wavelength (light.rs:157-160, physics.rs:74-84)
  float u = rng.next();
  p.wl = u < L.color.x ? 0u : (u < L.color.x + L.color.y ? 1u : 2u);
  // origin and direction are set once after the branches: stores to
  // different fields in each branch get merged into one store through a
  // chosen address, which keeps the photon state out of registers
  V3 o, d;
  if (L.kind == LIGHT_PARALLELOGRAM) {
    float t1 = rng.next();
    float t2 = rng.next();
    d = diffuse_reflection(mk(L.nvec), rng);
    o = mk(L.pos) + t1 * mk(L.d1) + t2 * mk(L.d2);
  } else if (L.kind == LIGHT_POINT) {
    // generate_random_dir (algebra.rs:211-223): rejection in the unit ball
    V3 v = mk(0.0f, 1.0f, 0.0f);
    for (int k = 0; k < 64; ++k) {
      float x = 2.0f * rng.next() - 1.0f;
      float y = 2.0f * rng.next() - 1.0f;
      float z = 2.0f * rng.next() - 1.0f;
      V3 c = mk(x, y, z);
      float len = sqrtf(dot(c, c));
      if (0.0f < len && len <= 1.0f) {
        v = c * (1.0f / len);
        break;
      }
    }
    d = v;
    o = mk(L.pos);
  } else {  // SunLight
    float t1 = rng.next();
    float t2 = rng.next();
    o = mk(L.pos) + t1 * mk(L.d1) + t2 * mk(L.d2);
    d = mk(L.dir);
  }
  p.o = o;
  p.d = d;
  p.med = sc.n_mats;  // M_AIR
  p.depth = 0;
  p.pid = pid;
  p.w = 1.0f;
}

// One step of trace_photon.  Returns whether the photon lives on; `store`
// reports a photon-map record (position, INCOMING direction, channel).
template <int BV, bool BVH = (BV != 0)>
__device__ __forceinline__ bool photon_bounce(const SceneView& sc, int uc, int maxd, PhotonState& p, Rng& rng,
                                              bool& store, V3& spos, int& ev_mat) {
  store = false;
  ev_mat = -1;
  if ((int)p.depth >= maxd) return false;  // tracer.rs:32
  Isect is;
  if (!intersect<BV>(sc, p.o, p.d, is)) return false;  // tracer.rs:36-39
  const DevMaterial M = material(sc, is.mat);
  // tracer.rs:77-79
  store = (uc == 0 || p.depth > 0) && surf_store(M);
  spos = is.pos;
  if (M.surface == SURF_SIMPLE) {
    // russian_roulette(&[roughness()]) where roughness() is diffuseness
    // (tracer.rs:51-53, surface.rs:358-367); check_under: 0 iff u <= p.
    float u = rng.next();
    if (u <= M.diff) {
      // reflect_diff (tracer.rs:83-92): survive with probability q (the
      // reference's c1 unless set otherwise) and weight c1/q
      float u2 = rng.next();
      const float qq = chan(M.q, p.wl), a = chan(M.c1, p.wl);
      if (!(u2 <= qq)) return false;
      if (a != qq) p.w *= a / qq;
      ev_mat = (int)is.mat;
      p.d = diffuse_reflection(is.n, rng);
      p.o = is.pos;
      p.depth += 1;
      return true;
    }
    // reflect_spec (tracer.rs:94-109)
    float cos1;
    V3 rdir = specular_reflection(is.n, p.d, cos1);
    float f = schlick(chan(M.c2, p.wl), cos1);
    float u2 = rng.next();
    if (u2 <= f) {
      p.d = rdir;
      p.o = is.pos;
      p.depth += 1;
      return true;
    }
    float ior2 = chan(M.ior, p.wl);
    if (ior2 == 0.0f) return false;
    // reflect_trans (tracer.rs:111-125)
    float eta = relative_ior(chan(material(sc, p.med).ior, p.wl), ior2);
    V3 t;
    float c2;
    if (!specular_refraction(is.n, p.d, eta, t, c2)) return false;
    p.med = dot(t, is.n) < 0.0f ? is.mat : sc.n_mats;
    p.d = t;
    p.o = is.pos;
    p.depth += 1;
    return true;
  }
  if (M.surface == SURF_TS) {
    // Surface::next_direction (surface.rs:211-258), medium rule tracer.rs:65-69
    float eta = relative_ior(chan(material(sc, p.med).ior, p.wl), chan(M.ior, p.wl));
    float cos1;
    V3 rdir0 = specular_reflection(is.n, p.d, cos1);
    V3 rdir = reflection_glossy(is.n, rdir0, power_glossy(M), rng);
    V3 hvec = normalize_or(rdir - p.d, is.n);
    V3 tdir;
    float cos2;
    bool has_t = specular_refraction(hvec, p.d, eta, tdir, cos2);
    float cs = cos1 < cos2 ? cos1 : cos2;
    float f = schlick(chan(M.c2, p.wl), cs);
    V3 nd;
    bool keep_medium = true;
    if (rng.next() <= f) {
      nd = rdir;
    } else if (!(rng.next() <= chan(M.c1, p.wl))) {
      return false;
    } else if (rng.next() <= M.diff) {
      nd = diffuse_reflection(is.n, rng);
    } else {
      if (!has_t) return false;
      nd = tdir;
      keep_medium = false;
    }
    if (!keep_medium) p.med = is.mat;
    p.d = nd;
    p.o = is.pos;
    p.depth += 1;
    return true;
  }
  return false;  // Nothing / DisneyBRDF / Brady: tracer.rs:75
}

// Photon tracer.  Wave w owns photons [w*per_wave, (w+1)*per_wave) and the
// output region rec + w*region_cap.  Each wave keeps 64 photons in flight:
// lanes whose photon died take the next indices of the wave's range (ballot
// + mbcnt, a wave-uniform counter in registers), and stored records go to
// consecutive slots of the region (coalesced 32-byte writes).  No global
// atomics: the record order depends only on the data, so it is reproducible.
// Adjoint replay (ADJ): the same walk, storing nothing; at each record slot
// the record's adjoint weight adj[slot] is spread over the lane's diffuse
// survivals so far (their materials kept in LDS): d(weight)/d(c1) =
// weight / c1 per survival, so each adds adj / c1 of its material and the
// record's channel.
struct AdjTrace {
  const float* adj;  // per record slot (trace layout)
  double* grad;      // 3 per material
};
constexpr uint32_t ADJ_LDS_MATS = 128;  // materials with per-block LDS partial sums
constexpr uint32_t ADJ_MAX_EV = MAX_DEPTH_LIMIT;

__device__ __forceinline__ void adj_add(float* part, double* grad, uint32_t m, uint32_t c, float v) {
  if (m < ADJ_LDS_MATS)
    atomicAdd(&part[3 * m + c], v);
  else
    atomicAdd(&grad[3 * m + c], (double)v);
}

__device__ __forceinline__ void adj_flush(float* part, double* grad, uint32_t n_mats) {
  __syncthreads();
  const uint32_t n = 3u * min(n_mats, ADJ_LDS_MATS);
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
    if (part[i] != 0.0f) atomicAdd(&grad[i], (double)part[i]);
}

template <int BV, bool ADJ = false, bool BVH = (BV != 0)>
__device__ __forceinline__ void k_trace_photons_body(SceneView sc, PassState* st, PhotonRec* rec, uint32_t* count,
                                                       uint32_t n_waves, uint32_t per_wave, uint32_t region_cap,
                                                       uint64_t seed, int uc, int maxd, Blk blk,
                                                       AdjTrace at = AdjTrace{nullptr, nullptr}) {
  __shared__ float part[ADJ ? 3 * ADJ_LDS_MATS : 1];
  __shared__ uint16_t evl[ADJ ? ADJ_MAX_EV * 256 : 1];
  scene_to_lds<BV>(sc);
  const uint32_t w = blk.b * (blockDim.x / 64) + (threadIdx.x >> 6);
  if (ADJ) {
    for (uint32_t i = threadIdx.x; i < 3 * ADJ_LDS_MATS; i += blockDim.x) part[i] = 0.0f;
    __syncthreads();
  }
  if (w >= n_waves) {
    if (ADJ) adj_flush(part, at.grad, sc.n_mats);
    return;
  }
  const uint32_t pass = st->pass;
  const uint32_t lane = lane_id();
  const uint64_t lo = (uint64_t)w * per_wave;
  const uint32_t p_begin = (uint32_t)min<uint64_t>(lo, sc.n_emit);
  const uint32_t p_end = (uint32_t)min<uint64_t>(lo + per_wave, sc.n_emit);
  PhotonRec* region = rec + (uint64_t)w * region_cap;
  uint32_t next = p_begin;  // wave-uniform
  uint32_t cnt = 0;         // wave-uniform
  bool alive = false, done = false;
  PhotonState p;
  Rng rng;
  p.pid = 0;
  uint32_t n_ev = 0;
  while (true) {
    const bool need = !alive && !done;
    const uint64_t m = __ballot(need);
    if (m) {
      if (need) {
        const uint32_t pid = next + (uint32_t)__popcll(m & lanemask_lt());
        if (pid < p_end) {
          rng.init(seed, pid, pass, TAG_PHOTON);
          emit_photon(sc, pid, rng, p);
          n_ev = 0;
          alive = true;
        } else {
          done = true;
        }
      }
      next += (uint32_t)__popcll(m);
    }
    if (!__any(alive)) break;
    bool store = false;
    V3 spos = mk(0.0f, 0.0f, 0.0f);
    const V3 sdir = p.d;
    const float sw = p.w;
    const uint32_t swl = p.wl;
    const uint32_t smeta = pack_meta(p.wl, p.depth, p.pid);
    int ev_mat = -1;
    if (alive) alive = photon_bounce<BV>(sc, uc, maxd, p, rng, store, spos, ev_mat);
    const uint64_t sm = __ballot(store);
    if (sm) {
      if (store) {
        const uint32_t slot = cnt + (uint32_t)__popcll(sm & lanemask_lt());
        if (slot < region_cap) {
          if (ADJ) {
            const float A = at.adj[(uint64_t)w * region_cap + slot];
            if (A != 0.0f)
              for (uint32_t k = 0; k < n_ev; ++k) {
                const uint32_t mi = evl[k * 256u + threadIdx.x];
                const float a = chan(material(sc, mi).c1, swl);
                if (a != 0.0f) adj_add(part, at.grad, mi, swl, A / a);
              }
          } else {
            PhotonRec r;
            r.pm = make_float4(spos.x, spos.y, spos.z, __uint_as_float(smeta));
            r.dm = make_float4(sdir.x, sdir.y, sdir.z, sw);
            region[slot] = r;
          }
        }
      }
      cnt += (uint32_t)__popcll(sm);
    }
    if (ADJ && ev_mat >= 0 && n_ev < ADJ_MAX_EV) evl[(n_ev++) * 256u + threadIdx.x] = (uint16_t)ev_mat;
  }
  if (ADJ) {
    adj_flush(part, at.grad, sc.n_mats);
    return;
  }
  if (lane == 0) {
    if (cnt > region_cap) atomicOr(&st->overflow, 1u);
    count[w] = min(cnt, region_cap);
  }
}

template <int BV, bool BVH = (BV != 0)>
__global__ __launch_bounds__(256) void k_trace_photons(SceneView sc, PassState* st, PhotonRec* rec, uint32_t* count,
                                                       uint32_t n_waves, uint32_t per_wave, uint32_t region_cap,
                                                       uint64_t seed, int uc, int maxd) {
  k_trace_photons_body<BV>(sc, st, rec, count, n_waves, per_wave, region_cap, seed, uc, maxd, Blk{blockIdx.x, gridDim.x});
}

// Exclusive scan of up to 65536 counts in one block; the total goes to
// *total (the pass's stored-photon or sorted-query count).
__device__ __forceinline__ void k_scan_counts_body(const uint32_t* count, uint32_t* offset, uint32_t n,
                                                   uint32_t* total, const uint32_t* extra, uint32_t extra_cap) {
  __shared__ uint32_t wsum[16];
  const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
  const uint32_t per = (n + 1023) / 1024;
  uint32_t sum = 0;
  for (uint32_t k = 0; k < per; ++k) {
    const uint32_t i = t * per + k;
    if (i < n) sum += count[i];
  }
  uint32_t incl = sum;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t v = __shfl_up(incl, o);
    if (lane >= (uint32_t)o) incl += v;
  }
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  if (t == 0) {
    uint32_t acc = 0;
    for (int i = 0; i < 16; ++i) {
      const uint32_t v = wsum[i];
      wsum[i] = acc;
      acc += v;
    }
    *total = acc + (extra ? min(*extra, extra_cap) : 0u);
  }
  __syncthreads();
  uint32_t run = wsum[w] + incl - sum;
  for (uint32_t k = 0; k < per; ++k) {
    const uint32_t i = t * per + k;
    if (i < n) {
      const uint32_t c = count[i];
      offset[i] = run;
      run += c;
    }
  }
}
__global__ __launch_bounds__(1024) void k_scan_counts(const uint32_t* count, uint32_t* offset, uint32_t n,
                                                      uint32_t* total, const uint32_t* extra, uint32_t extra_cap) {
  k_scan_counts_body(count, offset, n, total, extra, extra_cap);
}

uint32_t trace_waves(uint64_t n_photon, int n_cu) {
  uint64_t w = (n_photon + 255) / 256;
  const uint64_t maxw = (uint64_t)n_cu * TRACE_WAVES_PER_CU;
  if (w > maxw) w = maxw;
  if (w > 8192) w = 8192;
  if (w < 1) w = 1;
  return (uint32_t)w;
}

hipError_t launch_trace_photons(hipStream_t s, const SceneBuf& sc, PassState* st, PhotonBuf& pb, uint64_t seed,
                                int use_classic, int max_depth) {
  const uint32_t n_waves = pb.n_regions;
  const uint32_t per_wave = (uint32_t)((sc.n_emit_total + (uint64_t)n_waves - 1) / n_waves);
  switch (bvh_kind(sc)) {
    case 2:
      hipLaunchKernelGGL(k_trace_photons<2>, dim3((n_waves + 3) / 4), dim3(256), bvh_dyn_lds(sc), s, make_view(sc),
                       st, pb.rec,
                       pb.count, n_waves, per_wave, pb.region_cap, seed, use_classic, max_depth);
      break;
    case 1:
      hipLaunchKernelGGL(k_trace_photons<1>, dim3((n_waves + 3) / 4), dim3(256), 0, s, make_view(sc),
                       st, pb.rec,
                       pb.count, n_waves, per_wave, pb.region_cap, seed, use_classic, max_depth);
      break;
    default:
      hipLaunchKernelGGL(k_trace_photons<0>, dim3((n_waves + 3) / 4), dim3(256), 0, s, make_view(sc), st, pb.rec,
                       pb.count, n_waves, per_wave, pb.region_cap, seed, use_classic, max_depth);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, s, pb.count, pb.offset, n_waves, &st->n_stored,
                     (const uint32_t*)nullptr, 0u);
  return hipGetLastError();
}

// Region lookup: the last region whose offset is <= i.
__device__ __forceinline__ uint64_t region_index(const uint32_t* offset, uint32_t n_regions, uint32_t region_cap,
                                                 uint32_t i) {
  uint32_t lo = 0, hi = n_regions;  // offset[lo] <= i < offset[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (offset[mid] <= i) lo = mid;
    else hi = mid;
  }
  return (uint64_t)lo * region_cap + (i - offset[lo]);
}

__global__ __launch_bounds__(256) void k_compact(const PhotonRec* rec, const uint32_t* offset, uint32_t n_regions,
                                                 uint32_t region_cap, const PassState* st, PhotonRec* dense) {
  const uint32_t n = st->n_stored;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    dense[i] = rec[region_index(offset, n_regions, region_cap, i)];
}

hipError_t launch_compact(hipStream_t s, const PhotonBuf& pb, const PassState* st, PhotonRec* dense) {
  uint64_t blocks = (pb.cap + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(k_compact, dim3((uint32_t)blocks), dim3(256), 0, s, pb.rec, pb.offset, pb.n_regions,
                     pb.region_cap, st, dense);
  return hipGetLastError();
}

// ===========================================================================
// Photon map build (photonmap.rs:23-29).  The reference inserts photons one
// by one into the kdtree crate; here: Morton keys in the quantisation box ->
// LSD radix sort (8-bit digits, device-side count) -> leaves of 64
// consecutive photons -> balanced binary tree (heap order) of AABBs built
// bottom-up with arrival counters.

__device__ __forceinline__ uint32_t expand_bits10(uint32_t v) {
  v &= 0x3FFu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// Sort key of a position: a 2-bit distance level above a 30-bit Morton code.
// Level 0 = inside the quantisation box (1024 cells per axis); levels 1 and 2
// = inside boxes 16x and 256x as large around the same centre; level 3 =
// beyond (clamped).  Far photons (walks that leaked through a corner of the
// geometry travel arbitrarily far) thus sort into leaves of their own instead
// of sharing a leaf, and a huge bounding box, with photons near the scene.
__device__ __forceinline__ uint32_t morton_key(float px, float py, float pz, float4 bmin, float4 bscale, float fq,
                                                uint32_t level_shift = 30u) {
  float fx = (px - bmin.x) * bscale.x;
  float fy = (py - bmin.y) * bscale.y;
  float fz = (pz - bmin.z) * bscale.z;
  const float n = fq + 1.0f, h = 0.5f * n;
  uint32_t level = 0;
  float s = 1.0f;
  while (level < 3 && !((fx - h) * s + h >= 0.0f && (fx - h) * s + h < n && (fy - h) * s + h >= 0.0f &&
                        (fy - h) * s + h < n && (fz - h) * s + h >= 0.0f && (fz - h) * s + h < n)) {
    ++level;
    s *= (1.0f / 16.0f);
  }
  fx = fminf(fmaxf((fx - h) * s + h, 0.0f), fq);
  fy = fminf(fmaxf((fy - h) * s + h, 0.0f), fq);
  fz = fminf(fmaxf((fz - h) * s + h, 0.0f), fq);
  return (level << level_shift) | (expand_bits10((uint32_t)fx) << 2) | (expand_bits10((uint32_t)fy) << 1) |
         expand_bits10((uint32_t)fz);
}

float4 morton_scale(float4 bmin, float4 bmax, uint32_t qmax) {
  float4 scale;
  scale.x = (float)((double)(qmax + 1) / fmax((double)bmax.x - (double)bmin.x, 1e-30));
  scale.y = (float)((double)(qmax + 1) / fmax((double)bmax.y - (double)bmin.y, 1e-30));
  scale.z = (float)((double)(qmax + 1) / fmax((double)bmax.z - (double)bmin.z, 1e-30));
  scale.w = 0.0f;
  return scale;
}


// Keys of the stored photons (morton_key) and their dense -> region index.
__device__ __forceinline__ void k_morton_body(const PhotonRec* ph, const uint32_t* offset, uint32_t n_regions,
                                                uint32_t region_cap, const PassState* st, uint32_t* keys,
                                                uint32_t* vals, float4 bmin, float4 bscale, uint32_t qmax,
                                                uint32_t* nonfinite, Blk blk) {
  const uint32_t n = st->n_stored;
  const float fq = (float)qmax;
  for (uint32_t i = blk.b * blockDim.x + threadIdx.x; i < n; i += blk.n * blockDim.x) {
    const uint64_t src = region_index(offset, n_regions, region_cap, i);
    const float4 p = ph[src].pm;
    if (!(isfinite(p.x) && isfinite(p.y) && isfinite(p.z))) atomicOr(nonfinite, 1u);
    keys[i] = morton_key(p.x, p.y, p.z, bmin, bscale, fq);
    vals[i] = (uint32_t)src;
  }
}

__global__ __launch_bounds__(256) void k_morton(const PhotonRec* ph, const uint32_t* offset, uint32_t n_regions,
                                                uint32_t region_cap, const PassState* st, uint32_t* keys,
                                                uint32_t* vals, float4 bmin, float4 bscale, uint32_t qmax,
                                                uint32_t* nonfinite) {
  k_morton_body(ph, offset, n_regions, region_cap, st, keys, vals, bmin, bscale, qmax, nonfinite, Blk{blockIdx.x, gridDim.x});
}

// --- LSD radix sort, 8-bit digits, count on the device ---------------------
constexpr int RS_THREADS = 256;
constexpr int RS_TILE = RS_THREADS * 8;

__device__ __forceinline__ void rs_range(uint32_t n, uint32_t G, uint32_t b, uint32_t& beg, uint32_t& end) {
  uint32_t per = (n + G - 1) / G;
  per = (per + RS_TILE - 1) / RS_TILE * RS_TILE;
  uint64_t bb = (uint64_t)b * per;
  uint64_t ee = bb + per;
  beg = bb < n ? (uint32_t)bb : n;
  end = ee < n ? (uint32_t)ee : n;
}

__device__ __forceinline__ void k_rs_hist_body(const uint32_t* keys, const uint32_t* np,
                                                         uint32_t* ghist, uint32_t shift, Blk blk) {
  __shared__ uint32_t h[256];
  const uint32_t G = blk.n, b = blk.b;
  h[threadIdx.x] = 0;
  __syncthreads();
  uint32_t beg, end;
  rs_range(*np, G, b, beg, end);
  for (uint32_t i = beg + threadIdx.x; i < end; i += RS_THREADS) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
  __syncthreads();
  ghist[threadIdx.x * G + b] = h[threadIdx.x];
}

__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const uint32_t* keys, const uint32_t* np,
                                                         uint32_t* ghist, uint32_t shift) {
  k_rs_hist_body(keys, np, ghist, shift, Blk{blockIdx.x, gridDim.x});
}

// One block per digit: exclusive scan of that digit's per-block counts.
__device__ __forceinline__ void k_rs_scan_body(uint32_t* ghist, uint32_t* totals, uint32_t G, Blk blk) {
  __shared__ uint32_t wsum[16];
  const uint32_t d = blk.b;
  uint32_t* row = ghist + (uint64_t)d * G;
  const uint32_t t = threadIdx.x;
  const uint32_t per = (G + 1023) / 1024;  // entries per thread
  uint32_t local[4] = {0, 0, 0, 0};
  uint32_t sum = 0;
  for (uint32_t k = 0; k < per && k < 4; ++k) {
    uint32_t i = t * per + k;
    local[k] = i < G ? row[i] : 0u;
    sum += local[k];
  }
  // block exclusive scan of `sum`
  uint32_t lane = t & 63u, w = t >> 6;
  uint32_t incl = sum;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t v = __shfl_up(incl, o);
    if (lane >= (uint32_t)o) incl += v;
  }
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  if (t == 0) {
    uint32_t acc = 0;
    for (int i = 0; i < 16; ++i) {
      uint32_t v = wsum[i];
      wsum[i] = acc;
      acc += v;
    }
    totals[d] = acc;
  }
  __syncthreads();
  uint32_t run = wsum[w] + incl - sum;
  for (uint32_t k = 0; k < per && k < 4; ++k) {
    uint32_t i = t * per + k;
    if (i < G) row[i] = run;
    run += local[k];
  }
}

__global__ __launch_bounds__(1024) void k_rs_scan(uint32_t* ghist, uint32_t* totals, uint32_t G) {
  k_rs_scan_body(ghist, totals, G, Blk{blockIdx.x, gridDim.x});
}

// Stable scatter.  Each round moves 256 keys: lanes with equal digits are
// matched with 8 ballots, ranked by mbcnt, and placed after the same digit's
// keys of earlier waves and earlier rounds.
__device__ __forceinline__ void k_rs_scatter_body(const uint32_t* kin, const uint32_t* vin, uint32_t* kout,
                                                            uint32_t* vout, const uint32_t* np,
                                                            const uint32_t* ghist, const uint32_t* totals,
                                                            uint32_t shift, Blk blk) {
  __shared__ uint32_t run_off[256];
  __shared__ uint32_t wcnt[RS_THREADS / 64][256];
  __shared__ uint32_t dig[256];
  const uint32_t G = blk.n, b = blk.b, t = threadIdx.x;
  const uint32_t lane = t & 63u, w = t >> 6;
  // exclusive scan of the digit totals
  dig[t] = totals[t];
  __syncthreads();
  if (t < 64) {
    uint32_t a0 = dig[t * 4], a1 = dig[t * 4 + 1], a2 = dig[t * 4 + 2], a3 = dig[t * 4 + 3];
    uint32_t s = a0 + a1 + a2 + a3;
    uint32_t incl = s;
    for (int o = 1; o < 64; o <<= 1) {
      uint32_t v = __shfl_up(incl, o);
      if (lane >= (uint32_t)o) incl += v;
    }
    uint32_t ex = incl - s;
    dig[t * 4] = ex;
    dig[t * 4 + 1] = ex + a0;
    dig[t * 4 + 2] = ex + a0 + a1;
    dig[t * 4 + 3] = ex + a0 + a1 + a2;
  }
  __syncthreads();
  run_off[t] = dig[t] + ghist[t * G + b];
  for (uint32_t k = 0; k < RS_THREADS / 64; ++k) wcnt[k][t] = 0;
  __syncthreads();
  uint32_t beg, end;
  rs_range(*np, G, b, beg, end);
  for (uint32_t base = beg; base < end; base += RS_THREADS) {
    uint32_t i = base + t;
    bool valid = i < end;
    uint32_t key = 0, val = 0, d = 0;
    if (valid) {
      key = kin[i];
      val = vin[i];
      d = (key >> shift) & 255u;
    }
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 8; ++bit) {
      uint64_t bb = __ballot(valid && ((d >> bit) & 1u));
      peers &= ((d >> bit) & 1u) ? bb : ~bb;
    }
    uint32_t rank = (uint32_t)__popcll(peers & lanemask_lt());
    if (valid && rank == 0) wcnt[w][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pre = 0;
      for (uint32_t k = 0; k < w; ++k) pre += wcnt[k][d];
      uint32_t pos = run_off[d] + pre + rank;
      kout[pos] = key;
      vout[pos] = val;
    }
    __syncthreads();
    uint32_t s = 0;
    for (uint32_t k = 0; k < RS_THREADS / 64; ++k) {
      s += wcnt[k][t];
      wcnt[k][t] = 0;
    }
    run_off[t] += s;
    __syncthreads();
  }
}

__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(const uint32_t* kin, const uint32_t* vin, uint32_t* kout,
                                                            uint32_t* vout, const uint32_t* np,
                                                            const uint32_t* ghist, const uint32_t* totals,
                                                            uint32_t shift) {
  k_rs_scatter_body(kin, vin, kout, vout, np, ghist, totals, shift, Blk{blockIdx.x, gridDim.x});
}

// --- one-sweep LSD radix sort (the session's per-pass sorts) ----------------
// Look-back word of (tile, digit): 0 = not yet published; else bit 31 set
// when the count is the inclusive prefix over tiles 0..tile, clear when it
// is the tile's own count, and bits 0-30 = that count + 1.
constexpr uint32_t OS_INCL = 1u << 31;
constexpr uint32_t OS_COUNT = OS_INCL - 1u;
constexpr uint32_t OS_SPIN_LIMIT = 1u << 22;  // waits of a look-back step before it gives up
constexpr uint32_t OS_ERR_FLAG = 16u;         // PassState::overflow bit of a sort that gave up
constexpr uint32_t OS_LOOK_STEP = 16;         // look-back words read at once

__device__ __forceinline__ uint32_t os_load(uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void os_store(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t wave_incl_sum(uint32_t v, uint32_t lane) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t u = __shfl_up(v, o);
    if (lane >= (uint32_t)o) v += u;
  }
  return v;
}

// The kernel that writes a sort's keys clears its ctl words first (block 0).
__device__ __forceinline__ void os_clear_ctl(uint32_t* ctl, Blk blk) {
  if (ctl && blk.b == 0)
    for (uint32_t i = threadIdx.x; i < OS_CTL_WORDS; i += blockDim.x) ctl[i] = 0u;
}

// Counts of every digit of the first n = min(*np, cap) keys into ctl, and
// the look-back words of those keys' tiles cleared.  Each block counts in
// LDS, lane l into copy l % OS_HIST_COPIES (copies OS_HIST_STRIDE words
// apart, so that a bin's copies sit in different banks): query keys arrive
// in screen tiles whose lanes mostly share their high digits, and 64 lanes
// adding to one word serialise (31 us against 12 us for the grid's hashed
// keys with one copy).  Then one global atomic add per nonzero bin.
constexpr uint32_t OS_HIST_COPIES = 16;
constexpr uint32_t OS_HIST_STRIDE = OS_MAX_DIGITS * 256u + 1u;
__device__ __forceinline__ void k_os_hist_body(const uint32_t* keys, const uint32_t* np, uint32_t cap,
                                               uint32_t digits, uint32_t* ctl, uint32_t* look, uint32_t tiles,
                                               Blk blk) {
  __shared__ uint32_t h[OS_HIST_COPIES * OS_HIST_STRIDE];
  const uint32_t t = threadIdx.x;
  for (uint32_t i = t; i < OS_HIST_COPIES * OS_HIST_STRIDE; i += blockDim.x) h[i] = 0u;
  __syncthreads();
  const uint32_t n = min(*np, cap);
  const uint32_t stride = blk.n * blockDim.x;
  uint32_t* const mine = h + (t % OS_HIST_COPIES) * OS_HIST_STRIDE;
  for (uint32_t i = blk.b * blockDim.x + t; i < n; i += stride) {
    const uint32_t k = keys[i];
    for (uint32_t p = 0; p < digits; ++p) atomicAdd(&mine[p * 256u + ((k >> (8u * p)) & 255u)], 1u);
  }
  __syncthreads();
  for (uint32_t i = t; i < digits * 256u; i += blockDim.x) {
    uint32_t c = 0;
    for (uint32_t r = 0; r < OS_HIST_COPIES; ++r) c += h[r * OS_HIST_STRIDE + i];
    if (c != 0u) atomicAdd(&ctl[i], c);
  }
  const uint32_t words = (n + OS_TILE - 1) / OS_TILE * 256u;
  for (uint32_t p = 0; p < digits; ++p) {
    uint32_t* lk = look + (uint64_t)p * tiles * 256u;
    for (uint32_t i = blk.b * blockDim.x + t; i < words; i += stride) lk[i] = 0u;
  }
}

__global__ __launch_bounds__(1024) void k_os_hist(const uint32_t* keys, const uint32_t* np, uint32_t cap,
                                                  uint32_t digits, uint32_t* ctl, uint32_t* look, uint32_t tiles) {
  k_os_hist_body(keys, np, cap, digits, ctl, look, tiles, Blk{blockIdx.x, gridDim.x});
}

// Photon records moved with the last digit pass of the grid's sort: out[i]
// = grid_rec(in[v_i]) for the sorted values v_i (indices into in), double
// positions with them when in64 is given, and the GRID_READ_SLACK records
// after the last one zeroed.
struct OsRecords {
  const PhotonRec* in = nullptr;
  const double* in64 = nullptr;
  PhotonRec* out = nullptr;
  double* out64 = nullptr;
};

// One digit pass: kin/vin -> kout/vout, stable (REC: the records too).
// Persistent blocks take tiles by ticket, so every tile a block waits on
// belongs to a block that is already running.  Threads t < 256 own digit t
// for the tile's counts and look-back.  hist = this digit's 256 counts;
// look = this digit's words.
template <bool REC>
__device__ __forceinline__ void k_os_scatter_body(const uint32_t* kin, const uint32_t* vin, uint32_t* kout,
                                                  uint32_t* vout, const uint32_t* np, uint32_t cap,
                                                  const uint32_t* hist, uint32_t* ticket, uint32_t* look,
                                                  uint32_t shift, uint32_t* err, OsRecords rec, Blk blk) {
  constexpr uint32_t W = OS_THREADS / 64, KPT = OS_KEYS_PER_THREAD;
  static_assert(OS_THREADS >= 256 && OS_THREADS % 256 == 0, "digit owners are threads 0..255");
  __shared__ uint32_t sk[OS_TILE];
  __shared__ uint32_t sv[OS_TILE];
  __shared__ uint32_t wc[W][256];  // per wave: keys of each digit so far, then the wave's offset in the tile
  __shared__ uint32_t dbase[256];  // first output position of each digit
  __shared__ uint32_t loc[256];    // first tile position of each digit
  __shared__ uint32_t gout[256];   // output position of tile position 0 of each digit's run
  __shared__ uint32_t wsum[4];
  __shared__ uint32_t s_tile;
  const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
  const bool owner = t < 256u;
  const uint32_t n = min(*np, cap);
  const uint32_t n_tiles = (n + OS_TILE - 1) / OS_TILE;
  if (REC && blk.b == 0 && t < GRID_READ_SLACK)
    rec.out[n + t] = PhotonRec{make_float4(0.f, 0.f, 0.f, 0.f), make_float4(0.f, 0.f, 0.f, 0.f)};
  // exclusive scan of 256 counts held by the owners (all threads call it)
  auto owner_scan = [&](uint32_t c) -> uint32_t {
    const uint32_t incl = wave_incl_sum(c, lane);
    if (owner && lane == 63u) wsum[w] = incl;
    __syncthreads();
    uint32_t pre = 0;
    for (uint32_t k = 0; k < w && k < 4u; ++k) pre += wsum[k];
    return pre + incl - c;
  };
  {
    const uint32_t e = owner_scan(owner ? hist[t] : 0u);
    if (owner) dbase[t] = e;
  }
  for (;;) {
    __syncthreads();  // the last tile's LDS is read; wsum is free
    if (t == 0) s_tile = atomicAdd(ticket, 1u);
    for (uint32_t i = t; i < W * 256u; i += OS_THREADS) (&wc[0][0])[i] = 0u;
    __syncthreads();
    const uint32_t tile = s_tile;
    if (tile >= n_tiles) return;
    const uint32_t base = tile * OS_TILE;
    const uint32_t tn = min(OS_TILE, n - base);
    // wave w ranks tile positions wb + 64 k + lane, k = 0..KPT-1, in order
    const uint32_t wb = w * 64u * KPT;
    uint32_t key[KPT], val[KPT], rk[KPT];
#pragma unroll
    for (uint32_t k = 0; k < KPT; ++k) {
      const uint32_t i = wb + k * 64u + lane;
      key[k] = i < tn ? kin[base + i] : 0u;
      val[k] = i < tn ? vin[base + i] : 0u;
    }
#pragma unroll
    for (uint32_t k = 0; k < KPT; ++k) {
      const bool valid = wb + k * 64u + lane < tn;
      const uint32_t d = (key[k] >> shift) & 255u;
      uint64_t peers = __ballot(valid);
#pragma unroll
      for (int bit = 0; bit < 8; ++bit) {
        const uint64_t bb = __ballot(valid && ((d >> bit) & 1u));
        peers &= ((d >> bit) & 1u) ? bb : ~bb;
      }
      const uint32_t before = (uint32_t)__popcll(peers & lanemask_lt());
      const uint32_t run = wc[w][d];
      rk[k] = run + before;
      if (valid && before == 0u) wc[w][d] = run + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    // digit owners: the tile's count, each wave's offset in it
    uint32_t c = 0;
    uint32_t* const lw = look + (uint64_t)tile * 256u + (owner ? t : 0u);
    if (owner) {
      for (uint32_t k = 0; k < W; ++k) {
        const uint32_t v = wc[k][t];
        wc[k][t] = c;
        c += v;
      }
      os_store(lw, (tile == 0u ? OS_INCL : 0u) | (c + 1u));
    }
    const uint32_t lc = owner_scan(c);
    if (owner) loc[t] = lc;
    __syncthreads();
    // the tile in digit order, in LDS
#pragma unroll
    for (uint32_t k = 0; k < KPT; ++k) {
      if (wb + k * 64u + lane < tn) {
        const uint32_t d = (key[k] >> shift) & 255u;
        const uint32_t lp = loc[d] + wc[w][d] + rk[k];
        sk[lp] = key[k];
        sv[lp] = val[k];
      }
    }
    // keys of digit t in tiles 0..tile-1: walk back until an inclusive
    // count, OS_LOOK_STEP tiles per step (the loads cross XCDs, each costing
    // about a microsecond)
    if (owner) {
      uint32_t ex = 0;
      if (tile > 0u) {
        uint32_t j = tile, spins = 0;  // tiles [0, j) are still to be counted
        for (;;) {
          uint32_t v[OS_LOOK_STEP];
#pragma unroll
          for (uint32_t q = 0; q < OS_LOOK_STEP; ++q)
            v[q] = q < j ? os_load(look + (uint64_t)(j - 1u - q) * 256u + t) : 0u;
          uint32_t used = 0;
          bool found = false;
#pragma unroll
          for (uint32_t q = 0; q < OS_LOOK_STEP; ++q) {
            if (found || used < q || q >= j || v[q] == 0u) continue;  // stop at the first word not yet published
            ex += (v[q] & OS_COUNT) - 1u;
            used = q + 1u;
            found = (v[q] & OS_INCL) != 0u;
          }
          if (found) break;
          j -= used;
          if (used == 0u) {
            if (++spins > OS_SPIN_LIMIT) {
              atomicOr(err, OS_ERR_FLAG);
              break;
            }
            __builtin_amdgcn_s_sleep(1);
          }
        }
        os_store(lw, OS_INCL | (ex + c + 1u));
      }
      gout[t] = dbase[t] + ex - loc[t];
    }
    __syncthreads();
    for (uint32_t i = t; i < tn; i += OS_THREADS) {
      const uint32_t k = sk[i];
      const uint32_t pos = gout[(k >> shift) & 255u] + i;
      if (pos < n) {  // always, unless a look-back gave up
        const uint32_t v = sv[i];
        kout[pos] = k;
        vout[pos] = v;
        if constexpr (REC) {
          rec.out[pos] = grid_rec(rec.in[v]);
          if (rec.in64) {
            rec.out64[3 * (uint64_t)pos + 0] = rec.in64[3 * (uint64_t)v + 0];
            rec.out64[3 * (uint64_t)pos + 1] = rec.in64[3 * (uint64_t)v + 1];
            rec.out64[3 * (uint64_t)pos + 2] = rec.in64[3 * (uint64_t)v + 2];
          }
        }
      }
    }
  }
}

template <bool REC>
__global__ __launch_bounds__(OS_THREADS) void k_os_scatter(const uint32_t* kin, const uint32_t* vin, uint32_t* kout,
                                                           uint32_t* vout, const uint32_t* np, uint32_t cap,
                                                           const uint32_t* hist, uint32_t* ticket, uint32_t* look,
                                                           uint32_t shift, uint32_t* err, OsRecords rec) {
  k_os_scatter_body<REC>(kin, vin, kout, vout, np, cap, hist, ticket, look, shift, err, rec,
                         Blk{blockIdx.x, gridDim.x});
}

// The sort of n = *np <= o.cap keys (written by a kernel that cleared
// o.ctl): digits 8-bit digit passes, the result in keys[*cur] / vals[*cur]
// (with rec, the last pass writes the values to last_vals instead and moves
// the records).
static hipError_t onesweep_sort(hipStream_t s, const OneSweep& o, uint32_t* keys[2], uint32_t* vals[2],
                                const uint32_t* np, uint32_t key_bits, uint32_t* err, int* cur,
                                const OsRecords* rec = nullptr, uint32_t* last_vals = nullptr) {
  const uint32_t digits = (key_bits + 7u) / 8u;
  if (!o.ctl || !o.look || digits > OS_MAX_DIGITS) return hipErrorInvalidValue;
  const uint32_t cap = (uint32_t)o.cap;
  const uint32_t hblocks = std::max<uint32_t>(1u, std::min<uint32_t>(OS_HIST_BLOCKS, (cap + 4095u) / 4096u));
  hipLaunchKernelGGL(k_os_hist, dim3(hblocks), dim3(1024), 0, s, keys[0], np, cap, digits, o.ctl, o.look, o.tiles);
  const uint32_t blocks = o.blocks;
  *cur = 0;
  if (rec && !last_vals) return hipErrorInvalidValue;
  for (uint32_t p = 0; p < digits; ++p) {
    const bool last = p + 1u == digits;
    if (rec && last)
      hipLaunchKernelGGL(k_os_scatter<true>, dim3(blocks), dim3(OS_THREADS), 0, s, keys[*cur], vals[*cur],
                         keys[*cur ^ 1], last_vals, np, cap, o.ctl + 256u * p, o.ctl + OS_TICKETS + 32u * p,
                         o.look + (uint64_t)p * o.tiles * 256u, 8u * p, err, *rec);
    else
      hipLaunchKernelGGL(k_os_scatter<false>, dim3(blocks), dim3(OS_THREADS), 0, s, keys[*cur], vals[*cur],
                         keys[*cur ^ 1], vals[*cur ^ 1], np, cap, o.ctl + 256u * p, o.ctl + OS_TICKETS + 32u * p,
                         o.look + (uint64_t)p * o.tiles * 256u, 8u * p, err, OsRecords{});
    *cur ^= 1;
  }
  return hipGetLastError();
}

// Leaves + tree.  The tree is a complete binary tree in heap order over
// lpad = 2^depth leaves of 64 consecutive photons (empty leaves have empty
// boxes).  k_build_leaves: one block per 64 leaves; each wave gathers 16
// leaves' photons into tree order (coalesced 32-byte writes), reduces their
// boxes with cross-lane shuffles, and the block folds its 64 boxes up six
// levels in LDS.  k_build_top: one block folds the remaining top levels.  No
// cross-workgroup synchronisation (an agent-scope acquire/release would
// write back and invalidate the L2 of every XCD).
__device__ __forceinline__ float wmin(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ float wmax(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}

__device__ __forceinline__ void tree_shape(uint32_t n, uint32_t& L, uint32_t& depth) {
  L = (n + LEAF - 1) / LEAF;
  depth = 0;
  while ((1u << depth) < L) ++depth;
}

__device__ __forceinline__ Node node_union(const Node& a, const Node& c) {
  Node u;
  u.lo = make_float4(fminf(a.lo.x, c.lo.x), fminf(a.lo.y, c.lo.y), fminf(a.lo.z, c.lo.z), 0.0f);
  u.hi = make_float4(fmaxf(a.hi.x, c.hi.x), fmaxf(a.hi.y, c.hi.y), fmaxf(a.hi.z, c.hi.z), 0.0f);
  return u;
}

constexpr uint32_t SUB_LEAVES = 64;  // leaves per block of k_build_leaves
constexpr uint32_t SUB_LEVELS = 6;   // log2(SUB_LEAVES)

__device__ __forceinline__ void k_build_leaves_body(const PhotonRec* in, const double* in64, const uint32_t* order,
                                                      PhotonRec* out, double* out64, uint32_t* src_idx, Node* nodes,
                                                      PassState* st, Blk blk) {
  __shared__ Node box[SUB_LEAVES];
  const uint32_t n = st->n_stored;
  uint32_t L, depth;
  tree_shape(n, L, depth);
  const uint32_t lpad = 1u << depth;
  const uint32_t groups = (lpad + SUB_LEAVES - 1) / SUB_LEAVES;
  const uint32_t lane = lane_id(), wib = threadIdx.x >> 6;
  if (blk.b == 0 && threadIdx.x == 0) {
    st->n_leaves = L;
    st->depth = depth;
  }
  for (uint32_t g = blk.b; g < groups; g += blk.n) {
    for (uint32_t k = 0; k < SUB_LEAVES / 4; ++k) {
      const uint32_t jl = wib * (SUB_LEAVES / 4) + k;  // leaf within the group
      const uint32_t j = g * SUB_LEAVES + jl;
      float lx = INF_F, ly = INF_F, lz = INF_F, hx = -INF_F, hy = -INF_F, hz = -INF_F;
      const uint32_t e = j * LEAF + lane;
      if (lane < (uint32_t)LEAF && j < L && e < n) {
        const uint32_t src = order[e];
        const PhotonRec r = in[src];
        out[e] = r;
        src_idx[e] = src;
        if (in64) {
          out64[3 * (uint64_t)e + 0] = in64[3 * (uint64_t)src + 0];
          out64[3 * (uint64_t)e + 1] = in64[3 * (uint64_t)src + 1];
          out64[3 * (uint64_t)e + 2] = in64[3 * (uint64_t)src + 2];
        }
        lx = hx = r.pm.x;
        ly = hy = r.pm.y;
        lz = hz = r.pm.z;
      }
      lx = wmin(lx); ly = wmin(ly); lz = wmin(lz);
      hx = wmax(hx); hy = wmax(hy); hz = wmax(hz);
      if (lane == 0) {
        Node nd;
        nd.lo = make_float4(lx, ly, lz, 0.0f);
        nd.hi = make_float4(hx, hy, hz, 0.0f);
        box[jl] = nd;
        if (j < lpad) nodes[lpad - 1 + j] = nd;
      }
    }
    __syncthreads();
    // fold up to six levels inside the group
    uint32_t cnt = min(SUB_LEAVES, lpad);
    for (uint32_t lev = 1; lev <= SUB_LEVELS && cnt > 1; ++lev) {
      cnt >>= 1;
      Node u;
      const uint32_t t = threadIdx.x;
      if (t < cnt) u = node_union(box[2 * t], box[2 * t + 1]);
      __syncthreads();
      if (t < cnt) {
        box[t] = u;
        const uint32_t level = depth - lev;
        nodes[((1u << level) - 1u) + ((g * SUB_LEAVES) >> lev) + t] = u;
      }
      __syncthreads();
    }
  }
}

__global__ __launch_bounds__(256) void k_build_leaves(const PhotonRec* in, const double* in64, const uint32_t* order,
                                                      PhotonRec* out, double* out64, uint32_t* src_idx, Node* nodes,
                                                      PassState* st) {
  k_build_leaves_body(in, in64, order, out, out64, src_idx, nodes, st, Blk{blockIdx.x, gridDim.x});
}

__device__ __forceinline__ void k_build_top_body(Node* nodes, const PassState* st) {
  uint32_t L, depth;
  tree_shape(st->n_stored, L, depth);
  if (depth <= SUB_LEVELS) return;
  for (int level = (int)depth - (int)SUB_LEVELS - 1; level >= 0; --level) {
    const uint32_t first = (1u << level) - 1u;
    for (uint32_t i = threadIdx.x; i < (1u << level); i += blockDim.x) {
      const uint32_t nd = first + i;
      nodes[nd] = node_union(nodes[2 * nd + 1], nodes[2 * nd + 2]);
    }
    __threadfence_block();
    __syncthreads();
  }
}
__global__ __launch_bounds__(1024) void k_build_top(Node* nodes, const PassState* st) { k_build_top_body(nodes, st); }

// LSD radix sort of (key, value) pairs with the count on the device; the
// result lands in keys[*cur] / vals[*cur].
static hipError_t radix_sort(hipStream_t s, uint32_t* keys[2], uint32_t* vals[2], const uint32_t* np, uint32_t* ghist,
                             uint32_t* totals, uint32_t blocks, uint32_t key_bits, int* cur) {
  *cur = 0;
  for (uint32_t shift = 0; shift < key_bits; shift += 8) {
    hipLaunchKernelGGL(k_rs_hist, dim3(blocks), dim3(RS_THREADS), 0, s, keys[*cur], np, ghist, shift);
    hipLaunchKernelGGL(k_rs_scan, dim3(256), dim3(1024), 0, s, ghist, totals, blocks);
    hipLaunchKernelGGL(k_rs_scatter, dim3(blocks), dim3(RS_THREADS), 0, s, keys[*cur], vals[*cur], keys[*cur ^ 1],
                       vals[*cur ^ 1], np, ghist, totals, shift);
    *cur ^= 1;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_map_build(hipStream_t s, MapBuf& m, PassState* st, const PhotonBuf& in, float4 bmin, float4 bmax) {
  const uint32_t qmax = (1u << m.morton_bits) - 1u;
  const float4 scale = morton_scale(bmin, bmax, qmax);
  uint32_t mblocks = (uint32_t)((m.cap + 255) / 256);
  if (mblocks > 8192) mblocks = 8192;
  if (mblocks == 0) mblocks = 1;
  hipLaunchKernelGGL(k_morton, dim3(mblocks), dim3(256), 0, s, in.rec, in.offset, in.n_regions, in.region_cap, st,
                     m.keys[0], m.vals[0], bmin, scale, qmax, &st->nonfinite);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  int cur = 0;
  e = radix_sort(s, m.keys, m.vals, &st->n_stored, m.ghist, m.totals, m.sort_blocks, 32u, &cur);
  if (e != hipSuccess) return e;
  uint32_t groups = (m.lpad_max + SUB_LEAVES - 1) / SUB_LEAVES;
  if (groups > 16384) groups = 16384;
  hipLaunchKernelGGL(k_build_leaves, dim3(groups), dim3(256), 0, s, in.rec, in.rec64, m.vals[cur], m.sorted,
                     m.sorted64, m.src_idx, m.nodes, st);
  hipLaunchKernelGGL(k_build_top, dim3(1), dim3(1024), 0, s, m.nodes, st);
  return hipGetLastError();
}

// Next work item (query group, tile) for this wave from a counter, one
// atomic per item: the persistent kernels take items until they run out, so
// costly ones (dense caustics, glass) do not leave other waves idle.
__device__ __forceinline__ uint32_t wave_next(uint32_t* counter) {
  uint32_t g = 0;
  if (lane_id() == 0) g = atomicAdd(counter, 1u);
  return __builtin_amdgcn_readfirstlane(g);
}

// Next of n_items work items for a wave from the 8 counters ctr[c][0] of a
// pass (counter c hands out items c, c + 8, c + 16, ...): the wave starts on
// counter `c` (its index mod 8) and moves to the next one when that runs
// out; n_items when all are out.  Atomics on one address are serialised, and
// with one counter the start of a persistent kernel, every wave asking at
// once, held the last waves up; eight counters on their own lines spread
// that out while items are still taken in about the counters' joint order
// (costly ones first).
__device__ __forceinline__ uint32_t wave_next8(uint32_t (*ctr)[32], uint32_t& c, uint32_t& tries, uint32_t n_items) {
  while (tries < 8u) {
    const uint32_t k = wave_next(&ctr[c][0]);
    const uint32_t i = c + 8u * k;
    if (i < n_items) return i;
    c = (c + 1u) & 7u;
    ++tries;
  }
  return n_items;
}

// Next item for a wave from 8 partitions of n_items: counter c hands out
// items [c per, (c + 1) per) in order; a wave starts on counter c and moves
// to the next when that one runs out (the ends of the partitions are shared
// by every wave, so an unequal split costs nothing but locality).  The shade
// starts each block on counter (block index mod 8): workgroups are dispatched
// to the 8 XCDs in turn, so each XCD works through one eighth of the
// Morton-ordered queries and its L2 keeps that region's photons.  Groups
// handed out in turn over all XCDs had every XCD fetch every region: 620 MB
// of HBM reads per caustic shade against 257 MB partitioned, at the same
// speed.
__device__ __forceinline__ uint32_t wave_next_part(uint32_t (*ctr)[32], uint32_t& c, uint32_t& tries, uint32_t n_items,
                                                   uint32_t per) {
  while (tries < 8u) {
    const uint32_t k = wave_next(&ctr[c][0]);
    const uint32_t i = c * per + k;
    if (k < per && i < n_items) return i;
    c = (c + 1u) & 7u;
    ++tries;
  }
  return n_items;
}

// ===========================================================================
// Photon grid of a pass.  Every gather query of a pass has the same radius r
// (util/iterator.rb:33-38), so the pass's photons are binned into cubic cells
// of side c >= GRID_CELL_FACTOR * reach, reach a little above r: the box of a
// query's ball then meets at most two cells per axis.  Cells are hashed
// ((x p1) xor (y p2) xor (z p3) with the primes of Teschner et al. 2003,
// "Optimized Spatial Hashing for Collision Detection of Deformable Objects")
// into a power-of-two table sized from the pass's photon count; the photons
// are sorted by bucket with the stable radix sort, and cells[b] holds the
// range [start, end) of bucket b in that order.  A bucket may also hold
// photons of other cells (collisions): they fail the distance test like any
// photon outside the ball, so a query finds exactly { j : dist2_nc(q, p_j) <=
// r^2 }, the set KdTree::within returns in float (tracer.rs:180).
//
// Coverage: the cell coordinate floor(clamp((v - bmin) * inv_c, -lim, lim))
// is monotonic in v, and a photon the float test accepts has |p - q| <= r (1
// + 2^-21) per axis; a query visits the cells of q -+ reach per axis, with
// reach = r (1 + 2^-10) + 2.4e-7 |q| on that axis covering the rounding of
// q -+ reach as well, so the photon's own cell is always among them.
//
// The cells are hashed, so the grid need not end at the scene's bounds:
// photons outside keep cells of their own out to GRID_FAR_CELLS from the
// bounds (about 10 km at centimetre cells), the clamp only keeping the
// integer conversion defined.  Clamping them to the layer of cells just
// outside the bounds piled every photon that left through an open side
// (infinite Plain walls store them kilometres away) into the cells next to
// the scene, which queries there then scanned: 2.6 ms of gather per pass
// on the caustic box with plane walls.

constexpr uint32_t GRID_P1 = 73856093u, GRID_P2 = 19349663u, GRID_P3 = 83492791u;
constexpr float GRID_CELL_FACTOR = 2.1f;  // cell side / (2 * reach) >= 1.05
constexpr float GRID_MAX_CELLS = 32768.0f;  // cells per axis over the bounds, at most
constexpr float GRID_REL_MARGIN = 2.4e-7f;  // per-point margin, times max |coordinate|
constexpr int32_t GRID_FAR_CELLS = 1 << 20;  // cell coordinates clamp to [-this, this]

struct GridView {
  float4 bmin;
  float inv_c, reach;
  uint32_t mask;
  int32_t lim;
};

__device__ __forceinline__ GridView grid_view(const PassState* st, float4 bmin) {
  GridView g;
  g.bmin = bmin;
  g.inv_c = st->g_inv_c;
  g.reach = st->g_reach;
  g.mask = st->g_mask;
  g.lim = st->g_lim;
  return g;
}

__device__ __forceinline__ int grid_coord(float v, float b, float inv_c, int32_t lim) {
  const float f = (v - b) * inv_c;
  return (int)floorf(fminf(fmaxf(f, -(float)lim), (float)lim));
}

__device__ __forceinline__ uint32_t grid_bucket(int x, int y, int z, uint32_t mask) {
  return (((uint32_t)x * GRID_P1) ^ ((uint32_t)y * GRID_P2) ^ ((uint32_t)z * GRID_P3)) & mask;
}

// One thread: the pass's cell side, reach, bucket count (next power of two
// >= stored photons, at least 1024, at most the table) and clamp.
__device__ __forceinline__ void k_grid_setup_body(PassState* st, float4 bmin, float4 bmax, uint32_t cells_max,
                                                  Blk blk) {
  if (threadIdx.x != 0 || blk.b != 0) return;
  const float r = sqrtf(st->r2f);
  const float reach = r * (1.0f + 1.0f / 1024.0f);
  const float mb = fmaxf(fmaxf(fmaxf(fabsf(bmin.x), fabsf(bmin.y)), fmaxf(fabsf(bmin.z), fabsf(bmax.x))),
                         fmaxf(fabsf(bmax.y), fabsf(bmax.z)));
  const float ext = fmaxf(fmaxf(bmax.x - bmin.x, bmax.y - bmin.y), fmaxf(bmax.z - bmin.z, 1e-30f));
  const float c = fmaxf(GRID_CELL_FACTOR * (reach + GRID_REL_MARGIN * mb), ext / GRID_MAX_CELLS);
  st->g_inv_c = 1.0f / c;
  st->g_reach = reach;
  st->g_lim = GRID_FAR_CELLS;
  uint32_t t = 1024u;
  while (t < st->n_stored && t < cells_max) t <<= 1;
  st->g_mask = min(t, cells_max) - 1u;
}

__global__ void k_grid_setup(PassState* st, float4 bmin, float4 bmax, uint32_t cells_max) {
  k_grid_setup_body(st, bmin, bmax, cells_max, Blk{blockIdx.x, gridDim.x});
}

__device__ __forceinline__ void k_grid_keys_body(const PhotonRec* ph, const uint32_t* offset, uint32_t n_regions,
                                                 uint32_t region_cap, const PassState* st, uint32_t* keys,
                                                 uint32_t* vals, float4 bmin, uint32_t* nonfinite, uint32_t* os_ctl,
                                                 Blk blk) {
  os_clear_ctl(os_ctl, blk);
  const uint32_t n = st->n_stored;
  const GridView g = grid_view(st, bmin);
  for (uint32_t i = blk.b * blockDim.x + threadIdx.x; i < n; i += blk.n * blockDim.x) {
    const uint64_t src = region_index(offset, n_regions, region_cap, i);
    const float4 p = ph[src].pm;
    if (!(isfinite(p.x) && isfinite(p.y) && isfinite(p.z))) atomicOr(nonfinite, 1u);
    keys[i] = grid_bucket(grid_coord(p.x, g.bmin.x, g.inv_c, g.lim), grid_coord(p.y, g.bmin.y, g.inv_c, g.lim),
                          grid_coord(p.z, g.bmin.z, g.inv_c, g.lim), g.mask);
    vals[i] = (uint32_t)src;
  }
}

__global__ __launch_bounds__(256) void k_grid_keys(const PhotonRec* ph, const uint32_t* offset, uint32_t n_regions,
                                                   uint32_t region_cap, const PassState* st, uint32_t* keys,
                                                   uint32_t* vals, float4 bmin, uint32_t* nonfinite, uint32_t* os_ctl) {
  k_grid_keys_body(ph, offset, n_regions, region_cap, st, keys, vals, bmin, nonfinite, os_ctl,
                   Blk{blockIdx.x, gridDim.x});
}

// Every bucket of the pass empty (a kernel, not a memset, so that a captured
// pass graph replays it faithfully).
__device__ __forceinline__ void k_grid_clear_body(uint2* cells, const PassState* st, Blk blk) {
  const uint32_t nb = st->g_mask + 1u;
  for (uint32_t b = blk.b * blockDim.x + threadIdx.x; b < nb; b += blk.n * blockDim.x) cells[b] = make_uint2(0u, 0u);
}

__global__ __launch_bounds__(256) void k_grid_clear(uint2* cells, const PassState* st) {
  k_grid_clear_body(cells, st, Blk{blockIdx.x, gridDim.x});
}

// Bucket ranges from the sorted keys: the first photon of a bucket writes its
// start, the last its end (separate 32-bit words).
__device__ __forceinline__ void k_grid_cells_body(const uint32_t* keys, const PassState* st, uint2* cells, Blk blk) {
  const uint32_t n = st->n_stored;
  for (uint32_t i = blk.b * blockDim.x + threadIdx.x; i < n; i += blk.n * blockDim.x) {
    const uint32_t b = keys[i];
    if (i == 0 || keys[i - 1] != b) cells[b].x = i;
    if (i + 1 == n || keys[i + 1] != b) cells[b].y = i + 1;
  }
}

__global__ __launch_bounds__(256) void k_grid_cells(const uint32_t* keys, const PassState* st, uint2* cells) {
  k_grid_cells_body(keys, st, cells, Blk{blockIdx.x, gridDim.x});
}

hipError_t launch_grid_build(hipStream_t s, MapBuf& m, PassState* st, const PhotonBuf& in, float4 bmin,
                             float4 bmax, uint64_t expect) {
  if (!m.cells || m.cells_max == 0 || m.cells_max > GRID_MAX_BUCKETS || (m.cells_max & (m.cells_max - 1)) != 0)
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_grid_setup, dim3(1), dim3(64), 0, s, st, bmin, bmax, m.cells_max);
  const uint64_t items = expect != 0 ? std::min<uint64_t>(expect, m.cap) : m.cap;
  const uint32_t kblocks = grid_stride_blocks(items);
  hipLaunchKernelGGL(k_grid_keys, dim3(kblocks), dim3(256), 0, s, in.rec, in.offset, in.n_regions, in.region_cap,
                     (const PassState*)st, m.keys[0], m.vals[0], bmin, &st->nonfinite, m.os.ctl);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  int cur = 0;
  // the last digit pass also moves the records into bucket order (the
  // source index of each sorted photon goes straight to src_idx)
  const OsRecords rec{in.rec, in.rec64, m.sorted, m.sorted64};
  e = onesweep_sort(s, m.os, m.keys, m.vals, &st->n_stored, GRID_KEY_BITS, &st->overflow, &cur, &rec, m.src_idx);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_grid_clear, dim3(grid_stride_blocks(std::min<uint64_t>(m.cells_max, next_pow2_u64(items)))),
                     dim3(256), 0, s, m.cells, (const PassState*)st);
  hipLaunchKernelGGL(k_grid_cells, dim3(kblocks), dim3(256), 0, s, m.keys[cur], (const PassState*)st, m.cells);
  return hipGetLastError();
}

// Every photon of the pass's grid within the ball (q, r2): on_photon(j, pm,
// dm, d2) for each sorted index j with dist2_nc(q, p_j) <= r2, each once.
// Buckets are visited in cell order (x fastest, then y, then z), photons in
// sorted order within a bucket, so a query's sum is formed in a fixed order.
// Per lane: no cross-lane operations, so inactive lanes may skip it.
template <typename OnPhoton>
__device__ __forceinline__ void grid_range_scan(const PhotonRec* ph, uint2 se, float qx, float qy, float qz, float r2,
                                                OnPhoton& on_photon) {
  uint32_t j = se.x;
  // four whole records per step: their loads are in flight together
  for (; j + 3 < se.y; j += 4) {
    const PhotonRec p0 = ph[j], p1 = ph[j + 1], p2 = ph[j + 2], p3 = ph[j + 3];
    const float d0 = dist2_nc(qx, qy, qz, p0.pm.x, p0.pm.y, p0.pm.z);
    const float d1 = dist2_nc(qx, qy, qz, p1.pm.x, p1.pm.y, p1.pm.z);
    const float d2 = dist2_nc(qx, qy, qz, p2.pm.x, p2.pm.y, p2.pm.z);
    const float d3 = dist2_nc(qx, qy, qz, p3.pm.x, p3.pm.y, p3.pm.z);
    if (d0 <= r2) on_photon(j, p0.pm, p0.dm, d0);
    if (d1 <= r2) on_photon(j + 1, p1.pm, p1.dm, d1);
    if (d2 <= r2) on_photon(j + 2, p2.pm, p2.dm, d2);
    if (d3 <= r2) on_photon(j + 3, p3.pm, p3.dm, d3);
  }
  for (; j < se.y; ++j) {
    const PhotonRec p0 = ph[j];
    const float d0 = dist2_nc(qx, qy, qz, p0.pm.x, p0.pm.y, p0.pm.z);
    if (d0 <= r2) on_photon(j, p0.pm, p0.dm, d0);
  }
}

// The same scan calling on_photon(j, pm, dm, d2, inside, valid) for every
// photon of the range, inside = (d2 <= r2), valid = true (the walk below also
// calls it for slots past a range's end, with both false): for a callback
// whose work is a select
// (the shade's sum adds 0 for photons outside), with no branch per photon
// and so none of the exec-mask work that kept the scalar units half busy.
template <typename OnPhoton>
__device__ __forceinline__ void grid_range_scan_all(const PhotonRec* ph, uint2 se, float qx, float qy, float qz,
                                                    float r2, OnPhoton& on_photon) {
  uint32_t j = se.x;
  for (; j + 3 < se.y; j += 4) {
    const PhotonRec p0 = ph[j], p1 = ph[j + 1], p2 = ph[j + 2], p3 = ph[j + 3];
    const float d0 = dist2_nc(qx, qy, qz, p0.pm.x, p0.pm.y, p0.pm.z);
    const float d1 = dist2_nc(qx, qy, qz, p1.pm.x, p1.pm.y, p1.pm.z);
    const float d2 = dist2_nc(qx, qy, qz, p2.pm.x, p2.pm.y, p2.pm.z);
    const float d3 = dist2_nc(qx, qy, qz, p3.pm.x, p3.pm.y, p3.pm.z);
    on_photon(j, p0.pm, p0.dm, d0, d0 <= r2, true);
    on_photon(j + 1, p1.pm, p1.dm, d1, d1 <= r2, true);
    on_photon(j + 2, p2.pm, p2.dm, d2, d2 <= r2, true);
    on_photon(j + 3, p3.pm, p3.dm, d3, d3 <= r2, true);
  }
  for (; j < se.y; ++j) {
    const PhotonRec p0 = ph[j];
    const float d0 = dist2_nc(qx, qy, qz, p0.pm.x, p0.pm.y, p0.pm.z);
    on_photon(j, p0.pm, p0.dm, d0, d0 <= r2, true);
  }
}

// The photons of a query's (at most) eight cell ranges se[0..7] walked in
// one loop, as one sequence: a lane moves on to its next range when it has
// used up the current one.  Walked range by range, a wave ran each cell's
// loop as long as that cell's longest lane, and the sum of those exceeds
// the lane with the most photons over all its cells: on the first pass of
// the caustic workload (a model of its photons and Morton-ordered queries)
// 361 against 275 four-photon steps per wave of 64 queries.  Photons are
// visited in the same order as range by range, so every sum is formed as
// before.  Four consecutive records per step, loaded together from one
// address; the slots of a step that runs past its range's end (into the
// next bucket's records, or the GRID_READ_SLACK zeroed ones after the last)
// are marked not valid.  ALL: on_photon(j, pm, dm, d2, inside, valid) for
// every slot, inside only when valid; otherwise on_photon(j, pm, dm, d2) for
// the valid photons inside the ball.
template <bool ALL, typename OnPhoton>
__device__ __forceinline__ void grid_walk8(const PhotonRec* ph, const uint2 (&se)[8], float qx, float qy, float qz,
                                           float r2, OnPhoton& on_photon) {
  uint32_t k = 0, j = se[0].x, e = se[0].y;
  while (true) {
    while (j >= e) {
      if (++k >= 8u) return;
      // range k (k differs between lanes: a chain of selects, not an index)
      uint2 n = se[1];
#pragma unroll
      for (uint32_t t = 2; t < 8; ++t) n = k == t ? se[t] : n;
      j = n.x;
      e = n.y;
    }
    const PhotonRec* p = ph + j;
    const PhotonRec p0 = p[0], p1 = p[1], p2 = p[2], p3 = p[3];
    const uint32_t i1 = j + 1u, i2 = j + 2u, i3 = j + 3u;
    const bool v1 = i1 < e, v2 = i2 < e, v3 = i3 < e;
    const float d0 = dist2_nc(qx, qy, qz, p0.pm.x, p0.pm.y, p0.pm.z);
    const float d1 = dist2_nc(qx, qy, qz, p1.pm.x, p1.pm.y, p1.pm.z);
    const float d2 = dist2_nc(qx, qy, qz, p2.pm.x, p2.pm.y, p2.pm.z);
    const float d3 = dist2_nc(qx, qy, qz, p3.pm.x, p3.pm.y, p3.pm.z);
    if constexpr (ALL) {
      on_photon(j, p0.pm, p0.dm, d0, d0 <= r2, true);
      on_photon(i1, p1.pm, p1.dm, d1, v1 && d1 <= r2, v1);
      on_photon(i2, p2.pm, p2.dm, d2, v2 && d2 <= r2, v2);
      on_photon(i3, p3.pm, p3.dm, d3, v3 && d3 <= r2, v3);
    } else {
      if (d0 <= r2) on_photon(j, p0.pm, p0.dm, d0);
      if (v1 && d1 <= r2) on_photon(i1, p1.pm, p1.dm, d1);
      if (v2 && d2 <= r2) on_photon(i2, p2.pm, p2.dm, d2);
      if (v3 && d3 <= r2) on_photon(i3, p3.pm, p3.dm, d3);
    }
    j += 4u;
  }
}

template <typename OnPhoton>
__device__ __forceinline__ void grid_bucket_scan(const uint2* cells, const PhotonRec* ph, uint32_t b, float qx,
                                                 float qy, float qz, float r2, OnPhoton& on_photon) {
  grid_range_scan(ph, cells[b], qx, qy, qz, r2, on_photon);
}

// Returns false (and scans only the first three cells of an axis) if the
// query's cells span more than three per axis, which the cell side rules out
// for every finite point (the caller raises a flag).
// ALL = true: on_photon(j, pm, dm, d2, inside, valid) for every photon of the
// cells, in the same order, and for padding slots with valid = false.
template <bool ALL = false, typename OnPhoton>
__device__ __forceinline__ bool grid_query(const GridView& g, const uint2* cells, const PhotonRec* ph, float qx,
                                           float qy, float qz, float r2, OnPhoton&& on_photon) {
  const float rx = g.reach + GRID_REL_MARGIN * fabsf(qx);
  const float ry = g.reach + GRID_REL_MARGIN * fabsf(qy);
  const float rz = g.reach + GRID_REL_MARGIN * fabsf(qz);
  const int x0 = grid_coord(qx - rx, g.bmin.x, g.inv_c, g.lim), x1 = grid_coord(qx + rx, g.bmin.x, g.inv_c, g.lim);
  const int y0 = grid_coord(qy - ry, g.bmin.y, g.inv_c, g.lim), y1 = grid_coord(qy + ry, g.bmin.y, g.inv_c, g.lim);
  const int z0 = grid_coord(qz - rz, g.bmin.z, g.inv_c, g.lim), z1 = grid_coord(qz + rz, g.bmin.z, g.inv_c, g.lim);
  if (x1 - x0 <= 1 && y1 - y0 <= 1 && z1 - z0 <= 1) {
    // at most 2 x 2 x 2 cells (the cell side makes this the rule): buckets in
    // registers, a bucket met twice (two cells hashed alike) scanned once
    uint32_t b[8];
    bool v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int x = x0 + (k & 1), y = y0 + ((k >> 1) & 1), z = z0 + (k >> 2);
      b[k] = grid_bucket(x, y, z, g.mask);
      v[k] = x <= x1 && y <= y1 && z <= z1;
#pragma unroll
      for (int j = 0; j < k; ++j) v[k] = v[k] && !(v[j] && b[j] == b[k]);
    }
    // the cells' ranges are loaded together before any is scanned (an
    // unused or repeated cell's range is empty)
    uint2 se[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) se[k] = v[k] ? cells[b[k]] : make_uint2(0u, 0u);
    grid_walk8<ALL>(ph, se, qx, qy, qz, r2, on_photon);
    return true;
  }
  // three cells on some axis (float rounding at the edge of the bound): the
  // same order, duplicates found by recomputing the earlier cells' buckets
  const bool ok = x1 - x0 <= 2 && y1 - y0 <= 2 && z1 - z0 <= 2;
  const int xe = min(x1, x0 + 2), ye = min(y1, y0 + 2), ze = min(z1, z0 + 2);
  for (int z = z0; z <= ze; ++z)
    for (int y = y0; y <= ye; ++y)
      for (int x = x0; x <= xe; ++x) {
        const uint32_t b = grid_bucket(x, y, z, g.mask);
        bool dup = false;
        for (int zz = z0; zz <= z; ++zz)
          for (int yy = y0; yy <= ye; ++yy)
            for (int xx = x0; xx <= xe; ++xx) {
              const bool earlier = zz < z || (zz == z && (yy < y || (yy == y && xx < x)));
              dup = dup || (earlier && grid_bucket(xx, yy, zz, g.mask) == b);
            }
        if (!dup) {
          if constexpr (ALL)
            grid_range_scan_all(ph, cells[b], qx, qy, qz, r2, on_photon);
          else
            grid_bucket_scan(cells, ph, b, qx, qy, qz, r2, on_photon);
        }
      }
  return ok;
}

// ===========================================================================
// Scene BVH, built on the device once per scene.  The reference tests every
// object for every ray (calc_intersection, tracer.rs:306-335); here the
// bounded shapes (spheres, polygons, parallelograms) smaller than a quarter
// of the scene sit in a heap-ordered
// complete binary tree over leaves of bvh_leaf shapes in Morton order of
// their centroids, the same implicit layout as the photon map: node i has
// children 2i+1 and 2i+2 and needs no pointers.  Planes and the large shapes
// (walls, whose boxes would widen every node above them) stay in the linear
// list.

// Box of a bounded shape, padded by 1e-5 of its coordinate magnitude so that
// the float slab test never rejects a ray that the shape's own float test
// reports as a hit (flat boxes of axis-aligned quads included).
__device__ __forceinline__ Node prim_box(const DevShape& s) {
  V3 lo, hi;
  const V3 p = mk(s.p);
  if (s.kind == SHAPE_SPHERE) {
    const float r = fabsf(s.radius);
    lo = p - mk(r, r, r);
    hi = p + mk(r, r, r);
  } else {
    const V3 a = p + mk(s.d1), b = p + mk(s.d2);
    lo = mk(fminf(p.x, fminf(a.x, b.x)), fminf(p.y, fminf(a.y, b.y)), fminf(p.z, fminf(a.z, b.z)));
    hi = mk(fmaxf(p.x, fmaxf(a.x, b.x)), fmaxf(p.y, fmaxf(a.y, b.y)), fmaxf(p.z, fmaxf(a.z, b.z)));
    if (s.kind == SHAPE_PARALLELOGRAM) {
      const V3 c = a + mk(s.d2);
      lo = mk(fminf(lo.x, c.x), fminf(lo.y, c.y), fminf(lo.z, c.z));
      hi = mk(fmaxf(hi.x, c.x), fmaxf(hi.y, c.y), fmaxf(hi.z, c.z));
    }
  }
  const float m = fmaxf(fmaxf(fmaxf(fabsf(lo.x), fabsf(lo.y)), fmaxf(fabsf(lo.z), fabsf(hi.x))),
                        fmaxf(fabsf(hi.y), fabsf(hi.z)));
  const float pad = 1e-5f * (1.0f + m);
  Node b;
  b.lo = make_float4(lo.x - pad, lo.y - pad, lo.z - pad, 0.0f);
  b.hi = make_float4(hi.x + pad, hi.y + pad, hi.z + pad, 0.0f);
  return b;
}

__device__ __forceinline__ Node empty_node() {
  Node b;
  b.lo = make_float4(BVH_EMPTY, BVH_EMPTY, BVH_EMPTY, 0.0f);
  b.hi = b.lo;
  return b;
}

// Per BVH shape i (scene index ids[i]): its padded box and the 30-bit
// Morton key of its centroid in the centroid bounds.
__global__ __launch_bounds__(256) void k_bvh_keys(const DevShape* shapes, const uint32_t* ids, uint32_t n,
                                                  float4 cmin, float4 cscale, uint32_t* keys, uint32_t* vals,
                                                  Node* pbox) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Node b = prim_box(shapes[ids[i]]);
  pbox[i] = b;
  const float cx = 0.5f * (b.lo.x + b.hi.x), cy = 0.5f * (b.lo.y + b.hi.y), cz = 0.5f * (b.lo.z + b.hi.z);
  const uint32_t qx = (uint32_t)fminf(fmaxf((cx - cmin.x) * cscale.x, 0.0f), 1023.0f);
  const uint32_t qy = (uint32_t)fminf(fmaxf((cy - cmin.y) * cscale.y, 0.0f), 1023.0f);
  const uint32_t qz = (uint32_t)fminf(fmaxf((cz - cmin.z) * cscale.z, 0.0f), 1023.0f);
  keys[i] = (expand_bits10(qx) << 2) | (expand_bits10(qy) << 1) | expand_bits10(qz);
  vals[i] = i;
}

// One thread per leaf slot: the leaf's shapes in sorted order into bprims,
// the union of their boxes into the leaf node (slots past the shapes get the
// empty node).
__global__ __launch_bounds__(256) void k_bvh_leaves(const DevShape* shapes, const uint32_t* ids,
                                                    const uint32_t* order, const Node* pbox, uint32_t n,
                                                    uint32_t depth, uint32_t per_leaf, DevShape* bprims, Node* nodes) {
  const uint32_t leaf = blockIdx.x * blockDim.x + threadIdx.x;
  if (leaf >= (1u << depth)) return;
  const uint32_t base = leaf * per_leaf;
  Node u = empty_node();
  bool any = false;
  for (uint32_t j = 0; j < per_leaf; ++j) {
    const uint32_t k = base + j;
    if (k >= n) break;
    const uint32_t pi = order[k];
    const Node b = pbox[pi];
    u = any ? node_union(u, b) : b;
    any = true;
    bprims[k] = shapes[ids[pi]];
  }
  nodes[(1u << depth) - 1u + leaf] = u;
}

// Nodes of one level: the union of the two children, skipping empty ones.
__global__ __launch_bounds__(256) void k_bvh_level(Node* nodes, uint32_t level) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (1u << level)) return;
  const uint32_t i = (1u << level) - 1u + k;
  const Node a = nodes[2u * i + 1u], b = nodes[2u * i + 2u];
  const bool ea = a.lo.x == BVH_EMPTY, eb = b.lo.x == BVH_EMPTY;
  nodes[i] = ea ? b : (eb ? a : node_union(a, b));
}

hipError_t launch_bvh_build(hipStream_t s, const DevShape* shapes, const uint32_t* ids, uint32_t n, float4 cmin,
                            float4 cscale, uint32_t depth, uint32_t leaf, uint32_t* keys[2], uint32_t* vals[2],
                            const uint32_t* np,
                            uint32_t* ghist, uint32_t* totals, uint32_t sort_blocks, Node* pbox, DevShape* bprims,
                            Node* nodes) {
  hipLaunchKernelGGL(k_bvh_keys, dim3((n + 255) / 256), dim3(256), 0, s, shapes, ids, n, cmin, cscale, keys[0],
                     vals[0], pbox);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  int cur = 0;
  e = radix_sort(s, keys, vals, np, ghist, totals, sort_blocks, 32u, &cur);
  if (e != hipSuccess) return e;
  const uint32_t slots = 1u << depth;
  hipLaunchKernelGGL(k_bvh_leaves, dim3((slots + 255) / 256), dim3(256), 0, s, shapes, ids, vals[cur], pbox, n,
                     depth, leaf, bprims, nodes);
  for (int level = (int)depth - 1; level >= 0; --level)
    hipLaunchKernelGGL(k_bvh_level, dim3(((1u << level) + 255) / 256), dim3(256), 0, s, nodes, (uint32_t)level);
  return hipGetLastError();
}

// ===========================================================================
// Eye rays (tracer.rs:129-177).  The reference recurses over a binary tree
// (reflection child + refraction child).  The BSDF is linear in the child
// radiances, so the tree is walked depth-first with a per-lane throughput:
// every node adds  throughput * (emittance/(2 pi) + wd * direct)  to its
// pixel and emits one gather query with weight throughput * wd; children
// whose weight is exactly zero in all channels contribute nothing and are
// not walked.  The gather itself runs in k_gather over all queries.

struct EyeNode {
  V3 o, d;
  C3 w;
  uint32_t meta;  // depth (5 bits) | medium (11 bits) << 5 | path (16 bits) << 16
};

// Order in which the eye waves take tiles: tiles whose queries (last pass's
// count; zero before the first pass) exceed TILE_HEAVY first, in any order,
// the others after them.  The order changes which wave walks a pixel, never
// its results (a pixel's queries have fixed slots).
constexpr uint32_t TILE_HEAVY = 96;
__device__ __forceinline__ void k_tile_order_body(EyeBuf eb, PassState* st, Blk blk) {
  const uint32_t t = blk.b * blockDim.x + threadIdx.x;
  if (t >= eb.n_tiles) return;
  if (eb.tile_count[t] > TILE_HEAVY)
    eb.tile_order[atomicAdd(&st->tiles_front, 1u)] = t;
  else
    eb.tile_order[eb.n_tiles - 1u - atomicAdd(&st->tiles_back, 1u)] = t;
}

__global__ __launch_bounds__(256) void k_tile_order(EyeBuf eb, PassState* st) {
  k_tile_order_body(eb, st, Blk{blockIdx.x, gridDim.x});
}

// Lanes of the calling wave with `pred` whose 6-bit `key` equals this lane's
// (6 ballots; called by the whole wave).
__device__ __forceinline__ uint64_t peer_mask(uint32_t key, bool pred) {
  uint64_t m = __ballot(pred);
#pragma unroll
  for (int b = 0; b < 6; ++b) {
    const uint64_t bb = __ballot(((key >> b) & 1u) != 0u);
    m &= ((key >> b) & 1u) ? bb : ~bb;
  }
  return m;
}

// Persistent eye kernel.  A wave takes tiles in the order of tile_order from
// the pass's work counters and walks the eye trees of a tile's 64 pixels (one per
// lane to start with) together: each loop step processes one node per busy
// lane; a node with both a reflection and a refraction child continues with
// the reflection and puts the refraction child on the wave's pending pool
// (LIFO, in global memory), and idle lanes take the newest pool entries, so
// the deep trees of a tile behind glass are spread over all 64 lanes.  A
// node's results depend only on the node (its random stream is keyed by
// pixel, depth and path); the tile's per-pixel state lives in LDS: the
// count of in-place deeper query slots used and the emitted radiance.  Slots
// and sums are assigned in lane order within a step, and a wave holds one
// tile at a time, so the schedule, and with it every slot and sum, depends
// only on the tile: results are reproducible.  Deeper queries beyond a
// pixel's sec_max slots go to the overflow region (atomic slots, added in
// any order).
// CL: the classic ray tracer's tree instead (trace_ray_classic,
// tracer.rs:221-259): the mirror direction itself for the reflection (no
// glossy draw), the refraction about the surface normal, the child's medium
// by the side tdir leaves to (tracer.rs:247), and the BSDF at cos1; each
// node's query then carries the direct light and the camera's ambient
// instead of a photon estimate (k_shade_classic).
template <int BV, bool CL = false, bool BVH = (BV != 0)>
__device__ __forceinline__ void k_eye_body(SceneView sc, PassState* st, EyeBuf eb, uint64_t seed, int uc,
                                             int maxd, uint32_t n_waves, Blk blk) {
  __shared__ uint32_t s_nsec[4 * 64];  // per wave of the block and pixel of its tile
  __shared__ float s_acc[4 * 64 * 3];
  scene_to_lds<BV>(sc);
  const uint32_t wb = threadIdx.x >> 6;
  const uint32_t w = blk.b * (blockDim.x / 64) + wb;
  if (w >= n_waves) return;
  uint32_t* const nsec = s_nsec + wb * 64u;
  float* const acc3 = s_acc + wb * 192u;
  const uint32_t lane = lane_id();
  const uint32_t pass = st->pass;
  // draws of camera_ray on the pixel's root stream, which node 0 continues
  const uint32_t cam_draws = (sc.cam.blur ? 2u : 0u) + (sc.cam.progressive && sc.cam.antialias ? 2u : 0u);
  const uint32_t cap = 64u * eb.estack_depth;  // pool entries of this wave
  float4* const pool = eb.estack + (uint64_t)w * cap * 3u;
  const uint32_t tx = (uint32_t)sc.cam.tiles_x;
  const uint32_t ts = (uint32_t)sc.cam.tile_shift, tm = (1u << ts) - 1u;  // tile width 1 << ts
  uint32_t wc = w & 7u, wtries = 0;
  for (uint32_t k = wave_next8(st->work[WORK_EYE], wc, wtries, eb.n_tiles); k < eb.n_tiles;
       k = wave_next8(st->work[WORK_EYE], wc, wtries, eb.n_tiles)) {
    const uint32_t tile = eb.tile_order[k];
    const uint32_t x0 = (tile % tx) << ts, y0 = (tile / tx) << (6u - ts);
    const uint32_t t = tile * 64u + lane;
    const int x = (int)(x0 + (lane & tm)), y = (int)(y0 + (lane >> ts));
    const bool inside = x < sc.cam.xreso && y < sc.cam.yreso;
    const uint32_t my_pixel = inside ? (uint32_t)y * (uint32_t)sc.cam.xreso + (uint32_t)x : 0u;
    nsec[lane] = 0u;
    acc3[3u * lane] = 0.0f;
    acc3[3u * lane + 1u] = 0.0f;
    acc3[3u * lane + 2u] = 0.0f;
    bool alive = false, prim = false;
    uint32_t own = lane;  // the pixel (position in the tile) of the lane's node
    EyeNode cur;
    if (inside) {
      Rng rng;
      rng.init(seed, my_pixel, pass, TAG_EYE);
      camera_ray(sc.cam, x, y, rng, cur.o, cur.d);
      cur.w = C3{1.0f, 1.0f, 1.0f};
      cur.meta = (sc.n_mats << 5);
      alive = true;
    }
    uint32_t top = 0;  // wave-uniform: pool entries
    while (true) {
      // idle lanes take the newest pool entries, in lane order
      const uint64_t idle = __ballot(!alive);
      const uint32_t take = min((uint32_t)__popcll(idle), top);
      if (take) {
        if (!alive) {
          const uint32_t r = (uint32_t)__popcll(idle & lanemask_lt());
          if (r < take) {
            const float4* e = pool + (uint64_t)(top - 1u - r) * 3u;
            const float4 a = e[0], b = e[1], c = e[2];
            cur.o = mk(a.x, a.y, a.z);
            cur.meta = __float_as_uint(a.w);
            cur.d = mk(b.x, b.y, b.z);
            cur.w = C3{b.w, c.x, c.y};
            own = __float_as_uint(c.z);
            alive = true;
          }
        }
        top -= take;
      }
      if (!__any(alive)) break;
      const uint32_t px = (uint32_t)y0 + (own >> ts), qx = x0 + (own & tm);
      const uint32_t pixel = px * (uint32_t)sc.cam.xreso + qx;
      const uint32_t to = tile * 64u + own;  // the pixel's primary slot
      bool next_set = false, push = false, want_q = false;
      EyeNode nxt, pc;
      C3 emit{0.0f, 0.0f, 0.0f};
      Query qq;
      const uint32_t depth = cur.meta & 31u;
      if (alive) {
        const uint32_t med = (cur.meta >> 5) & 2047u;
        const uint32_t path = cur.meta >> 16;
        Isect is;
        if ((int)depth < maxd && intersect<BV>(sc, cur.o, cur.d, is)) {
          const DevMaterial M = material(sc, is.mat);
          const DevMaterial Mm = material(sc, med);
          emit = cur.w * (c3(M.em) * SR_HALF_F);
          float cos1;
          V3 rdir = specular_reflection(is.n, cur.d, cos1);
          V3 tdir = mk(0.0f, 0.0f, 0.0f);
          bool has_t = false;
          float cos2 = 0.0f;
          if constexpr (CL) {
            // tracer.rs:232,246
            has_t = specular_refraction(is.n, cur.d, relative_ior(Mm.ior_avg, M.ior_avg), tdir, cos2);
          } else {
            const bool spec_part = (M.surface == SURF_SIMPLE && M.diff != 1.0f) || M.surface == SURF_TS;
            if (spec_part) {
              // tracer.rs:150-161
              Rng nr;
              nr.init(seed, pixel, pass, TAG_EYE | (depth << 16) | path);
              if (depth == 0)
                for (uint32_t j = 0; j < cam_draws; ++j) (void)nr.next();
              rdir = reflection_glossy(is.n, rdir, power_glossy(M), nr);
              const float eta = relative_ior(Mm.ior_avg, M.ior_avg);
              const V3 hvec = normalize_or(rdir - cur.d, is.n);
              has_t = specular_refraction(hvec, cur.d, eta, tdir, cos2);
            }
          }
          const float cos0 = CL ? cos1 : (cos1 < cos2 ? cos1 : cos2);  // tracer.rs:173, :257
          const BsdfW W = bsdf_weights(M, cos0);
          const C3 wd = cur.w * W.wd;
          if (nonzero(wd)) {
            // di = direct light (when uc) + photon estimate (tracer.rs:136-145):
            // both are evaluated per query by k_shade
            want_q = true;
            qq.a = make_float4(is.pos.x, is.pos.y, is.pos.z, __uint_as_float(pixel));
            qq.b = make_float4(is.n.x, is.n.y, is.n.z, __uint_as_float(is.mat));
            qq.w = make_float4(wd.r, wd.g, wd.b, 0.0f);
          }
          if ((int)depth + 1 < maxd) {
            const C3 ws = cur.w * W.ws;
            const C3 wt = cur.w * W.wt;
            const bool rc = surf_reflect(M, cos1) && nonzero(ws);
            const bool tc = has_t && surf_refract(M, cos1) && nonzero(wt);
            if (tc) {
              // tracer.rs:164-167; classic: tracer.rs:247
              const uint32_t tmed = CL ? (dot(tdir, is.n) < 0.0f ? is.mat : sc.n_mats)
                                       : (is.inside_to_out ? sc.n_mats : is.mat);
              const EyeNode c{is.pos, tdir, wt, (depth + 1) | (tmed << 5) | ((path | (1u << depth)) << 16)};
              if (rc) {
                pc = c;
                push = true;
              } else {
                nxt = c;
                next_set = true;
              }
            }
            if (rc) {
              nxt = EyeNode{is.pos, rdir, ws, (depth + 1) | (med << 5) | (path << 16)};
              next_set = true;
            }
          }
        }
      }
      // emitted radiance: each pixel's contributions of this step, in lane
      // order, added by the pixel's lowest lane to the pixel's sum
      const bool has_e = alive && nonzero(emit);
      if (__any(has_e)) {
        const uint64_t pe = peer_mask(own, has_e);
        const bool lead = has_e && (pe & lanemask_lt()) == 0u;
        uint64_t rest = pe;
        float ar = 0.0f, ag = 0.0f, ab = 0.0f;
        if (lead) {
          ar = acc3[3u * own];
          ag = acc3[3u * own + 1u];
          ab = acc3[3u * own + 2u];
        }
        while (__any(lead && rest != 0u)) {
          const bool go = lead && rest != 0u;
          const int src = go ? __ffsll((unsigned long long)rest) - 1 : (int)lane;
          const float er = __shfl(emit.r, src), eg = __shfl(emit.g, src), eb_ = __shfl(emit.b, src);
          if (go) {
            ar += er;
            ag += eg;
            ab += eb_;
            rest &= rest - 1u;
          }
        }
        if (lead) {
          acc3[3u * own] = ar;
          acc3[3u * own + 1u] = ag;
          acc3[3u * own + 2u] = ab;
        }
      }
      // query slots: the root's is the pixel's primary slot; deeper ones take
      // the pixel's next in-place slots in lane order, then the overflow region
      const bool need = want_q && depth > 0;
      uint32_t slot = want_q ? to : 0xFFFFFFFFu;
      if (__any(need)) {
        const uint64_t pq = peer_mask(own, need);
        const uint32_t rank = (uint32_t)__popcll(pq & lanemask_lt());
        const uint32_t base = need ? nsec[own] : 0u;
        __asm__ volatile("" ::: "memory");  // every lane reads before the leader writes
        if (need) {
          const uint32_t idx = base + rank;
          if (idx < eb.sec_max) {
            slot = eb.q_primary * (1u + idx) + to;
          } else {
            const uint32_t j = atomicAdd(&st->n_queries, 1u);
            slot = j < eb.ovf_cap ? eb.q_primary * (1u + eb.sec_max) + j : 0xFFFFFFFFu;
            if (j >= eb.ovf_cap) atomicOr(&st->overflow, 2u);
          }
          if (rank == 0u) nsec[own] = min(base + (uint32_t)__popcll(pq), eb.sec_max);
        }
      }
      if (want_q && slot != 0xFFFFFFFFu) {
        eb.q[slot] = qq;
        eb.qpos[slot] = qq.a;
      }
      prim = prim || (want_q && depth == 0u);  // a root is its own lane's first node
      // pending refraction children onto the pool, in lane order
      const uint64_t pm = __ballot(push);
      if (pm) {
        if (push) {
          const uint32_t e = top + (uint32_t)__popcll(pm & lanemask_lt());
          if (e < cap) {
            float4* q = pool + (uint64_t)e * 3u;
            q[0] = make_float4(pc.o.x, pc.o.y, pc.o.z, __uint_as_float(pc.meta));
            q[1] = make_float4(pc.d.x, pc.d.y, pc.d.z, pc.w.r);
            q[2] = make_float4(pc.w.g, pc.w.b, __uint_as_float(own), 0.0f);
          } else {
            atomicOr(&st->overflow, 8u);
          }
        }
        top = min(top + (uint32_t)__popcll(pm), cap);
      }
      if (next_set)
        cur = nxt;
      else
        alive = false;
    }
    // the tile is done: each lane writes its pixel's emitted radiance and
    // in-place query count
    if (inside) eb.pix[my_pixel] = make_float4(acc3[3u * lane], acc3[3u * lane + 1u], acc3[3u * lane + 2u], 0.0f);
    eb.sec_count[t] = inside ? nsec[lane] | (prim ? SEC_PRIMARY : 0u) : 0u;
  }
}

template <int BV, bool BVH = (BV != 0)>
__global__ __launch_bounds__(256) void k_eye(SceneView sc, PassState* st, EyeBuf eb, uint64_t seed, int uc,
                                             int maxd, uint32_t n_waves) {
  k_eye_body<BV>(sc, st, eb, seed, uc, maxd, n_waves, Blk{blockIdx.x, gridDim.x});
}
template <int BV>
__global__ __launch_bounds__(256) void k_eye_classic(SceneView sc, PassState* st, EyeBuf eb, uint64_t seed,
                                                     uint32_t n_waves) {
  k_eye_body<BV, true>(sc, st, eb, seed, 1, CLASSIC_DEPTH, n_waves, Blk{blockIdx.x, gridDim.x});
}

// Queries kept in place per tile (primary + deeper), for the sort's offsets.
__device__ __forceinline__ void k_tile_count_body(EyeBuf eb, Blk blk) {
  const uint32_t t = blk.b * blockDim.x + threadIdx.x;
  const uint32_t tile = t >> 6;
  uint32_t c = 0;
  if (t < eb.q_primary) {
    const uint32_t sc = eb.sec_count[t];
    c = (sc >> 31) + (sc & ~SEC_PRIMARY);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if ((threadIdx.x & 63u) == 0 && tile < eb.n_tiles) eb.tile_count[tile] = c;
}

__global__ __launch_bounds__(256) void k_tile_count(EyeBuf eb) {
  k_tile_count_body(eb, Blk{blockIdx.x, gridDim.x});
}

hipError_t launch_eye(hipStream_t s, const SceneBuf& sc, PassState* st, const EyeBuf& eb, uint64_t seed,
                      int use_classic, int max_depth, int n_cu) {
  hipLaunchKernelGGL(k_tile_order, dim3((eb.n_tiles + 255) / 256), dim3(256), 0, s, eb, st);
  // 6 persistent waves per CU: the eye's 123 registers allow 4 waves per
  // SIMD, and a smaller grid leaves room beside it for the photon branch
  // and the other pipeline's kernels (16 per CU filled every slot: caustic
  // 0.855 -> 0.816 ms per pass, mesh 95.6 -> 87.8 ms; 4 and 8: 0.815-0.834
  // and 90.6-94.5 ms)
  uint32_t n_waves = std::min<uint32_t>(std::min<uint32_t>(eb.n_tiles, (uint32_t)n_cu * EYE_WAVES_PER_CU), eb.estack_lanes / 64u);
  if (n_waves < 1) n_waves = 1;
  const uint32_t blocks = (n_waves + 3) / 4;
  switch (bvh_kind(sc)) {
    case 2:
      hipLaunchKernelGGL(k_eye<2>, dim3(blocks), dim3(256), bvh_dyn_lds(sc), s, make_view(sc), st, eb, seed,
                       use_classic,
                       max_depth, n_waves);
      break;
    case 1:
      hipLaunchKernelGGL(k_eye<1>, dim3(blocks), dim3(256), 0, s, make_view(sc), st, eb, seed,
                       use_classic,
                       max_depth, n_waves);
      break;
    default:
      hipLaunchKernelGGL(k_eye<0>, dim3(blocks), dim3(256), 0, s, make_view(sc), st, eb, seed, use_classic,
                       max_depth, n_waves);
  }
  hipLaunchKernelGGL(k_tile_count, dim3((eb.q_primary + 255) / 256), dim3(256), 0, s, eb);
  return hipGetLastError();
}

// ===========================================================================
// Wave-coherent radius query over the heap-ordered tree.  Stackless DFS:
// descend to the left child while any lane's ball overlaps the node box;
// otherwise step to the next sibling, climbing while at a right child.  At a
// leaf, the 64 photons are staged once into this wave's LDS slice with
// coalesced loads and every lane tests them from LDS (broadcast reads).

struct LeafLds {
  float4 pm[LEAF];
  float4 dm[LEAF];
};

template <typename OnPhoton>
__device__ __forceinline__ void wave_query(const Node* nodes, const PhotonRec* ph, uint32_t n, uint32_t depth,
                                           bool active, float qx, float qy, float qz, float r2box, LeafLds& lds,
                                           OnPhoton&& on_photon) {
  const uint32_t lane = lane_id();
  const uint32_t first_leaf = (1u << depth) - 1u;
  uint32_t node = 0;
  if (n == 0) return;
  while (true) {
    node = __builtin_amdgcn_readfirstlane(node);
    const Node nd = ldc(nodes, node);
    const bool ov = active && box_dist2_nc(qx, qy, qz, nd.lo, nd.hi) <= r2box;
    if (__ballot(ov)) {
      if (node < first_leaf) {
        node = 2 * node + 1;
        continue;
      }
      const uint32_t base = (node - first_leaf) * LEAF;
      const uint32_t cnt = base < n ? min((uint32_t)LEAF, n - base) : 0u;
      if (lane < cnt) {
        const PhotonRec r = ph[base + lane];
        lds.pm[lane] = r.pm;
        lds.dm[lane] = r.dm;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (ov) {
        for (uint32_t j = 0; j < cnt; ++j) on_photon(base + j, lds.pm[j], lds.dm[j]);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    while (node != 0 && (node & 1u) == 0) node = (node - 1) >> 1;
    if (node == 0) break;
    node += 1;
  }
}

// Gather of one query: Σ filter * max(0, -n.dir) per channel over the ball
// (estimate_radiance, tracer.rs:179-195; photon_to_radiance, optics.rs:224-233).
struct GatherAcc {
  float r, g, b;
};

__device__ __forceinline__ GatherAcc gather_query(const Node* nodes, const PhotonRec* ph, uint32_t n, uint32_t depth,
                                                  bool active, const Query& qq, float r2, int pf, LeafLds& lds) {
  GatherAcc a{0.0f, 0.0f, 0.0f};
  const float qx = qq.a.x, qy = qq.a.y, qz = qq.a.z;
  const float nx = qq.b.x, ny = qq.b.y, nz = qq.b.z;
  const float inv_r2 = 1.0f / r2;
  wave_query(nodes, ph, n, depth, active, qx, qy, qz, r2, lds, [&](uint32_t, float4 pm, float4 dm) {
    const float d2 = dist2_nc(qx, qy, qz, pm.x, pm.y, pm.z);
    if (d2 <= r2) {
      const float c = nx * dm.x + ny * dm.y + nz * dm.z;
      const float v = c < 0.0f ? -c * filter_weight(pf, d2 * inv_r2) * dm.w : 0.0f;
      const uint32_t wl = __float_as_uint(pm.w) & 3u;
      a.r += wl == 0 ? v : 0.0f;
      a.g += wl == 1 ? v : 0.0f;
      a.b += wl == 2 ? v : 0.0f;
    }
  });
  return a;
}

// Sort keys of the queries: one wave per tile; lane i writes its pixel's
// in-place queries (primary, then deeper ones) at the tile's offset plus the
// prefix of the earlier lanes' counts.  Threads past the tiles key the
// overflow queries, which follow all in-place ones.  Key = Morton code of
// the position.
__device__ __forceinline__ void k_query_keys_body(EyeBuf eb, const PassState* st, float4 bmin, float4 bscale,
                                                    float fq, uint32_t lshift, uint32_t* keys, uint32_t* vals,
                                                    uint32_t* os_ctl, Blk blk) {
  os_clear_ctl(os_ctl, blk);
  const uint32_t t = blk.b * blockDim.x + threadIdx.x;
  const uint32_t g = t >> 6, lane = lane_id();
  const uint64_t ovf_base = (uint64_t)eb.q_primary * (1u + eb.sec_max);
  if (g >= eb.n_tiles) {
    // the threads past the tiles key the overflow queries, striding
    const uint32_t n_ovf = min(st->n_queries, eb.ovf_cap);
    const uint32_t stride = blk.n * blockDim.x - eb.n_tiles * 64u;
    for (uint32_t i = t - eb.n_tiles * 64u; i < n_ovf; i += stride) {
      const uint32_t pos = st->n_qsort - n_ovf + i;
      const float4 a = eb.qpos[ovf_base + i];
      keys[pos] = morton_key(a.x, a.y, a.z, bmin, bscale, fq, lshift);
      vals[pos] = (uint32_t)(ovf_base + i);
    }
    return;
  }
  const uint32_t sc = eb.sec_count[t];
  const bool primary = (sc & SEC_PRIMARY) != 0u;
  const uint32_t nsec = sc & ~SEC_PRIMARY;
  const uint32_t c = (primary ? 1u : 0u) + nsec;
  uint32_t incl = c;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t v = __shfl_up(incl, o);
    if (lane >= (uint32_t)o) incl += v;
  }
  uint32_t pos = eb.tile_off[g] + incl - c;
  // slot q_primary * k + t for k = 0 (the primary, when there is one) ..
  // nsec, four positions loaded at once (a lane behind glass has a dozen)
  for (uint32_t k0 = primary ? 0u : 1u; k0 <= nsec; k0 += 4u) {
    float4 a[4];
#pragma unroll
    for (uint32_t j = 0; j < 4u; ++j)
      if (k0 + j <= nsec) a[j] = eb.qpos[eb.q_primary * (k0 + j) + t];
#pragma unroll
    for (uint32_t j = 0; j < 4u; ++j) {
      if (k0 + j <= nsec) {
        keys[pos] = morton_key(a[j].x, a[j].y, a[j].z, bmin, bscale, fq, lshift);
        vals[pos] = eb.q_primary * (k0 + j) + t;
        ++pos;
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_query_keys(EyeBuf eb, const PassState* st, float4 bmin, float4 bscale,
                                                    float fq, uint32_t lshift, uint32_t* keys, uint32_t* vals,
                                                    uint32_t* os_ctl) {
  k_query_keys_body(eb, st, bmin, bscale, fq, lshift, keys, vals, os_ctl, Blk{blockIdx.x, gridDim.x});
}

// Shading of the queries in Morton order, one query per lane, 64 spatially
// adjacent queries per wave (their photons mostly share cache lines); waves
// take groups of 64 from the pass's work counters, each XCD from its own
// eighth of the groups first (wave_next_part).  Each lane gathers over the cells
// of the pass's photon grid that its ball meets:
//   res[slot] = w * (scale * sum filter * max(0,-n.dir) + direct light (uc))
// per channel: the di of tracer.rs:136-145 weighted by the node's diffuse
// BSDF factor and path throughput.  Gather and shadow rays share the kernel
// so that the waves of one hide the memory latency of the other.  A query
// whose cells exceed the grid's span rule raises flag 4 of the pass state.
// The gather of one query: the summed -n.dir * filter * weight per channel
// into ar, ag, ab.  PF >= 0 fixes the photon filter at compile time (no test
// of it in the per-photon loop, and only that filter's registers); PF = -1
// takes it from pf (batches, whose scenes may differ).
template <int PF>
__device__ __forceinline__ bool shade_gather(const GridView& g, const uint2* cells, const PhotonRec* ph, const Query& qq,
                                             float r2, float inv_r2, int pf, float& ar, float& ag, float& ab) {
  const float nx = qq.b.x, ny = qq.b.y, nz = qq.b.z;
  const int f = PF >= 0 ? PF : pf;
  // grid records (grid_rec): the direction carries the weight, pm.w and dm.w
  // flag channels 0 and 1; v * 1 and v * 0 are exact, so each channel's sum
  // is formed as with a select
  return grid_query<true>(g, cells, ph, qq.a.x, qq.a.y, qq.a.z, r2,
                          [&](uint32_t, float4 pm, float4 dm, float d2, bool in, bool) {
    const float c = nx * dm.x + ny * dm.y + nz * dm.z;  // optics.rs:225
    const float v = (in && c < 0.0f) ? -c * filter_weight(f, d2 * inv_r2) : 0.0f;
    ar = fmaf(v, pm.w, ar);
    ag = fmaf(v, dm.w, ag);
    ab = fmaf(v, 1.0f - pm.w - dm.w, ab);
  });
}

template <int BV, int PF, bool BVH = (BV != 0)>
__device__ __forceinline__ void k_shade_body(SceneView sc, const uint2* cells, const PhotonRec* ph, PassState* st,
                                               EyeBuf eb, const uint32_t* order, int pf, int uc, Blk blk) {
  scene_to_lds<BV>(sc);
  const float r2 = st->r2f, scale = st->scale, inv_r2 = 1.0f / r2;
  const uint32_t nq = st->n_qsort;
  const GridView g = grid_view(st, sc.bmin);
  const uint32_t groups = (nq + 63) / 64;
  bool complete = true;
  uint32_t wc = blk.b & 7u, wtries = 0;  // the block's XCD under in-turn dispatch
  const uint32_t per = (groups + 7u) / 8u;
  for (uint32_t gi = wave_next_part(st->work[WORK_SHADE], wc, wtries, groups, per); gi < groups;
       gi = wave_next_part(st->work[WORK_SHADE], wc, wtries, groups, per)) {
    const uint32_t i = gi * 64 + lane_id();
    if (i >= nq) continue;
    const uint32_t slot = order[i];
    const Query qq = eb.q[slot];
    float ar = 0.0f, ag = 0.0f, ab = 0.0f;
    complete &= shade_gather<PF>(g, cells, ph, qq, r2, inv_r2, pf, ar, ag, ab);
    C3 di = C3{ar * scale, ag * scale, ab * scale};
    if (uc) di = di + direct_light<BV>(sc, mk(qq.a), mk(qq.b));
    eb.res[slot] = make_float4(qq.w.x * di.r, qq.w.y * di.g, qq.w.z * di.b, 0.0f);
  }
  if (!complete) atomicOr(&st->overflow, 4u);
}

template <int BV, int PF>
__global__ __launch_bounds__(256) void k_shade(SceneView sc, const uint2* cells, const PhotonRec* ph, PassState* st,
                                               EyeBuf eb, const uint32_t* order, int pf, int uc) {
  k_shade_body<BV, PF>(sc, cells, ph, st, eb, order, pf, uc, Blk{blockIdx.x, gridDim.x});
}

// k_shade for the pass's photon filter
template <int BV>
static void launch_shade(hipStream_t s, uint32_t blocks, size_t shmem, const SceneView& v, const uint2* cells,
                         const PhotonRec* ph, PassState* st, const EyeBuf& eb, const uint32_t* order, int pf, int uc) {
  if (pf == FILTER_CONE)
    hipLaunchKernelGGL((k_shade<BV, FILTER_CONE>), dim3(blocks), dim3(256), shmem, s, v, cells, ph, st, eb, order, pf,
                       uc);
  else if (pf == FILTER_GAUSS)
    hipLaunchKernelGGL((k_shade<BV, FILTER_GAUSS>), dim3(blocks), dim3(256), shmem, s, v, cells, ph, st, eb, order,
                       pf, uc);
  else
    hipLaunchKernelGGL((k_shade<BV, FILTER_NONE>), dim3(blocks), dim3(256), shmem, s, v, cells, ph, st, eb, order, pf,
                       uc);
}

// The same shading over a box-tree map (ppm_trace_eye on a PhotonMap, whose
// radius comes with the call): a wave takes 64 consecutive sorted queries
// and traverses the tree for the union of their balls; each overlapped leaf
// is staged once into LDS with coalesced loads and every lane whose ball
// overlaps the leaf box tests its photons from LDS (broadcast reads).
template <int BV, bool BVH = (BV != 0)>
__device__ __forceinline__ void k_shade_tree_body(SceneView sc, const Node* nodes, const PhotonRec* ph,
                                               const PassState* st, EyeBuf eb, const uint32_t* order, int pf,
                                               int uc, Blk blk) {
  __shared__ LeafLds lds[4];
  const uint32_t wib = threadIdx.x >> 6, lane = lane_id();
  const uint32_t n = st->n_stored, depth = st->depth;
  const float r2 = st->r2f, scale = st->scale, inv_r2 = 1.0f / r2;
  const uint32_t nq = st->n_qsort;
  const uint32_t groups = (nq + 63) / 64;
  const uint32_t waves = blk.n * 4;
  const uint32_t first_leaf = (1u << depth) - 1u;
  for (uint32_t g = blk.b * 4 + wib; g < groups; g += waves) {
    const uint32_t i = g * 64 + lane;
    const bool active = i < nq;
    uint32_t slot = 0;
    Query qq{};
    if (active) {
      slot = order[i];
      qq = eb.q[slot];
    }
    const float qx = qq.a.x, qy = qq.a.y, qz = qq.a.z;
    const float nx = qq.b.x, ny = qq.b.y, nz = qq.b.z;
    float ar = 0.0f, ag = 0.0f, ab = 0.0f;
    uint32_t node = 0;
    while (n > 0) {
      node = __builtin_amdgcn_readfirstlane(node);
      const Node nd = ldc(nodes, node);
      const bool ov = active && box_dist2_nc(qx, qy, qz, nd.lo, nd.hi) <= r2;
      if (__ballot(ov)) {
        if (node < first_leaf) {
          node = 2 * node + 1;
          continue;
        }
        const uint32_t base = (node - first_leaf) * LEAF;
        const uint32_t cnt = base < n ? min((uint32_t)LEAF, n - base) : 0u;
        if (lane < cnt) {
          const PhotonRec r = ph[base + lane];
          lds[wib].pm[lane] = r.pm;
          lds[wib].dm[lane] = r.dm;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (ov) {
          for (uint32_t j = 0; j < cnt; ++j) {
            const float4 pm = lds[wib].pm[j];
            const float d2 = dist2_nc(qx, qy, qz, pm.x, pm.y, pm.z);
            if (d2 <= r2) {
              const float4 dm = lds[wib].dm[j];
              const float c = nx * dm.x + ny * dm.y + nz * dm.z;  // optics.rs:225
              const float v = c < 0.0f ? -c * filter_weight(pf, d2 * inv_r2) * dm.w : 0.0f;
              const uint32_t wl = __float_as_uint(pm.w) & 3u;
              ar += wl == 0 ? v : 0.0f;
              ag += wl == 1 ? v : 0.0f;
              ab += wl == 2 ? v : 0.0f;
            }
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
      while (node != 0 && (node & 1u) == 0) node = (node - 1) >> 1;
      if (node == 0) break;
      node += 1;
    }
    C3 di = C3{ar * scale, ag * scale, ab * scale};
    if (uc) di = di + direct_light<BV>(sc, mk(qq.a), mk(qq.b));
    if (active) eb.res[slot] = make_float4(qq.w.x * di.r, qq.w.y * di.g, qq.w.z * di.b, 0.0f);
  }
}

template <int BV, bool BVH = (BV != 0)>
__global__ __launch_bounds__(256) void k_shade_tree(SceneView sc, const Node* nodes, const PhotonRec* ph,
                                               const PassState* st, EyeBuf eb, const uint32_t* order, int pf,
                                               int uc) {
  k_shade_tree_body<BV>(sc, nodes, ph, st, eb, order, pf, uc, Blk{blockIdx.x, gridDim.x});
}

// ===========================================================================
// Adjoint of one pass with respect to the Simple materials' c1, for
// gimg = dL/d(this pass's image).  Per sorted query (pixel p, material m,
// weight w = path throughput * c1/pi * diffuseness, result res = w * di):
//   eye side:    dL/dc1[m] += gimg[p] * res / c1[m]  (res is linear in c1[m]);
//   photon side: each photon j in the ball adds to its adjoint weight
//                adj[j] += gimg[p][wl_j] * w[wl_j] * scale * max(0,-n.dir_j)
//                          * filter * weight_j,
// the photon's share of dL; the replay of the photon walks then turns
// adj[j] into dL/dc1 over the diffuse survivals before record j.
__global__ __launch_bounds__(256) void k_adj_gather(SceneView sc, const uint2* cells, const PhotonRec* ph,
                                                    PassState* st, EyeBuf eb, const uint32_t* order, int pf,
                                                    const float4* gimg, float* adj, double* grad) {
  __shared__ float part[3 * ADJ_LDS_MATS];
  for (uint32_t i = threadIdx.x; i < 3 * ADJ_LDS_MATS; i += blockDim.x) part[i] = 0.0f;
  __syncthreads();
  const uint32_t wib = threadIdx.x >> 6, lane = lane_id();
  const float r2 = st->r2f, scale = st->scale, inv_r2 = 1.0f / r2;
  const uint32_t nq = st->n_qsort;
  const GridView g = grid_view(st, sc.bmin);
  const uint32_t groups = (nq + 63) / 64;
  const uint32_t waves = gridDim.x * 4;
  bool complete = true;
  for (uint32_t gi = blockIdx.x * 4 + wib; gi < groups; gi += waves) {
    const uint32_t i = gi * 64 + lane;
    const bool active = i < nq;
    Query qq{};
    float4 gp = make_float4(0.f, 0.f, 0.f, 0.f), res = gp;
    if (active) {
      const uint32_t slot = order[i];
      qq = eb.q[slot];
      res = eb.res[slot];
      gp = gimg[__float_as_uint(qq.a.w)];
      const uint32_t m = __float_as_uint(qq.b.w);
      const DevMaterial& M = sc.mats[m];
      if (M.surface == SURF_SIMPLE) {
        if (M.c1.x != 0.0f) adj_add(part, grad, m, 0, gp.x * res.x / M.c1.x);
        if (M.c1.y != 0.0f) adj_add(part, grad, m, 1, gp.y * res.y / M.c1.y);
        if (M.c1.z != 0.0f) adj_add(part, grad, m, 2, gp.z * res.z / M.c1.z);
      }
    }
    // per channel: gimg * w * scale, the factor on each photon's term
    const float fr = gp.x * qq.w.x * scale, fg = gp.y * qq.w.y * scale, fb = gp.z * qq.w.z * scale;
    const bool want = active && (fr != 0.0f || fg != 0.0f || fb != 0.0f);
    if (want) {
      const float nx = qq.b.x, ny = qq.b.y, nz = qq.b.z;
      complete &= grid_query(g, cells, ph, qq.a.x, qq.a.y, qq.a.z, r2,
                             // the channel factors by value: a conditional over referenced
                             // floats selects an address, which puts the closure in scratch
                             [&, fr, fg, fb](uint32_t j, float4 pm, float4 dm, float d2) {
                               // grid records (grid_rec): weighted direction, channel flags
                               const float c = nx * dm.x + ny * dm.y + nz * dm.z;
                               if (c < 0.0f) {
                                 const float f = pm.w != 0.0f ? fr : (dm.w != 0.0f ? fg : fb);
                                 const float a = f * (-c * filter_weight(pf, d2 * inv_r2));
                                 if (a != 0.0f) atomicAdd(&adj[j], a);
                               }
                             });
    }
  }
  if (!complete) atomicOr(&st->overflow, 4u);
  adj_flush(part, grad, sc.n_mats);
}

// Photon adjoint weights from map order to trace-slot order.
__global__ __launch_bounds__(256) void k_adj_permute(const uint32_t* src_idx, const float* adj_sorted,
                                                     float* adj_slot, const PassState* st) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < st->n_stored) adj_slot[src_idx[e]] = adj_sorted[e];
}

template <int BV, bool BVH = (BV != 0)>
__global__ __launch_bounds__(256) void k_adj_trace(SceneView sc, PassState* st, uint32_t n_waves, uint32_t per_wave,
                                                   uint32_t region_cap, uint64_t seed, int uc, int maxd,
                                                   const float* adj_slot, double* grad) {
  k_trace_photons_body<BV, true>(sc, st, nullptr, nullptr, n_waves, per_wave, region_cap, seed, uc, maxd,
                                  Blk{blockIdx.x, gridDim.x}, AdjTrace{adj_slot, grad});
}

hipError_t launch_adjoint(hipStream_t s, const SceneBuf& sc, const MapBuf& m, PassState* st, const PhotonBuf& pb,
                          const EyeBuf& eb, int pfilter, int n_cu, uint64_t seed, int use_classic, int max_depth,
                          const float4* gimg, float* adj_sorted, float* adj_slot, double* grad) {
  hipError_t e;
  if ((e = hipMemsetAsync(adj_sorted, 0, sizeof(float) * m.cap, s)) != hipSuccess) return e;
  if ((e = hipMemsetAsync(adj_slot, 0, sizeof(float) * pb.cap, s)) != hipSuccess) return e;
  const SceneView v = make_view(sc);
  const uint32_t blocks = (uint32_t)n_cu * 8u;
  hipLaunchKernelGGL(k_adj_gather, dim3(blocks), dim3(256), 0, s, v, m.cells, m.sorted, st, eb, eb.qvals[0],
                     pfilter, gimg, adj_sorted, grad);
  hipLaunchKernelGGL(k_adj_permute, dim3((uint32_t)((m.cap + 255) / 256)), dim3(256), 0, s, m.src_idx, adj_sorted,
                     adj_slot, (const PassState*)st);
  const uint32_t n_waves = pb.n_regions;
  const uint32_t per_wave = (uint32_t)((sc.n_emit_total + (uint64_t)n_waves - 1) / n_waves);
  switch (bvh_kind(sc)) {
    case 2:
      hipLaunchKernelGGL(k_adj_trace<2>, dim3((n_waves + 3) / 4), dim3(256), bvh_dyn_lds(sc), s, v, st, n_waves,
                       per_wave,
                       pb.region_cap, seed, use_classic, max_depth, (const float*)adj_slot, grad);
      break;
    case 1:
      hipLaunchKernelGGL(k_adj_trace<1>, dim3((n_waves + 3) / 4), dim3(256), 0, s, v, st, n_waves,
                       per_wave,
                       pb.region_cap, seed, use_classic, max_depth, (const float*)adj_slot, grad);
      break;
    default:
      hipLaunchKernelGGL(k_adj_trace<0>, dim3((n_waves + 3) / 4), dim3(256), 0, s, v, st, n_waves, per_wave,
                       pb.region_cap, seed, use_classic, max_depth, (const float*)adj_slot, grad);
  }
  return hipGetLastError();
}

// Overflow results into their pixels (float atomics; few per pass).
__device__ __forceinline__ void k_overflow_add_body(EyeBuf eb, const PassState* st, Blk blk) {
  const uint32_t n = min(st->n_queries, eb.ovf_cap);
  for (uint32_t i = blk.b * blockDim.x + threadIdx.x; i < n; i += blk.n * blockDim.x) {
    const uint64_t slot = (uint64_t)eb.q_primary * (1u + eb.sec_max) + i;
    const uint32_t p = __float_as_uint(eb.q[slot].a.w);
    const float4 r = eb.res[slot];
    atomicAdd(&eb.pix[p].x, r.x);
    atomicAdd(&eb.pix[p].y, r.y);
    atomicAdd(&eb.pix[p].z, r.z);
  }
}

__global__ __launch_bounds__(256) void k_overflow_add(EyeBuf eb, const PassState* st) {
  k_overflow_add_body(eb, st, Blk{blockIdx.x, gridDim.x});
}

// Per-pixel pass radiance = pix (emission + overflow results) + res of the pixel's primary and deeper queries, in a fixed
// order; accumulated into the FP64 sums (util/averager2.rb:58-60) when sum
// is given, else stored back into pix.
__device__ __forceinline__ void k_accumulate_body(EyeBuf eb, DevCamera cam, double* sum, double* sumsq, Blk blk) {
  const uint32_t t = blk.b * blockDim.x + threadIdx.x;
  if (t >= eb.q_primary) return;
  const uint32_t g = t >> 6, wi = t & 63u, ts = (uint32_t)cam.tile_shift;
  const int x = (int)(((g % (uint32_t)cam.tiles_x) << ts) + (wi & ((1u << ts) - 1u)));
  const int y = (int)(((g / (uint32_t)cam.tiles_x) << (6u - ts)) + (wi >> ts));
  if (x >= cam.xreso || y >= cam.yreso) return;
  const uint32_t p = (uint32_t)y * (uint32_t)cam.xreso + (uint32_t)x;
  float4 c = eb.pix[p];
  const uint32_t sc = eb.sec_count[t];
  const uint32_t nsec = sc & ~SEC_PRIMARY;
  // results of slots q_primary * k + t, k = 0 (the primary, when there is
  // one) .. nsec, added in that order, four loaded at once
  for (uint32_t k0 = (sc & SEC_PRIMARY) != 0u ? 0u : 1u; k0 <= nsec; k0 += 4u) {
    float4 r[4];
#pragma unroll
    for (uint32_t j = 0; j < 4u; ++j)
      if (k0 + j <= nsec) r[j] = eb.res[eb.q_primary * (k0 + j) + t];
#pragma unroll
    for (uint32_t j = 0; j < 4u; ++j) {
      if (k0 + j <= nsec) {
        c.x += r[j].x;
        c.y += r[j].y;
        c.z += r[j].z;
      }
    }
  }
  if (sum) {
    const double r = c.x, gg = c.y, b = c.z;
    double* s3 = sum + 3 * (uint64_t)p;
    double* q3 = sumsq + 3 * (uint64_t)p;
    s3[0] += r;
    s3[1] += gg;
    s3[2] += b;
    q3[0] += r * r;
    q3[1] += gg * gg;
    q3[2] += b * b;
  } else {
    eb.pix[p] = c;
  }
}

__global__ __launch_bounds__(256) void k_accumulate(EyeBuf eb, DevCamera cam, double* sum, double* sumsq) {
  k_accumulate_body(eb, cam, sum, sumsq, Blk{blockIdx.x, gridDim.x});
}

hipError_t launch_query_prep(hipStream_t s, const SceneBuf& sc, const MapBuf& m, PassState* st, const EyeBuf& eb) {
  hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, s, eb.tile_count, eb.tile_off, eb.n_tiles,
                     &st->n_qsort, (const uint32_t*)&st->n_queries, eb.ovf_cap);
  const bool fine = eb.q_fine != 0u;
  const uint32_t qmax = fine ? QUERY_FINE_MORTON_MAX : QUERY_MORTON_MAX;
  const uint32_t lshift = fine ? QUERY_FINE_LEVEL_SHIFT : QUERY_LEVEL_SHIFT;
  const uint32_t kbits = fine ? QUERY_FINE_KEY_BITS : QUERY_KEY_BITS;
  const int first = (int)((kbits / 8u) & 1u);
  hipLaunchKernelGGL(k_query_keys, dim3((eb.q_primary + 255) / 256 + OVF_BLOCKS), dim3(256), 0, s, eb, st, sc.bmin,
                     morton_scale(sc.bmin, sc.bmax, qmax), (float)qmax, lshift, eb.qkeys[first], eb.qvals[first],
                     eb.qos.ctl);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  uint32_t* qk[2] = {eb.qkeys[first], eb.qkeys[first ^ 1]};
  uint32_t* qv[2] = {eb.qvals[first], eb.qvals[first ^ 1]};
  int cur = 0;
  e = onesweep_sort(s, eb.qos, qk, qv, &st->n_qsort, kbits, &st->overflow, &cur);
  if (e == hipSuccess && (cur ^ first) != 0) e = hipErrorInvalidValue;  // the sorted order is read from qvals[0]
  return e;
}

hipError_t launch_gather(hipStream_t s, const SceneBuf& sc, const MapBuf& m, PassState* st, const EyeBuf& eb,
                         int pfilter, int use_classic, int n_cu, double* sum, double* sumsq, hipEvent_t* ev,
                         bool prepared) {
  const uint32_t blocks = (uint32_t)n_cu * SHADE_BLOCKS_PER_CU;
  hipError_t e;
  if (!prepared && (e = launch_query_prep(s, sc, m, st, eb)) != hipSuccess) return e;
  uint32_t* const* qv = eb.qvals;
  const int cur = 0;
  if (ev && (e = hipEventRecord(ev[0], s)) != hipSuccess) return e;
  const SceneView v = make_view(sc);
  if (m.cells) {
    // persistent: at least as many waves as stay resident (8 blocks of 4
    // waves per CU); waves beyond that find the groups taken and leave
    switch (bvh_kind(sc)) {
      case 2:
        launch_shade<2>(s, blocks, bvh_dyn_lds(sc), v, m.cells, m.sorted, st, eb, qv[cur], pfilter, use_classic);
        break;
      case 1:
        launch_shade<1>(s, blocks, 0, v, m.cells, m.sorted, st, eb, qv[cur], pfilter, use_classic);
        break;
      default:
        launch_shade<0>(s, blocks, 0, v, m.cells, m.sorted, st, eb, qv[cur], pfilter, use_classic);
    }
  } else {
    switch (bvh_kind(sc)) {
      case 2:
        hipLaunchKernelGGL(k_shade_tree<2>, dim3(blocks), dim3(256), bvh_dyn_lds(sc), s, v, m.nodes, m.sorted, st,
                         eb, qv[cur],
                         pfilter, use_classic);
        break;
      case 1:
        hipLaunchKernelGGL(k_shade_tree<1>, dim3(blocks), dim3(256), 0, s, v, m.nodes, m.sorted, st,
                         eb, qv[cur],
                         pfilter, use_classic);
        break;
      default:
        hipLaunchKernelGGL(k_shade_tree<0>, dim3(blocks), dim3(256), 0, s, v, m.nodes, m.sorted, st, eb, qv[cur],
                         pfilter, use_classic);
    }
  }
  if (ev && (e = hipEventRecord(ev[1], s)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_overflow_add, dim3(OVF_BLOCKS), dim3(256), 0, s, eb, (const PassState*)st);
  hipLaunchKernelGGL(k_accumulate, dim3((eb.q_primary + 255) / 256), dim3(256), 0, s, eb, sc.cam, sum, sumsq);
  return hipGetLastError();
}

// ===========================================================================
// The classic ray tracer (rtc.rs:17-47, trace_ray_classic): the eye trees of
// k_eye_classic, then per query di = the direct light of every light
// (tracer.rs:234-237) plus the camera's ambient (:238), no photon map;
// res = w * di, summed per pixel as a pass's radiance.
template <int BV, bool BVH = (BV != 0)>
__global__ __launch_bounds__(256) void k_shade_classic(SceneView sc, const PassState* st, EyeBuf eb,
                                                       const uint32_t* order, float4 ambient) {
  scene_to_lds<BV>(sc);
  const uint32_t nq = st->n_qsort;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t base = blockIdx.x * blockDim.x; base < nq; base += stride) {
    const uint32_t i = base + threadIdx.x;
    if (i >= nq) continue;  // after the loop's last wave-wide step only
    const uint32_t slot = order[i];
    const Query qq = eb.q[slot];
    const C3 di = direct_light<BV>(sc, mk(qq.a), mk(qq.b)) + C3{ambient.x, ambient.y, ambient.z};
    eb.res[slot] = make_float4(qq.w.x * di.r, qq.w.y * di.g, qq.w.z * di.b, 0.0f);
  }
}

hipError_t launch_classic(hipStream_t s, const SceneBuf& sc, PassState* st, const EyeBuf& eb, uint64_t seed,
                          float4 ambient, int n_cu) {
  hipLaunchKernelGGL(k_tile_order, dim3((eb.n_tiles + 255) / 256), dim3(256), 0, s, eb, st);
  uint32_t n_waves = std::min<uint32_t>(std::min<uint32_t>(eb.n_tiles, (uint32_t)n_cu * EYE_WAVES_PER_CU), eb.estack_lanes / 64u);
  if (n_waves < 1) n_waves = 1;
  const uint32_t blocks = (n_waves + 3) / 4;
  const SceneView v = make_view(sc);
  switch (bvh_kind(sc)) {
    case 2:
      hipLaunchKernelGGL(k_eye_classic<2>, dim3(blocks), dim3(256), bvh_dyn_lds(sc), s, v, st, eb, seed, n_waves);
      break;
    case 1:
      hipLaunchKernelGGL(k_eye_classic<1>, dim3(blocks), dim3(256), 0, s, v, st, eb, seed, n_waves);
      break;
    default:
      hipLaunchKernelGGL(k_eye_classic<0>, dim3(blocks), dim3(256), 0, s, v, st, eb, seed, n_waves);
  }
  hipLaunchKernelGGL(k_tile_count, dim3((eb.q_primary + 255) / 256), dim3(256), 0, s, eb);
  hipError_t e = launch_query_prep(s, sc, MapBuf{}, st, eb);
  if (e != hipSuccess) return e;
  const uint32_t sblocks = (uint32_t)n_cu * 8u;
  switch (bvh_kind(sc)) {
    case 2:
      hipLaunchKernelGGL(k_shade_classic<2>, dim3(sblocks), dim3(256), bvh_dyn_lds(sc), s, v, (const PassState*)st,
                         eb, eb.qvals[0], ambient);
      break;
    case 1:
      hipLaunchKernelGGL(k_shade_classic<1>, dim3(sblocks), dim3(256), 0, s, v, (const PassState*)st, eb,
                         eb.qvals[0], ambient);
      break;
    default:
      hipLaunchKernelGGL(k_shade_classic<0>, dim3(sblocks), dim3(256), 0, s, v, (const PassState*)st, eb,
                         eb.qvals[0], ambient);
  }
  hipLaunchKernelGGL(k_overflow_add, dim3(OVF_BLOCKS), dim3(256), 0, s, eb, (const PassState*)st);
  hipLaunchKernelGGL(k_accumulate, dim3((eb.q_primary + 255) / 256), dim3(256), 0, s, eb, sc.cam, (double*)nullptr,
                     (double*)nullptr);
  return hipGetLastError();
}

// ===========================================================================
// Test hooks: radius query (KdTree::within) and gather (estimate_radiance) on
// host-given query points.  refine64 decides photons whose float squared
// distance lies within the float error band of r^2 again in double, from the
// double positions kept by ppm_photons_upload, which reproduces the FP64
// predicate of the reference exactly.

// Membership of photon j; *t_out receives d2 / r2 for the filter weight,
// formed from the double distance when refining.
__device__ __forceinline__ bool in_ball(float d2f, float r2f, float band, bool refine, const double* p64,
                                        uint32_t j, double qx, double qy, double qz, double r2, float* t_out) {
  if (!refine) {
    *t_out = d2f / r2f;
    return d2f <= r2f;
  }
  const double dx = qx - p64[3 * (uint64_t)j + 0];
  const double dy = qy - p64[3 * (uint64_t)j + 1];
  const double dz = qz - p64[3 * (uint64_t)j + 2];
  double s;
  {
#pragma clang fp contract(off)
    s = dx * dx + dy * dy;
    s = s + dz * dz;
  }
  *t_out = (float)(s / r2);
  if (d2f < r2f - band) return true;
  if (d2f > r2f + band) return false;
  return s <= r2;
}

__global__ __launch_bounds__(256) void k_query(const Node* nodes, const PhotonRec* ph, const double* ph64,
                                               const uint32_t* src_idx, const PassState* st, const float4* qf,
                                               const double* q64, const float4* qn, uint32_t nq, double r2,
                                               int refine, int mode, int pf, uint32_t* counts,
                                               const uint64_t* offsets, uint32_t* idx_out, float4* rgb_out) {
  __shared__ LeafLds lds[4];
  const uint32_t w_in_b = threadIdx.x >> 6;
  const uint32_t lane = lane_id();
  const uint32_t n = st->n_stored;
  const uint32_t depth = st->depth;
  const float r2f = (float)r2;
  const uint32_t groups = (nq + 63) / 64;
  const uint32_t waves = gridDim.x * 4;
  for (uint32_t g = blockIdx.x * 4 + w_in_b; g < groups; g += waves) {
    const uint32_t i = g * 64 + lane;
    const bool active = i < nq;
    const float4 qp = active ? qf[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 nn = (active && mode == 2) ? qn[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    double qx = 0, qy = 0, qz = 0;
    float band = 0.0f;
    if (active && refine) {
      qx = q64[3 * (uint64_t)i];
      qy = q64[3 * (uint64_t)i + 1];
      qz = q64[3 * (uint64_t)i + 2];
      // generous bound on |fl32(d2) - d2| for photons near the sphere
      const double sc = fmax(fmax(fabs(qx), fabs(qy)), fabs(qz)) + sqrt(r2);
      band = (float)(64.0 * 5.9604644775390625e-8 * (sc * sqrt(r2) + r2) + 1e-30);
    }
    const float r2box = refine ? r2f + band : r2f;
    uint32_t cnt = 0;
    const uint64_t off = (mode == 1 && active) ? offsets[i] : 0;
    float ar = 0.f, ag = 0.f, ab = 0.f;
    wave_query(nodes, ph, n, depth, active, qp.x, qp.y, qp.z, r2box, lds[w_in_b],
               [&](uint32_t j, float4 pm, float4 dm) {
                 const float d2 = dist2_nc(qp.x, qp.y, qp.z, pm.x, pm.y, pm.z);
                 float t;
                 if (in_ball(d2, r2f, band, refine != 0, ph64, j, qx, qy, qz, r2, &t)) {
                   if (mode == 1) idx_out[off + cnt] = src_idx[j];
                   ++cnt;
                   if (mode == 2) {
                     const float c = nn.x * dm.x + nn.y * dm.y + nn.z * dm.z;
                     const float v = c < 0.0f ? -c * filter_weight(pf, t) * dm.w : 0.0f;
                     const uint32_t wl = __float_as_uint(pm.w) & 3u;
                     ar += wl == 0 ? v : 0.0f;
                     ag += wl == 1 ? v : 0.0f;
                     ab += wl == 2 ? v : 0.0f;
                   }
                 }
               });
    if (active) {
      if (mode == 0) counts[i] = cnt;
      if (mode == 2) rgb_out[i] = make_float4(ar, ag, ab, 0.0f);
    }
  }
}

hipError_t launch_query(hipStream_t s, const MapBuf& m, PassState* st, const float4* qf, const double* q64,
                        const float4* qn, uint32_t nq, double r2, int refine64, int mode, int pfilter,
                        uint32_t* counts, const uint64_t* offsets, uint32_t* idx_out, float4* rgb_out, int n_cu) {
  uint32_t groups = (nq + 63) / 64;
  uint32_t blocks = (groups + 3) / 4;
  uint32_t maxb = (uint32_t)n_cu * 8u;
  if (blocks > maxb) blocks = maxb;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(k_query, dim3(blocks), dim3(256), 0, s, m.nodes, m.sorted, m.sorted64, m.src_idx, st, qf, q64,
                     qn, nq, r2, refine64, mode, pfilter, counts, offsets, idx_out, rgb_out);
  return hipGetLastError();
}

// The same index sets over a pass's photon grid (a map built by
// launch_grid_build for a radius >= sqrt(r2)): one query per thread through
// grid_query, the cell walk of the shade kernel's gather, which calls back
// for every photon of the query's cells in the shade's order.  Membership is
// the shade's float predicate, or, when refining, decided again in double
// near the sphere (the cells cover the double ball as well: the grid's reach
// exceeds r by r/1024 plus the rounding of float coordinates).  flags bit 0:
// a query whose cells exceeded the grid's span rule.
__global__ __launch_bounds__(256) void k_grid_query(const uint2* cells, const PhotonRec* ph, const double* ph64,
                                                    const uint32_t* src_idx, const PassState* st, float4 bmin,
                                                    const float4* qf, const double* q64, uint32_t nq, double r2,
                                                    int refine, int mode, uint32_t* counts, const uint64_t* offsets,
                                                    uint32_t* idx_out, uint32_t* flags) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  const GridView g = grid_view(st, bmin);
  const float4 qp = qf[i];
  const float r2f = (float)r2;
  double qx = 0, qy = 0, qz = 0;
  float band = 0.0f;
  if (refine) {
    qx = q64[3 * (uint64_t)i];
    qy = q64[3 * (uint64_t)i + 1];
    qz = q64[3 * (uint64_t)i + 2];
    const double sc = fmax(fmax(fabs(qx), fabs(qy)), fabs(qz)) + sqrt(r2);
    band = (float)(64.0 * 5.9604644775390625e-8 * (sc * sqrt(r2) + r2) + 1e-30);
  }
  uint32_t cnt = 0;
  const uint64_t off = mode == 1 ? offsets[i] : 0;
  const bool ok = grid_query<true>(g, cells, ph, qp.x, qp.y, qp.z, r2f,
                                   [&](uint32_t j, float4, float4, float d2, bool in, bool valid) {
                                     float t;
                                     const bool keep =
                                         refine ? valid && in_ball(d2, r2f, band, true, ph64, j, qx, qy, qz, r2, &t)
                                                : in;
                                     if (keep) {
                                       if (mode == 1) idx_out[off + cnt] = src_idx[j];
                                       ++cnt;
                                     }
                                   });
  if (!ok) atomicOr(flags, 1u);
  if (mode == 0) counts[i] = cnt;
}

hipError_t launch_grid_query(hipStream_t s, const MapBuf& m, const PassState* st, float4 bmin, const float4* qf,
                             const double* q64, uint32_t nq, double r2, int refine64, int mode, uint32_t* counts,
                             const uint64_t* offsets, uint32_t* idx_out, uint32_t* flags) {
  if (nq == 0) return hipSuccess;
  if (!m.cells || (refine64 && !m.sorted64)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_grid_query, dim3((nq + 255) / 256), dim3(256), 0, s, m.cells, m.sorted, m.sorted64,
                     m.src_idx, st, bmin, qf, q64, nq, r2, refine64, mode, counts, offsets, idx_out, flags);
  return hipGetLastError();
}

// The gather of host-given queries through the shade kernel itself
// (k_shade, direct light off): eb.q holds the queries (position, normal,
// weight), order their order, st the pass state of the grid with r2f,
// scale and n_qsort set and the work counters zeroed; eb.res receives
// w * scale * Σ filter * max(0, -n.dir) * weight per channel.
hipError_t launch_grid_shade(hipStream_t s, const SceneBuf& sc, const MapBuf& m, PassState* st, const EyeBuf& eb,
                             const uint32_t* order, int pfilter, int n_cu) {
  if (!m.cells || !sc.mats || sc.bvh) return hipErrorInvalidValue;
  launch_shade<0>(s, (uint32_t)n_cu * 8u, 0, make_view(sc), m.cells, m.sorted, st, eb, order, pfilter, 0);
  return hipGetLastError();
}

}  // namespace ppm

namespace ppm {

// Test hook: nearest hit of host-given rays (calc_intersection,
// tracer.rs:306-335): t and scene index of the shape, or (inf, -1).  The
// shapes are read as the pass kernels read them: a small tree, the shapes
// tested one by one and the materials from their LDS copies.
template <int BV, bool BVH = (BV != 0)>
__global__ __launch_bounds__(256) void k_intersect(SceneView sc, const float4* o, const float4* d, uint32_t n,
                                                   float* t_out, int32_t* obj_out) {
  scene_to_lds<BV>(sc);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Hit h = nearest<BV>(sc, mk(o[i]), mk(d[i]));
  t_out[i] = h.obj >= 0 ? h.t : INF_F;
  obj_out[i] = h.obj;
}

hipError_t launch_intersect(hipStream_t s, const SceneBuf& sc, const float4* o, const float4* d, uint32_t n,
                            float* t_out, int32_t* obj_out) {
  if (n == 0) return hipSuccess;
  switch (bvh_kind(sc)) {
    case 2:
      hipLaunchKernelGGL(k_intersect<2>, dim3((n + 255) / 256), dim3(256), bvh_dyn_lds(sc), s, make_view(sc), o, d,
                         n, t_out, obj_out);
      break;
    case 1:
      hipLaunchKernelGGL(k_intersect<1>, dim3((n + 255) / 256), dim3(256), 0, s, make_view(sc), o, d, n, t_out,
                       obj_out);
      break;
    default:
      hipLaunchKernelGGL(k_intersect<0>, dim3((n + 255) / 256), dim3(256), 0, s, make_view(sc), o, d, n, t_out,
                       obj_out);
  }
  return hipGetLastError();
}

}  // namespace ppm

// ===========================================================================
// Batched passes: many scenes, each with its own buffers and schedule, in one
// launch per stage.  A stage's grid is the concatenation of every scene's
// own grid for that stage: scene s owns blocks [first[s], first[s+1]) and
// its blocks run the same kernel body as a single-scene launch, with their
// block index taken within that range, so a scene computes exactly what its
// own session would.
namespace ppm {

// Scene owning block b: the last s with first[s] <= b (uniform per block,
// so the search runs on scalar loads).
__device__ __forceinline__ uint32_t batch_find(const uint32_t* first, uint32_t n, uint32_t b) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (ldc(first, mid) <= b) lo = mid; else hi = mid;
  }
  return lo;
}

// The scene's item, read at a uniform address through the constant space.
__device__ __forceinline__ BatchItem load_item(const BatchItem* items, uint32_t s) {
  return ldc(items, s);
}

#define BATCH_PROLOGUE(kind)                                        \
  const uint32_t* first = tab + (size_t)(kind) * (n_all + 1u);      \
  const uint32_t s = batch_find(first, n, blockIdx.x);              \
  const uint32_t b0 = ldc(first, s);                                \
  const Blk blk{blockIdx.x - b0, ldc(first, s + 1u) - b0};          \
  const BatchItem it = load_item(items, s);

__global__ void kb_pass_begin(const BatchItem* items, uint32_t n) {
  if (blockIdx.x >= n) return;
  const BatchItem it = load_item(items, blockIdx.x);
  k_pass_begin_body(it.st, it.first_pass, it.r2_first, it.alpha, it.power, it.pass_stride, Blk{0u, 1u});
}

__global__ void kb_pass_end(const BatchItem* items, uint32_t n) {
  if (blockIdx.x >= n) return;
  const BatchItem it = load_item(items, blockIdx.x);
  k_pass_end_body(it.st, Blk{0u, 1u});
}

template <int BV, bool BVH = (BV != 0)>
__global__ __launch_bounds__(256) void kb_trace(const BatchItem* items, const uint32_t* tab, uint32_t n,
                                                uint32_t n_all) {
  BATCH_PROLOGUE(BK_TRACE)
  k_trace_photons_body<BV>(make_view(it.sc), it.st, it.pb.rec, it.pb.count, it.pb.n_regions, it.trace_per_wave,
                            it.pb.region_cap, it.seed, it.uc, it.maxd, blk);
}

// one block per scene: the photon counts of the tracing waves, then (second
// use) the per-tile query counts
__global__ __launch_bounds__(1024) void kb_scan_photons(const BatchItem* items, uint32_t n) {
  if (blockIdx.x >= n) return;
  const BatchItem it = load_item(items, blockIdx.x);
  k_scan_counts_body(it.pb.count, it.pb.offset, it.pb.n_regions, &it.st->n_stored, nullptr, 0u);
}

__global__ __launch_bounds__(1024) void kb_scan_tiles(const BatchItem* items, uint32_t n) {
  if (blockIdx.x >= n) return;
  const BatchItem it = load_item(items, blockIdx.x);
  k_scan_counts_body(it.eb.tile_count, it.eb.tile_off, it.eb.n_tiles, &it.st->n_qsort, &it.st->n_queries,
                     it.eb.ovf_cap);
}

__global__ void kb_grid_setup(const BatchItem* items, uint32_t n) {
  if (blockIdx.x >= n) return;
  const BatchItem it = load_item(items, blockIdx.x);
  k_grid_setup_body(it.st, it.sc.bmin, it.sc.bmax, it.m.cells_max, Blk{0u, 1u});
}

__global__ __launch_bounds__(256) void kb_grid_keys(const BatchItem* items, const uint32_t* tab, uint32_t n,
                                                    uint32_t n_all) {
  BATCH_PROLOGUE(BK_MORTON)
  k_grid_keys_body(it.pb.rec, it.pb.offset, it.pb.n_regions, it.pb.region_cap, it.st, it.m.keys[0], it.m.vals[0],
                   it.sc.bmin, &it.st->nonfinite, it.m.os.ctl, blk);
}

__global__ __launch_bounds__(256) void kb_grid_clear(const BatchItem* items, const uint32_t* tab, uint32_t n,
                                                     uint32_t n_all) {
  BATCH_PROLOGUE(BK_CLEAR)
  k_grid_clear_body(it.m.cells, it.st, blk);
}

// the sorted keys / values after GRID_KEY_BITS / 8 digit passes
constexpr uint32_t GRID_SORTED = (GRID_KEY_BITS / 8u) & 1u;

__global__ __launch_bounds__(256) void kb_grid_cells(const BatchItem* items, const uint32_t* tab, uint32_t n,
                                                     uint32_t n_all) {
  BATCH_PROLOGUE(BK_LEAVES)
  k_grid_cells_body(it.m.keys[GRID_SORTED], it.st, it.m.cells, blk);
}

// The sorts' histograms: the photon keys (which = 0, in buffer 0) or the
// query keys (which = 1, in buffer QUERY_KEYS_FIRST); selects, not a runtime
// index into the item's arrays (which would put the whole item in scratch
// memory).
__global__ __launch_bounds__(1024) void kb_os_hist(const BatchItem* items, const uint32_t* tab, uint32_t n,
                                                   uint32_t n_all, int which) {
  BATCH_PROLOGUE(which ? BK_QHIST : BK_PHIST)
  if (which)
    k_os_hist_body(QUERY_KEYS_FIRST != 0u ? it.eb.qkeys[1] : it.eb.qkeys[0], &it.st->n_qsort, (uint32_t)it.eb.qos.cap,
                   QUERY_KEY_BITS / 8u, it.eb.qos.ctl, it.eb.qos.look, it.eb.qos.tiles, blk);
  else
    k_os_hist_body(it.m.keys[0], &it.st->n_stored, (uint32_t)it.m.os.cap, GRID_KEY_BITS / 8u, it.m.os.ctl,
                   it.m.os.look, it.m.os.tiles, blk);
}

// Digit pass p of a sort; REC: the last pass of the photon sort, which also
// moves the records into bucket order and writes the sorted source indices
// to src_idx.
template <bool REC>
__global__ __launch_bounds__(OS_THREADS) void kb_os_scatter(const BatchItem* items, const uint32_t* tab, uint32_t n,
                                                            uint32_t n_all, int which, uint32_t p) {
  BATCH_PROLOGUE(which ? BK_QOS : BK_POS)
  const bool odd = ((p + (which ? QUERY_KEYS_FIRST : 0u)) & 1u) != 0;
  if (!REC && which) {  // (one body per kernel: each instance has its own LDS)
    k_os_scatter_body<false>(odd ? it.eb.qkeys[1] : it.eb.qkeys[0], odd ? it.eb.qvals[1] : it.eb.qvals[0],
                             odd ? it.eb.qkeys[0] : it.eb.qkeys[1], odd ? it.eb.qvals[0] : it.eb.qvals[1],
                             &it.st->n_qsort, (uint32_t)it.eb.qos.cap, it.eb.qos.ctl + 256u * p,
                             it.eb.qos.ctl + OS_TICKETS + 32u * p, it.eb.qos.look + (uint64_t)p * it.eb.qos.tiles * 256u,
                             8u * p, &it.st->overflow, OsRecords{}, blk);
  } else {
    const OsRecords rec = REC ? OsRecords{it.pb.rec, it.pb.rec64, it.m.sorted, it.m.sorted64} : OsRecords{};
    k_os_scatter_body<REC>(odd ? it.m.keys[1] : it.m.keys[0], odd ? it.m.vals[1] : it.m.vals[0],
                           odd ? it.m.keys[0] : it.m.keys[1],
                           REC ? it.m.src_idx : (odd ? it.m.vals[0] : it.m.vals[1]), &it.st->n_stored,
                           (uint32_t)it.m.os.cap, it.m.os.ctl + 256u * p, it.m.os.ctl + OS_TICKETS + 32u * p,
                           it.m.os.look + (uint64_t)p * it.m.os.tiles * 256u, 8u * p, &it.st->overflow, rec, blk);
  }
}

__global__ __launch_bounds__(256) void kb_tile_order(const BatchItem* items, const uint32_t* tab, uint32_t n,
                                                     uint32_t n_all) {
  BATCH_PROLOGUE(BK_PIX)
  k_tile_order_body(it.eb, it.st, blk);
}

template <int BV, bool BVH = (BV != 0)>
__global__ __launch_bounds__(256) void kb_eye(const BatchItem* items, const uint32_t* tab, uint32_t n,
                                              uint32_t n_all) {
  BATCH_PROLOGUE(BK_EYE)
  k_eye_body<BV>(make_view(it.sc), it.st, it.eb, it.seed, it.uc, it.maxd, it.eye_waves, blk);
}

__global__ __launch_bounds__(256) void kb_tile_count(const BatchItem* items, const uint32_t* tab, uint32_t n,
                                                     uint32_t n_all) {
  BATCH_PROLOGUE(BK_PIX)
  k_tile_count_body(it.eb, blk);
}

__global__ __launch_bounds__(256) void kb_query_keys(const BatchItem* items, const uint32_t* tab, uint32_t n,
                                                     uint32_t n_all) {
  BATCH_PROLOGUE(BK_QKEYS)
  // batches sort on the coarse keys whatever their rooms' sizes
  k_query_keys_body(it.eb, it.st, it.sc.bmin, it.mscale, (float)QUERY_MORTON_MAX, QUERY_LEVEL_SHIFT,
                    QUERY_KEYS_FIRST != 0u ? it.eb.qkeys[1] : it.eb.qkeys[0],
                    QUERY_KEYS_FIRST != 0u ? it.eb.qvals[1] : it.eb.qvals[0], it.eb.qos.ctl, blk);
}

template <int BV, bool BVH = (BV != 0)>
__global__ __launch_bounds__(256) void kb_shade(const BatchItem* items, const uint32_t* tab, uint32_t n,
                                                uint32_t n_all) {
  BATCH_PROLOGUE(BK_SHADE)
  k_shade_body<BV, -1>(make_view(it.sc), it.m.cells, it.m.sorted, it.st, it.eb, it.eb.qvals[0], it.pfilter, it.uc,
                    blk);
}

__global__ __launch_bounds__(256) void kb_overflow_add(const BatchItem* items, const uint32_t* tab, uint32_t n,
                                                       uint32_t n_all) {
  BATCH_PROLOGUE(BK_OVF)
  k_overflow_add_body(it.eb, it.st, blk);
}

__global__ __launch_bounds__(256) void kb_accumulate(const BatchItem* items, const uint32_t* tab, uint32_t n,
                                                     uint32_t n_all) {
  BATCH_PROLOGUE(BK_PIX)
  k_accumulate_body(it.eb, it.sc.cam, it.sum, it.sumsq, blk);
}

#undef BATCH_PROLOGUE

hipError_t launch_batch_pass(hipStream_t s, hipStream_t s2, hipEvent_t fork, hipEvent_t join,
                             const BatchItem* items, const uint32_t* tab, const uint32_t* blocks, uint32_t n,
                             uint32_t n_all, bool bvh, uint32_t dyn) {
  // blocks[k] = total blocks of kind k over the first n scenes.  The eye
  // branch (tile order, eye trace, query sort) runs on s2 beside the photon
  // branch (trace, grid build): the two touch disjoint buffers and counters.
  const dim3 T256(256), T1024(1024), T64(64), TOS(OS_THREADS);
  hipError_t e;
  hipLaunchKernelGGL(kb_pass_begin, dim3(n), T64, 0, s, items, n);
  if ((e = hipEventRecord(fork, s)) != hipSuccess) return e;
  if ((e = hipStreamWaitEvent(s2, fork, 0)) != hipSuccess) return e;
  hipLaunchKernelGGL(kb_tile_order, dim3(blocks[BK_PIX]), T256, 0, s2, items, tab, n, n_all);
  switch ((bvh ? (dyn != 0u ? 2 : 1) : 0)) {
    case 2:
      hipLaunchKernelGGL(kb_eye<2>, dim3(blocks[BK_EYE]), T256, dyn, s2, items, tab, n, n_all);
      break;
    case 1:
      hipLaunchKernelGGL(kb_eye<1>, dim3(blocks[BK_EYE]), T256, 0, s2, items, tab, n, n_all);
      break;
    default:
      hipLaunchKernelGGL(kb_eye<0>, dim3(blocks[BK_EYE]), T256, 0, s2, items, tab, n, n_all);
  }
  hipLaunchKernelGGL(kb_tile_count, dim3(blocks[BK_PIX]), T256, 0, s2, items, tab, n, n_all);
  hipLaunchKernelGGL(kb_scan_tiles, dim3(n), T1024, 0, s2, items, n);
  hipLaunchKernelGGL(kb_query_keys, dim3(blocks[BK_QKEYS]), T256, 0, s2, items, tab, n, n_all);
  hipLaunchKernelGGL(kb_os_hist, dim3(blocks[BK_QHIST]), T1024, 0, s2, items, tab, n, n_all, 1);
  for (uint32_t p = 0; p < QUERY_KEY_BITS / 8u; ++p)
    hipLaunchKernelGGL(kb_os_scatter<false>, dim3(blocks[BK_QOS]), TOS, 0, s2, items, tab, n, n_all, 1, p);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = hipEventRecord(join, s2)) != hipSuccess) return e;
  switch ((bvh ? (dyn != 0u ? 2 : 1) : 0)) {
    case 2:
      hipLaunchKernelGGL(kb_trace<2>, dim3(blocks[BK_TRACE]), T256, dyn, s, items, tab, n, n_all);
      break;
    case 1:
      hipLaunchKernelGGL(kb_trace<1>, dim3(blocks[BK_TRACE]), T256, 0, s, items, tab, n, n_all);
      break;
    default:
      hipLaunchKernelGGL(kb_trace<0>, dim3(blocks[BK_TRACE]), T256, 0, s, items, tab, n, n_all);
  }
  hipLaunchKernelGGL(kb_scan_photons, dim3(n), T1024, 0, s, items, n);
  hipLaunchKernelGGL(kb_grid_setup, dim3(n), T64, 0, s, items, n);
  hipLaunchKernelGGL(kb_grid_keys, dim3(blocks[BK_MORTON]), T256, 0, s, items, tab, n, n_all);
  // the last digit pass also moves the records into bucket order
  hipLaunchKernelGGL(kb_os_hist, dim3(blocks[BK_PHIST]), T1024, 0, s, items, tab, n, n_all, 0);
  for (uint32_t p = 0; p + 1u < GRID_KEY_BITS / 8u; ++p)
    hipLaunchKernelGGL(kb_os_scatter<false>, dim3(blocks[BK_POS]), TOS, 0, s, items, tab, n, n_all, 0, p);
  hipLaunchKernelGGL(kb_os_scatter<true>, dim3(blocks[BK_POS]), TOS, 0, s, items, tab, n, n_all, 0,
                     GRID_KEY_BITS / 8u - 1u);
  hipLaunchKernelGGL(kb_grid_clear, dim3(blocks[BK_CLEAR]), T256, 0, s, items, tab, n, n_all);
  hipLaunchKernelGGL(kb_grid_cells, dim3(blocks[BK_LEAVES]), T256, 0, s, items, tab, n, n_all);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = hipStreamWaitEvent(s, join, 0)) != hipSuccess) return e;
  switch ((bvh ? (dyn != 0u ? 2 : 1) : 0)) {
    case 2:
      hipLaunchKernelGGL(kb_shade<2>, dim3(blocks[BK_SHADE]), T256, dyn, s, items, tab, n, n_all);
      break;
    case 1:
      hipLaunchKernelGGL(kb_shade<1>, dim3(blocks[BK_SHADE]), T256, 0, s, items, tab, n, n_all);
      break;
    default:
      hipLaunchKernelGGL(kb_shade<0>, dim3(blocks[BK_SHADE]), T256, 0, s, items, tab, n, n_all);
  }
  hipLaunchKernelGGL(kb_overflow_add, dim3(blocks[BK_OVF]), T256, 0, s, items, tab, n, n_all);
  hipLaunchKernelGGL(kb_accumulate, dim3(blocks[BK_PIX]), T256, 0, s, items, tab, n, n_all);
  hipLaunchKernelGGL(kb_pass_end, dim3(n), T64, 0, s, items, n);
  return hipGetLastError();
}

}  // namespace ppm

// ===========================================================================
// The gather without allocation (ppm_gather_noalloc), for callers that cannot
// allocate device memory.  Its buffers live in `scratch`, a fixed area of the
// library's code object: one copy per GPU, held until the process ends,
// whose device address noalloc_init() takes once per device (the array's
// name is only a host-side stand-in; kernels get the address).  Every call
// on a GPU reuses that GPU's area, so two calls that overlap on one GPU
// race; whatever must start at zero (the shade's work counters, its flags)
// is cleared by the call's own kernels, never assumed from the last call.
namespace ppm {

// Layout: positions, normals and results as doubles (NA_D3 each), then
// either the tree query's float records (NA_F4 each) or the shade's queries,
// order, results and pass state.
constexpr size_t na_align(size_t v) { return (v + 255) & ~(size_t)255; }
constexpr size_t NA_D3 = na_align(3ull * NOALLOC_MAX_QUERIES * sizeof(double));
constexpr size_t NA_F4 = na_align((size_t)NOALLOC_MAX_QUERIES * sizeof(float4));
constexpr size_t NA_TREE = 3 * NA_F4;
constexpr size_t NA_GRID = na_align((size_t)NOALLOC_MAX_QUERIES * sizeof(Query)) +
                           na_align((size_t)NOALLOC_MAX_QUERIES * sizeof(uint32_t)) + NA_F4 +
                           na_align(sizeof(PassState));
constexpr size_t NA_BYTES = 3 * NA_D3 + (NA_TREE > NA_GRID ? NA_TREE : NA_GRID);
static_assert(NA_BYTES <= NOALLOC_SCRATCH_MAX_BYTES, "the no-allocation area outgrew its stated bound");

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

static std::mutex g_scratch_mu;
static unsigned char* g_scratch_dev[NOALLOC_MAX_DEVICES] = {};

hipError_t noalloc_init(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device < 0 || device >= (int)NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
  if (g_scratch_dev[device]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (!p) return hipErrorInvalidSymbol;
  g_scratch_dev[device] = static_cast<unsigned char*>(p);
  return hipSuccess;
}

namespace {

struct NaLayout {
  double* q64;    // query positions (host doubles)
  double* n64;    // query normals
  double* out64;  // results (doubles, copied back)
  float4* qf;     // tree: float positions
  float4* qn;     // tree: float normals
  float4* rgb;    // tree: float sums
  Query* q;       // grid: shade queries
  uint32_t* order;
  float4* res;    // grid: shade results
  PassState* st;  // grid: the shade's pass state
};

NaLayout na_layout(unsigned char* base) {
  NaLayout L;
  L.q64 = reinterpret_cast<double*>(base);
  L.n64 = reinterpret_cast<double*>(base + NA_D3);
  L.out64 = reinterpret_cast<double*>(base + 2 * NA_D3);
  unsigned char* r = base + 3 * NA_D3;
  L.qf = reinterpret_cast<float4*>(r);
  L.qn = reinterpret_cast<float4*>(r + NA_F4);
  L.rgb = reinterpret_cast<float4*>(r + 2 * NA_F4);
  L.q = reinterpret_cast<Query*>(r);
  unsigned char* o = r + na_align((size_t)NOALLOC_MAX_QUERIES * sizeof(Query));
  L.order = reinterpret_cast<uint32_t*>(o);
  o += na_align((size_t)NOALLOC_MAX_QUERIES * sizeof(uint32_t));
  L.res = reinterpret_cast<float4*>(o);
  L.st = reinterpret_cast<PassState*>(o + NA_F4);
  return L;
}

}  // namespace

// Host doubles -> the float records of the tree query (as run_query forms
// them: each coordinate rounded to float).
__global__ __launch_bounds__(256) void k_na_tree_prep(const double* q64, const double* n64, float4* qf, float4* qn,
                                                      uint32_t nq, int with_nrm) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  qf[i] = make_float4((float)q64[3 * i], (float)q64[3 * i + 1], (float)q64[3 * i + 2], 0.0f);
  if (with_nrm) qn[i] = make_float4((float)n64[3 * i], (float)n64[3 * i + 1], (float)n64[3 * i + 2], 0.0f);
}

// Host doubles -> shade queries (weight 1, identity order), as run_grid_gather forms them.
__global__ __launch_bounds__(256) void k_na_grid_prep(const double* q64, const double* n64, Query* q,
                                                      uint32_t* order, uint32_t nq) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  Query r;
  r.a = make_float4((float)q64[3 * i], (float)q64[3 * i + 1], (float)q64[3 * i + 2], __uint_as_float(i));
  r.b = make_float4((float)n64[3 * i], (float)n64[3 * i + 1], (float)n64[3 * i + 2], __uint_as_float(0u));
  r.w = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
  q[i] = r;
  order[i] = i;
}

// The shade's pass state: the map's, with this call's radius, scale and
// query count, the flags and work counters cleared.
__global__ __launch_bounds__(256) void k_na_grid_state(const PassState* src, PassState* dst, double r2, float scale,
                                                       uint32_t nq) {
  constexpr uint32_t words = sizeof(PassState) / 4u;
  const uint32_t* s = reinterpret_cast<const uint32_t*>(src);
  uint32_t* d = reinterpret_cast<uint32_t*>(dst);
  for (uint32_t w = threadIdx.x; w < words; w += blockDim.x) d[w] = s[w];
  __syncthreads();
  uint32_t* work = &dst->work[0][0][0];
  for (uint32_t w = threadIdx.x; w < 2u * 8u * 32u; w += blockDim.x) work[w] = 0u;
  if (threadIdx.x == 0) {
    dst->r2 = r2;
    dst->r2f = (float)r2;
    dst->scale = scale;
    dst->n_qsort = nq;
    dst->overflow = 0u;
  }
}

// Float sums -> doubles times k (the tree's power / pi / r2, formed on the
// host exactly as ppm_gather forms it; 1 for the grid, already scaled).
__global__ __launch_bounds__(256) void k_na_finish(const float4* rgb, double* out, uint32_t nq, double k) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  const float4 v = rgb[i];
  out[3 * i] = (double)v.x * k;
  out[3 * i + 1] = (double)v.y * k;
  out[3 * i + 2] = (double)v.z * k;
}

hipError_t launch_gather_noalloc(hipStream_t s, int device, const MapBuf& m, const PassState* map_st, bool grid,
                                 const SceneBuf& grid_scene, const double* q_xyz, const double* q_nrm, uint32_t nq,
                                 double r2, double k, float scale, int refine64, int pfilter, int n_cu,
                                 double* rgb_out, uint32_t* overflow_out) {
  if (device < 0 || device >= (int)NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
  unsigned char* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    base = g_scratch_dev[device];
  }
  if (!base) return hipErrorNotInitialized;
  if (nq == 0 || nq > NOALLOC_MAX_QUERIES) return hipErrorInvalidValue;
  const NaLayout L = na_layout(base);
  const dim3 blocks((nq + 255) / 256);
  hipError_t e;
  if ((e = hipMemcpyAsync(L.q64, q_xyz, sizeof(double) * 3 * nq, hipMemcpyHostToDevice, s)) != hipSuccess) return e;
  if ((e = hipMemcpyAsync(L.n64, q_nrm, sizeof(double) * 3 * nq, hipMemcpyHostToDevice, s)) != hipSuccess) return e;
  *overflow_out = 0u;
  if (grid) {
    hipLaunchKernelGGL(k_na_grid_prep, blocks, dim3(256), 0, s, L.q64, L.n64, L.q, L.order, nq);
    hipLaunchKernelGGL(k_na_grid_state, dim3(1), dim3(256), 0, s, map_st, L.st, r2, scale, nq);
    EyeBuf eb;
    eb.q = L.q;
    eb.res = L.res;
    if ((e = launch_grid_shade(s, grid_scene, m, L.st, eb, L.order, pfilter, n_cu)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_na_finish, blocks, dim3(256), 0, s, L.res, L.out64, nq, 1.0);
    if ((e = hipMemcpyAsync(overflow_out, &L.st->overflow, sizeof(uint32_t), hipMemcpyDeviceToHost, s)) !=
        hipSuccess)
      return e;
  } else {
    hipLaunchKernelGGL(k_na_tree_prep, blocks, dim3(256), 0, s, L.q64, L.n64, L.qf, L.qn, nq, 1);
    if ((e = launch_query(s, m, const_cast<PassState*>(map_st), L.qf, refine64 ? L.q64 : nullptr, L.qn, nq, r2,
                          refine64, 2, pfilter, nullptr, nullptr, nullptr, L.rgb, n_cu)) != hipSuccess)
      return e;
    hipLaunchKernelGGL(k_na_finish, blocks, dim3(256), 0, s, L.rgb, L.out64, nq, k);
  }
  if ((e = hipGetLastError()) != hipSuccess) return e;
  return hipMemcpyAsync(rgb_out, L.out64, sizeof(double) * 3 * nq, hipMemcpyDeviceToHost, s);
}

}  // namespace ppm
