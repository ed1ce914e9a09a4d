// ppm_api.cpp — the C ABI of include/ppm.h over the gfx950 kernels.
//
// Host responsibilities: validate inputs (the reference panics where this
// returns a status), lay scenes out for the device (ppm_layout.hpp), size
// buffers, order launches on the context's stream and capture one PPM-PA
// pass into a hipGraph that is replayed for every pass.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ppm.h"
#include "ppm_launch.hpp"
#include "ppm_layout.hpp"

using namespace ppm;

// Streams a context makes for its sessions and batches on first use and
// keeps until ppm_destroy, so that creating and destroying sessions creates
// and destroys no streams: AUX_PIPE + j - 1 runs pass pipeline j >= 1 (its
// graph launches), AUX_EYE + j the eye branch of pipeline j (forked into a
// pass graph's capture, or used directly without graphs), AUX_CAPTURE and
// AUX_CAPTURE2 the streams pass graphs are captured on.  Calls on a context
// never overlap (each returns once its work is done), so its sessions and
// batches may share them.
constexpr int AUX_PIPE = 0, AUX_EYE = 3, AUX_CAPTURE = 7, AUX_CAPTURE2 = 8, AUX_N = 9;

struct ppm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int n_cu = 256;
  std::string err;
  hipStream_t aux[AUX_N] = {};
};

struct ppm_scene {
  ppm_ctx* ctx = nullptr;
  SceneBuf buf;
  std::vector<DevLight> lights;  // host copy (photon counts are filled per call)
  double total_flux = 0.0;
  double bmin[3] = {0, 0, 0}, bmax[3] = {1, 1, 1};
  // as created, for ppm_scene_set_params
  std::vector<DevMaterial> mats0;
  std::vector<float> flux0;
  double total_flux0 = 0.0;
  double light_scale = 1.0;
  std::vector<DevShape> lin_host;  // the shapes tested one by one, as on the device
};

struct ppm_photons {
  ppm_ctx* ctx = nullptr;
  uint64_t cap = 0;
  PhotonBuf pb;
  PassState* st = nullptr;
  bool has64 = false;
  bool has_bounds = false;
  double bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};
  double power = 1.0;
};

struct ppm_map {
  ppm_ctx* ctx = nullptr;
  MapBuf m;
  PassState* st = nullptr;
  double power = 1.0;
  bool has64 = false;
  bool built = false;
  // built by ppm_map_build_grid: the passes' hashed grid for radius^2
  // grid_r2 over the box [gbmin, gbmax] (cells allocated on first use)
  bool grid = false;
  double grid_r2 = 0.0;
  float4 gbmin{}, gbmax{};
  DevMaterial* hook_mat = nullptr;  // M_AIR alone: the scene of the shade kernel run by ppm_gather
};

// One pass pipeline of a session: the buffers, pass state, FP64 sums and
// streams of every K-th pass (K = the session's pipelines), and its captured
// pass graph.  Passes of different pipelines run at once, as the reference's
// iterator runs NPARA pm | rt processes at once (util/iterator.rb:18,90-117).
struct Pipe {
  PhotonBuf pb;
  MapBuf map;
  EyeBuf eb;
  double* sum = nullptr;
  double* sumsq = nullptr;
  PassState* st = nullptr;
  hipStream_t stream = nullptr;  // this pipeline's passes, in order
  hipStream_t s2 = nullptr;      // the eye branch of its passes
  hipEvent_t evf = nullptr, evj = nullptr, done = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

struct ppm_session {
  ppm_ctx* ctx = nullptr;
  const ppm_scene* scene = nullptr;
  ppm_pass_cfg cfg{};
  SceneBuf sbuf;
  DevLight* lights = nullptr;
  double power = 0.0;
  uint32_t passes_done = 0;
  std::vector<Pipe> pipes;  // pass p (from cfg.first_pass) runs on pipes[p % K]
  hipEvent_t ev[10] = {};
  ppm_timings t{};
};

namespace {

// The context's stream i (made on first use).
hipError_t ctx_stream(ppm_ctx* c, int i, hipStream_t* out) {
  if (!c->aux[i]) {
    const hipError_t e = hipStreamCreateWithFlags(&c->aux[i], hipStreamNonBlocking);
    if (e != hipSuccess) {
      c->aux[i] = nullptr;
      return e;
    }
  }
  *out = c->aux[i];
  return hipSuccess;
}

// Replace the context's stream i (after a capture that failed part-way, which
// may leave a stream forked into it in an unknown capture state).
void ctx_stream_renew(ppm_ctx* c, int i) {
  if (c->aux[i]) {
    (void)hipStreamSynchronize(c->aux[i]);
    (void)hipStreamDestroy(c->aux[i]);
  }
  c->aux[i] = nullptr;
}

ppm_status fail(ppm_ctx* c, ppm_status s, const std::string& msg) {
  if (c) c->err = msg;
  return s;
}

#define HIP_TRY(ctx, expr)                                                                  \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return fail((ctx), e_ == hipErrorOutOfMemory ? PPM_E_OOM : PPM_E_HIP,                 \
                  std::string(#expr) + ": " + hipGetErrorString(e_));                       \
  } while (0)

template <typename T>
hipError_t dalloc(T** p, uint64_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  return hipMalloc((void**)p, count * sizeof(T));
}

template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

// One-sweep sort state for up to cap keys (the look-back words count keys + 1
// in 31 bits).
hipError_t onesweep_alloc(OneSweep& o, uint64_t cap, int n_cu) {
  if (cap >= 0x7FFFFFFFull) return hipErrorInvalidValue;
  o.cap = cap;
  o.tiles = (uint32_t)std::max<uint64_t>(1, (cap + OS_TILE - 1) / OS_TILE);
  o.blocks = std::max<uint32_t>(1, std::min<uint32_t>((uint32_t)n_cu * OS_BLOCKS_PER_CU, o.tiles));
  hipError_t e;
  if ((e = dalloc(&o.ctl, OS_CTL_WORDS)) != hipSuccess) return e;
  if ((e = hipMemset(o.ctl, 0, sizeof(uint32_t) * OS_CTL_WORDS)) != hipSuccess) return e;
  return dalloc(&o.look, (uint64_t)OS_MAX_DIGITS * o.tiles * 256u);
}

void onesweep_free(OneSweep& o) {
  dfree(o.ctl);
  dfree(o.look);
}

float4 f4(const double v[3], float w = 0.0f) { return make_float4((float)v[0], (float)v[1], (float)v[2], w); }

bool finite3(const double v[3]) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }

void cross3(const double a[3], const double b[3], double o[3]) {
  o[0] = a[1] * b[2] - b[1] * a[2];
  o[1] = a[2] * b[0] - b[2] * a[0];
  o[2] = a[0] * b[1] - b[0] * a[1];
}
bool norm3(const double a[3], double o[3]) {
  double n = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  if (n == 0.0) return false;
  double inv = 1.0 / n;
  o[0] = a[0] * inv;
  o[1] = a[1] * inv;
  o[2] = a[2] * inv;
  return true;
}

uint32_t next_pow2(uint64_t v) {
  uint32_t p = 1;
  while (p < v) p <<= 1;
  return p;
}

// per-light photon counts (ppmpa.rs:30-31,65-72): power = Σ flux / n; n_l = round(flux_l / power)
uint64_t assign_photons(std::vector<DevLight>& L, double total_flux, uint64_t n_photon, double* power) {
  double pw = total_flux / (double)n_photon;
  *power = pw;
  uint64_t first = 0;
  for (auto& l : L) {
    double nl = std::round((double)l.flux / pw);
    uint64_t c = (nl > 0 && std::isfinite(nl)) ? (uint64_t)nl : 0;
    l.first_photon = (uint32_t)first;
    l.n_photon = (uint32_t)c;
    first += c;
  }
  return first;
}

ppm_status check_cfg_depth(ppm_ctx* c, int max_depth) {
  if (max_depth < 1 || max_depth > MAX_DEPTH_LIMIT)
    return fail(c, PPM_E_INVALID, "max_depth must be in [1, 16]");
  return PPM_OK;
}

// A map for cap photons: the box tree (grid = false) or the per-pass grid.
hipError_t map_alloc(MapBuf& m, uint64_t cap, int n_cu, bool grid) {
  m.cap = cap;
  m.sort_blocks = (uint32_t)std::min<uint64_t>(4096, std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)n_cu * 4,
                                                                                                (cap + 2047) / 2048)));
  hipError_t e;
  // the grid's gathers read up to GRID_READ_SLACK records past the last photon
  if ((e = dalloc(&m.sorted, cap + GRID_READ_SLACK)) != hipSuccess) return e;
  if ((e = dalloc(&m.src_idx, cap)) != hipSuccess) return e;
  if (grid) {
    m.cells_max = (uint32_t)std::min<uint64_t>(GRID_MAX_BUCKETS, next_pow2(std::max<uint64_t>(cap, 1024)));
    if ((e = dalloc(&m.cells, m.cells_max)) != hipSuccess) return e;
    if ((e = onesweep_alloc(m.os, cap, n_cu)) != hipSuccess) return e;
  } else {
    const uint64_t leaves = (cap + LEAF - 1) / LEAF;
    m.lpad_max = next_pow2(leaves > 0 ? leaves : 1);
    if ((e = dalloc(&m.nodes, 2ull * m.lpad_max)) != hipSuccess) return e;
  }
  for (int i = 0; i < 2; ++i) {
    if ((e = dalloc(&m.keys[i], cap)) != hipSuccess) return e;
    if ((e = dalloc(&m.vals[i], cap)) != hipSuccess) return e;
  }
  if ((e = dalloc(&m.ghist, 256ull * m.sort_blocks)) != hipSuccess) return e;
  if ((e = dalloc(&m.totals, 256)) != hipSuccess) return e;
  return hipSuccess;
}

void map_free(MapBuf& m) {
  dfree(m.sorted);
  dfree(m.sorted64);
  dfree(m.src_idx);
  dfree(m.nodes);
  dfree(m.cells);
  for (int i = 0; i < 2; ++i) {
    dfree(m.keys[i]);
    dfree(m.vals[i]);
  }
  dfree(m.ghist);
  dfree(m.totals);
  onesweep_free(m.os);
}

hipError_t photons_alloc(PhotonBuf& pb, uint64_t cap, uint32_t max_regions) {
  pb.cap = cap;
  pb.max_regions = max_regions;
  hipError_t e;
  if ((e = dalloc(&pb.rec, cap)) != hipSuccess) return e;
  if ((e = dalloc(&pb.count, max_regions)) != hipSuccess) return e;
  if ((e = dalloc(&pb.offset, max_regions)) != hipSuccess) return e;
  return hipSuccess;
}

void photons_free(PhotonBuf& pb) {
  dfree(pb.rec);
  dfree(pb.rec64);
  dfree(pb.count);
  dfree(pb.offset);
}

// Region layout of a traced pass: one region per tracing wave.
bool photons_layout(PhotonBuf& pb, uint64_t n_photon, int max_depth, int n_cu) {
  const uint32_t w = trace_waves(n_photon, n_cu);
  const uint64_t per = (n_photon + w - 1) / w;
  const uint64_t rc = per * (uint64_t)max_depth;
  if (w > pb.max_regions || rc > 0xFFFFFFFFull || (uint64_t)w * rc > pb.cap) return false;
  pb.n_regions = w;
  pb.region_cap = (uint32_t)rc;
  return true;
}

uint64_t photons_needed(uint64_t n_photon, int max_depth, int n_cu) {
  const uint32_t w = trace_waves(n_photon, n_cu);
  const uint64_t per = (n_photon + w - 1) / w;
  return (uint64_t)w * per * (uint64_t)max_depth;
}

// Eye tiles of camera cam at tile shape `shift` ((1 << shift) x (64 >> shift)
// pixels); tiles per row in *tx.
uint32_t cam_tiles(const DevCamera& cam, uint32_t shift, uint32_t* tx = nullptr) {
  const uint32_t w = 1u << shift, h = 64u >> shift;
  const uint32_t x = ((uint32_t)cam.xreso + w - 1u) / w, y = ((uint32_t)cam.yreso + h - 1u) / h;
  if (tx) *tx = x;
  return x * y;
}
constexpr uint32_t TILE_SHIFT_MIN = 2, TILE_SHIFT_MAX = 5, TILE_SHIFT_DEFAULT = 3;
uint32_t cam_tiles_max(const DevCamera& cam) {
  uint32_t m = 0;
  for (uint32_t k = TILE_SHIFT_MIN; k <= TILE_SHIFT_MAX; ++k) m = std::max(m, cam_tiles(cam, k));
  return m;
}

// The eye buffers at tile shape `shift` (sized for any shape by eye_alloc);
// the tiles' query counts of earlier passes, which only order the eye's
// work, start again from zero.
hipError_t eye_set_tiles(EyeBuf& eb, DevCamera& cam, uint32_t shift) {
  uint32_t tx = 0;
  const uint32_t tiles = cam_tiles(cam, shift, &tx);
  if (tiles > eb.tiles_cap) return hipErrorInvalidValue;
  cam.tile_shift = (int32_t)shift;
  cam.tiles_x = (int32_t)tx;
  eb.n_tiles = tiles;
  eb.q_primary = tiles * 64u;
  eb.q_fine = eb.q_primary >= QUERY_FINE_PIXELS ? 1u : 0u;
  return hipMemset(eb.tile_count, 0, sizeof(uint32_t) * eb.tiles_cap);
}

hipError_t eye_alloc(EyeBuf& eb, const DevCamera& cam, int max_depth, int n_cu) {
  // buffers for the most tiles any tile shape needs, so that a batch can
  // tune the shape without reallocating
  const uint32_t tiles = cam_tiles_max(cam);
  eb.tiles_cap = tiles;
  eb.q_primary = tiles * 64u;
  // deeper queries kept per pixel: a glass sphere yields about one per depth
  eb.sec_max = (uint32_t)std::max(1, std::min(2 * max_depth, 32));
  eb.n_pix = (uint32_t)cam.xreso * (uint32_t)cam.yreso;
  eb.ovf_cap = 2u * eb.n_pix + 1024u;
  const uint64_t inplace = (uint64_t)eb.q_primary * (1u + eb.sec_max);
  hipError_t e;
  if ((e = dalloc(&eb.q, inplace + eb.ovf_cap)) != hipSuccess) return e;
  if ((e = dalloc(&eb.qpos, inplace + eb.ovf_cap)) != hipSuccess) return e;
  if ((e = dalloc(&eb.sec_count, eb.q_primary)) != hipSuccess) return e;
  if ((e = dalloc(&eb.pix, eb.n_pix)) != hipSuccess) return e;
  if ((e = dalloc(&eb.tile_count, tiles)) != hipSuccess) return e;
  if ((e = hipMemset(eb.tile_count, 0, sizeof(uint32_t) * tiles)) != hipSuccess) return e;
  if ((e = dalloc(&eb.tile_off, tiles)) != hipSuccess) return e;
  if ((e = dalloc(&eb.tile_order, tiles)) != hipSuccess) return e;
  for (int i = 0; i < 2; ++i) {
    if ((e = dalloc(&eb.qkeys[i], inplace + eb.ovf_cap)) != hipSuccess) return e;
    if ((e = dalloc(&eb.qvals[i], inplace + eb.ovf_cap)) != hipSuccess) return e;
  }
  if ((e = onesweep_alloc(eb.qos, inplace + eb.ovf_cap, n_cu)) != hipSuccess) return e;
  if ((e = dalloc(&eb.res, inplace + eb.ovf_cap)) != hipSuccess) return e;
  eb.estack_lanes = 64u * std::max<uint32_t>(1, std::min<uint32_t>(tiles, (uint32_t)n_cu * EYE_WAVES_PER_CU_MAX));
  eb.estack_depth = (uint32_t)std::max(1, std::min(max_depth, MAX_DEPTH_LIMIT));
  if ((e = dalloc(&eb.estack, 3ull * eb.estack_depth * eb.estack_lanes)) != hipSuccess) return e;
  DevCamera c = cam;
  return eye_set_tiles(eb, c, (uint32_t)cam.tile_shift);
}

void eye_free(EyeBuf& eb) {
  dfree(eb.q);
  dfree(eb.qpos);
  dfree(eb.sec_count);
  dfree(eb.pix);
  dfree(eb.tile_count);
  dfree(eb.tile_off);
  dfree(eb.tile_order);
  for (int i = 0; i < 2; ++i) {
    dfree(eb.qkeys[i]);
    dfree(eb.qvals[i]);
  }
  onesweep_free(eb.qos);
  dfree(eb.res);
  dfree(eb.estack);
}

// What the overflow flags of a pass state say (device-side bits 1 photon
// region, 2 overflow queries, 4 a query's cells past the grid's span, 8 the
// eye's pending pool, 16 a sort's look-back that gave up waiting).
std::string overflow_text(uint32_t f) {
  static const char* what[] = {"a photon region's capacity", "the overflow query slots",
                               "a query's cells past the grid's span", "the eye's pending-node pool",
                               "a sort's look-back wait (results of that pass are wrong)"};
  std::string t;
  for (int b = 0; b < 5; ++b)
    if (f & (1u << b)) t += (t.empty() ? "" : ", ") + std::string(what[b]);
  return t.empty() ? "flags " + std::to_string(f) : t;
}

PassState host_state() {
  PassState s;
  std::memset(&s, 0, sizeof(s));
  return s;
}

}  // namespace

// The map's buffers as a box tree (no grid cells, so the gather launch picks
// the tree kernels) or as the passes' grid (no tree nodes).
static MapBuf tree_view(const ppm_map* m) {
  MapBuf b = m->m;
  b.cells = nullptr;
  b.cells_max = 0;
  return b;
}
static MapBuf grid_view_of(const ppm_map* m) {
  MapBuf b = m->m;
  b.nodes = nullptr;
  b.lpad_max = 0;
  return b;
}
static MapBuf map_view(const ppm_map* m) { return m->grid ? grid_view_of(m) : tree_view(m); }

namespace {}  // namespace

extern "C" {

const char* ppm_status_string(ppm_status s) {
  switch (s) {
    case PPM_OK: return "ok";
    case PPM_E_INVALID: return "invalid argument";
    case PPM_E_NONFINITE: return "non-finite coordinate";
    case PPM_E_OOM: return "out of memory";
    case PPM_E_HIP: return "HIP runtime error";
    case PPM_E_RANGE: return "size beyond capacity";
    case PPM_E_PARSE: return "malformed photon map";
    case PPM_E_NODEVICE: return "no usable device";
    default: return "unknown status";
  }
}

ppm_status ppm_create(ppm_ctx** out, int device, void* hip_stream) {
  if (!out) return PPM_E_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return PPM_E_NODEVICE;
  if (device < 0 || device >= n) return PPM_E_NODEVICE;
  if (hipSetDevice(device) != hipSuccess) return PPM_E_NODEVICE;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return PPM_E_NODEVICE;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return PPM_E_NODEVICE;
  // the device address of this GPU's area for ppm_gather_noalloc (once per device)
  if (noalloc_init(device) != hipSuccess) return PPM_E_HIP;
  ppm_ctx* c = new ppm_ctx();
  c->device = device;
  c->n_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  if (hip_stream) {
    c->stream = (hipStream_t)hip_stream;
  } else {
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
      delete c;
      return PPM_E_HIP;
    }
    c->own_stream = true;
  }
  *out = c;
  return PPM_OK;
}

void ppm_destroy(ppm_ctx* c) {
  if (!c) return;
  for (hipStream_t& a : c->aux)
    if (a) {
      (void)hipStreamSynchronize(a);
      (void)hipStreamDestroy(a);
      a = nullptr;
    }
  if (c->own_stream && c->stream) {
    (void)hipStreamSynchronize(c->stream);
    (void)hipStreamDestroy(c->stream);
  }
  delete c;
}

const char* ppm_last_error(const ppm_ctx* c) { return c ? c->err.c_str() : "null context"; }

ppm_status ppm_sync(ppm_ctx* c) {
  if (!c) return PPM_E_INVALID;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return PPM_OK;
}

ppm_status ppm_camera_setup(ppm_camera_desc* cam, int32_t xreso, int32_t yreso, const double eye[3],
                            const double target[3], const double upper[3], double focus, double focal_len,
                            double f_number) {
  if (!cam || !eye || !target || !upper || xreso <= 0 || yreso <= 0 || focal_len == 0.0 || f_number == 0.0)
    return PPM_E_INVALID;
  // camera.rs:151-164
  double d[3] = {target[0] - eye[0], target[1] - eye[1], target[2] - eye[2]};
  double ez[3], ex[3], ey[3], t[3];
  if (!norm3(d, ez)) return PPM_E_INVALID;
  cross3(upper, ez, t);
  if (!norm3(t, ex)) return PPM_E_INVALID;
  cross3(ex, ez, t);
  if (!norm3(t, ey)) return PPM_E_INVALID;
  const double step = (focus * 0.035 / focal_len) / (double)xreso;
  const double ea = focal_len / f_number;
  const double lx = (double)(xreso / 2), ly = (double)(yreso / 2);
  cam->xreso = xreso;
  cam->yreso = yreso;
  for (int i = 0; i < 3; ++i) {
    cam->eye_pos[i] = eye[i];
    cam->esx[i] = step * ex[i];
    cam->esy[i] = step * ey[i];
    cam->eex[i] = ea * ex[i];
    cam->eey[i] = ea * ey[i];
  }
  for (int i = 0; i < 3; ++i) cam->origin[i] = focus * ez[i] - (lx - 0.5) * cam->esx[i] - (ly - 0.5) * cam->esy[i];
  return PPM_OK;
}

namespace {

// Shapes tested one by one and, when the scene asks for one (or has at least
// BVH_MIN_PRIMS bounded shapes in automatic mode), the device-built BVH over
// the bounded ones.  shapes[] is already on the device as sc->buf.shapes.
// Box of a bounded shape (sphere, polygon, parallelogram) in double.
static void shape_box(const DevShape& s, double lo[3], double hi[3]) {
  const double p[3] = {s.p.x, s.p.y, s.p.z};
  if (s.kind == SHAPE_SPHERE) {
    const double r = std::fabs((double)s.radius);
    for (int a = 0; a < 3; ++a) {
      lo[a] = p[a] - r;
      hi[a] = p[a] + r;
    }
    return;
  }
  const double d1[3] = {s.d1.x, s.d1.y, s.d1.z}, d2[3] = {s.d2.x, s.d2.y, s.d2.z};
  const double q = s.kind == SHAPE_PARALLELOGRAM ? 1.0 : 0.0;  // the fourth corner p + d1 + d2
  for (int a = 0; a < 3; ++a) {
    lo[a] = std::min({p[a], p[a] + d1[a], p[a] + d2[a], p[a] + q * (d1[a] + d2[a])});
    hi[a] = std::max({p[a], p[a] + d1[a], p[a] + d2[a], p[a] + q * (d1[a] + d2[a])});
  }
}

static ppm_status scene_layout_bvh(ppm_ctx* c, ppm_scene* sc, const std::vector<DevShape>& shapes, int flags);

// Padded float boxes of the shapes tested one by one, for the shadow test's
// candidate culling: a plane is marked by lo.w = 1, a point gets an empty box.
static hipError_t upload_lin_boxes(SceneBuf& b, const std::vector<DevShape>& lin) {
  std::vector<Node> boxes(std::max<size_t>(lin.size(), 1));
  for (size_t i = 0; i < lin.size(); ++i) {
    const DevShape& s = lin[i];
    Node& nd = boxes[i];
    if (s.kind == SHAPE_PLAIN) {
      nd.lo = make_float4(0.f, 0.f, 0.f, 1.f);
      nd.hi = make_float4(0.f, 0.f, 0.f, 0.f);
      continue;
    }
    if (s.kind != SHAPE_SPHERE && s.kind != SHAPE_POLYGON && s.kind != SHAPE_PARALLELOGRAM) {
      nd.lo = make_float4(INFINITY, INFINITY, INFINITY, 0.f);
      nd.hi = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.f);
      continue;
    }
    double lo[3], hi[3];
    shape_box(s, lo, hi);
    double m = 0.0;
    for (int a = 0; a < 3; ++a) m = std::max({m, std::fabs(lo[a]), std::fabs(hi[a])});
    const double pad = 2e-5 * (1.0 + m);
    nd.lo = make_float4(std::nextafter((float)(lo[0] - pad), -INFINITY), std::nextafter((float)(lo[1] - pad), -INFINITY),
                        std::nextafter((float)(lo[2] - pad), -INFINITY), 0.f);
    nd.hi = make_float4(std::nextafter((float)(hi[0] + pad), INFINITY), std::nextafter((float)(hi[1] + pad), INFINITY),
                        std::nextafter((float)(hi[2] + pad), INFINITY), 0.f);
  }
  hipError_t e = dalloc(&b.lin_box, boxes.size());
  if (e == hipSuccess) e = hipMemcpy(b.lin_box, boxes.data(), boxes.size() * sizeof(Node), hipMemcpyHostToDevice);
  return e;
}

// Shapes tested one by one and, when the scene asks for one (or has at least
// BVH_MIN_PRIMS bounded shapes in automatic mode), the device-built BVH over
// the bounded shapes whose box sides are all within a quarter of the scene's
// extent; planes and larger shapes (walls) are tested one by one.
// shapes[] is already on the device as sc->buf.shapes.
static ppm_status scene_layout(ppm_ctx* c, ppm_scene* sc, const std::vector<DevShape>& shapes, int flags) {
  const ppm_status r = scene_layout_bvh(c, sc, shapes, flags);
  if (r != PPM_OK) return r;
  const hipError_t e = upload_lin_boxes(sc->buf, sc->lin_host);
  if (e != hipSuccess) return fail(c, e == hipErrorOutOfMemory ? PPM_E_OOM : PPM_E_HIP, hipGetErrorString(e));
  return PPM_OK;
}

static ppm_status scene_layout_bvh(ppm_ctx* c, ppm_scene* sc, const std::vector<DevShape>& shapes, int flags) {
  SceneBuf& b = sc->buf;
  std::vector<uint32_t> bounded;
  for (size_t i = 0; i < shapes.size(); ++i) {
    const uint32_t k = shapes[i].kind;
    if (k == SHAPE_SPHERE || k == SHAPE_POLYGON || k == SHAPE_PARALLELOGRAM) bounded.push_back((uint32_t)i);
  }
  const bool use_bvh = flags == PPM_SCENE_BVH_ON || (flags == PPM_SCENE_BVH_AUTO && bounded.size() >= BVH_MIN_PRIMS);
  if (!use_bvh || bounded.empty()) {
    b.lin = b.shapes;
    b.n_lin = b.n_shapes;
    sc->lin_host = shapes;
    return PPM_OK;
  }
  double smin[3] = {INFINITY, INFINITY, INFINITY}, smax[3] = {-INFINITY, -INFINITY, -INFINITY};
  std::vector<double> boxes(6 * bounded.size());
  for (size_t j = 0; j < bounded.size(); ++j) {
    double* lo = &boxes[6 * j];
    double* hi = lo + 3;
    shape_box(shapes[bounded[j]], lo, hi);
    for (int a = 0; a < 3; ++a) {
      smin[a] = std::min(smin[a], lo[a]);
      smax[a] = std::max(smax[a], hi[a]);
    }
  }
  const double size = std::max({smax[0] - smin[0], smax[1] - smin[1], smax[2] - smin[2]});
  std::vector<uint32_t> in_bvh;
  std::vector<DevShape> lin;
  double cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
  std::vector<char> large(shapes.size(), 0);
  auto wide = [&](size_t j) {
    const double* lo = &boxes[6 * j];
    const double* hi = lo + 3;
    const double side = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]});
    return side > 0.25 * size && bounded.size() > 1;
  };
  size_t n_wide = 0;
  for (size_t j = 0; j < bounded.size(); ++j) n_wide += wide(j) ? 1 : 0;
  // A ParallelogramLight's own panel (a parallelogram with exactly its
  // corner and sides) stays in the list as well: there the shadow test knows
  // that a ray to one of the light's samples meets it, and no walk of the
  // tree has to reach the panel at the end of every shadow ray (64 rooms of
  // 2-512 spheres: 3.11 -> 2.91 s; the 327,744-shape mesh scene 80.2 ->
  // 78.5 ms per pass).  Beside a large tree the panel's centre still bounds
  // the Morton box: without it the box, and with it the tree over the mesh,
  // changed, and that scene got slower in every stage (79.4 -> 83.1 ms).
  const bool small_tree = bounded.size() - n_wide <= BVH_SMALL_PRIMS;
  auto lamp = [&](const DevShape& sh) {
    if (sh.kind != SHAPE_PARALLELOGRAM) return false;
    for (const DevLight& l : sc->lights)
      if (l.kind == LIGHT_PARALLELOGRAM && sh.p.x == l.pos.x && sh.p.y == l.pos.y && sh.p.z == l.pos.z &&
          sh.d1.x == l.d1.x && sh.d1.y == l.d1.y && sh.d1.z == l.d1.z && sh.d2.x == l.d2.x && sh.d2.y == l.d2.y &&
          sh.d2.z == l.d2.z)
        return true;
    return false;
  };
  for (size_t j = 0; j < bounded.size(); ++j) {
    const double* lo = &boxes[6 * j];
    const double* hi = lo + 3;
    const bool is_lamp = !wide(j) && lamp(shapes[bounded[j]]);
    if (is_lamp && !small_tree)
      for (int a = 0; a < 3; ++a) {
        const double cc = 0.5 * (lo[a] + hi[a]);
        cmin[a] = std::min(cmin[a], cc);
        cmax[a] = std::max(cmax[a], cc);
      }
    if (wide(j) || is_lamp) {
      large[bounded[j]] = 1;
      continue;
    }
    in_bvh.push_back(bounded[j]);
    for (int a = 0; a < 3; ++a) {
      const double cc = 0.5 * (lo[a] + hi[a]);
      cmin[a] = std::min(cmin[a], cc);
      cmax[a] = std::max(cmax[a], cc);
    }
  }
  for (size_t i = 0; i < shapes.size(); ++i) {
    const uint32_t k = shapes[i].kind;
    if (large[i] || !(k == SHAPE_SPHERE || k == SHAPE_POLYGON || k == SHAPE_PARALLELOGRAM)) lin.push_back(shapes[i]);
  }
  if (in_bvh.empty()) {
    b.lin = b.shapes;
    b.n_lin = b.n_shapes;
    sc->lin_host = shapes;
    return PPM_OK;
  }
  const uint64_t n = in_bvh.size();
  const uint32_t per_leaf = n <= BVH_SMALL_PRIMS ? BVH_LEAF_SMALL : 1u;
  const uint64_t leaves = (n + per_leaf - 1) / per_leaf;
  uint32_t depth = 0;
  while ((1ull << depth) < leaves) ++depth;
  if (depth > BVH_MAX_DEPTH) return fail(c, PPM_E_RANGE, "too many shapes for the BVH (at most 16M)");
  float4 fmin = make_float4((float)cmin[0], (float)cmin[1], (float)cmin[2], 0.0f);
  float4 fscale;
  float* sc3[3] = {&fscale.x, &fscale.y, &fscale.z};
  for (int a = 0; a < 3; ++a) {
    const double ext = cmax[a] - cmin[a];
    *sc3[a] = ext > 0.0 ? (float)(1023.0 / ext) : 0.0f;
  }
  fscale.w = 0.0f;
  const uint32_t sort_blocks =
      (uint32_t)std::min<uint64_t>(4096, std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)c->n_cu * 4, (n + 2047) / 2048)));
  uint32_t* ids = nullptr;
  uint32_t *keys[2] = {nullptr, nullptr}, *vals[2] = {nullptr, nullptr};
  uint32_t *np = nullptr, *ghist = nullptr, *totals = nullptr;
  Node* pbox = nullptr;
  const uint32_t n32 = (uint32_t)n;
  hipError_t e;
  if ((e = dalloc(&ids, n)) == hipSuccess && (e = dalloc(&keys[0], n)) == hipSuccess &&
      (e = dalloc(&keys[1], n)) == hipSuccess && (e = dalloc(&vals[0], n)) == hipSuccess &&
      (e = dalloc(&vals[1], n)) == hipSuccess && (e = dalloc(&np, 1)) == hipSuccess &&
      (e = dalloc(&ghist, 256ull * sort_blocks)) == hipSuccess && (e = dalloc(&totals, 256)) == hipSuccess &&
      (e = dalloc(&pbox, n)) == hipSuccess && (e = dalloc(&b.bprims, n)) == hipSuccess &&
      (e = dalloc(&b.bvh, (2ull << depth) - 1)) == hipSuccess &&
      (e = dalloc(&b.lin, lin.size())) == hipSuccess &&
      (e = hipMemcpy(ids, in_bvh.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice)) == hipSuccess &&
      (e = hipMemcpy(np, &n32, sizeof(n32), hipMemcpyHostToDevice)) == hipSuccess &&
      (lin.empty() ||
       (e = hipMemcpy(b.lin, lin.data(), lin.size() * sizeof(DevShape), hipMemcpyHostToDevice)) == hipSuccess) &&
      (e = launch_bvh_build(c->stream, b.shapes, ids, n32, fmin, fscale, depth, per_leaf, keys, vals, np, ghist,
                            totals, sort_blocks, pbox, b.bprims, b.bvh)) == hipSuccess)
    e = hipStreamSynchronize(c->stream);
  dfree(ids);
  dfree(keys[0]);
  dfree(keys[1]);
  dfree(vals[0]);
  dfree(vals[1]);
  dfree(np);
  dfree(ghist);
  dfree(totals);
  dfree(pbox);
  if (e != hipSuccess) return fail(c, e == hipErrorOutOfMemory ? PPM_E_OOM : PPM_E_HIP, hipGetErrorString(e));
  b.n_lin = (uint32_t)lin.size();
  b.n_bprims = n32;
  b.bvh_depth = depth;
  b.bvh_leaf = per_leaf;
  sc->lin_host = lin;
  return PPM_OK;
}

}  // namespace

ppm_status ppm_scene_create(ppm_ctx* c, const ppm_scene_desc* d, ppm_scene** out) {
  if (!c || !d || !out) return fail(c, PPM_E_INVALID, "null argument");
  *out = nullptr;
  if (d->n_shapes < 0 || d->n_materials < 0 || d->n_lights < 0 || d->n_materials > 2046)
    return fail(c, PPM_E_INVALID, "bad counts (at most 2046 materials)");
  if ((d->n_shapes && !d->shapes) || (d->n_materials && !d->materials) || (d->n_lights && !d->lights))
    return fail(c, PPM_E_INVALID, "null array");
  if (d->camera.xreso <= 0 || d->camera.yreso <= 0) return fail(c, PPM_E_INVALID, "bad camera resolution");
  if ((uint64_t)((d->camera.xreso + 7) / 8) * (uint64_t)((d->camera.yreso + 7) / 8) > 65536)
    return fail(c, PPM_E_RANGE, "camera resolution beyond 65536 8x8 tiles");
  std::vector<DevShape> shapes(d->n_shapes);
  for (int i = 0; i < d->n_shapes; ++i) {
    const ppm_shape_desc& s = d->shapes[i];
    DevShape& o = shapes[i];
    std::memset(&o, 0, sizeof(o));
    if (s.kind < 0 || s.kind > 4) return fail(c, PPM_E_INVALID, "bad shape kind");
    if (s.material < 0 || s.material >= d->n_materials) return fail(c, PPM_E_INVALID, "shape material out of range");
    if (!finite3(s.pos) || !finite3(s.nvec) || !finite3(s.dir1) || !finite3(s.dir2) || !std::isfinite(s.dist) ||
        !std::isfinite(s.radius))
      return fail(c, PPM_E_NONFINITE, "non-finite shape parameter");
    o.kind = (uint32_t)s.kind;
    o.mat = (uint32_t)s.material;
    o.dist = (float)s.dist;
    o.radius = (float)s.radius;
    o.p = f4(s.pos);
    o.d1 = f4(s.dir1);
    o.d2 = f4(s.dir2);
    double n[3] = {s.nvec[0], s.nvec[1], s.nvec[2]};
    if ((s.kind == PPM_SHAPE_POLYGON || s.kind == PPM_SHAPE_PARALLELOGRAM) && n[0] == 0 && n[1] == 0 && n[2] == 0) {
      // Shape::new_polygon / new_parallelogram (geometry.rs:93-115)
      double cr[3];
      cross3(s.dir1, s.dir2, cr);
      if (!norm3(cr, n)) return fail(c, PPM_E_INVALID, "degenerate polygon");
    }
    o.n = f4(n);
    const uint32_t idx = (uint32_t)i;
    std::memcpy(&o.p.w, &idx, sizeof(idx));
  }
  if (d->flags < PPM_SCENE_BVH_AUTO || d->flags > PPM_SCENE_BVH_ON) return fail(c, PPM_E_INVALID, "bad scene flags");
  std::vector<DevMaterial> mats(d->n_materials + 1);
  for (int i = 0; i <= d->n_materials; ++i) {
    DevMaterial& o = mats[i];
    std::memset(&o, 0, sizeof(o));
    if (i == d->n_materials) {  // M_AIR (scene.rs:13-18)
      o.surface = SURF_NOTHING;
      o.ior = make_float4(1.f, 1.f, 1.f, 0.f);
      o.ior_avg = 1.0f;
      continue;
    }
    const ppm_material_desc& m = d->materials[i];
    if (m.surface < 0 || m.surface > 4) return fail(c, PPM_E_INVALID, "bad surface kind");
    o.surface = (uint32_t)m.surface;
    o.diff = (float)m.diffuseness;
    o.metal = (float)m.metalness;
    o.rough = (float)m.roughness;
    o.density_pow = (float)(1.0 / (std::pow(10.0, 5.0 * (1.0 - std::sqrt(m.roughness))) + 1.0));
    o.alpha = (float)(m.roughness * m.roughness * m.roughness * m.roughness);
    o.ior_avg = (float)((m.ior[0] + m.ior[1] + m.ior[2]) / 3.0);
    o.em = f4(m.emittance);
    o.ior = f4(m.ior);
    o.c1 = f4(m.color1);
    o.c2 = f4(m.color2);
    o.q = o.c1;
  }
  ppm_scene* sc = new ppm_scene();
  sc->ctx = c;
  sc->lights.resize(d->n_lights);
  for (int i = 0; i < d->n_lights; ++i) {
    const ppm_light_desc& l = d->lights[i];
    DevLight& o = sc->lights[i];
    std::memset(&o, 0, sizeof(o));
    if (l.kind < 0 || l.kind > 2 || !(l.flux >= 0.0)) {
      delete sc;
      return fail(c, PPM_E_INVALID, "bad light");
    }
    o.kind = (uint32_t)l.kind;
    o.flux = (float)l.flux;
    o.color = f4(l.color);
    o.pos = f4(l.pos);
    o.nvec = f4(l.nvec);
    o.d1 = f4(l.dir1);
    o.d2 = f4(l.dir2);
    double dn[3];
    o.dir = norm3(l.dir, dn) ? f4(dn) : f4(l.dir);
    if (o.kind == LIGHT_PARALLELOGRAM) {
      static const float uv[5] = {0.1f, 0.3f, 0.5f, 0.7f, 0.9f};
      for (int k = 0; k < 25; ++k) {
        const float u = uv[k / 5], v = uv[k % 5];
        o.tss[k] = make_float4(std::fmaf(v, o.d2.x, std::fmaf(u, o.d1.x, o.pos.x)),
                               std::fmaf(v, o.d2.y, std::fmaf(u, o.d1.y, o.pos.y)),
                               std::fmaf(v, o.d2.z, std::fmaf(u, o.d1.z, o.pos.z)), 0.0f);
      }
    }
    sc->total_flux += l.flux;
  }
  SceneBuf& b = sc->buf;
  b.n_shapes = (uint32_t)d->n_shapes;
  b.n_mats = (uint32_t)d->n_materials;
  b.n_lights = (uint32_t)d->n_lights;
  const ppm_camera_desc& cd = d->camera;
  DevCamera& cam = b.cam;
  std::memset(&cam, 0, sizeof(cam));
  cam.xreso = cd.xreso;
  cam.yreso = cd.yreso;
  cam.progressive = cd.progressive;
  cam.antialias = cd.antialias;
  cam.blur = cd.blur;
  cam.pfilter = cd.pfilter;
  cam.tiles_x = (cd.xreso + 7) / 8;
  cam.tile_shift = (int32_t)TILE_SHIFT_DEFAULT;
  cam.eye = f4(cd.eye_pos);
  cam.origin = f4(cd.origin);
  cam.esx = f4(cd.esx);
  cam.esy = f4(cd.esy);
  cam.eex = f4(cd.eex);
  cam.eey = f4(cd.eey);
  for (int i = 0; i < 3; ++i) {
    sc->bmin[i] = d->bounds_min[i];
    sc->bmax[i] = d->bounds_max[i];
    if (!(sc->bmax[i] > sc->bmin[i]) || !std::isfinite(sc->bmin[i]) || !std::isfinite(sc->bmax[i])) {
      delete sc;
      return fail(c, PPM_E_INVALID, "bad photon-map bounds");
    }
  }
  b.bmin = f4(sc->bmin);
  b.bmax = f4(sc->bmax);
  hipError_t e = hipSuccess;
  if ((e = dalloc(&b.shapes, shapes.size())) == hipSuccess && (e = dalloc(&b.mats, mats.size())) == hipSuccess &&
      (e = dalloc(&b.lights, sc->lights.size())) == hipSuccess) {
    if (!shapes.empty()) e = hipMemcpy(b.shapes, shapes.data(), sizeof(DevShape) * shapes.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b.mats, mats.data(), sizeof(DevMaterial) * mats.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && !sc->lights.empty())
      e = hipMemcpy(b.lights, sc->lights.data(), sizeof(DevLight) * sc->lights.size(), hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) {
    ppm_scene_destroy(sc);
    return fail(c, e == hipErrorOutOfMemory ? PPM_E_OOM : PPM_E_HIP, hipGetErrorString(e));
  }
  sc->mats0 = mats;
  for (const DevLight& l : sc->lights) sc->flux0.push_back(l.flux);
  sc->total_flux0 = sc->total_flux;
  const ppm_status bs = scene_layout(c, sc, shapes, d->flags);
  if (bs != PPM_OK) {
    ppm_scene_destroy(sc);
    return bs;
  }  // a ParallelogramLight's own shape: a parallelogram of the shapes tested
  // one by one with exactly the light's corner and sides (pos.w holds its
  // index there, else ~0); the shadow test knows it is hit, not an occluder
  for (DevLight& l : sc->lights) {
    uint32_t self = 0xFFFFFFFFu;
    if (l.kind == LIGHT_PARALLELOGRAM)
      for (size_t k = 0; k < sc->lin_host.size() && k < 64; ++k) {
        const DevShape& sh = sc->lin_host[k];
        if (sh.kind == SHAPE_PARALLELOGRAM && sh.p.x == l.pos.x && sh.p.y == l.pos.y && sh.p.z == l.pos.z &&
            sh.d1.x == l.d1.x && sh.d1.y == l.d1.y && sh.d1.z == l.d1.z && sh.d2.x == l.d2.x &&
            sh.d2.y == l.d2.y && sh.d2.z == l.d2.z) {
          self = (uint32_t)k;
          break;
        }
      }
    std::memcpy(&l.pos.w, &self, sizeof(self));
  }
  if (!sc->lights.empty()) {
    e = hipMemcpy(b.lights, sc->lights.data(), sizeof(DevLight) * sc->lights.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      ppm_scene_destroy(sc);
      return fail(c, PPM_E_HIP, hipGetErrorString(e));
    }
  }

  *out = sc;
  return PPM_OK;
}

void ppm_scene_destroy(ppm_scene* sc) {
  if (!sc) return;
  if (sc->ctx) (void)hipStreamSynchronize(sc->ctx->stream);
  if (sc->buf.lin != sc->buf.shapes) dfree(sc->buf.lin);
  sc->buf.lin = nullptr;
  dfree(sc->buf.shapes);
  dfree(sc->buf.mats);
  dfree(sc->buf.lights);
  dfree(sc->buf.bvh);
  dfree(sc->buf.bprims);
  dfree(sc->buf.lin_box);
  delete sc;
}

ppm_status ppm_scene_set_params(ppm_ctx* c, ppm_scene* sc, uint32_t n, const double* albedo, const double* survival,
                                double light_scale) {
  if (!c || !sc) return fail(c, PPM_E_INVALID, "null argument");
  const uint32_t nm = (uint32_t)sc->mats0.size() - 1;  // without M_AIR
  if (n > nm) return fail(c, PPM_E_RANGE, "more parameters than materials");
  if (!(light_scale >= 0.0) || !std::isfinite(light_scale)) return fail(c, PPM_E_INVALID, "bad light scale");
  std::vector<DevMaterial> mats = sc->mats0;
  for (uint32_t i = 0; i < nm; ++i) {
    DevMaterial& m = mats[i];
    if (i < n && albedo) {
      const double* a = albedo + 3 * i;
      if (!finite3(a)) return fail(c, PPM_E_NONFINITE, "non-finite albedo");
      m.c1 = f4(a);
    } else {
      m.c1 = sc->mats0[i].c1;
    }
    m.q = m.c1;
    if (i < n && survival) {
      const double* q = survival + 3 * i;
      for (int k = 0; k < 3; ++k)
        if (!(q[k] > 0.0 && q[k] <= 1.0)) return fail(c, PPM_E_INVALID, "survival probabilities must lie in (0, 1]");
      m.q = f4(q);
    }
    m.em = make_float4(sc->mats0[i].em.x * (float)light_scale, sc->mats0[i].em.y * (float)light_scale,
                       sc->mats0[i].em.z * (float)light_scale, 0.0f);
  }
  for (size_t i = 0; i < sc->lights.size(); ++i) sc->lights[i].flux = (float)(sc->flux0[i] * light_scale);
  sc->total_flux = sc->total_flux0 * light_scale;
  sc->light_scale = light_scale;
  HIP_TRY(c, hipMemcpy(sc->buf.mats, mats.data(), sizeof(DevMaterial) * mats.size(), hipMemcpyHostToDevice));
  if (!sc->lights.empty())
    HIP_TRY(c, hipMemcpy(sc->buf.lights, sc->lights.data(), sizeof(DevLight) * sc->lights.size(),
                         hipMemcpyHostToDevice));
  return PPM_OK;
}

ppm_status ppm_scene_stats(const ppm_scene* sc, uint32_t* n_linear, uint32_t* n_bvh, uint32_t* bvh_depth) {
  if (!sc) return PPM_E_INVALID;
  if (n_linear) *n_linear = sc->buf.n_lin;
  if (n_bvh) *n_bvh = sc->buf.n_bprims;
  if (bvh_depth) *bvh_depth = sc->buf.bvh_depth;
  return PPM_OK;
}

ppm_status ppm_intersect(ppm_ctx* c, const ppm_scene* sc, const double* origins, const double* dirs, uint64_t n,
                         double* t_out, int32_t* obj_out) {
  if (!c || !sc || (n && (!origins || !dirs || !t_out || !obj_out))) return fail(c, PPM_E_INVALID, "null argument");
  if (n > (1ull << 26)) return fail(c, PPM_E_RANGE, "at most 2^26 rays per call");
  if (n == 0) return PPM_OK;
  std::vector<float4> ho(n), hd(n);
  for (uint64_t i = 0; i < n; ++i) {
    const double* o = origins + 3 * i;
    double dn[3];
    if (!finite3(o) || !finite3(dirs + 3 * i)) return fail(c, PPM_E_NONFINITE, "non-finite ray");
    if (!norm3(dirs + 3 * i, dn)) return fail(c, PPM_E_INVALID, "zero-length ray direction");
    ho[i] = f4(o);
    hd[i] = f4(dn);
  }
  float4 *d_o = nullptr, *dD = nullptr;
  float* dt = nullptr;
  int32_t* dobj = nullptr;
  std::vector<float> ht(n);
  hipError_t e;
  if ((e = dalloc(&d_o, n)) == hipSuccess && (e = dalloc(&dD, n)) == hipSuccess && (e = dalloc(&dt, n)) == hipSuccess &&
      (e = dalloc(&dobj, n)) == hipSuccess &&
      (e = hipMemcpyAsync(d_o, ho.data(), n * sizeof(float4), hipMemcpyHostToDevice, c->stream)) == hipSuccess &&
      (e = hipMemcpyAsync(dD, hd.data(), n * sizeof(float4), hipMemcpyHostToDevice, c->stream)) == hipSuccess &&
      (e = launch_intersect(c->stream, sc->buf, d_o, dD, (uint32_t)n, dt, dobj)) == hipSuccess &&
      (e = hipMemcpyAsync(ht.data(), dt, n * sizeof(float), hipMemcpyDeviceToHost, c->stream)) == hipSuccess &&
      (e = hipMemcpyAsync(obj_out, dobj, n * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream)) == hipSuccess)
    e = hipStreamSynchronize(c->stream);
  dfree(d_o);
  dfree(dD);
  dfree(dt);
  dfree(dobj);
  if (e != hipSuccess) return fail(c, e == hipErrorOutOfMemory ? PPM_E_OOM : PPM_E_HIP, hipGetErrorString(e));
  for (uint64_t i = 0; i < n; ++i) t_out[i] = (double)ht[i];
  return PPM_OK;
}

// ---------------------------------------------------------------------------
// photons

ppm_status ppm_photons_create(ppm_ctx* c, uint64_t capacity, ppm_photons** out) {
  if (!c || !out) return fail(c, PPM_E_INVALID, "null argument");
  *out = nullptr;
  if (capacity == 0 || capacity > 0x7FFFFFFFull) return fail(c, PPM_E_RANGE, "capacity must be in [1, 2^31)");
  ppm_photons* p = new ppm_photons();
  p->ctx = c;
  p->cap = capacity;
  hipError_t e = photons_alloc(p->pb, capacity, 8192);
  if (e == hipSuccess) e = dalloc(&p->st, 1);
  if (e != hipSuccess) {
    ppm_photons_destroy(p);
    return fail(c, PPM_E_OOM, hipGetErrorString(e));
  }
  PassState s = host_state();
  uint32_t zero = 0;
  if ((e = hipMemcpy(p->st, &s, sizeof(s), hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(p->pb.count, &zero, sizeof(zero), hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(p->pb.offset, &zero, sizeof(zero), hipMemcpyHostToDevice)) != hipSuccess) {
    ppm_photons_destroy(p);
    return fail(c, PPM_E_HIP, hipGetErrorString(e));
  }
  p->pb.n_regions = 1;
  p->pb.region_cap = (uint32_t)capacity;
  *out = p;
  return PPM_OK;
}

void ppm_photons_destroy(ppm_photons* p) {
  if (!p) return;
  if (p->ctx) (void)hipStreamSynchronize(p->ctx->stream);
  photons_free(p->pb);
  dfree(p->st);
  delete p;
}

ppm_status ppm_trace_photons(ppm_ctx* c, const ppm_scene* sc, uint64_t seed, uint32_t pass, uint64_t n_photon,
                             int32_t use_classic, int32_t max_depth, ppm_photons* out) {
  if (!c || !sc || !out) return fail(c, PPM_E_INVALID, "null argument");
  if (ppm_status s = check_cfg_depth(c, max_depth)) return s;
  if (n_photon == 0 || n_photon > 0x7FFFFFFFull) return fail(c, PPM_E_RANGE, "n_photon must be in [1, 2^31)");
  if (sc->lights.empty() || !(sc->total_flux > 0.0)) return fail(c, PPM_E_INVALID, "scene has no emitting light");
  ppm_scene* s = const_cast<ppm_scene*>(sc);
  double power = 0;
  uint64_t total = assign_photons(s->lights, s->total_flux, n_photon, &power);
  if (total == 0 || total > PHOTON_INDEX_MAX)
    return fail(c, PPM_E_RANGE, "photons per pass must be in [1, 2^25] (the record's photon index)");
  if (!photons_layout(out->pb, total, max_depth, c->n_cu))
    return fail(c, PPM_E_RANGE, "photon set capacity too small: need " +
                                    std::to_string(photons_needed(total, max_depth, c->n_cu)));
  HIP_TRY(c, hipMemcpyAsync(s->buf.lights, s->lights.data(), sizeof(DevLight) * s->lights.size(),
                            hipMemcpyHostToDevice, c->stream));
  SceneBuf b = s->buf;
  b.n_emit_total = (uint32_t)total;
  dfree(out->pb.rec64);
  out->has64 = false;
  out->power = power;
  out->has_bounds = true;
  std::memcpy(out->bmin, s->bmin, sizeof(out->bmin));
  std::memcpy(out->bmax, s->bmax, sizeof(out->bmax));
  PassState hs = host_state();
  hs.pass = pass;
  hs.power = power;
  HIP_TRY(c, hipMemcpyAsync(out->st, &hs, sizeof(hs), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(c, launch_trace_photons(c->stream, b, out->st, out->pb, seed, use_classic, max_depth));
  PassState h;
  HIP_TRY(c, hipMemcpyAsync(&h, out->st, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (h.overflow) return fail(c, PPM_E_RANGE, "photon region capacity exceeded");
  return PPM_OK;
}

ppm_status ppm_photons_count(ppm_ctx* c, const ppm_photons* p, uint64_t* n) {
  if (!c || !p || !n) return fail(c, PPM_E_INVALID, "null argument");
  PassState h;
  HIP_TRY(c, hipMemcpyAsync(&h, p->st, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  *n = h.n_stored;
  return PPM_OK;
}

ppm_status ppm_photons_upload(ppm_ctx* c, ppm_photons* p, uint64_t n, const double* pos, const double* dir,
                              const uint8_t* wl) {
  if (!c || !p || (n && (!pos || !dir || !wl))) return fail(c, PPM_E_INVALID, "null argument");
  if (n > p->cap) return fail(c, PPM_E_RANGE, "photon set capacity exceeded");
  if (n > PHOTON_INDEX_MAX) return fail(c, PPM_E_RANGE, "at most 2^25 photons per set (the record's photon index)");
  std::vector<PhotonRec> recs(n);
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint64_t i = 0; i < n; ++i) {
    const double* pp = pos + 3 * i;
    const double* dd = dir + 3 * i;
    if (!finite3(pp) || !finite3(dd)) return fail(c, PPM_E_NONFINITE, "non-finite photon " + std::to_string(i));
    if (wl[i] > 2) return fail(c, PPM_E_INVALID, "bad channel code at photon " + std::to_string(i));
    double dn[3];
    if (!norm3(dd, dn)) return fail(c, PPM_E_INVALID, "zero direction at photon " + std::to_string(i));
    recs[i].pm = make_float4((float)pp[0], (float)pp[1], (float)pp[2],
                             __builtin_bit_cast(float, pack_meta(wl[i], 0, (uint32_t)i)));
    recs[i].dm = make_float4((float)dn[0], (float)dn[1], (float)dn[2], 1.0f);
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], pp[k]);
      hi[k] = std::max(hi[k], pp[k]);
    }
  }
  if (!p->pb.rec64) HIP_TRY(c, dalloc(&p->pb.rec64, 3 * p->cap));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (n) {
    HIP_TRY(c, hipMemcpy(p->pb.rec, recs.data(), sizeof(PhotonRec) * n, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(p->pb.rec64, pos, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
  }
  uint32_t cnt = (uint32_t)n, zero = 0;
  HIP_TRY(c, hipMemcpy(p->pb.count, &cnt, sizeof(cnt), hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(p->pb.offset, &zero, sizeof(zero), hipMemcpyHostToDevice));
  p->pb.n_regions = 1;
  p->pb.region_cap = (uint32_t)p->cap;
  PassState s = host_state();
  s.n_stored = (uint32_t)n;
  HIP_TRY(c, hipMemcpy(p->st, &s, sizeof(s), hipMemcpyHostToDevice));
  p->has64 = true;
  p->has_bounds = true;
  for (int k = 0; k < 3; ++k) {
    if (n == 0) {
      lo[k] = 0.0;
      hi[k] = 1.0;
    }
    double pad = std::max(1e-6, (hi[k] - lo[k]) * 1e-6);
    p->bmin[k] = lo[k] - pad;
    p->bmax[k] = hi[k] + pad;
  }
  return PPM_OK;
}

ppm_status ppm_photons_download(ppm_ctx* c, const ppm_photons* p, uint64_t cap, double* pos, double* dir,
                                uint8_t* wl, uint32_t* meta, uint64_t* n) {
  if (!c || !p || !n) return fail(c, PPM_E_INVALID, "null argument");
  uint64_t cnt = 0;
  if (ppm_status s = ppm_photons_count(c, p, &cnt)) return s;
  *n = cnt;
  if (cnt > cap) return fail(c, PPM_E_RANGE, "output capacity too small");
  if (cnt && (!pos || !dir || !wl)) return fail(c, PPM_E_INVALID, "null output array");
  std::vector<PhotonRec> recs(cnt);
  if (cnt) {
    PhotonRec* dense = nullptr;
    HIP_TRY(c, dalloc(&dense, cnt));
    hipError_t e = launch_compact(c->stream, p->pb, p->st, dense);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(recs.data(), dense, sizeof(PhotonRec) * cnt, hipMemcpyDeviceToHost);
    dfree(dense);
    if (e != hipSuccess) return fail(c, PPM_E_HIP, hipGetErrorString(e));
  }
  for (uint64_t i = 0; i < cnt; ++i) {
    uint32_t m = __builtin_bit_cast(uint32_t, recs[i].pm.w);
    pos[3 * i] = recs[i].pm.x;
    pos[3 * i + 1] = recs[i].pm.y;
    pos[3 * i + 2] = recs[i].pm.z;
    dir[3 * i] = recs[i].dm.x;
    dir[3 * i + 1] = recs[i].dm.y;
    dir[3 * i + 2] = recs[i].dm.z;
    wl[i] = (uint8_t)(m & 3u);
    if (meta) meta[i] = m;
  }
  return PPM_OK;
}

// ---------------------------------------------------------------------------
// map

ppm_status ppm_map_create(ppm_ctx* c, uint64_t capacity, ppm_map** out) {
  if (!c || !out) return fail(c, PPM_E_INVALID, "null argument");
  *out = nullptr;
  if (capacity == 0 || capacity > 0x7FFFFFFFull) return fail(c, PPM_E_RANGE, "capacity must be in [1, 2^31)");
  ppm_map* m = new ppm_map();
  m->ctx = c;
  hipError_t e = map_alloc(m->m, capacity, c->n_cu, false);
  if (e == hipSuccess) e = dalloc(&m->st, 1);
  if (e != hipSuccess) {
    ppm_map_destroy(m);
    return fail(c, PPM_E_OOM, hipGetErrorString(e));
  }
  *out = m;
  return PPM_OK;
}

void ppm_map_destroy(ppm_map* m) {
  if (!m) return;
  if (m->ctx) (void)hipStreamSynchronize(m->ctx->stream);
  map_free(m->m);
  dfree(m->st);
  dfree(m->hook_mat);
  delete m;
}

ppm_status ppm_map_build(ppm_ctx* c, const ppm_scene* sc, const ppm_photons* ph, double power, ppm_map* m) {
  if (!c || !ph || !m) return fail(c, PPM_E_INVALID, "null argument");
  if (ph->cap > m->m.cap) return fail(c, PPM_E_RANGE, "map capacity smaller than the photon set's");
  const double* lo = sc ? sc->bmin : ph->bmin;
  const double* hi = sc ? sc->bmax : ph->bmax;
  if (!sc && !ph->has_bounds) return fail(c, PPM_E_INVALID, "no bounds for the Morton keys");
  if (ph->has64 && !m->m.sorted64) HIP_TRY(c, dalloc(&m->m.sorted64, 3 * m->m.cap));
  HIP_TRY(c, hipMemcpyAsync(m->st, ph->st, sizeof(PassState), hipMemcpyDeviceToDevice, c->stream));
  PhotonBuf in = ph->pb;
  if (!ph->has64) in.rec64 = nullptr;
  m->built = false;
  m->grid = false;
  MapBuf tv = tree_view(m);
  HIP_TRY(c, launch_map_build(c->stream, tv, m->st, in, f4(lo), f4(hi)));
  PassState h;
  HIP_TRY(c, hipMemcpyAsync(&h, m->st, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  m->power = power;
  m->has64 = ph->has64;
  m->built = true;
  if (h.nonfinite) return fail(c, PPM_E_NONFINITE, "non-finite photon position (photonmap.rs:26)");
  return PPM_OK;
}

ppm_status ppm_map_build_grid(ppm_ctx* c, const ppm_scene* sc, const ppm_photons* ph, double power, double r2,
                              ppm_map* m) {
  if (!c || !ph || !m) return fail(c, PPM_E_INVALID, "null argument");
  if (ph->cap > m->m.cap) return fail(c, PPM_E_RANGE, "map capacity smaller than the photon set's");
  if (!(r2 > 0.0) || !std::isfinite(r2) || !std::isfinite((float)r2)) return fail(c, PPM_E_INVALID, "bad radius");
  const double* lo = sc ? sc->bmin : ph->bmin;
  const double* hi = sc ? sc->bmax : ph->bmax;
  if (!sc && !ph->has_bounds) return fail(c, PPM_E_INVALID, "no bounds for the grid");
  m->built = false;
  if (!m->m.cells) {
    m->m.cells_max = (uint32_t)std::min<uint64_t>(GRID_MAX_BUCKETS, next_pow2(std::max<uint64_t>(m->m.cap, 1024)));
    HIP_TRY(c, dalloc(&m->m.cells, m->m.cells_max));
  }
  if (!m->m.os.ctl) HIP_TRY(c, onesweep_alloc(m->m.os, m->m.cap, c->n_cu));
  if (!m->hook_mat) {
    DevMaterial air;
    std::memset(&air, 0, sizeof(air));
    HIP_TRY(c, dalloc(&m->hook_mat, 1));
    HIP_TRY(c, hipMemcpy(m->hook_mat, &air, sizeof(air), hipMemcpyHostToDevice));
  }
  if (ph->has64 && !m->m.sorted64) HIP_TRY(c, dalloc(&m->m.sorted64, 3 * m->m.cap));
  // the pass state of a pass with this radius (k_grid_setup sizes the cells from r2f)
  PassState hs;
  HIP_TRY(c, hipMemcpyAsync(&hs, ph->st, sizeof(hs), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  hs.r2 = r2;
  hs.r2f = (float)r2;
  hs.power = power;
  hs.overflow = 0;
  hs.nonfinite = 0;
  HIP_TRY(c, hipMemcpy(m->st, &hs, sizeof(hs), hipMemcpyHostToDevice));
  PhotonBuf in = ph->pb;
  if (!ph->has64) in.rec64 = nullptr;
  m->gbmin = f4(lo);
  m->gbmax = f4(hi);
  MapBuf g = grid_view_of(m);
  HIP_TRY(c, launch_grid_build(c->stream, g, m->st, in, m->gbmin, m->gbmax));
  PassState h;
  HIP_TRY(c, hipMemcpyAsync(&h, m->st, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  m->power = power;
  m->has64 = ph->has64;
  m->grid = true;
  m->grid_r2 = r2;
  m->built = true;
  if (h.nonfinite) return fail(c, PPM_E_NONFINITE, "non-finite photon position (photonmap.rs:26)");
  return PPM_OK;
}

static ppm_status run_query(ppm_ctx* c, const ppm_map* m, const double* q_xyz, const double* q_nrm, uint64_t nq,
                            double r2, int refine64, int mode, int pfilter, std::vector<uint32_t>* counts,
                            const std::vector<uint64_t>* offsets, uint32_t* idx_out, uint64_t idx_total,
                            std::vector<float4>* rgb) {
  std::vector<float4> qf(nq), qn;
  for (uint64_t i = 0; i < nq; ++i) {
    if (!finite3(q_xyz + 3 * i)) return fail(c, PPM_E_NONFINITE, "non-finite query point");
    qf[i] = make_float4((float)q_xyz[3 * i], (float)q_xyz[3 * i + 1], (float)q_xyz[3 * i + 2], 0.0f);
  }
  if (q_nrm) {
    qn.resize(nq);
    for (uint64_t i = 0; i < nq; ++i)
      qn[i] = make_float4((float)q_nrm[3 * i], (float)q_nrm[3 * i + 1], (float)q_nrm[3 * i + 2], 0.0f);
  }
  float4 *dqf = nullptr, *dqn = nullptr, *drgb = nullptr;
  double* dq64 = nullptr;
  uint32_t *dcnt = nullptr, *didx = nullptr, *dflag = nullptr, hflag = 0;
  uint64_t* doff = nullptr;
  hipError_t e = hipSuccess;
  auto cleanup = [&]() {
    dfree(dqf); dfree(dqn); dfree(drgb); dfree(dq64); dfree(dcnt); dfree(didx); dfree(doff); dfree(dflag);
  };
  do {
    if ((e = dalloc(&dqf, nq)) != hipSuccess) break;
    if ((e = hipMemcpy(dqf, qf.data(), sizeof(float4) * nq, hipMemcpyHostToDevice)) != hipSuccess) break;
    if (refine64) {
      if ((e = dalloc(&dq64, 3 * nq)) != hipSuccess) break;
      if ((e = hipMemcpy(dq64, q_xyz, sizeof(double) * 3 * nq, hipMemcpyHostToDevice)) != hipSuccess) break;
    }
    if (q_nrm) {
      if ((e = dalloc(&dqn, nq)) != hipSuccess) break;
      if ((e = hipMemcpy(dqn, qn.data(), sizeof(float4) * nq, hipMemcpyHostToDevice)) != hipSuccess) break;
    }
    if (mode == 0 && (e = dalloc(&dcnt, nq)) != hipSuccess) break;
    if (mode == 1) {
      if ((e = dalloc(&doff, nq)) != hipSuccess) break;
      if ((e = hipMemcpy(doff, offsets->data(), sizeof(uint64_t) * nq, hipMemcpyHostToDevice)) != hipSuccess) break;
      if ((e = dalloc(&didx, idx_total)) != hipSuccess) break;
    }
    if (mode == 2 && (e = dalloc(&drgb, nq)) != hipSuccess) break;
    if (m->grid) {
      if ((e = dalloc(&dflag, 1)) != hipSuccess) break;
      if ((e = hipMemsetAsync(dflag, 0, sizeof(uint32_t), c->stream)) != hipSuccess) break;
      if ((e = launch_grid_query(c->stream, grid_view_of(m), m->st, m->gbmin, dqf, dq64, (uint32_t)nq, r2, refine64,
                                 mode, dcnt, doff, didx, dflag)) != hipSuccess)
        break;
      if ((e = hipMemcpyAsync(&hflag, dflag, sizeof(hflag), hipMemcpyDeviceToHost, c->stream)) != hipSuccess) break;
    } else if ((e = launch_query(c->stream, tree_view(m), m->st, dqf, dq64, dqn, (uint32_t)nq, r2, refine64, mode,
                                 pfilter, dcnt, doff, didx, drgb, c->n_cu)) != hipSuccess) {
      break;
    }
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) break;
    if (mode == 0) {
      counts->resize(nq);
      e = hipMemcpy(counts->data(), dcnt, sizeof(uint32_t) * nq, hipMemcpyDeviceToHost);
    } else if (mode == 1) {
      if (idx_total) e = hipMemcpy(idx_out, didx, sizeof(uint32_t) * idx_total, hipMemcpyDeviceToHost);
    } else {
      rgb->resize(nq);
      e = hipMemcpy(rgb->data(), drgb, sizeof(float4) * nq, hipMemcpyDeviceToHost);
    }
  } while (0);
  cleanup();
  if (e != hipSuccess) return fail(c, e == hipErrorOutOfMemory ? PPM_E_OOM : PPM_E_HIP, hipGetErrorString(e));
  if (hflag) return fail(c, PPM_E_RANGE, "a query's cells exceed the grid's span");
  return PPM_OK;
}

// ppm_gather on a grid map: the queries run through the passes' shade
// kernel (direct light off), in the given order, with the pass scale
// power / (pi r2) applied in float as in a pass.
static ppm_status run_grid_gather(ppm_ctx* c, const ppm_map* m, const double* q_xyz, const double* q_nrm,
                                  uint64_t nq, double r2, int pfilter, double* rgb_out) {
  std::vector<Query> hq(nq);
  std::vector<uint32_t> order(nq);
  for (uint64_t i = 0; i < nq; ++i) {
    if (!finite3(q_xyz + 3 * i) || !finite3(q_nrm + 3 * i)) return fail(c, PPM_E_NONFINITE, "non-finite query");
    hq[i].a = make_float4((float)q_xyz[3 * i], (float)q_xyz[3 * i + 1], (float)q_xyz[3 * i + 2],
                          __builtin_bit_cast(float, (uint32_t)i));
    hq[i].b = make_float4((float)q_nrm[3 * i], (float)q_nrm[3 * i + 1], (float)q_nrm[3 * i + 2],
                          __builtin_bit_cast(float, 0u));
    hq[i].w = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    order[i] = (uint32_t)i;
  }
  PassState hs;
  HIP_TRY(c, hipMemcpyAsync(&hs, m->st, sizeof(hs), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  hs.r2 = r2;
  hs.r2f = (float)r2;
  hs.scale = (float)(m->power / (M_PI * r2));
  hs.n_qsort = (uint32_t)nq;
  hs.overflow = 0;
  std::memset(hs.work, 0, sizeof(hs.work));
  EyeBuf eb;
  PassState* st = nullptr;
  uint32_t* dorder = nullptr;
  std::vector<float4> res(nq);
  hipError_t e = hipSuccess;
  do {
    if ((e = dalloc(&st, 1)) != hipSuccess) break;
    if ((e = dalloc(&eb.q, nq)) != hipSuccess) break;
    if ((e = dalloc(&eb.res, nq)) != hipSuccess) break;
    if ((e = dalloc(&dorder, nq)) != hipSuccess) break;
    if ((e = hipMemcpy(st, &hs, sizeof(hs), hipMemcpyHostToDevice)) != hipSuccess) break;
    if ((e = hipMemcpy(eb.q, hq.data(), sizeof(Query) * nq, hipMemcpyHostToDevice)) != hipSuccess) break;
    if ((e = hipMemcpy(dorder, order.data(), sizeof(uint32_t) * nq, hipMemcpyHostToDevice)) != hipSuccess) break;
    SceneBuf sb;
    sb.mats = m->hook_mat;
    sb.n_mats = 0;
    sb.bmin = m->gbmin;
    sb.bmax = m->gbmax;
    if ((e = launch_grid_shade(c->stream, sb, grid_view_of(m), st, eb, dorder, pfilter, c->n_cu)) != hipSuccess)
      break;
    if ((e = hipMemcpyAsync(res.data(), eb.res, sizeof(float4) * nq, hipMemcpyDeviceToHost, c->stream)) !=
        hipSuccess)
      break;
    if ((e = hipMemcpyAsync(&hs, st, sizeof(hs), hipMemcpyDeviceToHost, c->stream)) != hipSuccess) break;
    e = hipStreamSynchronize(c->stream);
  } while (0);
  dfree(st);
  dfree(eb.q);
  dfree(eb.res);
  dfree(dorder);
  if (e != hipSuccess) return fail(c, e == hipErrorOutOfMemory ? PPM_E_OOM : PPM_E_HIP, hipGetErrorString(e));
  if (hs.overflow) return fail(c, PPM_E_RANGE, "a query's cells exceed the grid's span");
  for (uint64_t i = 0; i < nq; ++i) {
    rgb_out[3 * i] = res[i].x;
    rgb_out[3 * i + 1] = res[i].y;
    rgb_out[3 * i + 2] = res[i].z;
  }
  return PPM_OK;
}

ppm_status ppm_radius_query(ppm_ctx* c, const ppm_map* m, const double* q_xyz, uint64_t nq, double r2,
                            int32_t refine64, uint64_t idx_cap, uint32_t* idx_out, uint64_t* offsets_out) {
  if (!c || !m || (nq && !q_xyz) || !offsets_out) return fail(c, PPM_E_INVALID, "null argument");
  if (!m->built) return fail(c, PPM_E_INVALID, "map not built");
  if (refine64 && !m->has64) return fail(c, PPM_E_INVALID, "refine64 needs a set loaded by ppm_photons_upload");
  if (!(r2 >= 0.0) || !std::isfinite(r2)) return fail(c, PPM_E_INVALID, "bad radius");
  if (m->grid && r2 > m->grid_r2) return fail(c, PPM_E_INVALID, "r2 exceeds the radius the grid was built for");
  if (nq > 0x7FFFFFFFull) return fail(c, PPM_E_RANGE, "too many queries");
  offsets_out[0] = 0;
  if (nq == 0) return PPM_OK;
  std::vector<uint32_t> counts;
  if (ppm_status s = run_query(c, m, q_xyz, nullptr, nq, r2, refine64, 0, 0, &counts, nullptr, nullptr, 0, nullptr))
    return s;
  std::vector<uint64_t> off(nq);
  uint64_t tot = 0;
  for (uint64_t i = 0; i < nq; ++i) {
    off[i] = tot;
    offsets_out[i] = tot;
    tot += counts[i];
  }
  offsets_out[nq] = tot;
  if (tot > idx_cap) return fail(c, PPM_E_RANGE, "idx_cap too small: need " + std::to_string(tot));
  if (tot && !idx_out) return fail(c, PPM_E_INVALID, "null idx_out");
  if (ppm_status s = run_query(c, m, q_xyz, nullptr, nq, r2, refine64, 1, 0, nullptr, &off, idx_out, tot, nullptr))
    return s;
  for (uint64_t i = 0; i < nq; ++i) std::sort(idx_out + off[i], idx_out + off[i] + counts[i]);
  return PPM_OK;
}

// The argument checks of ppm_gather (and ppm_gather_noalloc).
static ppm_status check_gather(ppm_ctx* c, const ppm_map* m, const double* q_xyz, const double* q_nrm, uint64_t nq,
                               double r2, int32_t pfilter, int32_t refine64, double* rgb_out) {
  if (!c || !m || (nq && (!q_xyz || !q_nrm || !rgb_out))) return fail(c, PPM_E_INVALID, "null argument");
  if (!m->built) return fail(c, PPM_E_INVALID, "map not built");
  if (refine64 && !m->has64) return fail(c, PPM_E_INVALID, "refine64 needs a set loaded by ppm_photons_upload");
  if (!(r2 > 0.0) || !std::isfinite(r2)) return fail(c, PPM_E_INVALID, "bad radius");
  if (pfilter < 0 || pfilter > 2) return fail(c, PPM_E_INVALID, "bad photon filter");
  if (nq > 0x7FFFFFFFull) return fail(c, PPM_E_RANGE, "too many queries");
  if (m->grid) {
    if (refine64) return fail(c, PPM_E_INVALID, "the grid's gather decides membership in float (refine64 = 0)");
    if (r2 > m->grid_r2) return fail(c, PPM_E_INVALID, "r2 exceeds the radius the grid was built for");
  }
  return PPM_OK;
}

ppm_status ppm_gather(ppm_ctx* c, const ppm_map* m, const double* q_xyz, const double* q_nrm, uint64_t nq, double r2,
                      int32_t pfilter, int32_t refine64, double* rgb_out) {
  if (ppm_status s = check_gather(c, m, q_xyz, q_nrm, nq, r2, pfilter, refine64, rgb_out)) return s;
  if (nq == 0) return PPM_OK;
  if (m->grid) return run_grid_gather(c, m, q_xyz, q_nrm, nq, r2, pfilter, rgb_out);
  std::vector<float4> rgb;
  if (ppm_status s = run_query(c, m, q_xyz, q_nrm, nq, r2, refine64, 2, pfilter, nullptr, nullptr, nullptr, 0, &rgb))
    return s;
  // tracer.rs:191-193: Σ (wt * power) * max(0, -cos) then * (1/pi / r2)
  const double k = m->power * (1.0 / M_PI) / r2;
  for (uint64_t i = 0; i < nq; ++i) {
    rgb_out[3 * i] = (double)rgb[i].x * k;
    rgb_out[3 * i + 1] = (double)rgb[i].y * k;
    rgb_out[3 * i + 2] = (double)rgb[i].z * k;
  }
  return PPM_OK;
}

static_assert(PPM_NOALLOC_MAX_QUERIES == NOALLOC_MAX_QUERIES, "ppm.h and the kernels disagree on the bound");

ppm_status ppm_gather_noalloc(ppm_ctx* c, const ppm_map* m, const double* q_xyz, const double* q_nrm, uint64_t nq,
                              double r2, int32_t pfilter, int32_t refine64, double* rgb_out) {
  if (ppm_status s = check_gather(c, m, q_xyz, q_nrm, nq, r2, pfilter, refine64, rgb_out)) return s;
  if (nq > PPM_NOALLOC_MAX_QUERIES)
    return fail(c, PPM_E_RANGE, "nq exceeds PPM_NOALLOC_MAX_QUERIES (" + std::to_string(PPM_NOALLOC_MAX_QUERIES) + ")");
  if (nq == 0) return PPM_OK;
  for (uint64_t i = 0; i < nq; ++i) {
    if (!finite3(q_xyz + 3 * i)) return fail(c, PPM_E_NONFINITE, "non-finite query point");
    if (m->grid && !finite3(q_nrm + 3 * i)) return fail(c, PPM_E_NONFINITE, "non-finite query");
  }
  // the same scale factors as ppm_gather: the tree's sums times power/pi/r2
  // in double, the grid's scaled in float by the shade kernel
  const double k = m->power * (1.0 / M_PI) / r2;
  const float scale = (float)(m->power / (M_PI * r2));
  SceneBuf sb;
  sb.mats = m->hook_mat;
  sb.n_mats = 0;
  sb.bmin = m->gbmin;
  sb.bmax = m->gbmax;
  uint32_t overflow = 0;
  HIP_TRY(c, launch_gather_noalloc(c->stream, c->device, map_view(m), m->st, m->grid, sb, q_xyz, q_nrm,
                                   (uint32_t)nq, r2, k, scale, refine64, pfilter, c->n_cu, rgb_out, &overflow));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (overflow) return fail(c, PPM_E_RANGE, "a query's cells exceed the grid's span");
  return PPM_OK;
}

ppm_status ppm_trace_eye(ppm_ctx* c, const ppm_scene* sc, const ppm_map* m, uint64_t seed, uint32_t pass, double r2,
                         int32_t use_classic, int32_t max_depth, int32_t pfilter, double* rgb_out) {
  if (!c || !sc || !m || !rgb_out) return fail(c, PPM_E_INVALID, "null argument");
  if (ppm_status s = check_cfg_depth(c, max_depth)) return s;
  if (!m->built) return fail(c, PPM_E_INVALID, "map not built");
  if (!(r2 > 0.0) || !std::isfinite(r2)) return fail(c, PPM_E_INVALID, "bad radius");
  if (pfilter < 0 || pfilter > 2) return fail(c, PPM_E_INVALID, "bad photon filter");
  if (m->grid && r2 > m->grid_r2) return fail(c, PPM_E_INVALID, "r2 exceeds the radius the grid was built for");
  EyeBuf eb;
  PassState* st = nullptr;
  hipError_t e = hipSuccess;
  std::vector<float4> h;
  PassState hs;
  do {
    if ((e = eye_alloc(eb, sc->buf.cam, max_depth, c->n_cu)) != hipSuccess) break;
    if ((e = dalloc(&st, 1)) != hipSuccess) break;
    if ((e = hipMemcpyAsync(&hs, m->st, sizeof(PassState), hipMemcpyDeviceToHost, c->stream)) != hipSuccess) break;
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) break;
    hs.pass = pass;
    hs.r2 = r2;
    hs.power = m->power;
    hs.r2f = (float)r2;
    hs.scale = (float)(m->power / (M_PI * r2));
    hs.n_queries = 0;
    hs.overflow = 0;
    hs.tiles_front = hs.tiles_back = 0;
    std::memset(hs.work, 0, sizeof(hs.work));
    if ((e = hipMemcpy(st, &hs, sizeof(hs), hipMemcpyHostToDevice)) != hipSuccess) break;
    if ((e = launch_eye(c->stream, sc->buf, st, eb, seed, use_classic, max_depth, c->n_cu)) != hipSuccess) break;
    // a grid's cells are indexed from the box it was built over
    SceneBuf sb = sc->buf;
    if (m->grid) {
      sb.bmin = m->gbmin;
      sb.bmax = m->gbmax;
    }
    if ((e = launch_gather(c->stream, sb, map_view(m), st, eb, pfilter, use_classic, c->n_cu, nullptr, nullptr)) !=
        hipSuccess)
      break;
    h.resize(eb.n_pix);
    if ((e = hipMemcpyAsync(h.data(), eb.pix, sizeof(float4) * eb.n_pix, hipMemcpyDeviceToHost, c->stream)) !=
        hipSuccess)
      break;
    if ((e = hipMemcpyAsync(&hs, st, sizeof(hs), hipMemcpyDeviceToHost, c->stream)) != hipSuccess) break;
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) break;
  } while (0);
  eye_free(eb);
  dfree(st);
  if (e != hipSuccess) return fail(c, e == hipErrorOutOfMemory ? PPM_E_OOM : PPM_E_HIP, hipGetErrorString(e));
  if (hs.overflow) return fail(c, PPM_E_RANGE, "gather query buffer overflow");
  for (size_t p = 0; p < h.size(); ++p) {
    rgb_out[3 * p] = h[p].x;
    rgb_out[3 * p + 1] = h[p].y;
    rgb_out[3 * p + 2] = h[p].z;
  }
  return PPM_OK;
}

ppm_status ppm_trace_classic(ppm_ctx* c, const ppm_scene* sc, uint64_t seed, uint32_t pass, const double ambient[3],
                             double* rgb_out) {
  if (!c || !sc || !ambient || !rgb_out) return fail(c, PPM_E_INVALID, "null argument");
  for (int k = 0; k < 3; ++k)
    if (!std::isfinite(ambient[k])) return fail(c, PPM_E_INVALID, "bad ambient radiance");
  EyeBuf eb;
  PassState* st = nullptr;
  hipError_t e = hipSuccess;
  std::vector<float4> h;
  PassState hs = host_state();
  do {
    if ((e = eye_alloc(eb, sc->buf.cam, CLASSIC_DEPTH, c->n_cu)) != hipSuccess) break;
    if ((e = dalloc(&st, 1)) != hipSuccess) break;
    hs.pass = pass;
    if ((e = hipMemcpy(st, &hs, sizeof(hs), hipMemcpyHostToDevice)) != hipSuccess) break;
    const float4 amb = make_float4((float)ambient[0], (float)ambient[1], (float)ambient[2], 0.0f);
    if ((e = launch_classic(c->stream, sc->buf, st, eb, seed, amb, c->n_cu)) != hipSuccess) break;
    h.resize(eb.n_pix);
    if ((e = hipMemcpyAsync(h.data(), eb.pix, sizeof(float4) * eb.n_pix, hipMemcpyDeviceToHost, c->stream)) !=
        hipSuccess)
      break;
    if ((e = hipMemcpyAsync(&hs, st, sizeof(hs), hipMemcpyDeviceToHost, c->stream)) != hipSuccess) break;
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) break;
  } while (0);
  eye_free(eb);
  dfree(st);
  if (e != hipSuccess) return fail(c, e == hipErrorOutOfMemory ? PPM_E_OOM : PPM_E_HIP, hipGetErrorString(e));
  if (hs.overflow) return fail(c, PPM_E_RANGE, "gather query buffer overflow");
  for (size_t p = 0; p < h.size(); ++p) {
    rgb_out[3 * p] = h[p].x;
    rgb_out[3 * p + 1] = h[p].y;
    rgb_out[3 * p + 2] = h[p].z;
  }
  return PPM_OK;
}

// ---------------------------------------------------------------------------
// session

static void pipe_drop_graph(Pipe& p) {
  if (p.exec) (void)hipGraphExecDestroy(p.exec);
  if (p.graph) (void)hipGraphDestroy(p.graph);
  p.exec = nullptr;
  p.graph = nullptr;
}

void ppm_session_destroy(ppm_session* s) {
  if (!s) return;
  if (s->ctx) (void)hipStreamSynchronize(s->ctx->stream);
  for (Pipe& p : s->pipes) {
    // the streams are the context's (ctx_stream): waited for, not destroyed
    if (p.stream) (void)hipStreamSynchronize(p.stream);
    if (p.s2) (void)hipStreamSynchronize(p.s2);
    pipe_drop_graph(p);
    for (hipEvent_t e : {p.evf, p.evj, p.done})
      if (e) (void)hipEventDestroy(e);
    photons_free(p.pb);
    map_free(p.map);
    eye_free(p.eb);
    dfree(p.sum);
    dfree(p.sumsq);
    dfree(p.st);
  }
  for (auto& e : s->ev)
    if (e) (void)hipEventDestroy(e);
  dfree(s->lights);
  delete s;
}

// Pipelines of a session: cfg.pipelines, else 2 when a pass is large enough
// (at least 256k photons or 256k pixels) for two passes in flight to keep
// the GPU busier than one, else 1.
static uint32_t session_pipelines(const ppm_pass_cfg& cfg, uint64_t n_emit, uint64_t n_pix) {
  if (cfg.pipelines) return std::min<uint32_t>(cfg.pipelines, 4);
  return (n_emit >= (1u << 18) || n_pix >= (1u << 18)) ? 2u : 1u;
}

// own_streams = false: a batch's session, whose passes the batch launches on
// its own streams; the pipelines get no streams or events of their own.
static ppm_status session_create(ppm_ctx* c, const ppm_scene* sc, const ppm_pass_cfg* cfg, ppm_session** out,
                                 bool own_streams) {
  if (!c || !sc || !cfg || !out) return fail(c, PPM_E_INVALID, "null argument");
  *out = nullptr;
  if (ppm_status s = check_cfg_depth(c, cfg->max_depth)) return s;
  if (cfg->n_photon == 0 || cfg->n_photon > 0x7FFFFFFFull) return fail(c, PPM_E_RANGE, "n_photon out of range");
  if (!(cfg->radius0 > 0.0) || !std::isfinite(cfg->radius0) || !std::isfinite(cfg->alpha))
    return fail(c, PPM_E_INVALID, "bad radius0/alpha");
  if (cfg->pfilter < 0 || cfg->pfilter > 2) return fail(c, PPM_E_INVALID, "bad photon filter");
  if (cfg->pipelines > 4) return fail(c, PPM_E_INVALID, "at most 4 pipelines");
  if (sc->lights.empty() || !(sc->total_flux > 0.0)) return fail(c, PPM_E_INVALID, "scene has no emitting light");
  ppm_session* s = new ppm_session();
  s->ctx = c;
  s->scene = sc;
  s->cfg = *cfg;
  if (s->cfg.first_pass == 0) s->cfg.first_pass = 1;
  std::vector<DevLight> L = sc->lights;
  uint64_t total = assign_photons(L, sc->total_flux, cfg->n_photon, &s->power);
  if (total == 0 || total > PHOTON_INDEX_MAX) {
    delete s;
    return fail(c, PPM_E_RANGE, "photons per pass must be in [1, 2^25] (the record's photon index)");
  }
  const uint64_t cap = photons_needed(total, cfg->max_depth, c->n_cu);
  if (cap > 0x7FFFFFFFull) {
    delete s;
    return fail(c, PPM_E_RANGE, "n_photon * max_depth must stay below 2^31");
  }
  s->sbuf = sc->buf;
  s->sbuf.n_emit_total = (uint32_t)total;
  const uint64_t n_pix = (uint64_t)sc->buf.cam.xreso * (uint64_t)sc->buf.cam.yreso;
  s->cfg.pipelines = session_pipelines(s->cfg, total, n_pix);
  s->pipes.resize(s->cfg.pipelines);
  hipError_t e = hipSuccess;
  do {
    if ((e = dalloc(&s->lights, L.size())) != hipSuccess) break;
    if ((e = hipMemcpy(s->lights, L.data(), sizeof(DevLight) * L.size(), hipMemcpyHostToDevice)) != hipSuccess) break;
    s->sbuf.lights = s->lights;
    for (Pipe& p : s->pipes) {
      if ((e = photons_alloc(p.pb, cap, trace_waves(total, c->n_cu))) != hipSuccess) break;
      photons_layout(p.pb, total, cfg->max_depth, c->n_cu);
      if ((e = map_alloc(p.map, cap, c->n_cu, true)) != hipSuccess) break;
      if ((e = eye_alloc(p.eb, sc->buf.cam, cfg->max_depth, c->n_cu)) != hipSuccess) break;
      const uint64_t n3 = 3ull * p.eb.n_pix;
      if ((e = dalloc(&p.sum, n3)) != hipSuccess) break;
      if ((e = dalloc(&p.sumsq, n3)) != hipSuccess) break;
      if ((e = hipMemset(p.sum, 0, sizeof(double) * n3)) != hipSuccess) break;
      if ((e = hipMemset(p.sumsq, 0, sizeof(double) * n3)) != hipSuccess) break;
      if ((e = dalloc(&p.st, 1)) != hipSuccess) break;
      PassState h = host_state();
      if ((e = hipMemcpy(p.st, &h, sizeof(h), hipMemcpyHostToDevice)) != hipSuccess) break;
      // pipeline 0 runs on the context's stream: K pipelines then use 2K
      // streams in all, which HIP spreads over that many hardware queues
      // (GPU_MAX_HW_QUEUES, 4 by default) when K <= 2
      const int j = (int)(&p - &s->pipes[0]);
      if (j == 0 || !own_streams)
        p.stream = c->stream;
      else if ((e = ctx_stream(c, AUX_PIPE + j - 1, &p.stream)) != hipSuccess)
        break;
      if (!own_streams) continue;
      if ((e = ctx_stream(c, AUX_EYE + j, &p.s2)) != hipSuccess) break;
      if ((e = hipEventCreateWithFlags(&p.evf, hipEventDisableTiming)) != hipSuccess) break;
      if ((e = hipEventCreateWithFlags(&p.evj, hipEventDisableTiming)) != hipSuccess) break;
      if ((e = hipEventCreateWithFlags(&p.done, hipEventDisableTiming)) != hipSuccess) break;
    }
    if (e != hipSuccess) break;
    for (auto& ev : s->ev)
      if ((e = hipEventCreate(&ev)) != hipSuccess) break;
  } while (0);
  if (e != hipSuccess) {
    ppm_session_destroy(s);
    return fail(c, e == hipErrorOutOfMemory ? PPM_E_OOM : PPM_E_HIP, hipGetErrorString(e));
  }
  *out = s;
  return PPM_OK;
}

ppm_status ppm_session_create(ppm_ctx* c, const ppm_scene* sc, const ppm_pass_cfg* cfg, ppm_session** out) {
  return session_create(c, sc, cfg, out, true);
}

// Enqueue the kernels of one pass of pipeline p on stream st.  The eye
// branch (eye rays, query order) depends only on the pass state, so it runs
// on p.s2 beside the photon branch (trace, grid build) and joins before the
// gather; ev != nullptr (stage timing) runs every stage in order on st and
// records its events.
static hipError_t enqueue_pass(ppm_session* s, uint32_t j, hipStream_t st, hipEvent_t* ev) {
  const ppm_pass_cfg& cfg = s->cfg;
  const ppm_scene* sc = s->scene;
  Pipe& p = s->pipes[j];
  hipError_t e;
  const float4 bmin = make_float4((float)sc->bmin[0], (float)sc->bmin[1], (float)sc->bmin[2], 0.f);
  const float4 bmax = make_float4((float)sc->bmax[0], (float)sc->bmax[1], (float)sc->bmax[2], 0.f);
  const int n_cu = s->ctx->n_cu;
  const uint32_t K = (uint32_t)s->pipes.size();
  if (ev && (e = hipEventRecord(ev[0], st)) != hipSuccess) return e;
  if ((e = launch_pass_begin(st, p.st, cfg.first_pass + j, cfg.radius0 * cfg.radius0, cfg.alpha, s->power, K)) !=
      hipSuccess)
    return e;
  if (!ev) {
    if ((e = hipEventRecord(p.evf, st)) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(p.s2, p.evf, 0)) != hipSuccess) return e;
    if ((e = launch_eye(p.s2, s->sbuf, p.st, p.eb, cfg.seed, cfg.use_classic, cfg.max_depth, n_cu)) != hipSuccess)
      return e;
    if ((e = launch_query_prep(p.s2, s->sbuf, p.map, p.st, p.eb)) != hipSuccess) return e;
    if ((e = hipEventRecord(p.evj, p.s2)) != hipSuccess) return e;
  }
  if ((e = launch_trace_photons(st, s->sbuf, p.st, p.pb, cfg.seed, cfg.use_classic, cfg.max_depth)) != hipSuccess)
    return e;
  if (ev && (e = hipEventRecord(ev[1], st)) != hipSuccess) return e;
  if ((e = launch_grid_build(st, p.map, p.st, p.pb, bmin, bmax, 2ull * s->sbuf.n_emit_total)) != hipSuccess) return e;
  if (ev && (e = hipEventRecord(ev[2], st)) != hipSuccess) return e;
  if (ev) {
    if ((e = launch_eye(st, s->sbuf, p.st, p.eb, cfg.seed, cfg.use_classic, cfg.max_depth, n_cu)) != hipSuccess)
      return e;
    if ((e = hipEventRecord(ev[3], st)) != hipSuccess) return e;
    if ((e = launch_query_prep(st, s->sbuf, p.map, p.st, p.eb)) != hipSuccess) return e;
  } else if ((e = hipStreamWaitEvent(st, p.evj, 0)) != hipSuccess) {
    return e;
  }
  if ((e = launch_gather(st, s->sbuf, p.map, p.st, p.eb, cfg.pfilter, cfg.use_classic, n_cu, p.sum, p.sumsq,
                         ev ? ev + 4 : nullptr, true)) != hipSuccess)
    return e;
  if ((e = launch_pass_end(st, p.st)) != hipSuccess) return e;
  if (ev && (e = hipEventRecord(ev[6], st)) != hipSuccess) return e;
  return hipSuccess;
}

// The pass graph of pipeline j, captured once.  A capture that fails
// part-way leaves no graph behind, and the capture stream and the eye
// stream forked into it are replaced by fresh ones.
static hipError_t pipe_graph(ppm_session* s, uint32_t j) {
  Pipe& p = s->pipes[j];
  if (p.exec) return hipSuccess;
  ppm_ctx* c = s->ctx;
  hipStream_t cap;
  hipError_t e = ctx_stream(c, AUX_CAPTURE, &cap);
  if (e != hipSuccess) return e;
  e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    const hipError_t e2 = enqueue_pass(s, j, cap, nullptr);
    const hipError_t e3 = hipStreamEndCapture(cap, &p.graph);
    e = e2 != hipSuccess ? e2 : e3;
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&p.exec, p.graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    pipe_drop_graph(p);
    ctx_stream_renew(c, AUX_CAPTURE);
    ctx_stream_renew(c, AUX_EYE + (int)j);
    p.s2 = nullptr;
    const hipError_t e4 = ctx_stream(c, AUX_EYE + (int)j, &p.s2);
    (void)e4;  // a missing eye stream fails the next capture, not this report
  }
  return e;
}

ppm_status ppm_session_run(ppm_session* s, uint32_t n_pass) {
  if (!s) return PPM_E_INVALID;
  ppm_ctx* c = s->ctx;
  hipStream_t st = c->stream;
  const uint32_t K = (uint32_t)s->pipes.size();
  const bool timing = (s->cfg.flags & PPM_SESSION_STAGE_TIMING) != 0;
  const bool graph = !timing && (s->cfg.flags & PPM_SESSION_NO_GRAPH) == 0;
  if (timing) {
    for (uint32_t i = 0; i < n_pass; ++i) {
      HIP_TRY(c, enqueue_pass(s, (s->passes_done + i) % K, st, s->ev));
      HIP_TRY(c, hipEventSynchronize(s->ev[6]));
      float ms[6];
      for (int k = 0; k < 6; ++k) HIP_TRY(c, hipEventElapsedTime(&ms[k], s->ev[k], s->ev[k + 1]));
      s->t.ms_photon_trace += ms[0];
      s->t.ms_map_build += ms[1];
      s->t.ms_eye_trace += ms[2];
      s->t.ms_gather += ms[4];
      s->t.ms_accumulate += ms[3] + ms[5];
      s->t.ms_total += ms[0] + ms[1] + ms[2] + ms[3] + ms[4] + ms[5];
    }
  } else {
    if (graph)
      for (uint32_t j = 0; j < K; ++j) {
        const hipError_t e = pipe_graph(s, j);
        if (e != hipSuccess) return fail(c, PPM_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
      }
    // every pipeline starts after the context stream's earlier work and the
    // context stream waits for all of them: the caller sees one stream
    HIP_TRY(c, hipEventRecord(s->ev[8], st));
    for (Pipe& p : s->pipes) HIP_TRY(c, hipStreamWaitEvent(p.stream, s->ev[8], 0));
    for (uint32_t i = 0; i < n_pass; ++i) {
      const uint32_t j = (s->passes_done + i) % K;
      Pipe& p = s->pipes[j];
      if (graph)
        HIP_TRY(c, hipGraphLaunch(p.exec, p.stream));
      else
        HIP_TRY(c, enqueue_pass(s, j, p.stream, nullptr));
    }
    for (Pipe& p : s->pipes) {
      HIP_TRY(c, hipEventRecord(p.done, p.stream));
      HIP_TRY(c, hipStreamWaitEvent(st, p.done, 0));
    }
    HIP_TRY(c, hipEventRecord(s->ev[9], st));
    HIP_TRY(c, hipEventSynchronize(s->ev[9]));
    float ms = 0;
    HIP_TRY(c, hipEventElapsedTime(&ms, s->ev[8], s->ev[9]));
    s->t.ms_total += ms;
  }
  s->passes_done += n_pass;
  s->t.passes = s->passes_done;
  return PPM_OK;
}

// The pass states of all pipelines (after the context stream's work).
static hipError_t pipe_states(ppm_session* s, std::vector<PassState>& h) {
  h.resize(s->pipes.size());
  hipError_t e = hipStreamSynchronize(s->ctx->stream);
  for (size_t j = 0; j < s->pipes.size() && e == hipSuccess; ++j)
    e = hipMemcpy(&h[j], s->pipes[j].st, sizeof(PassState), hipMemcpyDeviceToHost);
  return e;
}

// Sum and sum of squares over all passes: the pipelines' FP64 sums added
// in pipeline order.
static hipError_t merged_sums(ppm_session* s, std::vector<double>& sum, std::vector<double>* sq) {
  const uint64_t n3 = 3ull * s->pipes[0].eb.n_pix;
  sum.assign(n3, 0.0);
  if (sq) sq->assign(n3, 0.0);
  std::vector<double> a(n3);
  for (Pipe& p : s->pipes) {
    hipError_t e = hipMemcpy(a.data(), p.sum, sizeof(double) * n3, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return e;
    for (uint64_t i = 0; i < n3; ++i) sum[i] += a[i];
    if (sq) {
      if ((e = hipMemcpy(a.data(), p.sumsq, sizeof(double) * n3, hipMemcpyDeviceToHost)) != hipSuccess) return e;
      for (uint64_t i = 0; i < n3; ++i) (*sq)[i] += a[i];
    }
  }
  return hipSuccess;
}

ppm_status ppm_session_timings(ppm_session* s, ppm_timings* t) {
  if (!s || !t) return PPM_E_INVALID;
  ppm_ctx* c = s->ctx;
  std::vector<PassState> h;
  HIP_TRY(c, pipe_states(s, h));
  *t = s->t;
  t->photons_stored = 0;
  t->gather_queries = 0;
  for (const PassState& x : h) {
    t->photons_stored += x.total_stored;
    t->gather_queries += x.total_queries;
  }
  t->photons_emitted = (uint64_t)s->sbuf.n_emit_total * s->passes_done;
  t->passes = s->passes_done;
  return PPM_OK;
}

ppm_status ppm_session_reset(ppm_session* s) {
  if (!s) return PPM_E_INVALID;
  ppm_ctx* c = s->ctx;
  const ppm_scene* sc = s->scene;
  // the scene's lights as they are now (ppm_scene_set_params may have
  // scaled them): same photon counts, new photon power
  std::vector<DevLight> L = sc->lights;
  double power = 0.0;
  const uint64_t total = assign_photons(L, sc->total_flux, s->cfg.n_photon, &power);
  if (total != s->sbuf.n_emit_total) return fail(c, PPM_E_INVALID, "the scene's photon counts changed; create a new session");
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  HIP_TRY(c, hipMemcpy(s->lights, L.data(), sizeof(DevLight) * L.size(), hipMemcpyHostToDevice));
  s->power = power;
  const PassState h = host_state();
  for (Pipe& p : s->pipes) {
    const uint64_t n3 = 3ull * p.eb.n_pix;
    HIP_TRY(c, hipMemset(p.sum, 0, sizeof(double) * n3));
    HIP_TRY(c, hipMemset(p.sumsq, 0, sizeof(double) * n3));
    HIP_TRY(c, hipMemcpy(p.st, &h, sizeof(h), hipMemcpyHostToDevice));
    // the captured pass has the photon power as an argument: capture again
    pipe_drop_graph(p);
  }
  s->passes_done = 0;
  s->t = ppm_timings{};
  return PPM_OK;
}

// The tile shift of a tile width (4, 8, 16 or 32 pixels; 0 = 8).
static bool tile_shift_of(uint32_t width, uint32_t* shift) {
  if (width == 0) width = 1u << TILE_SHIFT_DEFAULT;
  for (uint32_t k = TILE_SHIFT_MIN; k <= TILE_SHIFT_MAX; ++k)
    if (width == (1u << k)) {
      *shift = k;
      return true;
    }
  return false;
}

// Every pipeline's eye buffers and the session's camera at tile shape shift
// (the pipelines idle; their captured passes are captured again).
static hipError_t session_set_tiles(ppm_session* s, uint32_t shift) {
  for (Pipe& p : s->pipes) {
    hipError_t e;
    if (p.stream && (e = hipStreamSynchronize(p.stream)) != hipSuccess) return e;
    if (p.s2 && (e = hipStreamSynchronize(p.s2)) != hipSuccess) return e;
  }
  for (Pipe& p : s->pipes) {
    const hipError_t e = eye_set_tiles(p.eb, s->sbuf.cam, shift);
    if (e != hipSuccess) return e;
    pipe_drop_graph(p);
  }
  return hipSuccess;
}

ppm_status ppm_session_set_tile_width(ppm_session* s, uint32_t width) {
  if (!s) return PPM_E_INVALID;
  uint32_t shift = 0;
  if (!tile_shift_of(width, &shift)) return fail(s->ctx, PPM_E_INVALID, "tile width must be 4, 8, 16 or 32");
  HIP_TRY(s->ctx, session_set_tiles(s, shift));
  return PPM_OK;
}

ppm_status ppm_session_backward(ppm_session* s, const double* grad_image, double* grad_albedo, double* grad_light) {
  if (!s || !grad_image || !grad_albedo) return PPM_E_INVALID;
  ppm_ctx* c = s->ctx;
  const uint32_t K = s->passes_done;
  if (K == 0) return fail(c, PPM_E_INVALID, "backward before any pass");
  const ppm_pass_cfg& cfg = s->cfg;
  const ppm_scene* sc = s->scene;
  Pipe& p0 = s->pipes[0];
  const uint32_t n_pix = p0.eb.n_pix;
  const uint32_t nm = s->sbuf.n_mats;
  // dL/d(pass image) = dL/d(mean image) / K
  std::vector<float4> g(n_pix);
  for (uint32_t p = 0; p < n_pix; ++p) {
    const double* v = grad_image + 3ull * p;
    if (!finite3(v)) return fail(c, PPM_E_NONFINITE, "non-finite image gradient");
    g[p] = make_float4((float)(v[0] / K), (float)(v[1] / K), (float)(v[2] / K), 0.0f);
  }
  float4* gimg = nullptr;
  float *adj_sorted = nullptr, *adj_slot = nullptr;
  double* grad = nullptr;
  PassState* st = nullptr;
  const float4 bmin = make_float4((float)sc->bmin[0], (float)sc->bmin[1], (float)sc->bmin[2], 0.f);
  const float4 bmax = make_float4((float)sc->bmax[0], (float)sc->bmax[1], (float)sc->bmax[2], 0.f);
  hipStream_t q = c->stream;
  hipError_t e;
  if ((e = dalloc(&gimg, n_pix)) == hipSuccess && (e = dalloc(&adj_sorted, p0.map.cap)) == hipSuccess &&
      (e = dalloc(&adj_slot, p0.pb.cap)) == hipSuccess && (e = dalloc(&grad, 3ull * nm)) == hipSuccess &&
      (e = dalloc(&st, 1)) == hipSuccess &&
      (e = hipMemcpy(gimg, g.data(), sizeof(float4) * n_pix, hipMemcpyHostToDevice)) == hipSuccess &&
      (e = hipMemset(grad, 0, sizeof(double) * 3 * nm)) == hipSuccess) {
    const PassState h = host_state();
    e = hipMemcpy(st, &h, sizeof(h), hipMemcpyHostToDevice);
    // replay the session's passes from their seeds (the passes' only
    // checkpoint), each followed by its adjoint
    for (uint32_t k = 0; k < K && e == hipSuccess; ++k) {
      if ((e = launch_pass_begin(q, st, cfg.first_pass, cfg.radius0 * cfg.radius0, cfg.alpha, s->power)) != hipSuccess)
        break;
      // the eye branch beside the photon branch, as in a forward pass
      if ((e = hipEventRecord(p0.evf, q)) != hipSuccess || (e = hipStreamWaitEvent(p0.s2, p0.evf, 0)) != hipSuccess)
        break;
      if ((e = launch_eye(p0.s2, s->sbuf, st, p0.eb, cfg.seed, cfg.use_classic, cfg.max_depth, c->n_cu)) !=
          hipSuccess)
        break;
      if ((e = launch_query_prep(p0.s2, s->sbuf, p0.map, st, p0.eb)) != hipSuccess) break;
      if ((e = hipEventRecord(p0.evj, p0.s2)) != hipSuccess) break;
      if ((e = launch_trace_photons(q, s->sbuf, st, p0.pb, cfg.seed, cfg.use_classic, cfg.max_depth)) != hipSuccess)
        break;
      if ((e = launch_grid_build(q, p0.map, st, p0.pb, bmin, bmax, 2ull * s->sbuf.n_emit_total)) != hipSuccess) break;
      if ((e = hipStreamWaitEvent(q, p0.evj, 0)) != hipSuccess) break;
      if ((e = launch_gather(q, s->sbuf, p0.map, st, p0.eb, cfg.pfilter, cfg.use_classic, c->n_cu, nullptr,
                             nullptr, nullptr, true)) != hipSuccess)
        break;
      if ((e = launch_adjoint(q, s->sbuf, p0.map, st, p0.pb, p0.eb, cfg.pfilter, c->n_cu, cfg.seed,
                              cfg.use_classic, cfg.max_depth, gimg, adj_sorted, adj_slot, grad)) != hipSuccess)
        break;
      e = launch_pass_end(q, st);
    }
    // an error may leave the eye branch running: drain it before the frees
    const hipError_t e2 = hipStreamSynchronize(p0.s2);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipStreamSynchronize(q);
  }
  std::vector<double> hg(3ull * nm, 0.0);
  if (e == hipSuccess) e = hipMemcpy(hg.data(), grad, sizeof(double) * hg.size(), hipMemcpyDeviceToHost);
  dfree(gimg);
  dfree(adj_sorted);
  dfree(adj_slot);
  dfree(grad);
  dfree(st);
  if (e != hipSuccess) return fail(c, e == hipErrorOutOfMemory ? PPM_E_OOM : PPM_E_HIP, hipGetErrorString(e));
  for (size_t i = 0; i < hg.size(); ++i) grad_albedo[i] = hg[i];
  if (grad_light) {
    // the image is proportional to the light scale: dL/ds = sum g * mean / s
    std::vector<double> sum;
    HIP_TRY(c, merged_sums(s, sum, nullptr));
    double acc = 0.0;
    for (size_t i = 0; i < sum.size(); ++i) acc += grad_image[i] * (sum[i] / K);
    *grad_light = sc->light_scale > 0.0 ? acc / sc->light_scale : 0.0;
  }
  return PPM_OK;
}

ppm_status ppm_session_radius2(ppm_session* s, double* r2) {
  if (!s || !r2) return PPM_E_INVALID;
  ppm_ctx* c = s->ctx;
  std::vector<PassState> h;
  HIP_TRY(c, pipe_states(s, h));
  // the state of the pipeline that ran the last pass
  const uint32_t K = (uint32_t)s->pipes.size();
  *r2 = h[s->passes_done ? (s->passes_done - 1) % K : 0].r2;
  return PPM_OK;
}

ppm_status ppm_session_image(ppm_session* s, double* mean_out, double* var_out, uint32_t* passes_done) {
  if (!s) return PPM_E_INVALID;
  ppm_ctx* c = s->ctx;
  std::vector<PassState> h;
  HIP_TRY(c, pipe_states(s, h));
  if (passes_done) *passes_done = s->passes_done;
  for (const PassState& x : h) {
    if (x.overflow)
      return fail(c, PPM_E_RANGE, "capacity exceeded during a pass: " + overflow_text(x.overflow));
    if (x.nonfinite) return fail(c, PPM_E_NONFINITE, "non-finite photon position (photonmap.rs:26)");
  }
  std::vector<double> sum, sq;
  HIP_TRY(c, merged_sums(s, sum, &sq));
  const uint64_t n3 = sum.size();
  const double n = (double)s->passes_done;
  for (uint64_t i = 0; i < n3; ++i) {
    if (mean_out) mean_out[i] = n > 0 ? sum[i] / n : 0.0;
    if (var_out) var_out[i] = n > 1 ? std::max(0.0, (sq[i] - sum[i] * sum[i] / n) / (n - 1.0)) : 0.0;
  }
  return PPM_OK;
}

// ---------------------------------------------------------------------------
// Batches: many scenes with their own pass schedules, one launch per stage.

}  // extern "C"

struct ppm_batch {
  ppm_ctx* ctx = nullptr;
  std::vector<ppm_session*> sess;  // caller's order
  std::vector<uint32_t> n_pass;    // caller's order
  std::vector<uint32_t> order;     // caller indices by pass count, descending
  uint32_t K = 1;                  // pipelines: round r runs on pipeline r mod K
  BatchItem* items = nullptr;      // device, K x n (pipeline-major), each in `order`
  uint32_t* tab = nullptr;         // device, BK_KINDS x (n + 1)
  std::vector<uint32_t> htab;
  bool bvh = false;
  uint32_t dyn = 0;                // dynamic LDS of the BVH kernels: the most any scene asks (bvh_dyn_lds)
  ppm_launch_cfg lc{};
  std::vector<hipGraphExec_t> exec;  // per pipeline j and number of running scenes a: j * n + a - 1
  std::vector<hipGraph_t> graph;
  std::vector<hipStream_t> ps;       // pipeline j's stream (0: the context's)
  std::vector<hipEvent_t> pdone;     // pipeline j's last round
  hipEvent_t ev[2] = {};
  double ms_total = 0.0;
  uint32_t rounds_done = 0;
};

namespace {

void batch_drop_graphs(ppm_batch* b) {
  for (auto& e : b->exec)
    if (e) (void)hipGraphExecDestroy(e), e = nullptr;
  for (auto& g : b->graph)
    if (g) (void)hipGraphDestroy(g), g = nullptr;
}

// Per-scene block counts of every launch kind -> prefix table (host and
// device) for the current launch configuration.
hipError_t batch_layout(ppm_batch* b) {
  const uint32_t n = (uint32_t)b->order.size();
  const uint32_t n_cu = (uint32_t)b->ctx->n_cu;
  const uint32_t ew = b->lc.eye_waves_per_cu ? b->lc.eye_waves_per_cu : 16u;
  const uint32_t sb = b->lc.shade_blocks_per_cu ? b->lc.shade_blocks_per_cu : 8u;
  uint32_t shift = TILE_SHIFT_DEFAULT;
  if (!tile_shift_of(b->lc.tile_width, &shift)) return hipErrorInvalidValue;
  for (ppm_session* s : b->sess)
    if (s->sbuf.cam.tile_shift != (int32_t)shift) {
      const hipError_t e = session_set_tiles(s, shift);
      if (e != hipSuccess) return e;
    }
  std::vector<BatchItem> items((size_t)b->K * n);
  b->htab.assign((size_t)BK_KINDS * (n + 1), 0u);
  for (uint32_t j = 0; j < n; ++j) {
    const ppm_session* s = b->sess[b->order[j]];
    BatchItem& it = items[j];
    std::memset(&it, 0, sizeof(it));
    it.sc = s->sbuf;
    it.st = s->pipes[0].st;
    it.pb = s->pipes[0].pb;
    it.m = s->pipes[0].map;
    it.eb = s->pipes[0].eb;
    it.sum = s->pipes[0].sum;
    it.sumsq = s->pipes[0].sumsq;
    it.seed = s->cfg.seed;
    it.r2_first = s->cfg.radius0 * s->cfg.radius0;
    it.alpha = s->cfg.alpha;
    it.power = s->power;
    it.mscale = morton_scale(s->sbuf.bmin, s->sbuf.bmax, QUERY_MORTON_MAX);
    it.first_pass = s->cfg.first_pass;
    it.pass_stride = b->K;
    it.uc = s->cfg.use_classic;
    it.maxd = s->cfg.max_depth;
    it.pfilter = s->cfg.pfilter;
    it.trace_per_wave = (uint32_t)(((uint64_t)s->sbuf.n_emit_total + it.pb.n_regions - 1) / it.pb.n_regions);
    it.eye_waves = std::max(1u, std::min(std::min(it.eb.n_tiles, n_cu * ew), it.eb.estack_lanes / 64u));
    // the sorts' blocks from the expected key counts (about 2 stored photons
    // per emitted one, 1.5 queries per pixel): the histogram blocks stride
    // and the digit passes' persistent blocks take tiles until none is left,
    // so any count sorts every key, and a batch's sorts together need no
    // more blocks than fill the GPU
    const uint64_t pkeys = std::min<uint64_t>(it.m.cap, 2ull * s->sbuf.n_emit_total);
    const uint64_t qkeys = 3ull * it.eb.q_primary / 2;
    auto hist_blocks = [](uint64_t k) { return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(OS_HIST_BLOCKS, k / 16384)); };
    auto os_blocks = [](const OneSweep& o, uint64_t k) {
      return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(o.blocks, (k + OS_TILE - 1) / OS_TILE));
    };
    uint32_t cnt[BK_KINDS];
    cnt[BK_TRACE] = (it.pb.n_regions + 3) / 4;
    // the grid kernels over the photons the room is expected to store
    const uint64_t expect = std::min<uint64_t>(it.m.cap, 2ull * s->sbuf.n_emit_total);
    cnt[BK_MORTON] = grid_stride_blocks(expect);
    cnt[BK_PHIST] = hist_blocks(pkeys);
    cnt[BK_POS] = os_blocks(it.m.os, pkeys);
    cnt[BK_LEAVES] = grid_stride_blocks(expect);
    cnt[BK_CLEAR] = grid_stride_blocks(std::min<uint64_t>(it.m.cells_max, next_pow2_u64(expect)));
    cnt[BK_EYE] = (it.eye_waves + 3) / 4;
    cnt[BK_PIX] = (it.eb.q_primary + 255) / 256;
    cnt[BK_QKEYS] = (it.eb.q_primary + 255) / 256 + OVF_BLOCKS;
    cnt[BK_QHIST] = hist_blocks(qkeys);
    cnt[BK_QOS] = os_blocks(it.eb.qos, qkeys);
    cnt[BK_SHADE] = std::max(1u, std::min(n_cu * sb, it.eb.q_primary / 128u));
    cnt[BK_OVF] = OVF_BLOCKS;
    for (int k = 0; k < BK_KINDS; ++k) b->htab[(size_t)k * (n + 1) + j + 1] = b->htab[(size_t)k * (n + 1) + j] + cnt[k];
    // the other pipelines: the same sizes on their own buffers, passes
    // first_pass + q, + q + K, ...
    for (uint32_t q = 1; q < b->K; ++q) {
      BatchItem& iq = items[(size_t)q * n + j];
      const Pipe& p = s->pipes[q];
      iq = it;
      iq.st = p.st;
      iq.pb = p.pb;
      iq.m = p.map;
      iq.eb = p.eb;
      iq.sum = p.sum;
      iq.sumsq = p.sumsq;
      iq.first_pass = s->cfg.first_pass + q;
    }
  }
  hipError_t e;
  if ((e = hipMemcpy(b->items, items.data(), items.size() * sizeof(BatchItem), hipMemcpyHostToDevice)) != hipSuccess)
    return e;
  if ((e = hipMemcpy(b->tab, b->htab.data(), b->htab.size() * sizeof(uint32_t), hipMemcpyHostToDevice)) != hipSuccess)
    return e;
  batch_drop_graphs(b);
  b->exec.assign((size_t)b->K * n, nullptr);
  b->graph.assign((size_t)b->K * n, nullptr);
  return hipSuccess;
}

// One pass of the first `act` scenes (in `order`) on pipeline q, through a
// captured graph.
hipError_t batch_round(ppm_batch* b, uint32_t act, uint32_t q) {
  const uint32_t n = (uint32_t)b->order.size();
  const size_t gi = (size_t)q * n + act - 1;
  if (!b->exec[gi]) {
    uint32_t blocks[BK_KINDS];
    for (int k = 0; k < BK_KINDS; ++k) blocks[k] = b->htab[(size_t)k * (n + 1) + act];
    // captured on the context's capture streams; the fork/join events live
    // only while the graph is captured.  A capture that fails part-way
    // leaves no graph behind and replaces both capture streams.
    hipStream_t cap = nullptr, cap2 = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    hipError_t e = ctx_stream(b->ctx, AUX_CAPTURE, &cap);
    if (e == hipSuccess) e = ctx_stream(b->ctx, AUX_CAPTURE2, &cap2);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&join, hipEventDisableTiming);
    if (e == hipSuccess) e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      const hipError_t e2 =
          launch_batch_pass(cap, cap2, fork, join, b->items + (size_t)q * n, b->tab, blocks, act, n, b->bvh, b->dyn);
      const hipError_t e3 = hipStreamEndCapture(cap, &b->graph[gi]);
      e = e2 != hipSuccess ? e2 : e3;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&b->exec[gi], b->graph[gi], nullptr, nullptr, 0);
    if (join) (void)hipEventDestroy(join);
    if (fork) (void)hipEventDestroy(fork);
    if (e != hipSuccess) {
      if (b->exec[gi]) (void)hipGraphExecDestroy(b->exec[gi]);
      if (b->graph[gi]) (void)hipGraphDestroy(b->graph[gi]);
      b->exec[gi] = nullptr;
      b->graph[gi] = nullptr;
      ctx_stream_renew(b->ctx, AUX_CAPTURE);
      ctx_stream_renew(b->ctx, AUX_CAPTURE2);
      return e;
    }
  }
  return hipGraphLaunch(b->exec[gi], b->ps[q]);
}

// Every scene back to before its first pass.
hipError_t batch_reset(ppm_batch* b) {
  const PassState h = host_state();
  for (ppm_session* s : b->sess) {
    for (Pipe& p : s->pipes) {
      const uint64_t n3 = 3ull * p.eb.n_pix;
      hipError_t e;
      if ((e = hipMemsetAsync(p.sum, 0, sizeof(double) * n3, b->ctx->stream)) != hipSuccess) return e;
      if ((e = hipMemsetAsync(p.sumsq, 0, sizeof(double) * n3, b->ctx->stream)) != hipSuccess) return e;
      if ((e = hipMemcpyAsync(p.st, &h, sizeof(h), hipMemcpyHostToDevice, b->ctx->stream)) != hipSuccess) return e;
    }
    s->passes_done = 0;
  }
  b->rounds_done = 0;
  return hipStreamSynchronize(b->ctx->stream);
}

}  // namespace

extern "C" {

void ppm_batch_destroy(ppm_batch* b) {
  if (!b) return;
  if (b->ctx) (void)hipStreamSynchronize(b->ctx->stream);
  for (size_t q = 1; q < b->ps.size(); ++q)  // the context's streams: waited for, not destroyed
    if (b->ps[q]) (void)hipStreamSynchronize(b->ps[q]);
  for (hipEvent_t e : b->pdone)
    if (e) (void)hipEventDestroy(e);
  batch_drop_graphs(b);
  for (auto& e : b->ev)
    if (e) (void)hipEventDestroy(e);
  dfree(b->items);
  dfree(b->tab);
  for (ppm_session* s : b->sess) ppm_session_destroy(s);
  delete b;
}

ppm_status ppm_batch_create(ppm_ctx* c, uint32_t n, const ppm_scene* const* scenes, const ppm_pass_cfg* cfgs,
                            const uint32_t* n_passes, ppm_batch** out) {
  if (!c || !scenes || !cfgs || !n_passes || !out || n == 0) return fail(c, PPM_E_INVALID, "null argument or empty batch");
  *out = nullptr;
  if (n > 4096) return fail(c, PPM_E_RANGE, "at most 4096 scenes per batch");
  ppm_batch* b = new ppm_batch();
  b->ctx = c;
  b->n_pass.assign(n_passes, n_passes + n);
  // pipelines: the smallest cfgs[i].pipelines, 3 where it is 0 (automatic):
  // three rounds in flight by default, since a full round keeps the GPU busy
  // and a third pipeline fills the tails of each stage, where the rooms with
  // the most work are still running (64-room workload: 3.19 s with two,
  // 3.11 s with three, 3.18 s with four; a session gains nothing past two)
  b->K = 4;
  for (uint32_t i = 0; i < n; ++i) b->K = std::min<uint32_t>(b->K, cfgs[i].pipelines ? cfgs[i].pipelines : 3u);
  if (b->K == 0) b->K = 1;
  for (uint32_t i = 0; i < n; ++i) {
    if (!scenes[i]) {
      ppm_batch_destroy(b);
      return fail(c, PPM_E_INVALID, "null scene in batch");
    }
    ppm_pass_cfg cfg = cfgs[i];
    cfg.flags = 0;
    cfg.pipelines = b->K;
    ppm_session* s = nullptr;
    if (ppm_status st = session_create(c, scenes[i], &cfg, &s, false)) {
      ppm_batch_destroy(b);
      return st;
    }
    b->sess.push_back(s);
    if (scenes[i]->buf.bvh) b->bvh = true;
    b->dyn = std::max(b->dyn, bvh_dyn_lds(scenes[i]->buf));
  }
  b->order.resize(n);
  for (uint32_t i = 0; i < n; ++i) b->order[i] = i;
  std::stable_sort(b->order.begin(), b->order.end(),
                   [&](uint32_t a, uint32_t d) { return b->n_pass[a] > b->n_pass[d]; });
  hipError_t e;
  b->ps.assign(b->K, nullptr);
  b->pdone.assign(b->K, nullptr);
  b->ps[0] = c->stream;
  e = hipSuccess;
  for (uint32_t q = 1; q < b->K && e == hipSuccess; ++q) {
    e = ctx_stream(c, AUX_PIPE + (int)q - 1, &b->ps[q]);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->pdone[q], hipEventDisableTiming);
  }
  if (e != hipSuccess || (e = dalloc(&b->items, (uint64_t)b->K * n)) != hipSuccess ||
      (e = dalloc(&b->tab, (uint64_t)BK_KINDS * (n + 1))) != hipSuccess ||
      (e = hipEventCreate(&b->ev[0])) != hipSuccess || (e = hipEventCreate(&b->ev[1])) != hipSuccess ||
      (e = batch_layout(b)) != hipSuccess) {
    ppm_batch_destroy(b);
    return fail(c, e == hipErrorOutOfMemory ? PPM_E_OOM : PPM_E_HIP, hipGetErrorString(e));
  }
  *out = b;
  return PPM_OK;
}

ppm_status ppm_batch_set_launch(ppm_batch* b, const ppm_launch_cfg* lc) {
  if (!b || !lc) return PPM_E_INVALID;
  if (lc->eye_waves_per_cu > 64 || lc->shade_blocks_per_cu > 64)
    return fail(b->ctx, PPM_E_RANGE, "at most 64 eye waves / shade blocks per CU");
  uint32_t shift = 0;
  if (!tile_shift_of(lc->tile_width, &shift)) return fail(b->ctx, PPM_E_INVALID, "tile width must be 4, 8, 16 or 32");
  b->lc = *lc;
  HIP_TRY(b->ctx, batch_layout(b));
  return PPM_OK;
}

ppm_status ppm_batch_get_launch(const ppm_batch* b, ppm_launch_cfg* lc) {
  if (!b || !lc) return PPM_E_INVALID;
  *lc = b->lc;
  if (!lc->eye_waves_per_cu) lc->eye_waves_per_cu = 16;
  if (!lc->shade_blocks_per_cu) lc->shade_blocks_per_cu = 8;
  if (!lc->tile_width) lc->tile_width = 8;
  return PPM_OK;
}

// Passes until every scene has run n_passes[i] of them (or `max_rounds`
// more rounds, when it is not 0).
ppm_status ppm_batch_run(ppm_batch* b, uint32_t max_rounds) {
  if (!b) return PPM_E_INVALID;
  ppm_ctx* c = b->ctx;
  const uint32_t n = (uint32_t)b->order.size();
  const uint32_t total = b->n_pass[b->order[0]];
  uint32_t end = total;
  if (max_rounds && b->rounds_done + max_rounds < end) end = b->rounds_done + max_rounds;
  HIP_TRY(c, hipEventRecord(b->ev[0], c->stream));
  for (uint32_t q = 1; q < b->K; ++q) HIP_TRY(c, hipStreamWaitEvent(b->ps[q], b->ev[0], 0));
  for (uint32_t r = b->rounds_done; r < end; ++r) {
    uint32_t act = n;
    while (act > 0 && b->n_pass[b->order[act - 1]] <= r) --act;
    if (act == 0) break;
    HIP_TRY(c, batch_round(b, act, r % b->K));
    for (uint32_t j = 0; j < act; ++j) b->sess[b->order[j]]->passes_done += 1;
  }
  for (uint32_t q = 1; q < b->K; ++q) {
    HIP_TRY(c, hipEventRecord(b->pdone[q], b->ps[q]));
    HIP_TRY(c, hipStreamWaitEvent(c->stream, b->pdone[q], 0));
  }
  HIP_TRY(c, hipEventRecord(b->ev[1], c->stream));
  HIP_TRY(c, hipEventSynchronize(b->ev[1]));
  float ms = 0.0f;
  HIP_TRY(c, hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
  b->ms_total += ms;
  b->rounds_done = end;
  return PPM_OK;
}

ppm_status ppm_batch_tune(ppm_batch* b, uint32_t rounds, ppm_launch_cfg* best) {
  if (!b || rounds == 0) return PPM_E_INVALID;
  ppm_ctx* c = b->ctx;
  static const uint32_t eye_c[] = {4, 8, 16, 32};
  static const uint32_t shade_c[] = {2, 4, 8, 16};
  static const uint32_t tile_c[] = {4, 16, 32};
  ppm_launch_cfg pick = b->lc;
  double best_ms = INFINITY;
  auto trial = [&](const ppm_launch_cfg& lc) -> ppm_status {
    b->lc = lc;
    HIP_TRY(c, batch_layout(b));
    HIP_TRY(c, batch_reset(b));
    // the first round captures the graph; the next `rounds` are timed
    if (ppm_status st = ppm_batch_run(b, 1)) return st;
    const double t0 = b->ms_total;
    if (ppm_status st = ppm_batch_run(b, rounds)) return st;
    const double ms = b->ms_total - t0;
    if (ms < best_ms) {
      best_ms = ms;
      pick = lc;
    }
    return PPM_OK;
  };
  // launch sizes at 8-pixel-wide tiles, then the tile shape at the best sizes
  for (uint32_t ew : eye_c)
    for (uint32_t sb : shade_c) {
      ppm_launch_cfg lc{};
      lc.eye_waves_per_cu = ew;
      lc.shade_blocks_per_cu = sb;
      lc.tile_width = 8;
      if (ppm_status st = trial(lc)) return st;
    }
  const ppm_launch_cfg sizes = pick;
  for (uint32_t tw : tile_c) {
    ppm_launch_cfg lc = sizes;
    lc.tile_width = tw;
    if (ppm_status st = trial(lc)) return st;
  }
  b->lc = pick;
  HIP_TRY(c, batch_layout(b));
  HIP_TRY(c, batch_reset(b));
  b->ms_total = 0.0;
  if (best) ppm_batch_get_launch(b, best);
  return PPM_OK;
}

ppm_status ppm_batch_image(ppm_batch* b, uint32_t i, double* mean_out, double* var_out, uint32_t* passes_done) {
  if (!b || i >= b->sess.size()) return PPM_E_INVALID;
  return ppm_session_image(b->sess[i], mean_out, var_out, passes_done);
}

ppm_status ppm_batch_timings(ppm_batch* b, ppm_timings* t) {
  if (!b || !t) return PPM_E_INVALID;
  std::memset(t, 0, sizeof(*t));
  for (ppm_session* s : b->sess) {
    ppm_timings u;
    if (ppm_status st = ppm_session_timings(s, &u)) return st;
    t->photons_emitted += u.photons_emitted;
    t->photons_stored += u.photons_stored;
    t->gather_queries += u.gather_queries;
  }
  t->ms_total = b->ms_total;
  t->passes = b->rounds_done;
  return PPM_OK;
}

ppm_status ppm_run_passes(ppm_ctx* c, const ppm_scene* sc, const ppm_pass_cfg* cfg, uint32_t n_pass, double* mean_out,
                          double* var_out, ppm_timings* t) {
  ppm_session* s = nullptr;
  if (ppm_status st = ppm_session_create(c, sc, cfg, &s)) return st;
  ppm_status st = ppm_session_run(s, n_pass);
  if (st == PPM_OK) st = ppm_session_image(s, mean_out, var_out, nullptr);
  if (st == PPM_OK && t) st = ppm_session_timings(s, t);
  ppm_session_destroy(s);
  return st;
}

}  // extern "C"
