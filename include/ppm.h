/*
 * ppm.h — C ABI of the MI355X (gfx950) progressive-photon-mapping (PPM-PA)
 * hot path.
 *
 * The reference (eijian/ppmpa, a Rust crate) has no FFI layer: its GPU-side
 * replacement sits behind the crate's public call surface, which the CLIs in
 * src/bin/ppmpa.rs, pm.rs and rt.rs call directly.  Every entry point below names the reference
 * function it replaces.  Calls are batch-granular (one call per pass or per
 * batch of queries), never one call per photon.
 *
 * Conventions
 *  - Every call returns a ppm_status (0 = ok, < 0 = error); nothing aborts.
 *    ppm_last_error() returns a text description of the last failure on a
 *    context.  (The reference panics instead: photonmap.rs:26,71 unwrap on
 *    non-finite kd coordinates; algebra.rs:220, camera.rs:74 on zero-length
 *    normalisation; photonmap.rs:62-67 on malformed map text.)
 *  - The caller owns host buffers.  Device buffers are owned by the object
 *    that allocated them and are freed by the matching *_destroy.  Destroy
 *    every object made on a context before ppm_destroy on it: their
 *    *_destroy calls wait on the context's streams, which ppm_destroy frees.
 *  - All calls on one context are ordered on that context's HIP stream and
 *    are not thread-safe; use one context per host thread.
 *  - Geometry and colours cross the boundary as double; the device keeps
 *    traversal, intersection and gather in float and accumulates the
 *    per-pixel radiance and the radius schedule in double.
 */
#ifndef PPM_H
#define PPM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t ppm_status;

#define PPM_OK            0
#define PPM_E_INVALID    -1 /* bad argument or handle */
#define PPM_E_NONFINITE  -2 /* NaN/inf coordinate (photonmap.rs:26 panics) */
#define PPM_E_OOM        -3 /* device or host allocation failed */
#define PPM_E_HIP        -4 /* a HIP runtime call failed */
#define PPM_E_RANGE      -5 /* a size beyond a buffer's capacity */
#define PPM_E_PARSE      -6 /* malformed photon-map text */
#define PPM_E_NODEVICE   -7 /* no usable gfx950 device */

/* Shapes — geometry.rs:66-90 (Shape enum). */
#define PPM_SHAPE_POINT          0
#define PPM_SHAPE_PLAIN          1 /* infinite plane n.x + dist = 0 */
#define PPM_SHAPE_SPHERE         2
#define PPM_SHAPE_POLYGON        3 /* triangle p0, p0+dir1, p0+dir2 */
#define PPM_SHAPE_PARALLELOGRAM  4

/* Surfaces — surface.rs:16-41 (Surface enum). */
#define PPM_SURF_NOTHING  0
#define PPM_SURF_SIMPLE   1
#define PPM_SURF_TS       2
#define PPM_SURF_DISNEY   3
#define PPM_SURF_BRADY    4

/* Lights — light.rs:16-39 (Light enum). */
#define PPM_LIGHT_POINT          0
#define PPM_LIGHT_PARALLELOGRAM  1
#define PPM_LIGHT_SUN            2

/* Photon filters — optics.rs:17-21 (PhotonFilter enum), tracer.rs:197-216. */
#define PPM_FILTER_NONE   0
#define PPM_FILTER_CONE   1
#define PPM_FILTER_GAUSS  2

/* Wavelength channel codes — physics.rs:17-21. */
#define PPM_WL_RED    0
#define PPM_WL_GREEN  1
#define PPM_WL_BLUE   2

typedef struct ppm_ctx      ppm_ctx;
typedef struct ppm_scene    ppm_scene;
typedef struct ppm_photons  ppm_photons;
typedef struct ppm_map      ppm_map;
typedef struct ppm_session  ppm_session;

/* One object = shape + material index (object.rs:10-13). */
typedef struct {
  int32_t kind;       /* PPM_SHAPE_* */
  int32_t material;   /* index into ppm_scene_desc.materials */
  double  pos[3];     /* POINT position, SPHERE centre, POLYGON/PARALLELOGRAM p0 */
  double  nvec[3];    /* PLAIN, POLYGON, PARALLELOGRAM unit normal */
  double  dir1[3];    /* POLYGON/PARALLELOGRAM edge p1 - p0 */
  double  dir2[3];    /* POLYGON/PARALLELOGRAM edge p2 - p0 */
  double  dist;       /* PLAIN offset */
  double  radius;     /* SPHERE radius */
} ppm_shape_desc;

/* Material (material.rs:11-16) with its surface model (surface.rs:16-41). */
typedef struct {
  double  emittance[3];
  double  transmittance[3];
  double  ior[3];
  int32_t surface;      /* PPM_SURF_* */
  int32_t reserved;
  double  color1[3];    /* Simple: reflectance;    TS: albedo_diff */
  double  color2[3];    /* Simple: specular_refl;  TS: albedo_spec */
  double  diffuseness;  /* Simple: diffuseness;    TS: scatterness */
  double  metalness;
  double  roughness;
} ppm_material_desc;

/* Light (light.rs:16-39).  color is used as given (the reference normalises
 * it when the scene is built, scene.rs:24). */
typedef struct {
  int32_t kind;       /* PPM_LIGHT_* */
  int32_t reserved;
  double  color[3];
  double  flux;
  double  pos[3];
  double  nvec[3];
  double  dir1[3];
  double  dir2[3];
  double  dir[3];     /* SUN direction of travel */
} ppm_light_desc;

/* Camera (camera.rs:23-50): the derived screen vectors of read_camera
 * (camera.rs:151-164); ppm_camera_setup() computes them from the
 * user-facing parameters. */
typedef struct {
  int32_t xreso, yreso;
  int32_t progressive, antialias, blur;
  int32_t pfilter;      /* PPM_FILTER_* */
  double  eye_pos[3];
  double  origin[3];
  double  esx[3], esy[3];
  double  eex[3], eey[3];
} ppm_camera_desc;

/* ppm_scene_desc.flags: how rays find their nearest shape.  The reference
 * tests every object (calc_intersection, tracer.rs:306-335); a BVH built on
 * the device gives the same nearest hit (ties go to the lower shape index,
 * as the reference's stable sort keeps them). */
#define PPM_SCENE_BVH_AUTO 0 /* BVH when there are >= 64 spheres/polygons/parallelograms */
#define PPM_SCENE_BVH_OFF  1 /* every shape tested one by one */
#define PPM_SCENE_BVH_ON   2 /* BVH over the bounded shapes; planes one by one */

typedef struct {
  int32_t n_shapes, n_materials, n_lights;
  int32_t flags;                /* PPM_SCENE_BVH_* */
  const ppm_shape_desc*    shapes;
  const ppm_material_desc* materials;
  const ppm_light_desc*    lights;
  ppm_camera_desc          camera;
  /* Quantisation box of the photon-map Morton keys.  Photons outside it are
   * clamped to its faces (the tree stays exact; only its tightness suffers). */
  double bounds_min[3], bounds_max[3];
} ppm_scene_desc;

/* PPM-PA pass schedule (util/iterator.rb:26-38). */
typedef struct {
  uint64_t seed;
  uint64_t n_photon;     /* photons per pass (ppmpa.rs:30 'nphoton') */
  uint32_t first_pass;   /* 1-based index of the first pass to run */
  uint32_t pipelines;    /* passes in flight at once, each with its own buffers
                            (util/iterator.rb:18 NPARA = 4 parallel pm | rt
                            runs); 0 = automatic, at most 4 */
  double   radius0;      /* r_1, the radius of pass 1 (not squared) */
  double   alpha;        /* iterator.rb:17 ALPHA */
  int32_t  use_classic;  /* ppmpa.rs:16 DEF_USECLASSIC; "-nc" clears it */
  int32_t  max_depth;    /* tracer.rs:27 MAX_TRACE */
  int32_t  pfilter;      /* PPM_FILTER_* used by the gather */
  int32_t  flags;        /* PPM_SESSION_* */
} ppm_pass_cfg;

#define PPM_SESSION_STAGE_TIMING  1 /* hipEvents around every stage (no graph) */
#define PPM_SESSION_NO_GRAPH      2 /* launch the pass kernels directly */

typedef struct {
  double   ms_photon_trace;  /* emission + bouncing (H1, H2) */
  double   ms_map_build;     /* Morton keys + radix sort + tree (H4) */
  double   ms_eye_trace;     /* eye rays + direct light (H7) */
  double   ms_gather;        /* the gather kernel alone: radius query + radiance estimate (H5, H6) */
  double   ms_accumulate;    /* gather-query Morton sort + FP64 per-pixel accumulation (H8) */
  double   ms_total;         /* all passes, device time */
  uint64_t photons_emitted;  /* summed over passes */
  uint64_t photons_stored;   /* summed over passes */
  uint64_t gather_queries;   /* gather queries (all eye-tree nodes), summed over passes */
  uint32_t passes;
  uint32_t reserved;
} ppm_timings;

/* ---- context ------------------------------------------------------------ */

/* Open a context on HIP device `device`; hip_stream may be NULL (the context
 * then creates its own default-priority stream). */
ppm_status  ppm_create(ppm_ctx** out, int device, void* hip_stream);
void        ppm_destroy(ppm_ctx* ctx);
const char* ppm_last_error(const ppm_ctx* ctx);
const char* ppm_status_string(ppm_status s);
/* Block until all work queued on the context's stream is done. */
ppm_status  ppm_sync(ppm_ctx* ctx);

/* camera.rs:108-206 read_camera: fill the derived screen vectors of `cam`
 * from the camera parameters (focal_len in metres, as camera.rs:142). */
ppm_status  ppm_camera_setup(ppm_camera_desc* cam, int32_t xreso, int32_t yreso,
                             const double eye[3], const double target[3],
                             const double upper[3], double focus,
                             double focal_len, double f_number);

/* ---- scene -------------------------------------------------------------- */

/* Upload shapes, materials, lights and camera (scene.rs:20 read_scene,
 * camera.rs:108 read_camera). */
ppm_status  ppm_scene_create(ppm_ctx* ctx, const ppm_scene_desc* desc,
                             ppm_scene** out);
void        ppm_scene_destroy(ppm_scene* scene);

/* How the scene was laid out: shapes tested one by one, shapes in the BVH
 * and the BVH's leaf level (0 with no BVH).  No counterpart in the reference
 * (which has no BVH); for callers and tests that check the layout. */
ppm_status  ppm_scene_stats(const ppm_scene* scene, uint32_t* n_linear,
                            uint32_t* n_bvh, uint32_t* bvh_depth);

/* calc_intersection (tracer.rs:306-335) for n rays: origins and directions
 * are n x 3 doubles (directions are normalised first, as
 * Ray::new_from_elem does, geometry.rs:46-52); t_out receives the distance of the nearest hit with
 * t >= 1e-4 (tracer.rs:324) and obj_out the shape's index in the scene, or
 * (inf, -1) when nothing is hit.  Computed in float, like the passes. */
ppm_status  ppm_intersect(ppm_ctx* ctx, const ppm_scene* scene,
                          const double* origins, const double* dirs,
                          uint64_t n, double* t_out, int32_t* obj_out);

/* ---- photons (H1, H2) ---------------------------------------------------- */

/* A device-resident photon set of fixed capacity. */
ppm_status  ppm_photons_create(ppm_ctx* ctx, uint64_t capacity,
                               ppm_photons** out);
void        ppm_photons_destroy(ppm_photons* ph);

/* Σ_lights Σ_i trace_photon(uc, M_AIR, objs, 0, light.generate_photon())
 * (tracer.rs:31, light.rs:67, ppmpa.rs:86-92): emit the photons of one pass
 * and store their diffuse hits into `out` (overwriting it).  Per-light counts
 * are round(flux_l / (Σflux / n_photon)) (ppmpa.rs:30-31,70-72).  Photon i of
 * pass p draws its random numbers from the counter-based stream
 * (seed, p, i); max_depth is MAX_TRACE (tracer.rs:27). */
ppm_status  ppm_trace_photons(ppm_ctx* ctx, const ppm_scene* scene,
                              uint64_t seed, uint32_t pass, uint64_t n_photon,
                              int32_t use_classic, int32_t max_depth,
                              ppm_photons* out);

/* Stored-photon count (synchronises the stream). */
ppm_status  ppm_photons_count(ppm_ctx* ctx, const ppm_photons* ph, uint64_t* n);

/* Load a photon set from host memory: the content of the text map that
 * read_map parses (photonmap.rs:31-74).  pos and dir are n x 3 doubles, wl is
 * n channel codes.  Directions are re-normalised (geometry.rs:41-48 via
 * photonmap.rs:68).  The double positions are kept beside the float copy so
 * that radius queries can be decided exactly in double (refine64). */
ppm_status  ppm_photons_upload(ppm_ctx* ctx, ppm_photons* ph, uint64_t n,
                               const double* pos, const double* dir,
                               const uint8_t* wl);

/* Copy a photon set back to host memory (float values widened to double).
 * meta may be NULL; otherwise it receives the packed record tag:
 * bits 0-1 channel, bits 2-6 bounce depth, bits 7-31 photon index (the
 * photon's number in its pass, or its position in an uploaded set; hence
 * at most 2^25 photons per pass and per uploaded set, PPM_E_RANGE beyond). */
ppm_status  ppm_photons_download(ppm_ctx* ctx, const ppm_photons* ph,
                                 uint64_t cap, double* pos, double* dir,
                                 uint8_t* wl, uint32_t* meta, uint64_t* n);

/* ---- photon map (H4, H5, H6) --------------------------------------------- */

ppm_status  ppm_map_create(ppm_ctx* ctx, uint64_t capacity, ppm_map** out);
void        ppm_map_destroy(ppm_map* map);

/* build_photonmap (photonmap.rs:23-29): Morton keys within the scene's
 * bounds, radix sort, and a balanced tree of 32-photon leaves.  power is the
 * per-photon power (PhotonMap.power).  scene may be NULL: the bounds are then
 * those recorded with the photon set (the bounding box of an uploaded set, or
 * the bounds of the scene that traced it). */
ppm_status  ppm_map_build(ppm_ctx* ctx, const ppm_scene* scene,
                          const ppm_photons* ph, double power, ppm_map* map);

/* build_photonmap (photonmap.rs:23-29) in the form the PPM-PA passes build
 * it every pass (ppm_session_run): every query of a pass has one radius, so
 * the photons are binned into cubic cells of side >= 2.1 sqrt(r2), hashed
 * into a power-of-two bucket table and sorted by bucket.  ppm_radius_query
 * and ppm_gather on such a map run the passes' own code: the cell walk of
 * the shade kernel for the index sets, and the shade kernel itself (with
 * the direct light off, scale power / (pi r2) applied in float) for the
 * radiance; ppm_trace_eye shades through it as a pass does.  Their r2 must
 * not exceed the r2 given here, and ppm_gather on a grid takes refine64 = 0
 * (the passes decide membership in float).  scene may be NULL, as for
 * ppm_map_build. */
ppm_status  ppm_map_build_grid(ppm_ctx* ctx, const ppm_scene* scene,
                               const ppm_photons* ph, double power, double r2,
                               ppm_map* map);

/* KdTree::within(q, r2, squared_euclidean) (tracer.rs:180): for each of the
 * nq query points (nq x 3 doubles) the indices (into the photon set as it
 * was given to ppm_map_build) of the photons with squared distance <= r2.
 * offsets_out has nq + 1 entries; idx_out receives at most idx_cap indices,
 * each query's set sorted ascending by index.  refine64 = 0 decides in float
 * ((dx*dx + dy*dy) + dz*dz, no contraction); refine64 = 1 decides in double
 * as the reference does (needs a set loaded by ppm_photons_upload).
 * Returns PPM_E_RANGE (and the needed total in offsets_out[nq]) when
 * idx_cap is too small. */
ppm_status  ppm_radius_query(ppm_ctx* ctx, const ppm_map* map,
                             const double* q_xyz, uint64_t nq, double r2,
                             int32_t refine64, uint64_t idx_cap,
                             uint32_t* idx_out, uint64_t* offsets_out);

/* estimate_radiance (tracer.rs:179-195) for nq (position, normal) pairs:
 * (1/(pi r2)) Σ filter(d2, r2) * power * max(0, -n.dir) into the photon's
 * channel (optics.rs:224-233).  rgb_out is nq x 3 doubles. */
ppm_status  ppm_gather(ppm_ctx* ctx, const ppm_map* map, const double* q_xyz,
                       const double* q_nrm, uint64_t nq, double r2,
                       int32_t pfilter, int32_t refine64, double* rgb_out);

/* ppm_gather for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid allocation).
 * No counterpart in the reference.  Same arguments, checks and results as
 * ppm_gather on either map form (bit for bit), for nq <= the bound below
 * (PPM_E_RANGE beyond it).  It allocates nothing: queries, results and the
 * gather's state live in a fixed area of the library in device memory
 * (under 9 MiB, one per GPU, held until the process ends), so calls that
 * overlap on one GPU, from any contexts or threads, race.  Synchronises. */
#define PPM_NOALLOC_MAX_QUERIES 65536
ppm_status  ppm_gather_noalloc(ppm_ctx* ctx, const ppm_map* map,
                               const double* q_xyz, const double* q_nrm,
                               uint64_t nq, double r2, int32_t pfilter,
                               int32_t refine64, double* rgb_out);

/* ---- eye pass (H7) ------------------------------------------------------ */

/* One pass of per-pixel trace_ray (tracer.rs:129-177) over the camera's
 * screen (camera.rs:58-75, ppmpa.rs:81-82): rgb_out is yreso x xreso x 3
 * doubles, row-major (y, then x) as the reference prints it. */
ppm_status  ppm_trace_eye(ppm_ctx* ctx, const ppm_scene* scene,
                          const ppm_map* map, uint64_t seed, uint32_t pass,
                          double r2, int32_t use_classic, int32_t max_depth,
                          int32_t pfilter, double* rgb_out);

/* rtc (rtc.rs:17-47): trace_ray_classic (tracer.rs:221-259) for every pixel
 * of the camera's screen, a Whitted ray tracer with no photon map: at each
 * hit the direct light of every light (tracer.rs:263-290) plus the ambient
 * radiance (camera.rs:40, "ambient" of the camera file), mirror reflection
 * and refraction to depth 10.  The camera ray draws its blur / antialias
 * offsets from the pixel's stream (seed, pass) as ppm_trace_eye does.
 * rgb_out: yreso x xreso x 3 doubles, row-major. */
ppm_status  ppm_trace_classic(ppm_ctx* ctx, const ppm_scene* scene,
                              uint64_t seed, uint32_t pass,
                              const double ambient[3], double* rgb_out);

/* ---- PPM-PA passes (H8) -------------------------------------------------- */

/* util/iterator.rb + util/averager2.rb in one device-resident loop: pass i
 * uses r_i^2 = r_1^2 * prod_{k=1}^{i-1} (k + alpha) / (k + 1) (kept in
 * double on the device) and the per-pixel radiance is summed in double. */
ppm_status  ppm_session_create(ppm_ctx* ctx, const ppm_scene* scene,
                               const ppm_pass_cfg* cfg, ppm_session** out);
void        ppm_session_destroy(ppm_session* s);
/* Queue n_pass further passes (asynchronous, stream-ordered). */
ppm_status  ppm_session_run(ppm_session* s, uint32_t n_pass);
/* Per-pixel mean over the passes run so far (yreso x xreso x 3 doubles,
 * row-major) and, when var_out is not NULL, the unbiased per-pixel variance
 * of the per-pass values.  Synchronises. */
ppm_status  ppm_session_image(ppm_session* s, double* mean_out,
                              double* var_out, uint32_t* passes_done);
/* r_i^2 of the most recent pass, read back from the device (synchronises). */
ppm_status  ppm_session_radius2(ppm_session* s, double* r2_next);
ppm_status  ppm_session_timings(ppm_session* s, ppm_timings* t);
/* The eye rays' tiles: width x (64 / width) pixels, one wavefront each
 * (width 4, 8, 16 or 32; 8 when never set).  No counterpart in the
 * reference, which traces pixel by pixel (ppmpa.rs:70-80).  A pixel's
 * radiance is summed over its eye-tree nodes in the order the tile's wave
 * reaches them, so images at different widths agree to float rounding and
 * at equal widths bit for bit.  Takes effect from the next pass
 * (synchronises). */
ppm_status  ppm_session_set_tile_width(ppm_session* s, uint32_t width);

/* ---- differentiable form (no counterpart in the reference) ---------------- */

/* Set the reflectance c1 of materials 0..n-1 (n x 3 doubles; NULL keeps the
 * scene's), the survival probabilities q of their Simple diffuse bounce (n x
 * 3 in (0, 1]; NULL: q = c1), and scale every light's flux and every
 * emittance by light_scale, all relative to the scene as created.  A
 * diffuse bounce survives with probability q and weights the photon by c1/q
 * (tracer.rs:83-92 is the case q = c1, where every weight is 1).  Sessions
 * created afterwards see the new values. */
ppm_status  ppm_scene_set_params(ppm_ctx* ctx, ppm_scene* scene, uint32_t n,
                                 const double* albedo, const double* survival,
                                 double light_scale);
/* Rewind a session to before its first pass (sums cleared), taking up the
 * scene's current parameters (ppm_scene_set_params), so that a parameter
 * loop reuses the session's buffers.  No counterpart in the reference. */
ppm_status  ppm_session_reset(ppm_session* s);
/* Backward of the session's passes run so far: for grad_image = dL/d(mean
 * image) (yreso x xreso x 3), replays the passes from their seeds and
 * returns dL/dc1 of every material (n_materials x 3 doubles; non-zero for
 * Simple materials) and, when grad_light is not NULL, dL/d(light_scale). */
ppm_status  ppm_session_backward(ppm_session* s, const double* grad_image,
                                 double* grad_albedo, double* grad_light);

/* create + run(cfg->n_pass) + image + timings + destroy. */
ppm_status  ppm_run_passes(ppm_ctx* ctx, const ppm_scene* scene,
                           const ppm_pass_cfg* cfg, uint32_t n_pass,
                           double* mean_out, double* var_out, ppm_timings* t);

/* ---- batches (many scenes, one launch per stage) ------------------------- */

/* No counterpart in the reference, which renders one scene per process
 * (ppmpa.rs:48-92, iterator.rb:90-117): n scenes, each with its own pass
 * schedule (cfgs[i], n_passes[i] passes), run together.  Every stage of a
 * pass is one launch over all scenes still running (each scene's blocks run
 * the single-scene kernel body on the scene's own buffers), captured as one
 * hipGraph per number of running scenes.  Each scene's image equals what its
 * own session would produce.  Rounds (one pass of every running scene) are
 * spread over K pipelines as a session's passes are: round r runs on
 * pipeline r mod K, with its own buffers in every scene, so K rounds are in
 * flight at once.  K = the smallest cfgs[i].pipelines, 3 where that is 0. */
typedef struct ppm_batch ppm_batch;

/* Launch sizes of a batch (0 = the defaults 16, 8 and 8). */
typedef struct {
  uint32_t eye_waves_per_cu;     /* persistent eye waves per CU per scene */
  uint32_t shade_blocks_per_cu;  /* gather/shade blocks per CU per scene */
  uint32_t tile_width;           /* eye tiles of width x (64 / width) pixels:
                                    4, 8, 16 or 32 (ppm_session_set_tile_width) */
  uint32_t reserved;
} ppm_launch_cfg;

ppm_status  ppm_batch_create(ppm_ctx* ctx, uint32_t n,
                             const ppm_scene* const* scenes,
                             const ppm_pass_cfg* cfgs, const uint32_t* n_passes,
                             ppm_batch** out);
void        ppm_batch_destroy(ppm_batch* b);
/* Run passes until every scene has run its n_passes (max_rounds = 0), or at
 * most max_rounds more rounds.  Synchronises. */
ppm_status  ppm_batch_run(ppm_batch* b, uint32_t max_rounds);
ppm_status  ppm_batch_set_launch(ppm_batch* b, const ppm_launch_cfg* lc);
ppm_status  ppm_batch_get_launch(const ppm_batch* b, ppm_launch_cfg* lc);
/* Autotuner: time `rounds` passes of the batch under each candidate launch
 * configuration (eye waves 4/8/16/32 x shade blocks 2/4/8/16 per CU at
 * 8-pixel-wide tiles, then tile widths 4/16/32 at the fastest of those),
 * keep the fastest, and reset every scene to before its first pass. */
ppm_status  ppm_batch_tune(ppm_batch* b, uint32_t rounds, ppm_launch_cfg* best);
/* Scene i's per-pixel mean and variance, as ppm_session_image. */
ppm_status  ppm_batch_image(ppm_batch* b, uint32_t i, double* mean_out,
                            double* var_out, uint32_t* passes_done);
/* ms_total = device time of the batch's runs; counts summed over scenes. */
ppm_status  ppm_batch_timings(ppm_batch* b, ppm_timings* t);

#ifdef __cplusplus
}
#endif

#endif /* PPM_H */
