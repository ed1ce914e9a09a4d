// ppm_oracle.cpp — CPU FP64 restatement of the eijian/ppmpa PPM-PA path.
//
// TEST INFRASTRUCTURE ONLY.  This is the checker the GPU path is compared
// against (tests/, __graft_entry__.smoke(), and bench.py's cpu_baseline leg).
// It is never linked into, called by, or a fallback of the product library.
//
// It follows the reference's own structure and arithmetic: f64 everywhere
// (ray/mod.rs:14), the recursive trace_photon / trace_ray of tracer.rs, the
// inclusive squared-euclidean ball query of the kdtree crate, and no
// floating-point contraction (built with -ffp-contract=off, as Rust never
// contracts).  The only intended difference is the random source: the
// reference draws from the unseeded rand::thread_rng (physics.rs:324 etc.);
// here every draw comes from the same Philox4x32-10 counter streams the GPU
// uses, so both sides can be run on common random numbers.
//
// Pinning status of each part (see tests/test_oracle_*.py):
//   filters, normalize, target, normals, Color::normalize, decide_wavelength,
//   check_under: reference test values (tracer.rs:366-377, algebra.rs:268,
//   geometry.rs:217,227,233-242, physics.rs:372-377,409-415);
//   Philox: Random123 known-answer vectors;
//   intersections: closed forms on hand cases;
//   within: brute force + kdtreetest.rs fixture;
//   gather: uniform-carpet closed form;
//   radius schedule: Gamma-function closed form;
//   photon walk / eye tree: invariants (storage rule, albedo-0 and albedo-1
//   boxes, hemisphere moments, channel frequencies);
//   trace_ray_classic: closed forms (ambient on a Lambertian, two parallel
//   mirrors to the depth limit, glass slabs at 60 degrees).
//   The rand 0.6 stream contents and the kdtree crate's tie order are
//   unpinned by design (nothing in the reference fixes them).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/ppm.h"

namespace orc {

typedef double Flt;  // ray/mod.rs:14
const Flt NEARLY0 = 0.0001;          // ray/mod.rs:16
const Flt PI = 3.14159265358979323846;
const Flt PI4 = PI * 4.0;            // ray/mod.rs:18
const Flt ONE_PI = 1.0 / PI;         // tracer.rs:24
const Flt SR_HALF = 1.0 / (2.0 * PI);  // tracer.rs:25
const Flt PARA_DIV = 0.2;            // light.rs:162

// ---------------------------------------------------------------- algebra.rs
struct V {
  Flt x, y, z;
};
static inline V operator+(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline V operator-(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline V operator-(V a) { return {-a.x, -a.y, -a.z}; }
static inline V operator*(V a, Flt s) { return {a.x * s, a.y * s, a.z * s}; }
static inline V operator*(Flt s, V a) { return {s * a.x, s * a.y, s * a.z}; }
// algebra.rs:146-148
static inline Flt dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// algebra.rs:174-180
static inline V cross(V a, V b) { return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
// algebra.rs:134-136, 151-158
static inline bool normalize(V a, V& o) {
  Flt n = std::sqrt(dot(a, a));
  if (n == 0.0) return false;
  o = a * (1.0 / n);
  return true;
}
static inline V vd(const double* p) { return {p[0], p[1], p[2]}; }

struct C {
  Flt r, g, b;
};
static inline C operator+(C a, C b) { return {a.r + b.r, a.g + b.g, a.b + b.b}; }
static inline C operator*(C a, C b) { return {a.r * b.r, a.g * b.g, a.b * b.b}; }
static inline C operator*(C a, Flt s) { return {a.r * s, a.g * s, a.b * s}; }
static inline C operator*(Flt s, C a) { return {s * a.r, s * a.g, s * a.b}; }
static inline C cd(const double* p) { return {p[0], p[1], p[2]}; }
static inline Flt wlv(C c, int wl) { return wl == 0 ? c.r : (wl == 1 ? c.g : c.b); }  // physics.rs:87-93

// ---------------------------------------------------------------- RNG
static inline void philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int i = 0; i < 10; ++i) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

struct Rng {
  uint32_t key[2];
  uint32_t ctr[3];
  uint32_t n = 0;
  uint32_t blk[4];
  Rng(uint64_t seed, uint32_t elem, uint32_t pass, uint32_t tag) {
    key[0] = (uint32_t)seed;
    key[1] = (uint32_t)(seed >> 32);
    ctr[0] = elem;
    ctr[1] = pass;
    ctr[2] = tag;
  }
  // gen_range(0.0, 1.0) with 24 random bits
  Flt next() {
    uint32_t j = n & 3u;
    if (j == 0) {
      uint32_t c[4] = {ctr[0], ctr[1], ctr[2], n >> 2};
      philox(c, key, blk);
    }
    ++n;
    return (Flt)(blk[j] >> 8) * (1.0 / 16777216.0);
  }
};
const uint32_t TAG_PHOTON = 0x10000000u;
const uint32_t TAG_EYE = 0x20000000u;

// ---------------------------------------------------------------- physics.rs
static inline Flt relative_ior(Flt i1, Flt i2) { return i1 == 0.0 ? 1.0 : i2 / i1; }  // :200-205
// :183-189
static inline Flt relative_ior_wavelength(C i1, C i2, int wl) { return relative_ior(wlv(i1, wl), wlv(i2, wl)); }
// :192-196
static inline Flt relative_ior_average(C i1, C i2) {
  Flt a1 = (i1.r + i1.g + i1.b) / 3.0;
  Flt a2 = (i2.r + i2.g + i2.b) / 3.0;
  return relative_ior(a1, a2);
}
// :214-221
static inline V specular_reflection(V n, V v, Flt& cos1) {
  Flt c = -dot(v, n);
  if (c < 0.0) {
    cos1 = -c;
    return n;
  }
  cos1 = c;
  V r;
  if (!normalize(v + (2.0 * c) * n, r)) r = n;
  return r;
}
// :232-261
static V reflection_glossy(V n, V r, Flt pw, Rng& rng) {
  V u;
  if (!normalize(cross(V{0.00424, 1.0, 0.00764}, r), u)) {
    if (!normalize(cross(V{1.0, 0.00424, 0.00764}, r), u)) u = V{1, 0, 0};
  }
  V vv = cross(u, r);
  Flt c0 = dot(n, r);
  Flt xi0 = rng.next();
  Flt xi1 = std::pow(xi0, pw * c0);
  Flt xi2 = 2.0 * PI * rng.next();
  Flt x = std::cos(xi2) * std::sqrt(1.0 - xi1 * xi1);
  Flt y = xi1;
  Flt z = std::sin(xi2) * std::sqrt(1.0 - xi1 * xi1);
  V wi = x * u + y * r + z * vv;
  if (dot(n, wi) < 0.0) wi = (-x) * u + y * r - z * vv;
  V o;
  if (!normalize(wi, o)) o = V{1, 0, 0};
  return o;
}
// :270-285
static inline bool specular_refraction(V n, V v, Flt eta, V& t, Flt& cos2) {
  cos2 = 0.0;
  Flt c1 = -dot(v, n);
  if (c1 < 0.0) return false;
  Flt g0 = eta * eta + c1 * c1 - 1.0;
  if (g0 < 0.0) return false;
  Flt g = std::sqrt(g0);
  if (eta == 0.0) return false;  // 1/0 (only reachable when c1 == 1 exactly)
  if (!normalize((1.0 / eta) * (v + (c1 - g) * n), t)) return false;
  cos2 = g / eta;
  return true;
}
// :316-318
static inline Flt schlick(Flt f0, Flt c) { return f0 + (1.0 - f0) * std::pow(1.0 - c, 5.0); }
// :329-335
static inline int check_under(const Flt* ps, int len, Flt p) {
  int i = 0;
  while (i < len && p > ps[i]) ++i;
  return i;
}
// :323-327
static inline int russian_roulette1(Flt p0, Rng& rng) {
  Flt p = rng.next();
  return check_under(&p0, 1, p);
}
// :61-71 (clip negatives, normalise to sum 1)
static C color_normalize(C c) {
  Flt r = c.r < 0 ? 0 : c.r, g = c.g < 0 ? 0 : c.g, b = c.b < 0 ? 0 : c.b;
  Flt mag = r + g + b;
  if (mag == 0.0) return C{1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0};
  return C{r / mag, g / mag, b / mag};
}
// :74-84
static inline int decide_wavelength(C c, Flt p) { return p < c.r ? 0 : (p < c.r + c.g ? 1 : 2); }

// ---------------------------------------------------------------- algebra.rs:225-235 / surface.rs:408-417
static V generate_random_dir_by_angle(Rng& rng) {
  Flt phi = 2.0 * PI * rng.next();
  Flt xi = 2.0 * rng.next() - 1.0;
  Flt xi2 = std::sqrt(1.0 - std::pow(xi, 2.0));
  V v{xi2 * std::cos(phi), xi, xi2 * std::sin(phi)};
  V o;
  if (!normalize(v, o)) o = V{0, 1, 0};
  return o;
}
static V diffuse_reflection(V n, Rng& rng) {
  V d = generate_random_dir_by_angle(rng);
  return dot(n, d) > 0.0 ? d : -d;
}
// algebra.rs:211-223
static V generate_random_dir(Rng& rng) {
  for (int k = 0; k < 64; ++k) {
    Flt x = 2.0 * rng.next() - 1.0, y = 2.0 * rng.next() - 1.0, z = 2.0 * rng.next() - 1.0;
    V v{x, y, z};
    Flt len = std::sqrt(dot(v, v));
    if (0.0 < len && len <= 1.0) return v * (1.0 / len);
  }
  return V{0, 1, 0};
}

// ---------------------------------------------------------------- scene
struct Material {
  C em, ior, c1, c2;
  C q;  // survival probability of a diffuse bounce (= c1, the reference's rule, unless set)
  int surface;
  Flt diff, metal, rough, density_pow;
};
struct Shape {
  int kind, mat;
  V p, n, d1, d2;
  Flt dist, radius;
};
struct Light {
  int kind;
  C color;
  Flt flux;
  V pos, nvec, d1, d2, dir;
};
struct Camera {
  int xreso, yreso, progressive, antialias, blur, pfilter;
  V eye, origin, esx, esy, eex, eey;
};
struct Scene {
  std::vector<Shape> shapes;
  std::vector<Material> mats;  // last = M_AIR
  std::vector<Light> lights;
  Camera cam;
  int air;
};

// ---------------------------------------------------------------- geometry.rs
// distance_plain :170-177
static inline void d_plain(const Shape& s, V o, V d, Flt* ts, int& k) {
  Flt c = dot(s.n, d);
  if (c == 0.0) return;
  ts[k++] = (s.dist + dot(s.n, o)) / -c;
}
// distance_sphere :179-193
static inline void d_sphere(const Shape& s, V o, V d, Flt* ts, int& k) {
  V oc = s.p - o;
  Flt t0 = dot(oc, d);
  Flt t1 = s.radius * s.radius - (dot(oc, oc) - (t0 * t0));
  Flt t2 = std::sqrt(t1);
  if (t1 <= 0.0) return;
  if (t2 == 0.0) {
    ts[k++] = t0;
  } else {
    ts[k++] = t0 - t2;
    ts[k++] = t0 + t2;
  }
}
// method_moller :149-165
static inline bool method_moller(Flt l, V p0, V d1, V d2, V p, V d, Flt& u, Flt& v, Flt& t) {
  V re2 = cross(d, d2);
  Flt det = dot(re2, d1);
  V pp = p - p0;
  V te1 = cross(pp, d1);
  if (det == 0.0) return false;
  u = dot(re2, pp) / det;
  v = dot(te1, d) / det;
  t = dot(te1, d2) / det;
  return !(u < 0.0 || u > 1.0 || v < 0.0 || v > 1.0 || u + v > l);
}
static inline void distance(const Shape& s, V o, V d, Flt* ts, int& k) {
  Flt u, v, t;
  switch (s.kind) {
    case PPM_SHAPE_PLAIN: d_plain(s, o, d, ts, k); break;
    case PPM_SHAPE_SPHERE: d_sphere(s, o, d, ts, k); break;
    case PPM_SHAPE_POLYGON:
      if (method_moller(1.0, s.p, s.d1, s.d2, o, d, u, v, t)) ts[k++] = t;
      break;
    case PPM_SHAPE_PARALLELOGRAM:
      if (method_moller(2.0, s.p, s.d1, s.d2, o, d, u, v, t)) ts[k++] = t;
      break;
    default: break;
  }
}

struct Isect {
  V pos, n;
  int obj, mat;
  bool in;  // InOut::In
};

// calc_intersection (tracer.rs:306-350): collect t >= NEARLY0 of every
// object, stable sort, take the first; normal flipped to face the ray.
static bool calc_intersection(const Scene& sc, V o, V d, Isect& is) {
  Flt best = INFINITY;
  int bo = -1;
  for (size_t i = 0; i < sc.shapes.size(); ++i) {
    Flt ts[2];
    int k = 0;
    distance(sc.shapes[i], o, d, ts, k);
    for (int j = 0; j < k; ++j) {
      if (ts[j] < NEARLY0) continue;
      if (ts[j] < best) {  // first of equals kept, as the stable sort does
        best = ts[j];
        bo = (int)i;
      }
    }
  }
  if (bo < 0) return false;
  const Shape& s = sc.shapes[bo];
  V p = o + d * best;  // Ray::target
  V n;
  if (s.kind == PPM_SHAPE_SPHERE) {
    if (!normalize(p - s.p, n)) return false;
  } else {
    n = s.n;
  }
  if (dot(n, d) > 0.0) {
    is = Isect{p, -n, bo, s.mat, false};
  } else {
    is = Isect{p, n, bo, s.mat, true};
  }
  return true;
}

// ---------------------------------------------------------------- surface.rs
static bool is_black(C c) { return c.r == 0.0 && c.g == 0.0 && c.b == 0.0; }
static bool is_white(C c) { return c.r == 1.0 && c.g == 1.0 && c.b == 1.0; }
static bool s_reflect(const Material& m, Flt cos) {  // :68-100
  if (m.surface == PPM_SURF_SIMPLE) return !(m.diff == 1.0 || (cos == 1.0 && is_black(m.c2)));
  if (m.surface == PPM_SURF_TS) return m.metal == 0.0 ? true : (m.metal == 1.0 ? !is_black(m.c2) : true);
  return false;
}
static bool s_refract(const Material& m, Flt cos) {  // :102-133
  if (m.surface == PPM_SURF_SIMPLE) return !(cos == 0.0 && is_white(m.c2));
  if (m.surface == PPM_SURF_TS) return m.metal == 0.0 && m.diff < 1.0 && !is_black(m.c1);
  return false;
}
static bool s_store(const Material& m) {  // :289-310
  if (m.surface == PPM_SURF_SIMPLE) return m.diff > 0.0;
  if (m.surface == PPM_SURF_TS) return m.metal != 1.0 && m.diff != 0.0;
  return true;
}
static Flt s_roughness(const Material& m) {  // :358-379 (Simple returns diffuseness)
  if (m.surface == PPM_SURF_SIMPLE) return m.diff;
  if (m.surface == PPM_SURF_TS) return m.rough;
  return 0.0;
}
static Flt s_power_glossy(const Material& m) {  // :381-402
  return (m.surface == PPM_SURF_SIMPLE || m.surface == PPM_SURF_TS) ? m.density_pow : 0.0;
}
// reflection_index :512-515
static C reflection_index(C col, Flt c) {
  Flt c2 = std::pow(1.0 - c, 5.0);
  return C{col.r + (1.0 - col.r) * c2, col.g + (1.0 - col.g) * c2, col.b + (1.0 - col.b) * c2};
}
// Surface::bsdf :135-206
static C s_bsdf(const Material& m, Flt cos0, C di, C si, C ti) {
  if (m.surface == PPM_SURF_SIMPLE) {
    C f = reflection_index(m.c2, cos0);
    C f2{1.0 - f.r, 1.0 - f.g, 1.0 - f.b};  // Color::neg (physics.rs:116-122)
    return m.diff * ((m.c1 * ONE_PI) * di) + (1.0 - m.diff) * (f * si + (1.0 - m.metal) * f2 * ti);
  }
  if (m.surface == PPM_SURF_TS) {
    C f = reflection_index(m.c2, cos0);
    C f2{1.0 - f.r, 1.0 - f.g, 1.0 - f.b};
    C i_de{0, 0, 0};
    if (m.metal == 0.0) i_de = f2 * m.c1 * (m.diff * ONE_PI * di + (1.0 - m.diff) * ti);
    return i_de + f * si;
  }
  return C{0, 0, 0};
}
// The diffuse weight of the BSDF (the factor on di): zero means di is unused.
static C s_diffuse_weight(const Material& m, Flt cos0) {
  if (m.surface == PPM_SURF_SIMPLE) return (m.c1 * ONE_PI) * m.diff;
  if (m.surface == PPM_SURF_TS && m.metal == 0.0) {
    C f = reflection_index(m.c2, cos0);
    C f2{1.0 - f.r, 1.0 - f.g, 1.0 - f.b};
    return f2 * m.c1 * (m.diff * ONE_PI);
  }
  return C{0, 0, 0};
}

// Surface::next_direction :211-258.  Returns 0 absorbed, 1 reflected (medium
// kept), 2 transmitted (medium becomes the hit material).
static int s_next_direction(const Material& m, Flt eta, V n, V v, int wl, Rng& rng, V& out) {
  Flt cos1;
  V rdir0 = specular_reflection(n, v, cos1);
  V rdir = reflection_glossy(n, rdir0, s_power_glossy(m), rng);
  V hvec;
  if (!normalize(rdir - v, hvec)) hvec = n;
  V tdir;
  Flt cos2;
  bool has_t = specular_refraction(hvec, v, eta, tdir, cos2);
  Flt cs = cos1 < cos2 ? cos1 : cos2;
  if (m.surface != PPM_SURF_TS) return 0;
  Flt f = schlick(wlv(m.c2, wl), cs);
  if (russian_roulette1(f, rng) == 0) {
    out = rdir;
    return 1;
  }
  if (russian_roulette1(wlv(m.c1, wl), rng) == 1) return 0;
  if (russian_roulette1(m.diff, rng) == 0) {
    out = diffuse_reflection(n, rng);
    return 1;
  }
  if (!has_t) return 0;
  out = tdir;
  return 2;
}

// ---------------------------------------------------------------- photons
// w: the photon's weight, the product of c1/q over its diffuse survivals
// (1 when q = c1); ev: the materials of those survivals, in order.
constexpr int MAX_EV = 16;
struct PathW {
  Flt w = 1.0;
  int n = 0;
  uint16_t ev[MAX_EV] = {};
};
struct Photon {
  int wl;
  V pos, dir;
  uint32_t meta;
  PathW pw;
};

// Light::generate_photon (light.rs:67-91)
static void generate_photon(const Light& L, Rng& rng, int& wl, V& o, V& d) {
  wl = decide_wavelength(L.color, rng.next());  // select_wavelength :157-160
  if (L.kind == PPM_LIGHT_PARALLELOGRAM) {
    Flt t1 = rng.next(), t2 = rng.next();
    d = diffuse_reflection(L.nvec, rng);
    o = L.pos + t1 * L.d1 + t2 * L.d2;
  } else if (L.kind == PPM_LIGHT_POINT) {
    d = generate_random_dir(rng);
    o = L.pos;
  } else {
    Flt t1 = rng.next(), t2 = rng.next();
    o = L.pos + t1 * L.d1 + t2 * L.d2;
    d = L.dir;
  }
}

// trace_photon (tracer.rs:31-81), reflect_diff (:83-92), reflect_spec
// (:94-109), reflect_trans (:111-125).  Children are pushed before the
// parent's own record, as in the reference.
// With q != c1 a diffuse bounce survives with probability q and multiplies
// the photon's weight by c1/q (the same expectation; with q = c1 it is the
// reference's rule and every weight stays 1).
static void trace_photon(const Scene& sc, bool uc, int maxd, int m0, int l, int wl, V o, V d, uint32_t pid,
                         Rng& rng, std::vector<Photon>& pcs, const PathW& pw) {
  if (l >= maxd) return;
  Isect is;
  if (!calc_intersection(sc, o, d, is)) return;
  const Material& M = sc.mats[is.mat];
  if (M.surface == PPM_SURF_SIMPLE) {
    if (russian_roulette1(s_roughness(M), rng) == 0) {
      // reflect_diff
      if (russian_roulette1(wlv(M.q, wl), rng) == 0) {
        V dr = diffuse_reflection(is.n, rng);
        PathW p2 = pw;
        const Flt a = wlv(M.c1, wl), qq = wlv(M.q, wl);
        if (a != qq) p2.w = pw.w * (a / qq);
        if (p2.n < MAX_EV) p2.ev[p2.n++] = (uint16_t)is.mat;
        trace_photon(sc, uc, maxd, m0, l + 1, wl, is.pos, dr, pid, rng, pcs, p2);
      }
    } else {
      // reflect_spec
      Flt cos1;
      V rdir = specular_reflection(is.n, d, cos1);
      Flt f = schlick(wlv(M.c2, wl), cos1);
      if (russian_roulette1(f, rng) == 0) {
        trace_photon(sc, uc, maxd, m0, l + 1, wl, is.pos, rdir, pid, rng, pcs, pw);
      } else if (wlv(M.ior, wl) != 0.0) {
        // reflect_trans
        Flt eta = relative_ior_wavelength(sc.mats[m0].ior, M.ior, wl);
        V tdir;
        Flt c2;
        if (specular_refraction(is.n, d, eta, tdir, c2)) {
          int m02 = dot(tdir, is.n) < 0.0 ? is.mat : sc.air;
          trace_photon(sc, uc, maxd, m02, l + 1, wl, is.pos, tdir, pid, rng, pcs, pw);
        }
      }
    }
  } else if (M.surface == PPM_SURF_TS) {
    Flt eta = relative_ior_wavelength(sc.mats[m0].ior, M.ior, wl);
    V nd;
    int r = s_next_direction(M, eta, is.n, d, wl, rng, nd);
    if (r != 0) trace_photon(sc, uc, maxd, r == 1 ? m0 : is.mat, l + 1, wl, is.pos, nd, pid, rng, pcs, pw);
  }
  if ((!uc || l > 0) && s_store(M)) {
    uint32_t meta = (uint32_t)(wl & 3) | ((uint32_t)(l & 31) << 2) | (pid << 7);
    pcs.push_back(Photon{wl, is.pos, d, meta, pw});
  }
}

// ---------------------------------------------------------------- kd-tree
// A median-split kd-tree over the photon positions, restating what
// KdTree::within returns: every photon with squared_euclidean(q, p) =
// ((0 + dx^2) + dy^2) + dz^2 <= r2, sorted ascending by that distance
// (ties by index).
struct KdNode {
  int lo, hi;   // range in perm
  int axis;     // -1 leaf
  Flt split;
  int left, right;
};
struct Kd {
  std::vector<Photon> ph;
  std::vector<int> perm;
  std::vector<KdNode> nodes;
  Flt power = 1.0;
};

static int kd_build_rec(Kd& kd, int lo, int hi) {
  KdNode nd{lo, hi, -1, 0.0, -1, -1};
  int id = (int)kd.nodes.size();
  kd.nodes.push_back(nd);
  if (hi - lo <= 8) return id;
  Flt mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = lo; i < hi; ++i) {
    const V& p = kd.ph[kd.perm[i]].pos;
    Flt c[3] = {p.x, p.y, p.z};
    for (int a = 0; a < 3; ++a) {
      mn[a] = std::min(mn[a], c[a]);
      mx[a] = std::max(mx[a], c[a]);
    }
  }
  int axis = 0;
  for (int a = 1; a < 3; ++a)
    if (mx[a] - mn[a] > mx[axis] - mn[axis]) axis = a;
  if (!(mx[axis] > mn[axis])) return id;  // all equal: leaf
  int mid = (lo + hi) / 2;
  auto key = [&](int i) {
    const V& p = kd.ph[i].pos;
    return axis == 0 ? p.x : (axis == 1 ? p.y : p.z);
  };
  std::nth_element(kd.perm.begin() + lo, kd.perm.begin() + mid, kd.perm.begin() + hi,
                   [&](int a, int b) { return key(a) < key(b); });
  Flt split = key(kd.perm[mid]);
  int l = kd_build_rec(kd, lo, mid);
  int r = kd_build_rec(kd, mid, hi);
  kd.nodes[id].axis = axis;
  kd.nodes[id].split = split;
  kd.nodes[id].left = l;
  kd.nodes[id].right = r;
  return id;
}

static void kd_build(Kd& kd) {
  kd.perm.resize(kd.ph.size());
  std::iota(kd.perm.begin(), kd.perm.end(), 0);
  kd.nodes.clear();
  if (!kd.ph.empty()) kd_build_rec(kd, 0, (int)kd.ph.size());
}

// squared_euclidean of the kdtree crate: left fold over the axes.
static inline Flt sq_euclid(V a, V b) {
  Flt s = 0.0;
  Flt dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  s = s + dx * dx;
  s = s + dy * dy;
  s = s + dz * dz;
  return s;
}
static inline float sq_euclid_f(float ax, float ay, float az, float bx, float by, float bz) {
  float dx = ax - bx, dy = ay - by, dz = az - bz;
  float s = 0.0f;
  s = s + dx * dx;
  s = s + dy * dy;
  s = s + dz * dz;
  return s;
}

static void kd_within_rec(const Kd& kd, int id, V q, Flt r2, std::vector<std::pair<Flt, int>>& out) {
  const KdNode& nd = kd.nodes[id];
  if (nd.axis < 0) {
    for (int i = nd.lo; i < nd.hi; ++i) {
      int j = kd.perm[i];
      Flt d = sq_euclid(q, kd.ph[j].pos);
      if (d <= r2) out.push_back({d, j});
    }
    return;
  }
  Flt qa = nd.axis == 0 ? q.x : (nd.axis == 1 ? q.y : q.z);
  Flt diff = qa - nd.split;
  Flt d2 = diff * diff;
  // left holds coordinates <= split, right >= split
  if (diff <= 0.0) {
    kd_within_rec(kd, nd.left, q, r2, out);
    if (d2 <= r2) kd_within_rec(kd, nd.right, q, r2, out);
  } else {
    kd_within_rec(kd, nd.right, q, r2, out);
    if (d2 <= r2) kd_within_rec(kd, nd.left, q, r2, out);
  }
}

// KdTree::within (tracer.rs:180)
static void kd_within(const Kd& kd, V q, Flt r2, std::vector<std::pair<Flt, int>>& out) {
  out.clear();
  if (kd.nodes.empty()) return;
  kd_within_rec(kd, 0, q, r2, out);
  std::sort(out.begin(), out.end());
}

// filter_cone / filter_gauss (tracer.rs:197-216)
static Flt filter_cone(Flt d, Flt rmax) {
  const Flt K = 1.1, FAC = 1.0 - 2.0 / (3.0 * K);
  Flt d2 = std::sqrt(d / rmax) / K;
  return d2 > 1.0 ? 0.0 : (1.0 - d2) / FAC;
}
static Flt filter_gauss(Flt d, Flt rmax) {
  const Flt ALPHA = 0.918, BETA = 1.953, E_BETA = 1.0 - 0.14184788965323, CORR = 0.5;
  Flt e_r = 1.0 - std::exp(-BETA * d / (rmax * 2.0));
  return e_r > E_BETA ? 0.0 : ALPHA * (1.0 - e_r / E_BETA) + CORR;
}

// One photon's term of estimate_radiance (tracer.rs:185-192) with
// photon_to_radiance (optics.rs:224-233), added into its channel.
static inline void radiance_term(const Kd& kd, const Photon& p, Flt d2, Flt r2, int pf, V n, C& rad) {
  Flt wt = pf == PPM_FILTER_CONE ? filter_cone(d2, r2) : (pf == PPM_FILTER_GAUSS ? filter_gauss(d2, r2) : 1.0);
  Flt pw = wt * kd.power;
  Flt cos0 = dot(n, p.dir);
  Flt pw2 = cos0 < 0.0 ? pw * -cos0 * p.pw.w : 0.0;
  if (p.wl == 0) rad.r = rad.r + pw2;
  else if (p.wl == 1) rad.g = rad.g + pw2;
  else rad.b = rad.b + pw2;
}

// estimate_radiance (tracer.rs:179-195) with photon_to_radiance (optics.rs:224-233)
static C estimate_radiance(const Kd& kd, Flt r2, int pf, V pos, V n, std::vector<std::pair<Flt, int>>& buf) {
  kd_within(kd, pos, r2, buf);
  if (buf.empty()) return C{0, 0, 0};
  C rad{0, 0, 0};
  for (auto& e : buf) radiance_term(kd, kd.ph[e.second], e.first, r2, pf, n, rad);
  return rad * (ONE_PI / r2);
}

// The same estimate over the photons that the float predicate of the GPU
// passes keeps, fl32((dx^2 + dy^2) + dz^2) <= fl32(r2) on float coordinates
// (every photon tested; ascending distance, then index, as kd_within
// orders its result), each term still formed in double.
static C estimate_radiance_f32set(const Kd& kd, Flt r2, int pf, V pos, V n, std::vector<std::pair<Flt, int>>& buf) {
  buf.clear();
  const float qx = (float)pos.x, qy = (float)pos.y, qz = (float)pos.z, r2f = (float)r2;
  for (size_t j = 0; j < kd.ph.size(); ++j) {
    const Photon& p = kd.ph[j];
    const float d = sq_euclid_f(qx, qy, qz, (float)p.pos.x, (float)p.pos.y, (float)p.pos.z);
    if (d <= r2f) buf.push_back({sq_euclid(pos, p.pos), (int)j});
  }
  std::sort(buf.begin(), buf.end());
  C rad{0, 0, 0};
  for (auto& e : buf) radiance_term(kd, kd.ph[e.second], e.first, r2, pf, n, rad);
  return rad * (ONE_PI / r2);
}

// ---------------------------------------------------------------- direct light
// illuminated (tracer.rs:272-290) + get_radiance_from_light (:263-270) with
// Light::get_direction (light.rs:93-129) and get_radiance (:131-150).
static C radiance_from_light(const Scene& sc, V p, V n, const Light& L) {
  std::vector<V> lds;
  if (L.kind == PPM_LIGHT_POINT) {
    lds.push_back(L.pos - p);
  } else if (L.kind == PPM_LIGHT_PARALLELOGRAM) {
    static const Flt TS[5] = {0.1, 0.3, 0.5, 0.7, 0.9};
    for (int a = 0; a < 5; ++a)
      for (int b = 0; b < 5; ++b) {
        V v = (L.pos + TS[a] * L.d1 + TS[b] * L.d2) - p;
        if (dot(L.nvec, v) < 0.0) lds.push_back(v);
      }
  } else {
    V d = L.pos - p;
    if (!(dot(L.nvec, d) > 0.0)) {
      V dt2 = -L.dir;
      Flt u, v, t;
      if (method_moller(2.0, L.pos, L.d1, L.d2, p, dt2, u, v, t)) lds.push_back(t * dt2);
    }
  }
  std::vector<Flt> dists, coss;
  for (const V& ld : lds) {
    V ld3;
    if (!normalize(ld, ld3)) continue;
    Flt cos0 = dot(n, ld3);
    if (cos0 < 0.0) continue;
    Isect is;
    if (!calc_intersection(sc, p, ld3, is)) continue;
    Flt sq_l = dot(ld, ld);
    Flt sq_o = dot(is.pos - p, is.pos - p);
    if (sq_l - sq_o > 0.002) continue;
    dists.push_back(sq_l);
    coss.push_back(cos0 * cos0);
  }
  std::vector<C> rs;
  rs.push_back(C{0, 0, 0});  // RADIANCE0 first (light.rs:132)
  for (Flt d : dists) {
    if (L.kind == PPM_LIGHT_POINT) {
      Flt l0 = L.flux / (PI4 * d);
      rs.push_back(L.color * l0);
    } else if (L.kind == PPM_LIGHT_PARALLELOGRAM) {
      Flt l0 = (2.0 * L.flux * PARA_DIV * PARA_DIV) / (PI4 * d);
      rs.push_back(L.color * l0);
    } else {
      rs.push_back(L.color * L.flux);
    }
  }
  C rad{0, 0, 0};
  size_t k = std::min(rs.size(), coss.size());  // zip
  for (size_t i = 0; i < k; ++i) rad = rad + rs[i] * coss[i];
  return rad;
}

// ---------------------------------------------------------------- eye rays
struct EyeCtx {
  const Scene* sc;
  const Kd* kd;
  Flt r2;
  bool uc;
  int maxd;
  int pf;
  uint64_t seed;
  uint32_t pixel, pass;
  uint32_t cam_draws;
  const Rng* root;
  std::vector<std::pair<Flt, int>>* buf;
  const C* g = nullptr;        // dL/d(this pass's pixel), when the gradient is wanted
  double* grad = nullptr;      // dL/d(c1): 3 per material
};

// Gradient of one eye node's contribution wpath * wd * di to the pixel.
// Eye side: wd is proportional to c1 of the node's (Simple) material, so
// d/dc1 = contribution / c1.  Photon side: each photon in the ball carries
// the weight prod c1/q over its diffuse survivals, so its share of the
// contribution differentiates to share / c1 of every survival's material.
static void node_grad(const EyeCtx& ec, const Isect& is, const Material& mate, C wpath, C wd, C di, Flt r2) {
  const Scene& sc = *ec.sc;
  const C& g = *ec.g;
  double* grad = ec.grad;
  const C contrib{wpath.r * wd.r * di.r, wpath.g * wd.g * di.g, wpath.b * wd.b * di.b};
  if (mate.surface == PPM_SURF_SIMPLE) {
    if (mate.c1.r != 0.0) grad[3 * is.mat + 0] += g.r * contrib.r / mate.c1.r;
    if (mate.c1.g != 0.0) grad[3 * is.mat + 1] += g.g * contrib.g / mate.c1.g;
    if (mate.c1.b != 0.0) grad[3 * is.mat + 2] += g.b * contrib.b / mate.c1.b;
  }
  const Kd& kd = *ec.kd;
  for (auto& e : *ec.buf) {
    const Photon& p = kd.ph[e.second];
    if (p.pw.n == 0) continue;
    Flt wt = ec.pf == PPM_FILTER_CONE ? filter_cone(e.first, r2) : (ec.pf == PPM_FILTER_GAUSS ? filter_gauss(e.first, r2) : 1.0);
    Flt cos0 = dot(is.n, p.dir);
    if (!(cos0 < 0.0)) continue;
    const int c = p.wl;
    const Flt share = wt * kd.power * -cos0 * p.pw.w * (ONE_PI / r2) * wlv(wpath, c) * wlv(wd, c) * wlv(g, c);
    for (int k = 0; k < p.pw.n; ++k) {
      const int m = p.pw.ev[k];
      const Flt a = wlv(sc.mats[m].c1, c);
      if (a != 0.0) grad[3 * m + c] += share / a;
    }
  }
}

// trace_ray (tracer.rs:129-177).  Node (l, path) draws its random numbers
// from its own stream; children whose BSDF weight is zero in every channel
// are not traced (they would be multiplied by zero), nor is di when the
// diffuse weight is zero.
static C trace_ray(const EyeCtx& ec, int m0, int l, uint32_t path, V o, V d, C wpath = C{1, 1, 1}) {
  const Scene& sc = *ec.sc;
  if (l >= ec.maxd) return C{0, 0, 0};
  Isect is;
  if (!calc_intersection(sc, o, d, is)) return C{0, 0, 0};
  const Material& mate = sc.mats[is.mat];
  Rng rng(ec.seed, ec.pixel, ec.pass, TAG_EYE | ((uint32_t)l << 16) | path);
  if (l == 0) {
    rng = *ec.root;
  }
  // L_spec / L_trans geometry (tracer.rs:150-161)
  Flt cos1;
  V rdir0 = specular_reflection(is.n, d, cos1);
  V rdir = reflection_glossy(is.n, rdir0, s_power_glossy(mate), rng);
  Flt eta = relative_ior_average(sc.mats[m0].ior, mate.ior);
  V hvec;
  if (!normalize(rdir - d, hvec)) hvec = is.n;
  V tdir;
  Flt cos2;
  bool has_t = specular_refraction(hvec, d, eta, tdir, cos2);
  Flt cos = cos1 < cos2 ? cos1 : cos2;  // tracer.rs:173
  // L_diffuse (tracer.rs:136-145)
  C di{0, 0, 0};
  C wd = s_diffuse_weight(mate, cos);
  if (wd.r != 0.0 || wd.g != 0.0 || wd.b != 0.0) {
    if (ec.uc)
      for (const Light& lt : sc.lights) di = di + radiance_from_light(sc, is.pos, is.n, lt);
    di = di + estimate_radiance(*ec.kd, ec.r2, ec.pf, is.pos, is.n, *ec.buf);
    if (ec.grad) node_grad(ec, is, mate, wpath, wd, di, ec.r2);
  }
  C zero{0, 0, 0};
  C f = reflection_index(mate.c2, cos);
  C si{0, 0, 0}, ti{0, 0, 0};
  C ws = mate.surface == PPM_SURF_SIMPLE ? f * (1.0 - mate.diff) : (mate.surface == PPM_SURF_TS ? f : zero);
  if (s_reflect(mate, cos1) && (ws.r != 0.0 || ws.g != 0.0 || ws.b != 0.0))
    si = trace_ray(ec, m0, l + 1, path, is.pos, rdir, wpath * ws);
  C wt = zero;
  if (mate.surface == PPM_SURF_SIMPLE) {
    Flt k = (1.0 - mate.diff) * (1.0 - mate.metal);
    wt = C{(1.0 - f.r) * k, (1.0 - f.g) * k, (1.0 - f.b) * k};
  } else if (mate.surface == PPM_SURF_TS && mate.metal == 0.0) {
    wt = C{1.0 - f.r, 1.0 - f.g, 1.0 - f.b} * mate.c1 * (1.0 - mate.diff);
  }
  if (has_t && s_refract(mate, cos1) && (wt.r != 0.0 || wt.g != 0.0 || wt.b != 0.0)) {
    int m02 = is.in ? is.mat : sc.air;
    ti = trace_ray(ec, m02, l + 1, path | (1u << l), is.pos, tdir, wpath * wt);
  }
  return mate.em * SR_HALF + s_bsdf(mate, cos, di, si, ti);
}

// trace_ray_classic (tracer.rs:221-259): the Whitted tree of rtc, with no
// photon map and no random draw below the camera ray.  di = the direct
// light of every light plus the ambient radiance; the mirror direction for
// the reflection, the refraction about the normal (tdir's side decides the
// child's medium, :247), the BSDF at cos1.
static C trace_ray_classic(const Scene& sc, int m0, int l, V o, V d, C ambient) {
  if (l >= 10) return C{0, 0, 0};  // tracer.rs:222
  Isect is;
  if (!calc_intersection(sc, o, d, is)) return C{0, 0, 0};
  const Material& mate = sc.mats[is.mat];
  Flt cos1;
  V rdir = specular_reflection(is.n, d, cos1);
  C di{0, 0, 0};
  for (const Light& lt : sc.lights) di = di + radiance_from_light(sc, is.pos, is.n, lt);
  di = di + ambient;
  C si{0, 0, 0}, ti{0, 0, 0};
  if (s_reflect(mate, cos1)) si = trace_ray_classic(sc, m0, l + 1, is.pos, rdir, ambient);
  Flt eta = relative_ior_average(sc.mats[m0].ior, mate.ior);
  V tdir;
  Flt cos2;
  if (specular_refraction(is.n, d, eta, tdir, cos2) && s_refract(mate, cos1)) {
    int m02 = dot(tdir, is.n) < 0.0 ? is.mat : sc.air;
    ti = trace_ray_classic(sc, m02, l + 1, is.pos, tdir, ambient);
  }
  return mate.em * SR_HALF + s_bsdf(mate, cos1, di, si, ti);
}

// Camera::generate_ray (camera.rs:58-75)
static void generate_ray(const Camera& cam, int x, int y, Rng& rng, V& o, V& d) {
  V blur{0, 0, 0};
  if (cam.blur) {
    Flt r1 = rng.next() - 0.5, r2 = rng.next() - 0.5;
    blur = r1 * cam.eex + r2 * cam.eey;
  }
  Flt r3 = 0.0, r4 = 0.0;
  if (cam.progressive && cam.antialias) {
    r3 = rng.next() - 0.5;
    r4 = rng.next() - 0.5;
  }
  o = cam.eye + blur;
  V dir = cam.origin + ((Flt)x + r3) * cam.esx + ((Flt)y + r4) * cam.esy - blur;
  if (!normalize(dir, d)) d = V{0, 0, 1};
}

static C eye_pixel(const Scene& sc, const Kd& kd, uint64_t seed, uint32_t pass, Flt r2, bool uc, int maxd, int pf,
                   int x, int y, std::vector<std::pair<Flt, int>>& buf, const C* g = nullptr,
                   double* grad = nullptr) {
  uint32_t pixel = (uint32_t)y * (uint32_t)sc.cam.xreso + (uint32_t)x;
  Rng root(seed, pixel, pass, TAG_EYE);
  V o, d;
  generate_ray(sc.cam, x, y, root, o, d);
  EyeCtx ec{&sc, &kd, r2, uc, maxd, pf, seed, pixel, pass, root.n, &root, &buf, g, grad};
  return trace_ray(ec, sc.air, 0, 0u, o, d);
}

// ---------------------------------------------------------------- passes
static void assign(const Scene& sc, uint64_t n_photon, std::vector<uint64_t>& first, std::vector<uint64_t>& cnt,
                   Flt& power, uint64_t& total) {
  Flt tf = 0;
  for (auto& l : sc.lights) tf += l.flux;  // ppmpa.rs:30
  power = tf / (Flt)n_photon;
  total = 0;
  first.clear();
  cnt.clear();
  for (auto& l : sc.lights) {
    Flt nl = std::round(l.flux / power);  // ppmpa.rs:70-72
    uint64_t c = nl > 0 ? (uint64_t)nl : 0;
    first.push_back(total);
    cnt.push_back(c);
    total += c;
  }
}

static void trace_all_photons(const Scene& sc, uint64_t seed, uint32_t pass, uint64_t n_photon, bool uc, int maxd,
                              std::vector<Photon>& out, Flt& power) {
  std::vector<uint64_t> first, cnt;
  uint64_t total;
  assign(sc, n_photon, first, cnt, power, total);
  std::vector<std::vector<Photon>> per((size_t)total);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t i = 0; i < (int64_t)total; ++i) {
    size_t li = 0;
    for (size_t k = 0; k < sc.lights.size(); ++k)
      if ((uint64_t)i >= first[k]) li = k;
    Rng rng(seed, (uint32_t)i, pass, TAG_PHOTON);
    int wl;
    V o, d;
    generate_photon(sc.lights[li], rng, wl, o, d);
    trace_photon(sc, uc, maxd, sc.air, 0, wl, o, d, (uint32_t)i, rng, per[(size_t)i], PathW{});
  }
  out.clear();
  for (auto& v : per) out.insert(out.end(), v.begin(), v.end());
}

}  // namespace orc

using namespace orc;

struct orc_scene {
  Scene sc;
};
struct orc_map {
  Kd kd;
};

extern "C" {

// ---- primitives pinned by the reference's own tests --------------------------
void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) { philox(ctr, key, out); }
double orc_filter_cone(double d, double r2) { return filter_cone(d, r2); }
double orc_filter_gauss(double d, double r2) { return filter_gauss(d, r2); }
int orc_normalize(const double v[3], double out[3]) {
  V o;
  if (!normalize(vd(v), o)) return 0;
  out[0] = o.x; out[1] = o.y; out[2] = o.z;
  return 1;
}
// Ray::new_from_elem(p, d).target(t) (geometry.rs:30-37,56-58)
int orc_ray_target(const double p[3], const double d[3], double t, double out[3]) {
  V dn;
  if (!normalize(vd(d), dn)) return 0;
  V r = vd(p) + dn * t;
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
  return 1;
}
// Shape::new_polygon normal (geometry.rs:93-103)
int orc_polygon_normal(const double p0[3], const double p1[3], const double p2[3], double out[3]) {
  V n;
  if (!normalize(cross(vd(p1) - vd(p0), vd(p2) - vd(p0)), n)) return 0;
  out[0] = n.x; out[1] = n.y; out[2] = n.z;
  return 1;
}
void orc_color_normalize(const double c[3], double out[3]) {
  C r = color_normalize(cd(c));
  out[0] = r.r; out[1] = r.g; out[2] = r.b;
}
int orc_decide_wavelength(const double c[3], double p) { return decide_wavelength(cd(c), p); }
int orc_check_under(const double* ps, int n, double p) { return check_under(ps, n, p); }
double orc_schlick(double f0, double c) { return schlick(f0, c); }
double orc_relative_ior(double a, double b) { return relative_ior(a, b); }
// specular_refraction: returns 1 and t, cos2 when refracted
int orc_specular_refraction(const double n[3], const double v[3], double eta, double t[3], double* cos2) {
  V tt;
  Flt c2;
  bool ok = specular_refraction(vd(n), vd(v), eta, tt, c2);
  *cos2 = c2;
  if (ok) {
    t[0] = tt.x; t[1] = tt.y; t[2] = tt.z;
  }
  return ok ? 1 : 0;
}
// one uniform draw sequence of a stream (for statistical checks)
void orc_uniforms(uint64_t seed, uint32_t elem, uint32_t pass, uint32_t tag, int n, double* out) {
  Rng r(seed, elem, pass, tag);
  for (int i = 0; i < n; ++i) out[i] = r.next();
}
// diffuse_reflection samples about n, stream (seed, i, 0, tag)
void orc_diffuse_samples(const double n[3], uint64_t seed, int count, double* out) {
  for (int i = 0; i < count; ++i) {
    Rng r(seed, (uint32_t)i, 0, 0x30000000u);
    V d = diffuse_reflection(vd(n), r);
    out[3 * i] = d.x; out[3 * i + 1] = d.y; out[3 * i + 2] = d.z;
  }
}

// ---- scene ---------------------------------------------------------------------
static Material material_from_desc(const ppm_material_desc& m) {
  Material mm;
  mm.em = cd(m.emittance);
  mm.ior = cd(m.ior);
  mm.c1 = cd(m.color1);
  mm.c2 = cd(m.color2);
  mm.surface = m.surface;
  mm.diff = m.diffuseness;
  mm.metal = m.metalness;
  mm.rough = m.roughness;
  mm.density_pow = 1.0 / (std::pow(10.0, 5.0 * (1.0 - std::sqrt(m.roughness))) + 1.0);  // surface.rs:52
  mm.q = mm.c1;
  return mm;
}

// reflection_glossy (physics.rs:232-261) about the mirror direction r of a
// surface with normal n: `count` samples, sample i from stream (seed, i).
void orc_reflection_glossy(const double n[3], const double r[3], double pw, uint64_t seed, uint32_t count,
                           double* out) {
  for (uint32_t i = 0; i < count; ++i) {
    Rng rng(seed, i, 0, 0x30000000u);
    V o = reflection_glossy(vd(n), vd(r), pw, rng);
    out[3 * i] = o.x; out[3 * i + 1] = o.y; out[3 * i + 2] = o.z;
  }
}

// Surface::next_direction (surface.rs:211-258) of a photon of channel wl
// arriving along v at a surface of material md (normal n facing the photon,
// relative IOR eta): `count` draws, draw i from stream (seed, i); kind_out
// 0 absorbed, 1 reflected, 2 transmitted, with the direction in dirs_out.
void orc_next_direction(const ppm_material_desc* md, double eta, const double n[3], const double v[3], int wl,
                        uint64_t seed, uint32_t count, int32_t* kind_out, double* dirs_out) {
  const Material m = material_from_desc(*md);
  for (uint32_t i = 0; i < count; ++i) {
    Rng rng(seed, i, 0, 0x40000000u);
    V o{0, 0, 0};
    kind_out[i] = s_next_direction(m, eta, vd(n), vd(v), wl, rng, o);
    dirs_out[3 * i] = o.x; dirs_out[3 * i + 1] = o.y; dirs_out[3 * i + 2] = o.z;
  }
}

orc_scene* orc_scene_create(const ppm_scene_desc* d) {
  orc_scene* o = new orc_scene();
  Scene& sc = o->sc;
  for (int i = 0; i < d->n_shapes; ++i) {
    const ppm_shape_desc& s = d->shapes[i];
    Shape sh;
    sh.kind = s.kind;
    sh.mat = s.material;
    sh.p = vd(s.pos);
    sh.n = vd(s.nvec);
    sh.d1 = vd(s.dir1);
    sh.d2 = vd(s.dir2);
    sh.dist = s.dist;
    sh.radius = s.radius;
    if ((s.kind == PPM_SHAPE_POLYGON || s.kind == PPM_SHAPE_PARALLELOGRAM) && sh.n.x == 0 && sh.n.y == 0 && sh.n.z == 0)
      normalize(cross(sh.d1, sh.d2), sh.n);
    sc.shapes.push_back(sh);
  }
  for (int i = 0; i < d->n_materials; ++i) sc.mats.push_back(material_from_desc(d->materials[i]));
  Material air;  // M_AIR (scene.rs:13-18)
  air.em = C{0, 0, 0};
  air.ior = C{1, 1, 1};
  air.c1 = air.c2 = air.q = C{0, 0, 0};
  air.surface = PPM_SURF_NOTHING;
  air.diff = air.metal = air.rough = air.density_pow = 0.0;
  sc.air = (int)sc.mats.size();
  sc.mats.push_back(air);
  for (int i = 0; i < d->n_lights; ++i) {
    const ppm_light_desc& l = d->lights[i];
    Light L;
    L.kind = l.kind;
    L.color = cd(l.color);
    L.flux = l.flux;
    L.pos = vd(l.pos);
    L.nvec = vd(l.nvec);
    L.d1 = vd(l.dir1);
    L.d2 = vd(l.dir2);
    if (!normalize(vd(l.dir), L.dir)) L.dir = vd(l.dir);
    sc.lights.push_back(L);
  }
  const ppm_camera_desc& c = d->camera;
  sc.cam = Camera{c.xreso, c.yreso, c.progressive, c.antialias, c.blur, c.pfilter,
                  vd(c.eye_pos), vd(c.origin), vd(c.esx), vd(c.esy), vd(c.eex), vd(c.eey)};
  return o;
}
void orc_scene_destroy(orc_scene* s) { delete s; }

// Intersection of one ray (tracer.rs:306-350): returns the object index or -1.
int orc_intersect(const orc_scene* s, const double o[3], const double d[3], double pos[3], double nrm[3]) {
  Isect is;
  if (!calc_intersection(s->sc, vd(o), vd(d), is)) return -1;
  pos[0] = is.pos.x; pos[1] = is.pos.y; pos[2] = is.pos.z;
  nrm[0] = is.n.x; nrm[1] = is.n.y; nrm[2] = is.n.z;
  return is.obj;
}

// Direct light at (p, n) summed over the lights (tracer.rs:136-144).
void orc_direct_light(const orc_scene* s, const double p[3], const double n[3], double rgb[3]) {
  C r{0, 0, 0};
  for (const Light& L : s->sc.lights) r = r + radiance_from_light(s->sc, vd(p), vd(n), L);
  rgb[0] = r.r; rgb[1] = r.g; rgb[2] = r.b;
}

// One pass of photon tracing; returns the stored count (or -count when cap
// is too small, nothing written).
int64_t orc_trace_photons(const orc_scene* s, uint64_t seed, uint32_t pass, uint64_t n_photon, int uc, int maxd,
                          uint64_t cap, double* pos, double* dir, uint8_t* wl, uint32_t* meta, double* power) {
  std::vector<Photon> ph;
  Flt pw;
  trace_all_photons(s->sc, seed, pass, n_photon, uc != 0, maxd, ph, pw);
  if (power) *power = pw;
  if (ph.size() > cap) return -(int64_t)ph.size();
  for (size_t i = 0; i < ph.size(); ++i) {
    pos[3 * i] = ph[i].pos.x; pos[3 * i + 1] = ph[i].pos.y; pos[3 * i + 2] = ph[i].pos.z;
    dir[3 * i] = ph[i].dir.x; dir[3 * i + 1] = ph[i].dir.y; dir[3 * i + 2] = ph[i].dir.z;
    wl[i] = (uint8_t)ph[i].wl;
    if (meta) meta[i] = ph[i].meta;
  }
  return (int64_t)ph.size();
}

// Brute-force ball query: mode 0 = f64 (the reference), 1 = f32 with the
// GPU's operation order.  Index lists ascending; returns the total or -total
// when idx_cap is too small.
int64_t orc_within_brute(uint64_t n, const double* pos, uint64_t nq, const double* q, double r2, int mode,
                         uint64_t idx_cap, uint32_t* idx, uint64_t* offsets) {
  std::vector<std::vector<uint32_t>> res(nq);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t k = 0; k < (int64_t)nq; ++k) {
    const double* qq = q + 3 * k;
    for (uint64_t i = 0; i < n; ++i) {
      bool in;
      if (mode == 0) {
        in = sq_euclid(vd(qq), vd(pos + 3 * i)) <= r2;
      } else {
        float d = sq_euclid_f((float)qq[0], (float)qq[1], (float)qq[2], (float)pos[3 * i], (float)pos[3 * i + 1],
                              (float)pos[3 * i + 2]);
        in = d <= (float)r2;
      }
      if (in) res[k].push_back((uint32_t)i);
    }
  }
  uint64_t tot = 0;
  for (uint64_t k = 0; k < nq; ++k) {
    offsets[k] = tot;
    tot += res[k].size();
  }
  offsets[nq] = tot;
  if (tot > idx_cap) return -(int64_t)tot;
  for (uint64_t k = 0; k < nq; ++k) std::copy(res[k].begin(), res[k].end(), idx + offsets[k]);
  return (int64_t)tot;
}

// build_photonmap (photonmap.rs:23-29) over a host photon set.
orc_map* orc_map_build(uint64_t n, const double* pos, const double* dir, const uint8_t* wl, double power) {
  orc_map* m = new orc_map();
  m->kd.power = power;
  m->kd.ph.resize(n);
  for (uint64_t i = 0; i < n; ++i)
    m->kd.ph[i] = Photon{wl[i], vd(pos + 3 * i), vd(dir + 3 * i), 0u, PathW{}};
  kd_build(m->kd);
  return m;
}
void orc_map_destroy(orc_map* m) { delete m; }

// KdTree::within through the oracle's kd-tree (ascending by distance, then index).
int64_t orc_map_within(const orc_map* m, const double q[3], double r2, uint64_t cap, uint32_t* idx, double* d2) {
  std::vector<std::pair<Flt, int>> buf;
  kd_within(m->kd, vd(q), r2, buf);
  if (buf.size() > cap) return -(int64_t)buf.size();
  for (size_t i = 0; i < buf.size(); ++i) {
    idx[i] = (uint32_t)buf[i].second;
    if (d2) d2[i] = buf[i].first;
  }
  return (int64_t)buf.size();
}

// estimate_radiance for nq (position, normal) pairs.
void orc_gather(const orc_map* m, uint64_t nq, const double* q, const double* nrm, double r2, int pf, double* rgb) {
#pragma omp parallel
  {
    std::vector<std::pair<Flt, int>> buf;
#pragma omp for schedule(dynamic, 64)
    for (int64_t k = 0; k < (int64_t)nq; ++k) {
      C c = estimate_radiance(m->kd, r2, pf, vd(q + 3 * k), vd(nrm + 3 * k), buf);
      rgb[3 * k] = c.r; rgb[3 * k + 1] = c.g; rgb[3 * k + 2] = c.b;
    }
  }
}

// estimate_radiance for nq (position, normal) pairs over the float-decided
// ball of the GPU passes (estimate_radiance_f32set).
void orc_gather_f32set(const orc_map* m, uint64_t nq, const double* q, const double* nrm, double r2, int pf,
                       double* rgb) {
#pragma omp parallel
  {
    std::vector<std::pair<Flt, int>> buf;
#pragma omp for schedule(dynamic, 16)
    for (int64_t k = 0; k < (int64_t)nq; ++k) {
      C c = estimate_radiance_f32set(m->kd, r2, pf, vd(q + 3 * k), vd(nrm + 3 * k), buf);
      rgb[3 * k] = c.r; rgb[3 * k + 1] = c.g; rgb[3 * k + 2] = c.b;
    }
  }
}

// One eye pass over the camera's screen (ppmpa.rs:81-82): row-major image.
void orc_trace_eye(const orc_scene* s, const orc_map* m, uint64_t seed, uint32_t pass, double r2, int uc, int maxd,
                   int pf, double* rgb) {
  const Scene& sc = s->sc;
  const int W = sc.cam.xreso, H = sc.cam.yreso;
#pragma omp parallel
  {
    std::vector<std::pair<Flt, int>> buf;
#pragma omp for schedule(dynamic, 16)
    for (int64_t p = 0; p < (int64_t)W * H; ++p) {
      int y = (int)(p / W), x = (int)(p % W);
      C c = eye_pixel(sc, m->kd, seed, pass, r2, uc != 0, maxd, pf, x, y, buf);
      rgb[3 * p] = c.r; rgb[3 * p + 1] = c.g; rgb[3 * p + 2] = c.b;
    }
  }
}

// rtc (rtc.rs:17-47): trace_ray_classic for every pixel, the camera ray
// drawn from the pixel's stream (seed, pass): row-major image.
void orc_trace_classic(const orc_scene* s, uint64_t seed, uint32_t pass, const double ambient[3], double* rgb) {
  const Scene& sc = s->sc;
  const int W = sc.cam.xreso, H = sc.cam.yreso;
  const C amb = cd(ambient);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t p = 0; p < (int64_t)W * H; ++p) {
    int y = (int)(p / W), x = (int)(p % W);
    Rng root(seed, (uint32_t)p, pass, TAG_EYE);
    V o, d;
    generate_ray(sc.cam, x, y, root, o, d);
    C c = trace_ray_classic(sc, sc.air, 0, o, d, amb);
    rgb[3 * p] = c.r; rgb[3 * p + 1] = c.g; rgb[3 * p + 2] = c.b;
  }
}

// r_i^2 schedule (util/iterator.rb:33-38) for passes first..first+n-1.
void orc_radius_schedule(double r1, double alpha, uint32_t first, uint32_t n, double* r2_out) {
  Flt r2 = r1 * r1;
  for (uint32_t k = 1; k < first; ++k) r2 *= ((Flt)k + alpha) / ((Flt)k + 1.0);
  for (uint32_t i = 0; i < n; ++i) {
    r2_out[i] = r2;
    Flt k = (Flt)(first + i);
    r2 *= (k + alpha) / (k + 1.0);
  }
}

// PPM-PA passes (iterator.rb + averager2.rb): per-pixel mean and unbiased
// variance over the passes.  Returns stored photons summed over passes.
int64_t orc_run_passes(const orc_scene* s, const ppm_pass_cfg* cfg, uint32_t n_pass, double* mean, double* var,
                       int threads, double* ms_trace, double* ms_eye) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
  const Scene& sc = s->sc;
  const uint64_t P = (uint64_t)sc.cam.xreso * sc.cam.yreso;
  std::vector<double> sum(3 * P, 0.0), sq(3 * P, 0.0), img(3 * P);
  std::vector<double> r2s(n_pass);
  uint32_t first = cfg->first_pass ? cfg->first_pass : 1;
  orc_radius_schedule(cfg->radius0, cfg->alpha, first, n_pass, r2s.data());
  int64_t stored = 0;
  double t_trace = 0, t_eye = 0;
  for (uint32_t i = 0; i < n_pass; ++i) {
    uint32_t pass = first + i;
#ifdef _OPENMP
    double t0 = omp_get_wtime();
#endif
    orc_map m;
    Flt pw;
    trace_all_photons(sc, cfg->seed, pass, cfg->n_photon, cfg->use_classic != 0, cfg->max_depth, m.kd.ph, pw);
    m.kd.power = pw;
    kd_build(m.kd);
    stored += (int64_t)m.kd.ph.size();
#ifdef _OPENMP
    double t1 = omp_get_wtime();
#endif
    orc_trace_eye(s, &m, cfg->seed, pass, r2s[i], cfg->use_classic, cfg->max_depth, cfg->pfilter, img.data());
#ifdef _OPENMP
    double t2 = omp_get_wtime();
    t_trace += (t1 - t0) * 1e3;
    t_eye += (t2 - t1) * 1e3;
#endif
    for (uint64_t k = 0; k < 3 * P; ++k) {
      sum[k] += img[k];
      sq[k] += img[k] * img[k];
    }
  }
  const double n = (double)n_pass;
  for (uint64_t k = 0; k < 3 * P; ++k) {
    if (mean) mean[k] = n > 0 ? sum[k] / n : 0.0;
    if (var) var[k] = n > 1 ? std::max(0.0, (sq[k] - sum[k] * sum[k] / n) / (n - 1.0)) : 0.0;
  }
  if (ms_trace) *ms_trace = t_trace;
  if (ms_eye) *ms_eye = t_eye;
  return stored;
}

// Material parameters for the differentiable form: c1 (n_mat x 3, NULL:
// keep), the survival probabilities q of the diffuse bounce (NULL: q = c1),
// and a factor on every light's flux and every emittance.
void orc_scene_set_params(orc_scene* s, int n_mat, const double* c1, const double* q, double light_scale) {
  Scene& sc = s->sc;
  for (int i = 0; i < n_mat && i < sc.air; ++i) {
    Material& m = sc.mats[i];
    if (c1) m.c1 = cd(c1 + 3 * i);
    m.q = q ? cd(q + 3 * i) : m.c1;
    m.em = m.em * light_scale;
  }
  for (Light& L : sc.lights) L.flux *= light_scale;
}

// The passes of orc_run_passes with, for dL/d(mean image) g_img, the
// gradient dL/dc1 of the same estimator on the same random numbers
// (grad: 3 per material; pass k's image gets g_img / n_pass).
int64_t orc_run_passes_grad(const orc_scene* s, const ppm_pass_cfg* cfg, uint32_t n_pass, const double* g_img,
                            double* grad, double* mean, int threads) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
  const Scene& sc = s->sc;
  const int W = sc.cam.xreso, H = sc.cam.yreso;
  const uint64_t P = (uint64_t)W * H;
  const size_t NG = 3 * sc.mats.size();
  std::vector<double> sum(3 * P, 0.0), total(NG, 0.0);
  std::vector<double> r2s(n_pass);
  uint32_t first = cfg->first_pass ? cfg->first_pass : 1;
  orc_radius_schedule(cfg->radius0, cfg->alpha, first, n_pass, r2s.data());
  int64_t stored = 0;
  for (uint32_t i = 0; i < n_pass; ++i) {
    uint32_t pass = first + i;
    orc_map m;
    Flt pw;
    trace_all_photons(sc, cfg->seed, pass, cfg->n_photon, cfg->use_classic != 0, cfg->max_depth, m.kd.ph, pw);
    m.kd.power = pw;
    kd_build(m.kd);
    stored += (int64_t)m.kd.ph.size();
#pragma omp parallel
    {
      std::vector<std::pair<Flt, int>> buf;
      std::vector<double> gl(NG, 0.0);
#pragma omp for schedule(dynamic, 16)
      for (int64_t p = 0; p < (int64_t)P; ++p) {
        int y = (int)(p / W), x = (int)(p % W);
        const C g = C{g_img[3 * p], g_img[3 * p + 1], g_img[3 * p + 2]} * (1.0 / n_pass);
        C c = eye_pixel(sc, m.kd, cfg->seed, pass, r2s[i], cfg->use_classic != 0, cfg->max_depth, cfg->pfilter, x,
                        y, buf, &g, gl.data());
        sum[3 * p] += c.r; sum[3 * p + 1] += c.g; sum[3 * p + 2] += c.b;
      }
#pragma omp critical
      for (size_t k = 0; k < NG; ++k) total[k] += gl[k];
    }
  }
  const size_t nm = NG - 3;  // without M_AIR
  for (size_t k = 0; k < nm; ++k) grad[k] = total[k];
  if (mean)
    for (uint64_t k = 0; k < 3 * P; ++k) mean[k] = sum[k] / (double)n_pass;
  return stored;
}

int orc_max_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
