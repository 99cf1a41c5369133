// TEST INFRASTRUCTURE ONLY -- the oracle.  Nothing in the product links or
// calls this file; only tests/, the smoke check and the CPU-baseline leg of
// bench.py load it, as the checker.
//
// A CPU restatement of fmm-rainbow's p=1 Cartesian FMM step, written from the
// WGSL text of src/rainbow_fmm.wgsl and the planner/harness of src/lib.rs and
// native/main.rs.  Every kernel reads its Params block exactly as the WGSL
// entry point does (slot table of native/main.rs:89-128) and evaluates its
// f32 arithmetic in the order the WGSL expressions are written, left to
// right, with no FMA contraction (built with -ffp-contract=off):
//   dot(a, b)      = (a.x*b.x + a.y*b.y) + a.z*b.z
//   length(v)      = sqrt(dot(v, v))
//   inverseSqrt(x) = 1 / sqrt(x)          (both correctly rounded)
//   round(x)       = round-half-to-even   (nearbyint)
//   clamp(x,lo,hi) = min(max(x, lo), hi)
//   mix(a,b,t)     = a*(1-t) + b*t;  fract(x) = x - floor(x)
//   smoothstep(e0,e1,x): t = clamp((x-e0)/(e1-e0), 0, 1); t*t*(3-2t)
// i32 sums wrap (two's complement), as WGSL atomicAdd does.
//
// Parity pins (tests/): planner tests of src/lib.rs:163-217 and closed forms;
// FNV-1a textbook vectors; integer binning/deposit/upward against an
// independent numpy brute force and invariants; pair_field against its
// closed form in f64, antisymmetry and trace; the M2L/near stencils against
// the mass-partition invariant; the field solve against the f64 twin; the
// integrator against closed-form kick-drift and boundary cases; the galaxy
// seeding (initialize) by inverting hash32, the enclosed-mass model's
// closed values and the geometry of each seeded population (shares, radii,
// planes, arm phases, circular speeds, orthogonality); integrate's
// excitation and hue lanes and the control ("shepherd") branch
// (src/rainbow_fmm.wgsl:659-768) by one-term-at-a-time targets, geometric
// relaxation, the smoothstep's Hermite values, a Gaussian influence against
// exp in f64, exact attract/repel negation and swirl orthogonality.

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// src/rainbow_fmm.wgsl:9-16 and src/lib.rs:10-23.
constexpr uint32_t kAdaptiveLeafCapacity = 16u;
constexpr float kWorldHalfExtent = 1.60f;
const float kWorldEdge = kWorldHalfExtent * 0.985f;

// struct Params, src/rainbow_fmm.wgsl:18-50.
struct Params {
  uint32_t particle_count, depth, level, level_offset;
  uint32_t level_dim, parent_offset, parent_dim, total_cells;
  float dt, gravity, softening, point_size;
  float time, aspect, camera_yaw, camera_pitch;
  float viewport_width, viewport_height, exposure, speed;
  float control_point[4];
  float control_axis_mode[4];
  float control_settings[4];
};
static_assert(sizeof(Params) == 128, "Params is 128 bytes");

// struct Particle / Moment / Field, src/rainbow_fmm.wgsl:52-73.
struct Particle {
  float pos[4];  // xyz position, w hue
  float vel[4];  // xyz velocity, w excitation
};
struct Moment {
  uint32_t count;
  int32_t sx, sy, sz;
};
template <class R>
struct FieldT {
  R a[4];   // a.xyz, phi
  R j0[4];  // Jxx, Jxy, Jxz, Jyy
  R j1[4];  // Jyz, Jzz, 0, 0
};
static_assert(sizeof(Particle) == 32, "Particle is 32 bytes");
static_assert(sizeof(Moment) == 16, "Moment is 16 bytes");
static_assert(sizeof(FieldT<float>) == 48, "Field is 48 bytes");

int g_threads = 0;

#define ORACLE_PARALLEL_FOR _Pragma("omp parallel for schedule(static) num_threads(nthreads())")

int nthreads() {
#ifdef _OPENMP
  return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
  return 1;
#endif
}

// ---- planner restatement, src/lib.rs:25-69 ----

uint32_t level_cells(uint32_t level) { return 1u << (3u * level); }
uint32_t octree_offset(uint32_t level) { return (level_cells(level) - 1u) / 7u; }
uint32_t total_cells(uint32_t depth) { return octree_offset(depth + 1u); }

uint32_t depth_for(uint32_t n) {
  const uint32_t target = n / 8u + (n % 8u != 0u ? 1u : 0u);  // u32::div_ceil(n, 8)
  uint32_t d = 4u;
  while (d < 6u && level_cells(d) < target) ++d;
  return d;
}

float scale_for(uint32_t n) {
  const double worst = static_cast<double>(n) * static_cast<double>(kWorldHalfExtent);
  const double max_scale = 0.82 * 2147483647.0 / (worst > 1.0 ? worst : 1.0);
  return static_cast<float>(std::exp2(std::floor(std::log2(max_scale))));
}

// ---- geometry, src/rainbow_fmm.wgsl:112-152 ----

inline uint32_t linear_index(uint32_t x, uint32_t y, uint32_t z, uint32_t dim) {
  return x + dim * (y + dim * z);
}

inline void cell_coord(uint32_t index, uint32_t dim, uint32_t c[3]) {
  const uint32_t plane = dim * dim;
  const uint32_t z = index / plane;
  const uint32_t rest = index - z * plane;
  const uint32_t y = rest / dim;
  c[0] = rest - y * dim;
  c[1] = y;
  c[2] = z;
}

template <class R>
inline R cell_center1(uint32_t c, uint32_t dim) {
  return ((static_cast<R>(c) + R(0.5)) * (R(2.0) / static_cast<R>(dim)) - R(1.0)) *
         static_cast<R>(kWorldHalfExtent);
}

inline uint32_t position_cell1(float p, uint32_t dim) {
  float n = p / kWorldHalfExtent;
  n = std::fmin(std::fmax(n, -0.99999f), 0.99999f);
  const float s = std::floor((n * 0.5f + 0.5f) * static_cast<float>(dim));
  return static_cast<uint32_t>(s);
}

inline int32_t to_i32_sat(float x) {
  // WGSL i32(f32): truncation, saturated at the i32 range.
  if (!(x == x)) return 0;
  if (x >= 2147483648.0f) return 2147483647;
  if (x < -2147483648.0f) return static_cast<int32_t>(0x80000000u);
  return static_cast<int32_t>(x);
}

inline int32_t wrap_add(int32_t a, int32_t b) {
  return static_cast<int32_t>(static_cast<uint32_t>(a) + static_cast<uint32_t>(b));
}

template <class R>
inline R dot3(const R a[3], const R b[3]) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

template <class R>
inline void zero_field(FieldT<R>& f) {
  for (int k = 0; k < 4; ++k) f.a[k] = f.j0[k] = f.j1[k] = R(0);
}

// add_field, src/rainbow_fmm.wgsl:108-110.
template <class R>
inline void add_field(FieldT<R>& acc, const FieldT<R>& b) {
  for (int k = 0; k < 4; ++k) {
    acc.a[k] = acc.a[k] + b.a[k];
    acc.j0[k] = acc.j0[k] + b.j0[k];
    acc.j1[k] = acc.j1[k] + b.j1[k];
  }
}

// field_j_mul, src/rainbow_fmm.wgsl:112-118.
template <class R>
inline void field_j_mul(const FieldT<R>& f, const R v[3], R out[3]) {
  out[0] = (f.j0[0] * v[0] + f.j0[1] * v[1]) + f.j0[2] * v[2];
  out[1] = (f.j0[1] * v[0] + f.j0[3] * v[1]) + f.j1[0] * v[2];
  out[2] = (f.j0[2] * v[0] + f.j1[0] * v[1]) + f.j1[1] * v[2];
}

// moment_com_read, src/rainbow_fmm.wgsl:163-168.
template <class R>
inline void moment_com(const Moment& m, uint32_t count, float scale, R com[3]) {
  const R inv = R(1.0) / (static_cast<R>(scale) * static_cast<R>(count));
  com[0] = static_cast<R>(m.sx) * inv;
  com[1] = static_cast<R>(m.sy) * inv;
  com[2] = static_cast<R>(m.sz) * inv;
}

// pair_field, src/rainbow_fmm.wgsl:170-203.
template <class R>
inline void pair_field(const Moment* moments, uint32_t source_index, const R target[3],
                       R epsilon, const Params& p, FieldT<R>& out) {
  const uint32_t count = static_cast<uint32_t>(moments[source_index].count);
  if (count == 0u) {
    zero_field(out);
    return;
  }
  R source[3];
  moment_com<R>(moments[source_index], count, p.control_settings[2], source);
  const R r[3] = {source[0] - target[0], source[1] - target[1], source[2] - target[2]};
  const R r2 = dot3(r, r) + epsilon * epsilon;
  const R inv_r = R(1.0) / std::sqrt(r2);
  const R inv_r3 = inv_r * inv_r * inv_r;
  const R inv_r5 = inv_r3 * inv_r * inv_r;
  const R mass = static_cast<R>(count) / static_cast<R>(p.particle_count);
  const R gm = static_cast<R>(p.gravity) * mass;
  const R diagonal = -gm * inv_r3;
  const R tidal = R(3.0) * gm * inv_r5;
  out.a[0] = gm * r[0] * inv_r3;
  out.a[1] = gm * r[1] * inv_r3;
  out.a[2] = gm * r[2] * inv_r3;
  out.a[3] = -gm * inv_r;
  out.j0[0] = diagonal + tidal * r[0] * r[0];
  out.j0[1] = tidal * r[0] * r[1];
  out.j0[2] = tidal * r[0] * r[2];
  out.j0[3] = diagonal + tidal * r[1] * r[1];
  out.j1[0] = tidal * r[1] * r[2];
  out.j1[1] = diagonal + tidal * r[2] * r[2];
  out.j1[2] = R(0);
  out.j1[3] = R(0);
}

// level_of, src/rainbow_fmm.wgsl:380-392.
inline uint32_t level_of(uint32_t g) {
  uint32_t level = 1u, start = 1u;
  for (;;) {
    const uint32_t cells = 1u << (3u * level);
    if (g < start + cells) break;
    start += cells;
    level += 1u;
  }
  return level;
}

// The M2L stencil of m2l_all (src/rainbow_fmm.wgsl:483-517): children of the
// parent's 3x3x3 neighbourhood minus the target's own 3x3x3 neighbourhood,
// visited pz, py, px, then cz, cy, cx.
template <class Visit>
inline void m2l_stencil(uint32_t level, const uint32_t tc_u[3], Visit&& visit) {
  const uint32_t start = octree_offset(level);
  const uint32_t dim = 1u << level;
  const int32_t parent_dim = static_cast<int32_t>(dim >> 1u);
  const int32_t dim_i = static_cast<int32_t>(dim);
  const int32_t tc[3] = {static_cast<int32_t>(tc_u[0]), static_cast<int32_t>(tc_u[1]),
                         static_cast<int32_t>(tc_u[2])};
  const int32_t parent[3] = {tc[0] / 2, tc[1] / 2, tc[2] / 2};
  for (int pz = -1; pz <= 1; ++pz)
    for (int py = -1; py <= 1; ++py)
      for (int px = -1; px <= 1; ++px) {
        const int32_t sp[3] = {parent[0] + px, parent[1] + py, parent[2] + pz};
        if (sp[0] < 0 || sp[1] < 0 || sp[2] < 0 || sp[0] >= parent_dim ||
            sp[1] >= parent_dim || sp[2] >= parent_dim)
          continue;
        for (int cz = 0; cz < 2; ++cz)
          for (int cy = 0; cy < 2; ++cy)
            for (int cx = 0; cx < 2; ++cx) {
              const int32_t sc[3] = {sp[0] * 2 + cx, sp[1] * 2 + cy, sp[2] * 2 + cz};
              if (sc[0] < 0 || sc[1] < 0 || sc[2] < 0 || sc[0] >= dim_i || sc[1] >= dim_i ||
                  sc[2] >= dim_i)
                continue;
              if (std::abs(sc[0] - tc[0]) <= 1 && std::abs(sc[1] - tc[1]) <= 1 &&
                  std::abs(sc[2] - tc[2]) <= 1)
                continue;
              visit(start + linear_index(static_cast<uint32_t>(sc[0]), static_cast<uint32_t>(sc[1]),
                                         static_cast<uint32_t>(sc[2]), dim));
            }
      }
}

// The near stencil of near_all (src/rainbow_fmm.wgsl:551-570): the 26
// same-level neighbours, visited dz, dy, dx.
template <class Visit>
inline void near_stencil(uint32_t level, const uint32_t tc_u[3], Visit&& visit) {
  const uint32_t start = octree_offset(level);
  const uint32_t dim = 1u << level;
  const int32_t dim_i = static_cast<int32_t>(dim);
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        if (dx == 0 && dy == 0 && dz == 0) continue;
        const int32_t sc[3] = {static_cast<int32_t>(tc_u[0]) + dx, static_cast<int32_t>(tc_u[1]) + dy,
                               static_cast<int32_t>(tc_u[2]) + dz};
        if (sc[0] < 0 || sc[1] < 0 || sc[2] < 0 || sc[0] >= dim_i || sc[1] >= dim_i ||
            sc[2] >= dim_i)
          continue;
        visit(start + linear_index(static_cast<uint32_t>(sc[0]), static_cast<uint32_t>(sc[1]),
                                   static_cast<uint32_t>(sc[2]), dim));
      }
}

// Neumaier compensated accumulator for one component.
template <class R>
struct Comp {
  R s = R(0), c = R(0);
  void add(R t) {
    const R u = s + t;
    if (std::fabs(s) >= std::fabs(t))
      c = c + ((s - u) + t);
    else
      c = c + ((t - u) + s);
    s = u;
  }
  R value() const { return s + c; }
};

template <class R>
struct CompField {
  Comp<R> a[4], j0[4], j1[4];
  void reset(const FieldT<R>& f) {
    for (int k = 0; k < 4; ++k) {
      a[k].s = f.a[k];
      a[k].c = R(0);
      j0[k].s = f.j0[k];
      j0[k].c = R(0);
      j1[k].s = f.j1[k];
      j1[k].c = R(0);
    }
  }
  void add(const FieldT<R>& f) {
    for (int k = 0; k < 4; ++k) {
      a[k].add(f.a[k]);
      j0[k].add(f.j0[k]);
      j1[k].add(f.j1[k]);
    }
  }
  void store(FieldT<R>& f) const {
    for (int k = 0; k < 4; ++k) {
      f.a[k] = a[k].value();
      f.j0[k] = j0[k].value();
      f.j1[k] = j1[k].value();
    }
  }
};

// ---- entry points ----

// clear_cells, src/rainbow_fmm.wgsl:394-406.
void clear_cells(const Params& p, Moment* moments) {
  const int64_t n = p.total_cells;
  ORACLE_PARALLEL_FOR
  for (int64_t i = 0; i < n; ++i) moments[i] = Moment{0u, 0, 0, 0};
}

// deposit, src/rainbow_fmm.wgsl:408-422.  Integer sums are order-free, so
// per-thread partial sums merged afterwards give the atomics' result.  Each
// thread keeps its partial finest level between calls and lists the cells
// its bodies touched; only those are merged (relaxed atomic adds, which
// wrap as WGSL's atomicAdd does) and cleared again.
struct DepositScratch {
  std::vector<Moment> cells;
  std::vector<uint32_t> touched;
};

void deposit(const Params& p, const Particle* particles, Moment* moments) {
  static std::vector<DepositScratch> scratch;
  const uint32_t dim = p.level_dim;
  const uint32_t ncell = dim * dim * dim;
  const float fixed_point = p.control_settings[2];
  const int nt = nthreads();
  if (scratch.size() < static_cast<size_t>(nt)) scratch.resize(static_cast<size_t>(nt));
  const int64_t n = p.particle_count;
  Moment* finest = moments + p.level_offset;
#pragma omp parallel num_threads(nt)
  {
#ifdef _OPENMP
    const int t = omp_get_thread_num();
#else
    const int t = 0;
#endif
    DepositScratch& mine = scratch[static_cast<size_t>(t)];
    if (mine.cells.size() < ncell) mine.cells.assign(ncell, Moment{0u, 0, 0, 0});
    mine.touched.clear();
#pragma omp for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      const float* pos = particles[i].pos;
      const uint32_t c = linear_index(position_cell1(pos[0], dim), position_cell1(pos[1], dim),
                                      position_cell1(pos[2], dim), dim);
      Moment& m = mine.cells[c];
      if (m.count == 0u) mine.touched.push_back(c);
      m.count += 1u;
      m.sx = wrap_add(m.sx, to_i32_sat(std::nearbyint(pos[0] * fixed_point)));
      m.sy = wrap_add(m.sy, to_i32_sat(std::nearbyint(pos[1] * fixed_point)));
      m.sz = wrap_add(m.sz, to_i32_sat(std::nearbyint(pos[2] * fixed_point)));
    }
    for (const uint32_t c : mine.touched) {
      Moment& s = mine.cells[c];
      Moment& dst = finest[c];
      __atomic_fetch_add(&dst.count, s.count, __ATOMIC_RELAXED);
      __atomic_fetch_add(reinterpret_cast<uint32_t*>(&dst.sx), static_cast<uint32_t>(s.sx), __ATOMIC_RELAXED);
      __atomic_fetch_add(reinterpret_cast<uint32_t*>(&dst.sy), static_cast<uint32_t>(s.sy), __ATOMIC_RELAXED);
      __atomic_fetch_add(reinterpret_cast<uint32_t*>(&dst.sz), static_cast<uint32_t>(s.sz), __ATOMIC_RELAXED);
      s = Moment{0u, 0, 0, 0};
    }
  }
}

// upward, src/rainbow_fmm.wgsl:424-455 (Params = upward slot: level_dim is
// this level's side, parent_* name the child level).
void upward(const Params& p, Moment* moments) {
  const uint32_t dim = p.level_dim;
  const int64_t cell_count = static_cast<int64_t>(dim) * dim * dim;
  ORACLE_PARALLEL_FOR
  for (int64_t li = 0; li < cell_count; ++li) {
    uint32_t c[3];
    cell_coord(static_cast<uint32_t>(li), dim, c);
    uint32_t count = 0u;
    int32_t sx = 0, sy = 0, sz = 0;
    for (uint32_t dz = 0; dz < 2u; ++dz)
      for (uint32_t dy = 0; dy < 2u; ++dy)
        for (uint32_t dx = 0; dx < 2u; ++dx) {
          const uint32_t child = p.parent_offset + linear_index(c[0] * 2u + dx, c[1] * 2u + dy,
                                                                c[2] * 2u + dz, p.parent_dim);
          count += moments[child].count;
          sx = wrap_add(sx, moments[child].sx);
          sy = wrap_add(sy, moments[child].sy);
          sz = wrap_add(sz, moments[child].sz);
        }
    moments[p.level_offset + li] = Moment{count, sx, sy, sz};
  }
}

// m2l_all, src/rainbow_fmm.wgsl:457-519.
template <class R, bool kComp>
void m2l_all(const Params& p, const Moment* moments, FieldT<R>* fields) {
  const int64_t total = p.total_cells;
  ORACLE_PARALLEL_FOR
  for (int64_t gi = 1; gi < total; ++gi) {
    const uint32_t g = static_cast<uint32_t>(gi);
    const uint32_t level = level_of(g);
    const uint32_t start = octree_offset(level);
    const uint32_t dim = 1u << level;
    const uint32_t parent_dim = dim >> 1u;
    uint32_t tc[3];
    cell_coord(g - start, dim, tc);
    const uint32_t parent_index =
        octree_offset(level - 1u) + linear_index(tc[0] / 2u, tc[1] / 2u, tc[2] / 2u, parent_dim);
    if (static_cast<int32_t>(moments[g].count) == 0 ||
        static_cast<int32_t>(moments[parent_index].count) <= static_cast<int32_t>(kAdaptiveLeafCapacity)) {
      zero_field(fields[g]);
      continue;
    }
    const R center[3] = {cell_center1<R>(tc[0], dim), cell_center1<R>(tc[1], dim),
                         cell_center1<R>(tc[2], dim)};
    const R cell_epsilon =
        static_cast<R>(0.06f * kWorldHalfExtent) / static_cast<R>(dim);
    FieldT<R> result;
    zero_field(result);
    CompField<R> comp;
    if (kComp) comp.reset(result);
    m2l_stencil(level, tc, [&](uint32_t source_index) {
      FieldT<R> f;
      pair_field<R>(moments, source_index, center, cell_epsilon, p, f);
      if (kComp)
        comp.add(f);
      else
        add_field(result, f);
    });
    if (kComp) comp.store(result);
    fields[g] = result;
  }
}

// near_all, src/rainbow_fmm.wgsl:521-572.
template <class R, bool kComp>
void near_all(const Params& p, const Moment* moments, FieldT<R>* fields) {
  const int64_t total = p.total_cells;
  ORACLE_PARALLEL_FOR
  for (int64_t gi = 1; gi < total; ++gi) {
    const uint32_t g = static_cast<uint32_t>(gi);
    const uint32_t level = level_of(g);
    const uint32_t start = octree_offset(level);
    const uint32_t dim = 1u << level;
    const uint32_t parent_dim = dim >> 1u;
    const uint32_t target_count = moments[g].count;
    uint32_t tc[3];
    cell_coord(g - start, dim, tc);
    const uint32_t parent_index =
        octree_offset(level - 1u) + linear_index(tc[0] / 2u, tc[1] / 2u, tc[2] / 2u, parent_dim);
    const bool on_path =
        static_cast<int32_t>(moments[parent_index].count) > static_cast<int32_t>(kAdaptiveLeafCapacity);
    if (target_count == 0u || !on_path ||
        !(target_count <= kAdaptiveLeafCapacity || level == p.depth))
      continue;
    const R center[3] = {cell_center1<R>(tc[0], dim), cell_center1<R>(tc[1], dim),
                         cell_center1<R>(tc[2], dim)};
    FieldT<R> result = fields[g];
    const float eps_f = std::fmax(p.softening, (0.72f * kWorldHalfExtent) / static_cast<float>(dim));
    const R near_epsilon = sizeof(R) == sizeof(float)
                               ? static_cast<R>(eps_f)
                               : std::fmax(static_cast<R>(p.softening),
                                           static_cast<R>(0.72f * kWorldHalfExtent) / static_cast<R>(dim));
    CompField<R> comp;
    if (kComp) comp.reset(result);
    near_stencil(level, tc, [&](uint32_t source_index) {
      FieldT<R> f;
      pair_field<R>(moments, source_index, center, near_epsilon, p, f);
      if (kComp)
        comp.add(f);
      else
        add_field(result, f);
    });
    if (kComp) comp.store(result);
    fields[g] = result;
  }
}

// downward, src/rainbow_fmm.wgsl:574-611 (Params = level slot l).
template <class R>
void downward(const Params& p, const Moment* moments, FieldT<R>* fields) {
  const uint32_t dim = p.level_dim;
  const int64_t cell_count = static_cast<int64_t>(dim) * dim * dim;
  ORACLE_PARALLEL_FOR
  for (int64_t li = 0; li < cell_count; ++li) {
    const uint32_t local_index = static_cast<uint32_t>(li);
    const uint32_t target_index = p.level_offset + local_index;
    uint32_t c[3];
    cell_coord(local_index, dim, c);
    const uint32_t pc[3] = {c[0] / 2u, c[1] / 2u, c[2] / 2u};
    const uint32_t parent_index = p.parent_offset + linear_index(pc[0], pc[1], pc[2], p.parent_dim);
    bool on_path = true;  // target_is_on_adaptive_path, :574-583
    if (p.level != 0u)
      on_path = static_cast<int32_t>(moments[parent_index].count) >
                static_cast<int32_t>(kAdaptiveLeafCapacity);
    if (static_cast<int32_t>(moments[target_index].count) == 0 || !on_path) continue;
    const FieldT<R> parent_field = fields[parent_index];
    const R delta[3] = {cell_center1<R>(c[0], dim) - cell_center1<R>(pc[0], p.parent_dim),
                        cell_center1<R>(c[1], dim) - cell_center1<R>(pc[1], p.parent_dim),
                        cell_center1<R>(c[2], dim) - cell_center1<R>(pc[2], p.parent_dim)};
    FieldT<R> inherited = parent_field;
    R jd[3];
    field_j_mul(parent_field, delta, jd);
    inherited.a[0] = parent_field.a[0] + jd[0];
    inherited.a[1] = parent_field.a[1] + jd[1];
    inherited.a[2] = parent_field.a[2] + jd[2];
    inherited.a[3] = parent_field.a[3] - dot3(parent_field.a, delta);
    add_field(fields[target_index], inherited);
  }
}

// The acceleration integrate applies (src/rainbow_fmm.wgsl:622-657): adaptive
// descent, local expansion about the leaf centre, own-leaf softened monopole.
// Also returns the chosen cell and its count, and the body's potential from
// the same expansion, phi_c - a_c.d - d.(J d)/2 - G (count-1)/N / r_soft with
// d = x - c, which the reference does not form; it feeds the energy
// diagnostic.
template <class R, bool kComp>
void leaf_acceleration(const Params& p, const Moment* moments, const FieldT<R>* fields,
                       const float pos_f[3], R acc[3], R* phi, uint32_t* cell_out,
                       uint32_t* count_out, uint32_t* dim_out) {
  uint32_t chosen_dim = p.level_dim;
  uint32_t chosen_cell[3] = {position_cell1(pos_f[0], chosen_dim), position_cell1(pos_f[1], chosen_dim),
                             position_cell1(pos_f[2], chosen_dim)};
  uint32_t cell_index = p.level_offset + linear_index(chosen_cell[0], chosen_cell[1], chosen_cell[2], chosen_dim);
  for (uint32_t level = 1u; level <= p.depth; ++level) {
    const uint32_t dim = 1u << level;
    const uint32_t c[3] = {position_cell1(pos_f[0], dim), position_cell1(pos_f[1], dim),
                           position_cell1(pos_f[2], dim)};
    const uint32_t index = octree_offset(level) + linear_index(c[0], c[1], c[2], dim);
    const int32_t count_at_level = static_cast<int32_t>(moments[index].count);
    if (count_at_level <= static_cast<int32_t>(kAdaptiveLeafCapacity) || level == p.depth) {
      chosen_dim = dim;
      chosen_cell[0] = c[0];
      chosen_cell[1] = c[1];
      chosen_cell[2] = c[2];
      cell_index = index;
      break;
    }
  }
  const R pos[3] = {static_cast<R>(pos_f[0]), static_cast<R>(pos_f[1]), static_cast<R>(pos_f[2])};
  const R center[3] = {cell_center1<R>(chosen_cell[0], chosen_dim), cell_center1<R>(chosen_cell[1], chosen_dim),
                       cell_center1<R>(chosen_cell[2], chosen_dim)};
  const FieldT<R>& field = fields[cell_index];
  const R d[3] = {pos[0] - center[0], pos[1] - center[1], pos[2] - center[2]};
  const uint32_t count = moments[cell_index].count;
  R self[3] = {R(0), R(0), R(0)};
  R self_phi = R(0);
  bool has_self = false;
  if (count > 1u) {
    R com[3];
    moment_com<R>(moments[cell_index], count, p.control_settings[2], com);
    const R delta[3] = {com[0] - pos[0], com[1] - pos[1], com[2] - pos[2]};
    const R epsilon = sizeof(R) == sizeof(float)
                          ? static_cast<R>(std::fmax(p.softening, (0.62f * kWorldHalfExtent) /
                                                                      static_cast<float>(chosen_dim)))
                          : std::fmax(static_cast<R>(p.softening),
                                      static_cast<R>(0.62f * kWorldHalfExtent) / static_cast<R>(chosen_dim));
    const R inv_r = R(1.0) / std::sqrt(dot3(delta, delta) + epsilon * epsilon);
    const R self_mass = static_cast<R>(count - 1u) / static_cast<R>(p.particle_count);
    const R g = static_cast<R>(p.gravity);
    for (int k = 0; k < 3; ++k) self[k] = g * self_mass * delta[k] * inv_r * inv_r * inv_r;
    self_phi = -(g * self_mass) * inv_r;
    has_self = true;
  }
  if (!kComp) {
    R jd[3];
    field_j_mul(field, d, jd);
    for (int k = 0; k < 3; ++k) {
      acc[k] = field.a[k] + jd[k];
      if (has_self) acc[k] = acc[k] + self[k];
    }
  } else {
    const R row[3][3] = {{field.j0[0], field.j0[1], field.j0[2]},
                         {field.j0[1], field.j0[3], field.j1[0]},
                         {field.j0[2], field.j1[0], field.j1[1]}};
    for (int k = 0; k < 3; ++k) {
      Comp<R> s;
      s.s = field.a[k];
      s.add(row[k][0] * d[0]);
      s.add(row[k][1] * d[1]);
      s.add(row[k][2] * d[2]);
      if (has_self) s.add(self[k]);
      acc[k] = s.value();
    }
  }
  if (phi != nullptr) {
    R jd[3];
    field_j_mul(field, d, jd);
    *phi = ((field.a[3] - dot3(field.a, d)) - R(0.5) * dot3(d, jd)) + self_phi;
  }
  if (cell_out != nullptr) *cell_out = cell_index;
  if (count_out != nullptr) *count_out = count;
  if (dim_out != nullptr) *dim_out = chosen_dim;
}

inline float clampf(float x, float lo, float hi) { return std::fmin(std::fmax(x, lo), hi); }

inline float length3f(const float v[3]) { return std::sqrt(dot3(v, v)); }

// integrate, src/rainbow_fmm.wgsl:613-771.
template <bool kComp>
void integrate(const Params& p, const Moment* moments, const FieldT<float>* fields,
               Particle* particles) {
  const int64_t n = p.particle_count;
  const float edge_band = 0.12f * kWorldHalfExtent;
  ORACLE_PARALLEL_FOR
  for (int64_t i = 0; i < n; ++i) {
    Particle particle = particles[i];
    float pos[3] = {particle.pos[0], particle.pos[1], particle.pos[2]};
    float velocity[3] = {particle.vel[0], particle.vel[1], particle.vel[2]};
    float acceleration[3];
    uint32_t count = 0u;
    leaf_acceleration<float, kComp>(p, moments, fields, pos, acceleration, nullptr, nullptr, &count,
                                    nullptr);

    // Shepherd control field, src/rainbow_fmm.wgsl:659-689.
    float control_influence = 0.0f;
    if (p.control_point[3] > 0.5f) {
      const float control_delta[3] = {p.control_point[0] - pos[0], p.control_point[1] - pos[1],
                                      p.control_point[2] - pos[2]};
      const float distance2 = dot3(control_delta, control_delta);
      const float control_radius = std::fmax(p.control_settings[1], 0.02f);
      control_influence = std::exp(-2.0f * distance2 / (control_radius * control_radius));
      const float control_strength = p.control_settings[0];
      const float isq = 1.0f / std::sqrt(distance2 + control_radius * control_radius * 0.018f);
      const float direction[3] = {control_delta[0] * isq, control_delta[1] * isq, control_delta[2] * isq};
      const float mode = p.control_axis_mode[3];
      if (mode < 1.5f) {
        for (int k = 0; k < 3; ++k)
          acceleration[k] = acceleration[k] + direction[k] * control_strength * control_influence;
      } else if (mode < 2.5f) {
        for (int k = 0; k < 3; ++k)
          acceleration[k] = acceleration[k] - direction[k] * control_strength * control_influence;
      } else {
        const float axis_len = length3f(p.control_axis_mode);
        const float ax[3] = {p.control_axis_mode[0] / axis_len, p.control_axis_mode[1] / axis_len,
                             p.control_axis_mode[2] / axis_len};
        const float tangent[3] = {ax[1] * control_delta[2] - ax[2] * control_delta[1],
                                  ax[2] * control_delta[0] - ax[0] * control_delta[2],
                                  ax[0] * control_delta[1] - ax[1] * control_delta[0]};
        const float tangent_length = length3f(tangent);
        if (tangent_length > 0.00001f)
          for (int k = 0; k < 3; ++k)
            acceleration[k] =
                acceleration[k] + tangent[k] / tangent_length * control_strength * control_influence;
      }
    }

    const float dt = p.dt * p.speed;
    const float radius = std::fmax(length3f(pos), 0.0001f);
    const float velocity_length = std::fmax(length3f(velocity), 0.0001f);
    const float signed_radial_motion = clampf(dot3(pos, velocity) / (radius * velocity_length), -1.0f, 1.0f);
    const float radial_fraction = std::fabs(signed_radial_motion);
    const float radial_shock = radial_fraction * radial_fraction * radial_fraction;
    const float density_pulse =
        clampf((static_cast<float>(count) - static_cast<float>(kAdaptiveLeafCapacity)) / 72.0f, 0.0f, 1.0f);
    const float field_strength = clampf(length3f(acceleration) * (radius + 0.12f) * 0.58f, 0.0f, 1.0f);
    const float target_excitation =
        clampf(0.025f + radial_shock * 0.56f + density_pulse * 0.26f + field_strength * 0.13f +
                   control_influence * 0.42f,
               0.0f, 1.0f);
    float excitation = particle.vel[3];
    const float response = target_excitation > excitation ? 10.0f : 1.05f;
    const float w = clampf(dt * response, 0.0f, 1.0f);
    excitation = excitation * (1.0f - w) + target_excitation * w;

    for (int k = 0; k < 3; ++k) velocity[k] = velocity[k] + acceleration[k] * dt;
    for (int k = 0; k < 3; ++k) pos[k] = pos[k] + velocity[k] * dt;

    for (int axis = 0; axis < 3; ++axis) {
      const float over = std::fabs(pos[axis]) - (kWorldEdge - edge_band);
      if (over > 0.0f) {
        const float absorb = clampf(over / edge_band, 0.0f, 1.0f);
        velocity[axis] = velocity[axis] * (1.0f - absorb * 0.16f);
      }
      if (std::fabs(pos[axis]) > kWorldEdge) {
        pos[axis] = clampf(pos[axis], -kWorldEdge, kWorldEdge);
        velocity[axis] = velocity[axis] * -0.55f;
      }
    }

    const float t = clampf((excitation - 0.30f) / (0.70f - 0.30f), 0.0f, 1.0f);
    const float hue_drift = 1.0f - t * t * (3.0f - 2.0f * t);
    const float h = particle.pos[3] + dt * hue_drift *
                                          (0.012f + signed_radial_motion * 0.030f + excitation * 0.036f +
                                           control_influence * 0.080f);
    const float dynamical_hue = h - std::floor(h);
    particle.pos[0] = pos[0];
    particle.pos[1] = pos[1];
    particle.pos[2] = pos[2];
    particle.pos[3] = dynamical_hue;
    particle.vel[0] = velocity[0];
    particle.vel[1] = velocity[1];
    particle.vel[2] = velocity[2];
    particle.vel[3] = excitation;
    particles[i] = particle;
  }
}

const Params& slot(const uint8_t* slots, uint32_t index) {
  return *reinterpret_cast<const Params*>(slots + 256u * index);
}

// The field solve of native/main.rs:436-449: clear, deposit, upward (d-1..0),
// m2l_all, downward (1..d), near_all.
template <class R, bool kComp>
void solve_fields(const uint8_t* slots, const Particle* particles, Moment* moments, FieldT<R>* fields) {
  const uint32_t depth = slot(slots, 0).depth;
  clear_cells(slot(slots, depth), moments);
  deposit(slot(slots, depth), particles, moments);
  for (uint32_t l = depth; l-- > 0u;) upward(slot(slots, 8u + l), moments);
  m2l_all<R, kComp>(slot(slots, 0), moments, fields);
  for (uint32_t l = 1u; l <= depth; ++l) downward<R>(slot(slots, l), moments, fields);
  near_all<R, kComp>(slot(slots, 0), moments, fields);
}

// ---------------------------------------------------------------------------
// initialize, src/rainbow_fmm.wgsl:205-374: the reference's default scene.
// f32 throughout; sin/cos/pow/atan2 from the host libm (the WGSL leaves
// their accuracy to the implementation, so this restatement is compared
// within a tolerance).
// ---------------------------------------------------------------------------

struct V3 {
  float x, y, z;
};

V3 v3(float x, float y, float z) { return V3{x, y, z}; }
V3 scale3(float s, V3 v) { return v3(s * v.x, s * v.y, s * v.z); }
V3 cross3(V3 a, V3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
float len3(V3 v) { return std::sqrt((v.x * v.x + v.y * v.y) + v.z * v.z); }
V3 unit3(V3 v) {
  const float l = len3(v);
  return v3(v.x / l, v.y / l, v.z / l);
}
float clamp01f(float x) { return std::fmin(std::fmax(x, 0.0f), 1.0f); }

// hash32 / random01, src/rainbow_fmm.wgsl:205-214.
uint32_t hash32(uint32_t v) {
  uint32_t x = v;
  x = ((x >> 16u) ^ x) * 0x45d9f3bu;
  x = ((x >> 16u) ^ x) * 0x45d9f3bu;
  return (x >> 16u) ^ x;
}
float random01(uint32_t seed) { return static_cast<float>(hash32(seed)) * (1.0f / 4294967296.0f); }

// seeded_enclosed_mass, src/rainbow_fmm.wgsl:216-230.
float seeded_enclosed_mass(float r) {
  const float bulge = 0.05f * std::pow(clamp01f(r / 0.11f), 3.0f);
  const float halo = 0.70f * std::pow(clamp01f(r / 0.70f), 3.0f);
  const float disc = 0.18f * std::pow(clamp01f((r - 0.055f) / 0.60f), 1.0f / 0.72f);
  const float polar = 0.07f * std::pow(clamp01f((r - 0.22f) / 0.31f), 1.0f / 0.86f);
  return std::fmax(((bulge + halo) + disc) + polar, 0.0001f);
}

// One body of initialize, src/rainbow_fmm.wgsl:232-374.
Particle seed_body(uint32_t i, float gravity) {
  const float tau = 6.28318530718f;
  float u[6];
  for (uint32_t k = 0; k < 6u; ++k) u[k] = random01(i * 11u + 1u + k);
  V3 pos, vel;
  float excitation;
  if (u[5] < 0.75f) {  // :247-284
    const float z = u[0] * 2.0f - 1.0f;
    const float phi = tau * u[1];
    const float radius = (u[5] < 0.05f ? 0.11f : 0.70f) * std::pow(u[2], 0.3333333f);
    const float ring = std::sqrt(std::fmax(0.0f, 1.0f - z * z));
    const V3 dir = v3(ring * std::cos(phi), z, ring * std::sin(phi));
    pos = scale3(radius, dir);
    const float tz = u[3] * 2.0f - 1.0f;
    const float tphi = tau * u[4];
    const float tring = std::sqrt(std::fmax(0.0f, 1.0f - tz * tz));
    V3 tangent = cross3(dir, v3(tring * std::cos(tphi), tz, tring * std::sin(tphi)));
    if (len3(tangent) < 0.001f) tangent = cross3(dir, v3(0.0f, 1.0f, 0.0f));
    if (len3(tangent) < 0.001f) tangent = cross3(dir, v3(1.0f, 0.0f, 0.0f));
    const float orbital = 0.985f * std::sqrt(gravity * seeded_enclosed_mass(radius) / (radius + 0.018f));
    vel = scale3(orbital, unit3(tangent));
    excitation = u[5] < 0.05f ? 0.09f : 0.025f + u[4] * 0.025f;
  } else if (u[5] < 0.93f) {  // :285-318
    const float radius = 0.055f + 0.60f * std::pow(u[0], 0.72f);
    const float arm = static_cast<float>(i & 3u) * (tau / 4.0f);
    const float angle = (arm + radius * 7.8f) + (u[1] - 0.5f) * (0.045f + radius * 0.055f);
    const float thickness = (((u[2] + u[3]) + u[4]) - 1.5f) * (0.0035f + radius * 0.010f);
    const V3 flat = v3(radius * std::cos(angle), radius * std::sin(angle) * 0.72f, thickness);
    const float tilt = 0.20f;
    pos = v3(flat.x, flat.y * std::cos(tilt) - flat.z * std::sin(tilt),
             flat.y * std::sin(tilt) + flat.z * std::cos(tilt));
    const float orbital = 0.99f * std::sqrt(gravity * seeded_enclosed_mass(radius) / (radius + 0.018f));
    const V3 tangent = v3(-std::sin(angle), std::cos(angle) * 0.72f * std::cos(tilt),
                          std::cos(angle) * 0.72f * std::sin(tilt));
    const V3 pn = unit3(pos);
    const float wave = std::sin(angle * 3.0f - radius * 13.0f);
    const V3 radial = v3(((pn.x * orbital) * 0.035f) * wave, ((pn.y * orbital) * 0.035f) * wave,
                         ((pn.z * orbital) * 0.035f) * wave);
    const V3 tn = scale3(orbital, unit3(tangent));
    vel = v3(tn.x + radial.x, tn.y + radial.y, tn.z + radial.z);
    excitation = 0.08f + u[3] * 0.05f;
  } else {  // :319-351
    const float lineage = static_cast<float>(i & 3u);
    const float radius = (0.22f + 0.31f * std::pow(u[0], 0.86f)) + (lineage - 1.5f) * 0.007f;
    const float angle = tau * u[1] + radius * 3.5f;
    const float thickness = (((u[2] + u[3]) + u[4]) - 1.5f) * (0.004f + radius * 0.009f);
    const V3 flat = v3(radius * std::cos(angle), thickness, radius * std::sin(angle));
    const float tilt = -0.54f + lineage * 0.36f;
    pos = v3(flat.x * std::cos(tilt) - flat.y * std::sin(tilt), flat.x * std::sin(tilt) + flat.y * std::cos(tilt),
             flat.z);
    const V3 tangent = v3(-std::sin(angle) * std::cos(tilt), -std::sin(angle) * std::sin(tilt), std::cos(angle));
    const float orbital = 0.99f * std::sqrt(gravity * seeded_enclosed_mass(radius) / (radius + 0.018f));
    vel = scale3(-orbital, unit3(tangent));
    excitation = 0.20f + u[4] * 0.08f;
  }
  float hue = 0.0f;
  if (u[5] >= 0.75f) {  // :353-369
    const float speed = std::fmax(len3(vel), 0.0001f);
    const float radius = std::fmax(len3(pos), 0.0001f);
    const float radial_motion =
        std::fmin(std::fmax(((pos.x * vel.x + pos.y * vel.y) + pos.z * vel.z) / (radius * speed), -1.0f), 1.0f);
    const V3 am = cross3(pos, vel);
    const float spin = am.y / std::fmax(len3(am), 0.0001f);
    const float h = ((((std::atan2(vel.y, vel.x) / tau + (vel.z / speed) * 0.14f) + spin * 0.075f) +
                      radial_motion * 0.065f) +
                     excitation * 0.085f) +
                    static_cast<float>(i & 3u) * 0.018f;
    hue = h - std::floor(h);
  }
  Particle out;
  out.pos[0] = pos.x;
  out.pos[1] = pos.y;
  out.pos[2] = pos.z;
  out.pos[3] = hue;
  out.vel[0] = vel.x;
  out.vel[1] = vel.y;
  out.vel[2] = vel.z;
  out.vel[3] = excitation;
  return out;
}

}  // namespace

extern "C" {

void oracle_set_threads(int n) { g_threads = n; }

int oracle_threads(void) { return nthreads(); }

void oracle_plan(uint32_t n, uint32_t* depth, uint32_t* cells, float* scale) {
  const uint32_t d = depth_for(n);
  if (depth) *depth = d;
  if (cells) *cells = total_cells(d);
  if (scale) *scale = scale_for(n);
}

// Slot table of native/main.rs:89-128 with the harness constants of
// native/main.rs:57-82; returns the depth.
uint32_t oracle_build_plans(uint32_t n, uint8_t* slots) {
  std::memset(slots, 0, 4096);
  const uint32_t depth = depth_for(n);
  const uint32_t cells = total_cells(depth);
  const float scale = scale_for(n);
  auto write = [&](uint32_t index, uint32_t level, uint32_t parent_level) {
    Params p;
    std::memset(&p, 0, sizeof(p));
    p.particle_count = n;
    p.depth = depth;
    p.level = level;
    p.level_offset = octree_offset(level);
    p.level_dim = 1u << level;
    p.parent_offset = octree_offset(parent_level);
    p.parent_dim = 1u << parent_level;
    p.total_cells = cells;
    p.dt = 0.0009f;
    p.gravity = 0.34f;
    p.softening = 0.012f;
    p.point_size = 5.0f;
    p.aspect = 1.0f;
    p.exposure = 0.84f;
    p.speed = 1.0f;
    p.control_settings[0] = 0.7f;
    p.control_settings[1] = 0.22f;
    p.control_settings[2] = scale;
    std::memcpy(slots + 256u * index, &p, sizeof(p));
  };
  for (uint32_t l = 0; l <= depth; ++l) write(l, l, l > 0u ? l - 1u : 0u);
  for (uint32_t l = 0; l < depth; ++l) write(8u + l, l, l + 1u);
  return depth;
}

uint32_t oracle_fnv1a(const void* data, size_t bytes) {
  const uint8_t* b = static_cast<const uint8_t*>(data);
  uint32_t h = 2166136261u;
  for (size_t i = 0; i < bytes; ++i) h = (h ^ b[i]) * 16777619u;
  return h;
}

uint32_t oracle_position_cell(float p, uint32_t dim) { return position_cell1(p, dim); }

void oracle_clear_cells(const void* prm, void* moments) {
  clear_cells(*static_cast<const Params*>(prm), static_cast<Moment*>(moments));
}
void oracle_deposit(const void* prm, const void* particles, void* moments) {
  deposit(*static_cast<const Params*>(prm), static_cast<const Particle*>(particles),
          static_cast<Moment*>(moments));
}
void oracle_upward(const void* prm, void* moments) {
  upward(*static_cast<const Params*>(prm), static_cast<Moment*>(moments));
}
void oracle_m2l_all(const void* prm, const void* moments, void* fields) {
  m2l_all<float, false>(*static_cast<const Params*>(prm), static_cast<const Moment*>(moments),
                        static_cast<FieldT<float>*>(fields));
}
void oracle_downward(const void* prm, const void* moments, void* fields) {
  downward<float>(*static_cast<const Params*>(prm), static_cast<const Moment*>(moments),
                  static_cast<FieldT<float>*>(fields));
}
void oracle_near_all(const void* prm, const void* moments, void* fields) {
  near_all<float, false>(*static_cast<const Params*>(prm), static_cast<const Moment*>(moments),
                         static_cast<FieldT<float>*>(fields));
}
void oracle_integrate(const void* prm, const void* moments, const void* fields, void* particles) {
  integrate<false>(*static_cast<const Params*>(prm), static_cast<const Moment*>(moments),
                   static_cast<const FieldT<float>*>(fields), static_cast<Particle*>(particles));
}

// One field solve (FP32); compensated = 1 selects Neumaier sums.
void oracle_solve_fields(const uint8_t* slots, const void* particles, void* moments, void* fields,
                         int compensated) {
  if (compensated)
    solve_fields<float, true>(slots, static_cast<const Particle*>(particles),
                              static_cast<Moment*>(moments), static_cast<FieldT<float>*>(fields));
  else
    solve_fields<float, false>(slots, static_cast<const Particle*>(particles),
                               static_cast<Moment*>(moments), static_cast<FieldT<float>*>(fields));
}

// `steps` full steps (native/main.rs:412-421).
void oracle_step(const uint8_t* slots, void* particles, void* moments, void* fields, uint32_t steps,
                 int compensated) {
  const uint32_t depth = slot(slots, 0).depth;
  for (uint32_t s = 0; s < steps; ++s) {
    oracle_solve_fields(slots, particles, moments, fields, compensated);
    if (compensated)
      integrate<true>(slot(slots, depth), static_cast<const Moment*>(moments),
                      static_cast<const FieldT<float>*>(fields), static_cast<Particle*>(particles));
    else
      integrate<false>(slot(slots, depth), static_cast<const Moment*>(moments),
                       static_cast<const FieldT<float>*>(fields), static_cast<Particle*>(particles));
  }
}

// oracle_step (plain sums) with each phase's wall time added into
// seconds[7]: clear, deposit, upward, m2l_all, downward, near_all, integrate.
void oracle_step_timed(const uint8_t* slots, void* particles, void* moments, void* fields, uint32_t steps,
                       double* seconds) {
  using Clock = std::chrono::steady_clock;
  const uint32_t depth = slot(slots, 0).depth;
  Moment* m = static_cast<Moment*>(moments);
  FieldT<float>* f = static_cast<FieldT<float>*>(fields);
  Particle* parts = static_cast<Particle*>(particles);
  for (uint32_t s = 0; s < steps; ++s) {
    Clock::time_point t = Clock::now();
    auto lap = [&](int k) {
      const Clock::time_point now = Clock::now();
      seconds[k] += std::chrono::duration<double>(now - t).count();
      t = now;
    };
    clear_cells(slot(slots, depth), m);
    lap(0);
    deposit(slot(slots, depth), parts, m);
    lap(1);
    for (uint32_t l = depth; l-- > 0u;) upward(slot(slots, 8u + l), m);
    lap(2);
    m2l_all<float, false>(slot(slots, 0), m, f);
    lap(3);
    for (uint32_t l = 1u; l <= depth; ++l) downward<float>(slot(slots, l), m, f);
    lap(4);
    near_all<float, false>(slot(slots, 0), m, f);
    lap(5);
    integrate<false>(slot(slots, depth), m, f, parts);
    lap(6);
  }
}

// The same field solve in f64 (moments stay exact integers): separates the
// algorithm's error from f32 rounding.  fields64 holds 12 doubles per cell.
void oracle_solve_fields_f64(const uint8_t* slots, const void* particles, void* moments,
                             double* fields64) {
  solve_fields<double, false>(slots, static_cast<const Particle*>(particles),
                              static_cast<Moment*>(moments), reinterpret_cast<FieldT<double>*>(fields64));
}

// Per-body acceleration and potential integrate would apply (f32), 4 floats
// per body: ax, ay, az, phi.
void oracle_accelerations(const void* prm, const void* particles, const void* moments,
                          const void* fields, float* out, int compensated) {
  const Params& p = *static_cast<const Params*>(prm);
  const Particle* parts = static_cast<const Particle*>(particles);
  const int64_t n = p.particle_count;
  ORACLE_PARALLEL_FOR
  for (int64_t i = 0; i < n; ++i) {
    float acc[3], phi;
    if (compensated)
      leaf_acceleration<float, true>(p, static_cast<const Moment*>(moments),
                                     static_cast<const FieldT<float>*>(fields), parts[i].pos, acc, &phi,
                                     nullptr, nullptr, nullptr);
    else
      leaf_acceleration<float, false>(p, static_cast<const Moment*>(moments),
                                      static_cast<const FieldT<float>*>(fields), parts[i].pos, acc, &phi,
                                      nullptr, nullptr, nullptr);
    out[4 * i + 0] = acc[0];
    out[4 * i + 1] = acc[1];
    out[4 * i + 2] = acc[2];
    out[4 * i + 3] = phi;
  }
}

// The same in f64 from the f64 field solve.
void oracle_accelerations_f64(const void* prm, const void* particles, const void* moments,
                              const double* fields64, double* out) {
  const Params& p = *static_cast<const Params*>(prm);
  const Particle* parts = static_cast<const Particle*>(particles);
  const int64_t n = p.particle_count;
  ORACLE_PARALLEL_FOR
  for (int64_t i = 0; i < n; ++i) {
    double acc[3], phi;
    leaf_acceleration<double, false>(p, static_cast<const Moment*>(moments),
                                     reinterpret_cast<const FieldT<double>*>(fields64), parts[i].pos, acc,
                                     &phi, nullptr, nullptr, nullptr);
    out[4 * i + 0] = acc[0];
    out[4 * i + 1] = acc[1];
    out[4 * i + 2] = acc[2];
    out[4 * i + 3] = phi;
  }
}

// Leaf chosen by integrate's descent for every body (cell index).
void oracle_leaf_cells(const void* prm, const void* particles, const void* moments, uint32_t* out) {
  const Params& p = *static_cast<const Params*>(prm);
  const Particle* parts = static_cast<const Particle*>(particles);
  const int64_t n = p.particle_count;
  const Moment* m = static_cast<const Moment*>(moments);
  ORACLE_PARALLEL_FOR
  for (int64_t i = 0; i < n; ++i) {
    uint32_t dim = p.level_dim;
    const float* pos = parts[i].pos;
    uint32_t cell = p.level_offset + linear_index(position_cell1(pos[0], dim), position_cell1(pos[1], dim),
                                                  position_cell1(pos[2], dim), dim);
    for (uint32_t level = 1u; level <= p.depth; ++level) {
      dim = 1u << level;
      const uint32_t index = octree_offset(level) + linear_index(position_cell1(pos[0], dim),
                                                                 position_cell1(pos[1], dim),
                                                                 position_cell1(pos[2], dim), dim);
      if (static_cast<int32_t>(m[index].count) <= 16 || level == p.depth) {
        cell = index;
        break;
      }
    }
    out[i] = cell;
  }
}

// Ordered M2L / near source lists of global cell g (valid cells only,
// regardless of occupancy); returns the count written (<= 189 / 26).
uint32_t oracle_m2l_sources(uint32_t g, uint32_t* out) {
  const uint32_t level = level_of(g);
  uint32_t tc[3];
  cell_coord(g - octree_offset(level), 1u << level, tc);
  uint32_t n = 0;
  m2l_stencil(level, tc, [&](uint32_t s) { out[n++] = s; });
  return n;
}

uint32_t oracle_near_sources(uint32_t g, uint32_t* out) {
  const uint32_t level = level_of(g);
  uint32_t tc[3];
  cell_coord(g - octree_offset(level), 1u << level, tc);
  uint32_t n = 0;
  near_stencil(level, tc, [&](uint32_t s) { out[n++] = s; });
  return n;
}

// pair_field for one source moment at one target, in f32 and f64 (12 each).
void oracle_pair_field(const void* prm, const void* moment, const float* target, float epsilon,
                       float* out) {
  FieldT<float> f;
  pair_field<float>(static_cast<const Moment*>(moment), 0u, target, epsilon,
                    *static_cast<const Params*>(prm), f);
  std::memcpy(out, &f, sizeof(f));
}

// FP64 direct sum over all bodies (equal masses 1/N, Plummer softening):
// a_i = G sum_j m (x_j - x_i) / (|x_j - x_i|^2 + soft^2)^(3/2), phi_i the
// matching potential without the self term.  out: 4 doubles per target.
void oracle_direct_sum_f64(uint32_t n, const void* particles, double gravity, double softening,
                           uint32_t ntargets, const uint32_t* targets, double* out) {
  const Particle* parts = static_cast<const Particle*>(particles);
  const double m = 1.0 / static_cast<double>(n);
  const double e2 = softening * softening;
  ORACLE_PARALLEL_FOR
  for (int64_t t = 0; t < static_cast<int64_t>(ntargets); ++t) {
    const uint32_t i = targets[t];
    const double xi = parts[i].pos[0], yi = parts[i].pos[1], zi = parts[i].pos[2];
    double ax = 0.0, ay = 0.0, az = 0.0, phi = 0.0;
    for (uint32_t j = 0; j < n; ++j) {
      if (j == i) continue;
      const double dx = static_cast<double>(parts[j].pos[0]) - xi;
      const double dy = static_cast<double>(parts[j].pos[1]) - yi;
      const double dz = static_cast<double>(parts[j].pos[2]) - zi;
      const double r2 = dx * dx + dy * dy + dz * dz + e2;
      const double inv = 1.0 / std::sqrt(r2);
      const double inv3 = inv * inv * inv;
      ax += dx * inv3;
      ay += dy * inv3;
      az += dz * inv3;
      phi -= inv;
    }
    out[4 * t + 0] = gravity * m * ax;
    out[4 * t + 1] = gravity * m * ay;
    out[4 * t + 2] = gravity * m * az;
    out[4 * t + 3] = gravity * m * phi;
  }
}

// FP64 kinetic energy, linear and angular momentum of a particle buffer
// (equal masses 1/N): out[0] KE, out[2..4] P, out[5..7] L; out[1] = 0.
void oracle_conserved_f64(uint32_t n, const void* particles, double* out) {
  const Particle* parts = static_cast<const Particle*>(particles);
  const double m = 1.0 / static_cast<double>(n);
  double ke = 0, px = 0, py = 0, pz = 0, lx = 0, ly = 0, lz = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const double x = parts[i].pos[0], y = parts[i].pos[1], z = parts[i].pos[2];
    const double vx = parts[i].vel[0], vy = parts[i].vel[1], vz = parts[i].vel[2];
    ke += 0.5 * m * (vx * vx + vy * vy + vz * vz);
    px += m * vx;
    py += m * vy;
    pz += m * vz;
    lx += m * (y * vz - z * vy);
    ly += m * (z * vx - x * vz);
    lz += m * (x * vy - y * vx);
  }
  out[0] = ke;
  out[1] = 0.0;
  out[2] = px;
  out[3] = py;
  out[4] = pz;
  out[5] = lx;
  out[6] = ly;
  out[7] = lz;
}

// initialize (src/rainbow_fmm.wgsl:232-374) for bodies 0..n-1 into a
// reference-layout particle buffer.
void oracle_seed_galaxy(uint32_t n, float gravity, void* particles) {
  Particle* parts = static_cast<Particle*>(particles);
  for (uint32_t i = 0; i < n; ++i) parts[i] = seed_body(i, gravity);
}

// The Neumaier accumulator the compensated field solve uses, over `n` f32
// terms starting from `init` (the value already in the field, as near_all
// starts from the M2L and L2L result): returns s + c.
float oracle_comp_sum_f32(const float* terms, uint32_t n, float init) {
  Comp<float> acc;
  acc.s = init;
  for (uint32_t i = 0; i < n; ++i) acc.add(terms[i]);
  return acc.value();
}

uint32_t oracle_hash32(uint32_t x) { return hash32(x); }

float oracle_enclosed_mass(float r) { return seeded_enclosed_mass(r); }

}  // extern "C"
