// Procedural scenes for the unit-cube octree [-0.5,0.5]^3 and the pinhole
// camera helper.
//
// Leaf grid at depth D: voxel (i,j,k), i along x, j along y (up), k along z,
// centre (-0.5 + (i+0.5)/2^D, ...).  Every generator emits strictly
// increasing Morton codes plus {packed_normal, packed_color} per leaf and
// hands them to build_scene().
//
// Solids are stored as their surface shell: a filled voxel is kept when one of
// its six faces touches an empty voxel or the cube boundary.  A ray that
// starts outside the solid meets the same first voxel in the shell as in the
// solid, and the shell keeps the node count near the surface area.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "svo_internal.h"

using namespace svo_impl;

namespace {

uint32_t leaf_normal_up() { return pack_rgba(128, 255, 128, 0); }

// Colour of Menger leaves: RGBA8 = hash(seed, i, j, k), alpha 255.
uint32_t hash_color(uint64_t seed, uint32_t i, uint32_t j, uint32_t k) {
  return (uint32_t)(hash_xyz(seed, i, j, k) >> 32) | 0xFFu;
}

// Is the centre of voxel (i,j,k) of a 2^depth grid inside the Menger sponge
// of `level` iterations?  Digit n of a coordinate p in base 3 is
// floor(p*3^n) mod 3; with p = (2i+1)/2^(depth+1) the product is never an
// integer, so the test is exact in integers.
bool menger_inside(uint32_t i, uint32_t j, uint32_t k, uint32_t level, uint32_t depth) {
  const uint64_t den = 1ull << (depth + 1);
  uint64_t p3 = 1;
  for (uint32_t n = 1; n <= level; ++n) {
    p3 *= 3;
    int ones = 0;
    ones += ((((2ull * i + 1) * p3) / den) % 3) == 1;
    ones += ((((2ull * j + 1) * p3) / den) % 3) == 1;
    ones += ((((2ull * k + 1) * p3) / den) % 3) == 1;
    if (ones >= 2) return false;
  }
  return true;
}

// ------------------------------------------------------------ value noise
double lattice(uint64_t seed, uint32_t octave, int64_t ix, int64_t iz) {
  uint64_t h = mix64(mix64(seed ^ ((uint64_t)octave << 56)) ^ (((uint64_t)(uint32_t)ix << 32) | (uint32_t)iz));
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

double value_noise(uint64_t seed, uint32_t octave, double u, double v) {
  const double fu = std::floor(u), fv = std::floor(v);
  const int64_t iu = (int64_t)fu, iv = (int64_t)fv;
  double a = u - fu, b = v - fv;
  a = a * a * (3.0 - 2.0 * a);
  b = b * b * (3.0 - 2.0 * b);
  const double v00 = lattice(seed, octave, iu, iv), v10 = lattice(seed, octave, iu + 1, iv);
  const double v01 = lattice(seed, octave, iu, iv + 1), v11 = lattice(seed, octave, iu + 1, iv + 1);
  return (v00 * (1 - a) + v10 * a) * (1 - b) + (v01 * (1 - a) + v11 * a) * b;
}

// 8 octaves, base frequency 4 cells per cube edge, gain 0.5; result in
// [0.2, 0.8).
double terrain_h(uint64_t seed, double x, double z) {
  double sum = 0, norm = 0, amp = 1, f = 4;
  for (uint32_t o = 0; o < 8; ++o) {
    sum += amp * value_noise(seed, o, x * f, z * f);
    norm += amp;
    amp *= 0.5;
    f *= 2;
  }
  return 0.2 + 0.6 * (sum / norm);
}

uint32_t clamp_byte(int v) { return (uint32_t)std::min(255, std::max(0, v)); }

// Height bands (sand, grass, rock, snow) plus +-16 hash noise per channel.
uint32_t terrain_color(uint64_t seed, uint32_t i, uint32_t j, uint32_t k, uint32_t depth) {
  static const int band[4][3] = {{194, 178, 128}, {86, 125, 70}, {120, 110, 100}, {235, 235, 240}};
  const double y = (j + 0.5) / (double)(1u << depth);
  const int b = y < 0.35 ? 0 : y < 0.5 ? 1 : y < 0.65 ? 2 : 3;
  const uint64_t h = hash_xyz(seed ^ 0x5EEDull, i, j, k);
  return pack_rgba(clamp_byte(band[b][0] + (int)(h & 31u) - 16), clamp_byte(band[b][1] + (int)((h >> 8) & 31u) - 16),
                   clamp_byte(band[b][2] + (int)((h >> 16) & 31u) - 16), 255u);
}

svo_status finish(uint32_t depth, std::vector<uint64_t>& codes, std::vector<uint32_t>& attr,
                  uint32_t treelet_bytes, uint32_t flags, svo_scene** out) {
  return build_scene(depth, codes, attr, treelet_bytes, flags, out);
}

}  // namespace

extern "C" svo_status svo_scene_build_menger(uint32_t level, uint32_t depth, uint64_t seed,
                                             uint32_t treelet_bytes, uint32_t flags, svo_scene** out) {
  if (!out) { set_error("svo_scene_build_menger: out is null"); return SVO_ERR_ARG; }
  if (depth < 1 || depth > 10 || level > 8) { set_error("menger: depth %u / level %u out of range", depth, level); return SVO_ERR_ARG; }
  const uint32_t n = 1u << depth;
  // inside[] per axis digit pattern: filled iff centre inside and any face
  // neighbour outside (or on the cube boundary).
  auto inside = [&](int64_t i, int64_t j, int64_t k) {
    if (i < 0 || j < 0 || k < 0 || i >= n || j >= n || k >= n) return false;
    return menger_inside((uint32_t)i, (uint32_t)j, (uint32_t)k, level, depth);
  };
  // The Morton range is cut into ordered parts filled in parallel and joined
  // in order, so the codes stay sorted.
  const uint64_t total = 1ull << (3 * depth);
  const uint64_t parts = std::min<uint64_t>(total, 1024);
  std::vector<std::vector<uint64_t>> pc(parts);
  std::vector<std::vector<uint32_t>> pa(parts);
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t part = 0; part < (int64_t)parts; ++part) {
    const uint64_t c0 = total * part / parts, c1 = total * (part + 1) / parts;
    for (uint64_t c = c0; c < c1; ++c) {
      uint32_t i, j, k;
      morton_decode(c, depth, &i, &j, &k);
      if (!inside(i, j, k)) continue;
      const bool shell = !inside((int64_t)i - 1, j, k) || !inside((int64_t)i + 1, j, k) || !inside(i, (int64_t)j - 1, k) ||
                         !inside(i, (int64_t)j + 1, k) || !inside(i, j, (int64_t)k - 1) || !inside(i, j, (int64_t)k + 1);
      if (!shell) continue;
      pc[part].push_back(c);
      pa[part].push_back(leaf_normal_up());
      pa[part].push_back(hash_color(seed, i, j, k));
    }
  }
  std::vector<uint64_t> codes;
  std::vector<uint32_t> attr;
  for (uint64_t part = 0; part < parts; ++part) {
    codes.insert(codes.end(), pc[part].begin(), pc[part].end());
    attr.insert(attr.end(), pa[part].begin(), pa[part].end());
    std::vector<uint64_t>().swap(pc[part]);
    std::vector<uint32_t>().swap(pa[part]);
  }
  return finish(depth, codes, attr, treelet_bytes, flags, out);
}

extern "C" svo_status svo_terrain_height(uint64_t seed, double x, double z, double* h) {
  if (!h) { set_error("svo_terrain_height: null output"); return SVO_ERR_ARG; }
  *h = terrain_h(seed, x, z);
  return SVO_OK;
}

extern "C" svo_status svo_scene_build_terrain(uint32_t depth, uint64_t seed, uint32_t treelet_bytes,
                                              uint32_t flags, svo_scene** out) {
  if (!out) { set_error("svo_scene_build_terrain: out is null"); return SVO_ERR_ARG; }
  if (depth < 2 || depth > 14) { set_error("terrain: depth %u outside [2,14]", depth); return SVO_ERR_ARG; }
  const uint32_t n = 1u << depth;
  // Column fill heights: voxel j is filled iff its centre (j+0.5)/n < h.
  std::vector<uint32_t> H((size_t)n * n);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i)
    for (uint32_t k = 0; k < n; ++k) {
      const double h = terrain_h(seed, (i + 0.5) / n, (k + 0.5) / n);
      double c = std::ceil(h * n - 0.5);
      c = std::min((double)n, std::max(0.0, c));
      H[(size_t)i * n + k] = (uint32_t)c;
    }
  // Shell of each column: j = 0 (bottom face) and [lo, H-1], where lo is the
  // lowest voxel whose side faces an empty neighbour column (0 on the cube's
  // side walls).  Rows are counted, then filled, in parallel.
  auto shell = [&](uint32_t i, uint32_t k, uint32_t* lo_out) {
    const uint32_t h = H[(size_t)i * n + k];
    if (h == 0) return false;
    uint32_t lo = h - 1;
    if (i == 0 || k == 0 || i == n - 1 || k == n - 1) lo = 0;
    else {
      lo = std::min(lo, H[(size_t)(i - 1) * n + k]);
      lo = std::min(lo, H[(size_t)(i + 1) * n + k]);
      lo = std::min(lo, H[(size_t)i * n + k - 1]);
      lo = std::min(lo, H[(size_t)i * n + k + 1]);
    }
    *lo_out = lo;
    return true;
  };
  std::vector<uint64_t> row_start((size_t)n + 1, 0);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    uint64_t c = 0;
    for (uint32_t k = 0; k < n; ++k) {
      uint32_t lo;
      if (!shell((uint32_t)i, k, &lo)) continue;
      c += (lo > 0 ? 1 : 0) + (H[(size_t)i * n + k] - lo);
    }
    row_start[(size_t)i + 1] = c;
  }
  for (uint32_t i = 0; i < n; ++i) row_start[i + 1] += row_start[i];
  std::vector<uint64_t> codes(row_start[n]);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    uint64_t q = row_start[(size_t)i];
    for (uint32_t k = 0; k < n; ++k) {
      uint32_t lo;
      if (!shell((uint32_t)i, k, &lo)) continue;
      if (lo > 0) codes[q++] = morton_encode((uint32_t)i, 0, k, depth);
      for (uint32_t j = lo; j < H[(size_t)i * n + k]; ++j) codes[q++] = morton_encode((uint32_t)i, j, k, depth);
    }
  }
  radix_sort_u64(codes, 3 * depth);
  std::vector<uint32_t> attr(2 * codes.size());
#pragma omp parallel for schedule(static)
  for (int64_t q = 0; q < (int64_t)codes.size(); ++q) {
    uint32_t i, j, k;
    morton_decode(codes[(size_t)q], depth, &i, &j, &k);
    attr[2 * (size_t)q] = leaf_normal_up();
    attr[2 * (size_t)q + 1] = terrain_color(seed, i, j, k, depth);
  }
  return finish(depth, codes, attr, treelet_bytes, flags, out);
}

extern "C" svo_status svo_scene_build_spheres(uint32_t depth, uint32_t count, float r_min, float r_max, uint64_t seed,
                                              uint32_t treelet_bytes, uint32_t flags, svo_scene** out) {
  if (!out) { set_error("svo_scene_build_spheres: out is null"); return SVO_ERR_ARG; }
  if (depth < 2 || depth > 12 || count == 0 || !(r_min > 0) || !(r_max >= r_min) || r_max > 0.5f) {
    set_error("spheres: bad parameters (depth %u, count %u, r [%g,%g])", depth, count, r_min, r_max);
    return SVO_ERR_ARG;
  }
  const int64_t n = 1ll << depth;
  struct Sphere { double x, y, z, r; };
  std::vector<Sphere> sp(count);
  for (uint32_t s = 0; s < count; ++s) {
    auto u = [&](uint32_t f) { return (double)(mix64(mix64(seed) ^ (4ull * s + f)) >> 11) * (1.0 / 9007199254740992.0); };
    sp[s] = Sphere{u(0), u(1), u(2), r_min + (r_max - r_min) * u(3)};  // centre in [0,1)^3
  }
  // voxel (i,j,k) is filled by sphere S iff its centre lies strictly inside S
  auto in_sphere = [&](const Sphere& S, int64_t i, int64_t j, int64_t k) {
    const double x = (i + 0.5) / n - S.x, y = (j + 0.5) / n - S.y, z = (k + 0.5) / n - S.z;
    return x * x + y * y + z * z < S.r * S.r;
  };
  // uniform grid of sphere indices, cell edge >= 2 r_max
  const int gc = std::max(1, (int)std::floor(1.0 / (2.0 * r_max)));
  std::vector<std::vector<uint32_t>> grid((size_t)gc * gc * gc);
  auto cell = [&](double v) { return std::min(gc - 1, std::max(0, (int)std::floor(v * gc))); };
  for (uint32_t s = 0; s < count; ++s) grid[((size_t)cell(sp[s].x) * gc + cell(sp[s].y)) * gc + cell(sp[s].z)].push_back(s);
  auto filled = [&](int64_t i, int64_t j, int64_t k) {
    if (i < 0 || j < 0 || k < 0 || i >= n || j >= n || k >= n) return false;
    const int cx = cell((i + 0.5) / n), cy = cell((j + 0.5) / n), cz = cell((k + 0.5) / n);
    for (int a = std::max(0, cx - 1); a <= std::min(gc - 1, cx + 1); ++a)
      for (int b = std::max(0, cy - 1); b <= std::min(gc - 1, cy + 1); ++b)
        for (int c = std::max(0, cz - 1); c <= std::min(gc - 1, cz + 1); ++c)
          for (uint32_t s : grid[((size_t)a * gc + b) * gc + c])
            if (in_sphere(sp[s], i, j, k)) return true;
    return false;
  };
  // j-range of the voxels of column (i,k) inside S, exact w.r.t. in_sphere;
  // returns false if the column misses S
  auto column = [&](const Sphere& S, int64_t i, int64_t k, int64_t* lo, int64_t* hi) {
    if (i < 0 || k < 0 || i >= n || k >= n) return false;
    const double dx = (i + 0.5) / n - S.x, dz = (k + 0.5) / n - S.z;
    const double rem = S.r * S.r - dx * dx - dz * dz;
    if (!(rem > 0)) return false;
    const double sq = std::sqrt(rem);
    int64_t a = (int64_t)std::floor((S.y - sq) * n - 0.5), b = (int64_t)std::ceil((S.y + sq) * n - 0.5);
    a = std::max<int64_t>(0, a - 1);
    b = std::min<int64_t>(n - 1, b + 1);
    while (a <= b && !in_sphere(S, i, a, k)) ++a;
    while (b >= a && !in_sphere(S, i, b, k)) --b;
    if (a > b) return false;
    *lo = a;
    *hi = b;
    return true;
  };
  const int nthreads = std::max(1, omp_get_max_threads());
  std::vector<std::vector<uint64_t>> found(nthreads);
#pragma omp parallel num_threads(nthreads)
  {
    std::vector<uint64_t>& local = found[omp_get_thread_num()];
#pragma omp for schedule(dynamic, 1)
    for (int64_t s = 0; s < (int64_t)count; ++s) {
      const Sphere& S = sp[s];
      const int64_t i0 = std::max<int64_t>(0, (int64_t)std::floor((S.x - S.r) * n) - 1);
      const int64_t i1 = std::min<int64_t>(n - 1, (int64_t)std::ceil((S.x + S.r) * n) + 1);
      const int64_t k0 = std::max<int64_t>(0, (int64_t)std::floor((S.z - S.r) * n) - 1);
      const int64_t k1 = std::min<int64_t>(n - 1, (int64_t)std::ceil((S.z + S.r) * n) + 1);
      for (int64_t i = i0; i <= i1; ++i)
        for (int64_t k = k0; k <= k1; ++k) {
          int64_t lo, hi;
          if (!column(S, i, k, &lo, &hi)) continue;
          // voxels of this column whose horizontal neighbours all lie in S
          // are candidates only if a vertical neighbour leaves S (the ends)
          int64_t in_lo = lo + 1, in_hi = hi - 1;
          const int64_t nb[4][2] = {{i - 1, k}, {i + 1, k}, {i, k - 1}, {i, k + 1}};
          for (const auto& q : nb) {
            int64_t a, b;
            if (!column(S, q[0], q[1], &a, &b)) { in_lo = 1; in_hi = 0; break; }
            in_lo = std::max(in_lo, a);
            in_hi = std::min(in_hi, b);
          }
          for (int64_t j = lo; j <= hi; ++j) {
            if (j >= in_lo && j <= in_hi) { j = in_hi; continue; }
            const int64_t nbr[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k}, {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
            bool shell = false;
            for (const auto& v : nbr) {
              const bool inside_cube = v[0] >= 0 && v[1] >= 0 && v[2] >= 0 && v[0] < n && v[1] < n && v[2] < n;
              if (inside_cube && in_sphere(S, v[0], v[1], v[2])) continue;
              if (!filled(v[0], v[1], v[2])) { shell = true; break; }
            }
            if (shell) local.push_back(morton_encode((uint32_t)i, (uint32_t)j, (uint32_t)k, (uint32_t)depth));
          }
        }
    }
  }
  std::vector<uint64_t> codes;
  size_t total = 0;
  for (const auto& v : found) total += v.size();
  codes.reserve(total);
  for (auto& v : found) { codes.insert(codes.end(), v.begin(), v.end()); std::vector<uint64_t>().swap(v); }
  radix_sort_u64(codes, 3 * depth);
  codes.erase(std::unique(codes.begin(), codes.end()), codes.end());
  std::vector<uint32_t> attr(2 * codes.size());
  for (size_t q = 0; q < codes.size(); ++q) {
    uint32_t i, j, k;
    morton_decode(codes[q], depth, &i, &j, &k);
    attr[2 * q] = leaf_normal_up();
    attr[2 * q + 1] = hash_color(seed ^ 0xC0FFEEull, i, j, k);
  }
  return finish(depth, codes, attr, treelet_bytes, flags, out);
}

// ------------------------------------------------------------ camera

extern "C" svo_status svo_camera_look_at(const double eye[3], const double target[3], const double up[3],
                                         double fov_y_deg, uint32_t width, uint32_t height, double far_dist,
                                         double lod_multiplier, svo_camera* out) {
  if (!eye || !target || !up || !out) { set_error("svo_camera_look_at: null argument"); return SVO_ERR_ARG; }
  if (width == 0 || height == 0 || !(fov_y_deg > 0 && fov_y_deg < 180) || !(far_dist > 0)) {
    set_error("svo_camera_look_at: bad frame %ux%u / fov %g / far %g", width, height, fov_y_deg, far_dist);
    return SVO_ERR_ARG;
  }
  double f[3] = {target[0] - eye[0], target[1] - eye[1], target[2] - eye[2]};
  double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
  if (!(fl > 0)) { set_error("svo_camera_look_at: eye == target"); return SVO_ERR_ARG; }
  for (double& v : f) v /= fl;
  double r[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
  double rl = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
  if (!(rl > 1e-12)) { set_error("svo_camera_look_at: up parallel to view direction"); return SVO_ERR_ARG; }
  for (double& v : r) v /= rl;
  const double u[3] = {r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2], r[0] * f[1] - r[1] * f[0]};
  const double aspect = (double)width / (double)height;
  const double th = std::tan(fov_y_deg * M_PI / 360.0);
  const double hh = th * far_dist, hw = hh * aspect;
  for (int a = 0; a < 3; ++a) {
    const double c = eye[a] + f[a] * far_dist;
    out->far_lower_left[a] = (float)(c - r[a] * hw - u[a] * hh);
    out->pixel_h[a] = (float)(2.0 * hw * r[a] / width);
    out->pixel_v[a] = (float)(2.0 * hh * u[a] / height);
    out->position[a] = (float)eye[a];
  }
  out->far_lower_left[3] = out->pixel_h[3] = out->pixel_v[3] = out->position[3] = 0.0f;
  out->t_scale_ratio = (float)(lod_multiplier * th * aspect / (width * 0.5));
  out->width = width;
  out->height = height;
  out->view_mode = SVO_VIEW_COLOR;
  return SVO_OK;
}
