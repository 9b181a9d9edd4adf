// sw2d_gen.cpp -- seeded synthetic inputs (fractal DEM, rain fields,
// hyetographs, Manning rasters) for the benchmark configurations.
#include "sw2d_gen.h"

#include <algorithm>
#include <cmath>
#include <vector>

#include "sw2d.h"

namespace {

struct Rng {  // xoshiro256** seeded through splitmix64
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; i++) {
      x += 0x9E3779B97F4A7C15ull;
      uint64_t z = x;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      s[i] = z ^ (z >> 31);
    }
  }
  static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0,1)
  double normal() {  // Box-Muller, one value per call
    double u1 = uniform(), u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
  double gamma(double k) {
    if (k < 1.0) {
      const double u = uniform();
      return gamma(k + 1.0) * std::pow(u > 0 ? u : 1e-300, 1.0 / k);
    }
    const double d = k - 1.0 / 3.0, c = 1.0 / std::sqrt(9.0 * d);
    for (;;) {
      double x, v;
      do {
        x = normal();
        v = 1.0 + c * x;
      } while (v <= 0.0);
      v = v * v * v;
      const double u = uniform();
      if (u < 1.0 - 0.0331 * x * x * x * x) return d * v;
      if (std::log(u) < 0.5 * x * x + d * (1.0 - v + std::log(v))) return d * v;
    }
  }
};

}  // namespace

extern "C" {

int sw2d_gen_fractal_dem(int rows, int cols, uint64_t seed, double roughness, double zmin, double zmax,
                         double *out) {
  if (rows <= 0 || cols <= 0 || !out || !(roughness > 0) || !(zmax >= zmin)) return SW2D_E_INVALID;
  int k = 0;
  while ((1 << k) + 1 < std::max(rows, cols)) k++;
  const int n = (1 << k) + 1;
  if ((size_t)n * n > (size_t)1 << 34) return SW2D_E_INVALID;
  std::vector<float> a((size_t)n * n, 0.0f);
  Rng r(seed);
  auto A = [&](int i, int j) -> float & { return a[(size_t)i * n + j]; };
  A(0, 0) = (float)r.uniform();
  A(0, n - 1) = (float)r.uniform();
  A(n - 1, 0) = (float)r.uniform();
  A(n - 1, n - 1) = (float)r.uniform();
  double amp = 1.0;
  for (int step = n - 1; step > 1; step /= 2) {
    const int half = step / 2;
    for (int i = half; i < n; i += step)  // diamond: square centres
      for (int j = half; j < n; j += step) {
        const double avg = 0.25 * ((double)A(i - half, j - half) + A(i - half, j + half) + A(i + half, j - half) +
                                   A(i + half, j + half));
        A(i, j) = (float)(avg + (2.0 * r.uniform() - 1.0) * amp);
      }
    for (int i = 0; i < n; i += half)  // square: edge midpoints
      for (int j = ((i / half) % 2 == 0) ? half : 0; j < n; j += step) {
        double s = 0.0;
        int m = 0;
        if (i - half >= 0) { s += A(i - half, j); m++; }
        if (i + half < n) { s += A(i + half, j); m++; }
        if (j - half >= 0) { s += A(i, j - half); m++; }
        if (j + half < n) { s += A(i, j + half); m++; }
        A(i, j) = (float)(s / m + (2.0 * r.uniform() - 1.0) * amp);
      }
    amp *= roughness;
  }
  double lo = 1e300, hi = -1e300;
  for (int i = 0; i < rows; i++)
    for (int j = 0; j < cols; j++) {
      lo = std::min(lo, (double)A(i, j));
      hi = std::max(hi, (double)A(i, j));
    }
  const double sc = hi > lo ? (zmax - zmin) / (hi - lo) : 0.0;
  for (int i = 0; i < rows; i++)
    for (int j = 0; j < cols; j++) out[(size_t)i * cols + j] = zmin + ((double)A(i, j) - lo) * sc;
  return SW2D_OK;
}

int sw2d_gen_tilt(int rows, int cols, double dx, double slope, int edge, double *z) {
  if (rows <= 0 || cols <= 0 || !z) return SW2D_E_INVALID;
  for (int i = 0; i < rows; i++)
    for (int j = 0; j < cols; j++) {
      double d;
      switch (edge) {
        case SW2D_EDGE_NORTH: d = i; break;
        case SW2D_EDGE_SOUTH: d = rows - 1 - i; break;
        case SW2D_EDGE_WEST: d = j; break;
        case SW2D_EDGE_EAST: d = cols - 1 - j; break;
        default: return SW2D_E_INVALID;
      }
      z[(size_t)i * cols + j] += slope * dx * d;
    }
  return SW2D_OK;
}

int sw2d_gen_uniform(size_t n, uint64_t seed, double lo, double hi, double *out) {
  if (n && !out) return SW2D_E_INVALID;
  Rng r(seed);
  for (size_t k = 0; k < n; k++) out[k] = lo + (hi - lo) * r.uniform();
  return SW2D_OK;
}

int sw2d_gen_lognormal(size_t n, uint64_t seed, double mu, double sigma, double *out) {
  if (n && !out) return SW2D_E_INVALID;
  Rng r(seed);
  for (size_t k = 0; k < n; k++) out[k] = std::exp(mu + sigma * r.normal());
  return SW2D_OK;
}

int sw2d_gen_gamma(size_t n, uint64_t seed, double k, double theta, double *out) {
  if ((n && !out) || !(k > 0) || !(theta > 0)) return SW2D_E_INVALID;
  Rng r(seed);
  for (size_t i = 0; i < n; i++) out[i] = theta * r.gamma(k);
  return SW2D_OK;
}

}  // extern "C"
