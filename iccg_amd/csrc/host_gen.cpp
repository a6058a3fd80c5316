// host_gen.cpp -- synthetic SPD problems for the solve path (no reference
// counterpart: the reference reads SuiteSparse files and hard-codes b = 1).
//
// Deterministic on the host: splitmix64 stream -> 53-bit uniforms in [0,1);
// normals by Box-Muller in FP64 with glibc log/sqrt/cos/sin.  The CPU oracle
// and the GPU path receive the same bits.
#include <cmath>
#include <cstdint>
#include <vector>

#include "internal.hpp"

namespace iccg {

struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};

}  // namespace iccg

using namespace iccg;

extern "C" {

int iccg_gen_grid(int nx, int ny, int nz, double lo, double hi, double shift,
                  unsigned long long seed, int *n_out, long *nnz_out, int *rowptr, int *colind,
                  double *val) {
  return guard([&] {
    require(nx >= 1 && ny >= 1 && nz >= 1, ICCG_ERR_ARG, "grid sides must be positive");
    const long n = (long)nx * ny * nz;
    require(n <= 0x7fffffffL, ICCG_ERR_ARG, "grid too large for int32 indices");
    const bool hy = ny > 1, hz = nz > 1;
    const long nnz = n + 2L * ((long)(nx - 1) * ny * nz + (hy ? (long)nx * (ny - 1) * nz : 0) +
                               (hz ? (long)nx * ny * (nz - 1) : 0));
    require(nnz <= 0x7fffffffL, ICCG_ERR_ARG, "too many nonzeros for int32 indices");
    *n_out = (int)n;
    *nnz_out = nnz;
    if (!rowptr) return;
    // face coefficients, drawn x faces, y faces, z faces (x fastest in each)
    const bool vary = lo != hi;
    const double cconst = std::pow(10.0, lo);
    std::vector<double> cx, cy, cz;
    SplitMix64 rng(seed);
    auto draw = [&](std::vector<double> &c, long m) {
      c.resize((size_t)m);
      for (long t = 0; t < m; ++t)
        c[t] = vary ? std::pow(10.0, lo + (hi - lo) * rng.uniform()) : cconst;
    };
    draw(cx, (long)(nx + 1) * ny * nz);
    if (hy) draw(cy, (long)nx * (ny + 1) * nz);
    if (hz) draw(cz, (long)nx * ny * (nz + 1));
    auto CX = [&](int f, int j, int k) { return cx[(size_t)f + (size_t)(nx + 1) * ((size_t)j + (size_t)ny * k)]; };
    auto CY = [&](int i, int f, int k) { return cy[(size_t)i + (size_t)nx * ((size_t)f + (size_t)(ny + 1) * k)]; };
    auto CZ = [&](int i, int j, int f) { return cz[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * f)]; };
    long pos = 0;
    rowptr[0] = 0;
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          const long row = (long)i + (long)nx * ((long)j + (long)ny * k);
          double d = CX(i, j, k) + CX(i + 1, j, k);
          if (hy) d = (d + CY(i, j, k)) + CY(i, j + 1, k);
          if (hz) d = (d + CZ(i, j, k)) + CZ(i, j, k + 1);
          d = d + shift;
          if (hz && k > 0) { colind[pos] = (int)(row - (long)nx * ny); val[pos++] = -CZ(i, j, k); }
          if (hy && j > 0) { colind[pos] = (int)(row - nx); val[pos++] = -CY(i, j, k); }
          if (i > 0) { colind[pos] = (int)(row - 1); val[pos++] = -CX(i, j, k); }
          colind[pos] = (int)row;
          val[pos++] = d;
          if (i < nx - 1) { colind[pos] = (int)(row + 1); val[pos++] = -CX(i + 1, j, k); }
          if (hy && j < ny - 1) { colind[pos] = (int)(row + nx); val[pos++] = -CY(i, j + 1, k); }
          if (hz && k < nz - 1) { colind[pos] = (int)(row + (long)nx * ny); val[pos++] = -CZ(i, j, k + 1); }
          rowptr[row + 1] = (int)pos;
        }
    require(pos == nnz, ICCG_ERR_STATE, "grid generator size mismatch");
  });
}

int iccg_gen_normal(long n, unsigned long long seed, double *out) {
  return guard([&] {
    require(n >= 0 && (n == 0 || out), ICCG_ERR_ARG, "bad argument");
    SplitMix64 rng(seed);
    const double two_pi = 6.283185307179586;
    for (long m = 0; m < n; m += 2) {
      double u1 = rng.uniform(), u2 = rng.uniform();
      double r = std::sqrt(-2.0 * std::log(1.0 - u1));
      double t = two_pi * u2;
      out[m] = r * std::cos(t);
      if (m + 1 < n) out[m + 1] = r * std::sin(t);
    }
  });
}

}  // extern "C"
