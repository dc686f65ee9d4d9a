// keys.hip -- bounding box and 64-bit Hilbert keys on the device.
//
// box:  module_box.f90:48-93 (min/max, margin boxsize/10000 on both sides)
// keys: module_spacefilling.f90:224-257 (s = boxsize/2^20, truncating
//       conversion) and :380-440 (Kamata et al. 1999 3-D Hilbert map, 20
//       levels of 3 bits below a placeholder bit at bit 60).
// The arithmetic is the reference's FP64 arithmetic with contraction off, so
// keys are bit-identical to the reference's for the same coordinates.
#include "kernels.h"

namespace pepc {

namespace {

constexpr int kBoxThreads = 256;

__global__ __launch_bounds__(kBoxThreads) void box_partial_kernel(
    int64_t n, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, double* __restrict__ partial, DevScalars* __restrict__ sc) {
  double mn[3] = {INFINITY, INFINITY, INFINITY};
  double mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  bool nonfinite = false;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double v0 = x[i], v1 = y[i], v2 = z[i];
    nonfinite |= !isfinite(v0) || !isfinite(v1) || !isfinite(v2);
    mn[0] = fmin(mn[0], v0); mx[0] = fmax(mx[0], v0);
    mn[1] = fmin(mn[1], v1); mx[1] = fmax(mx[1], v1);
    mn[2] = fmin(mn[2], v2); mx[2] = fmax(mx[2], v2);
  }
  if (__ballot(nonfinite) != 0ull && lane_id() == 0) atomicOr(&sc->flags, kFlagBadBox);
  __shared__ double red[6][kBoxThreads / kWave];
  const int w = threadIdx.x / kWave;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double a = wave_min(mn[d]), b = wave_max(mx[d]);
    if (lane_id() == 0) { red[d][w] = a; red[3 + d][w] = b; }
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int d = threadIdx.x;
    double v = red[d][0];
    for (int k = 1; k < kBoxThreads / kWave; ++k) v = d < 3 ? fmin(v, red[d][k]) : fmax(v, red[d][k]);
    partial[blockIdx.x * 6 + d] = v;
  }
}

// module_box.f90:65-93 from the coordinate extrema: optional cube, margin
// boxsize/10000 on both sides, refinement cell, boxlength2 table.
__device__ void box_finalize(double mn[3], double mx[3], int force_cubic, DevBox* __restrict__ box,
                             DevScalars* __restrict__ sc) {
  if (force_cubic) {  // module_box.f90:71-74
    double lo = fmin(fmin(mn[0], mn[1]), mn[2]);
    double hi = fmax(fmax(mx[0], mx[1]), mx[2]);
    for (int d = 0; d < 3; ++d) { mn[d] = lo; mx[d] = hi; }
  }
  bool bad = false;
  for (int d = 0; d < 3; ++d) {
    double size0 = mx[d] - mn[d];
    double lo = mn[d] - size0 / 10000.0;
    double hi = mx[d] + size0 / 10000.0;
    box->boxmin[d] = lo;
    box->boxmax[d] = hi;
    box->boxsize[d] = hi - lo;
    box->cell[d] = box->boxsize[d] / 1048576.0;
    if (!isfinite(lo) || !isfinite(hi) || !(box->boxsize[d] > 0.0)) bad = true;
  }
  double m = fmax(fmax(box->boxsize[0], box->boxsize[1]), box->boxsize[2]);
  box->b2[0] = m * m;
  for (int l = 1; l <= kMaxLevel; ++l) box->b2[l] = box->b2[l - 1] / 4.0;
  if (bad) atomicOr(&sc->flags, kFlagBadBox);
}

__global__ __launch_bounds__(kWave) void box_final_kernel(int nblocks, const double* __restrict__ partial,
                                                         int force_cubic, DevBox* __restrict__ box,
                                                         DevScalars* __restrict__ sc) {
  double mn[3] = {INFINITY, INFINITY, INFINITY};
  double mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int b = threadIdx.x; b < nblocks; b += kWave) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      mn[d] = fmin(mn[d], partial[b * 6 + d]);
      mx[d] = fmax(mx[d], partial[b * 6 + 3 + d]);
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) { mn[d] = wave_min(mn[d]); mx[d] = wave_max(mx[d]); }
  if (threadIdx.x != 0) return;
  box_finalize(mn, mx, force_cubic, box, sc);
}

// Batched systems: one block per system computes that system's box over its
// particle range [soff[s], soff[s+1]) and tags the range with the system
// index.  An empty system gets a unit box that no particle refers to.
__global__ __launch_bounds__(kBoxThreads) void box_segment_kernel(
    int nsys, const int64_t* __restrict__ soff, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ z, int force_cubic,
    DevBox* __restrict__ boxes, uint32_t* __restrict__ sys, DevScalars* __restrict__ sc) {
  const int sid = blockIdx.x;
  const int64_t b = soff[sid], e = soff[sid + 1];
  double mn[3] = {INFINITY, INFINITY, INFINITY};
  double mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  bool nonfinite = false;
  for (int64_t i = b + threadIdx.x; i < e; i += blockDim.x) {
    double v0 = x[i], v1 = y[i], v2 = z[i];
    nonfinite |= !isfinite(v0) || !isfinite(v1) || !isfinite(v2);
    mn[0] = fmin(mn[0], v0); mx[0] = fmax(mx[0], v0);
    mn[1] = fmin(mn[1], v1); mx[1] = fmax(mx[1], v1);
    mn[2] = fmin(mn[2], v2); mx[2] = fmax(mx[2], v2);
    sys[i] = (uint32_t)sid;
  }
  if (__ballot(nonfinite) != 0ull && lane_id() == 0) atomicOr(&sc->flags, kFlagBadBox);
  __shared__ double red[6][kBoxThreads / kWave];
  const int w = threadIdx.x / kWave;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double lo = wave_min(mn[d]), hi = wave_max(mx[d]);
    if (lane_id() == 0) { red[d][w] = lo; red[3 + d][w] = hi; }
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int d = 0; d < 3; ++d) {
    mn[d] = red[d][0];
    mx[d] = red[3 + d][0];
    for (int k = 1; k < kBoxThreads / kWave; ++k) {
      mn[d] = fmin(mn[d], red[d][k]);
      mx[d] = fmax(mx[d], red[3 + d][k]);
    }
  }
  if (e <= b) {
    for (int d = 0; d < 3; ++d) { mn[d] = 0.0; mx[d] = 1.0; }
  }
  box_finalize(mn, mx, force_cubic, boxes + sid, sc);
}

// Hilbert tables packed 3 bits per entry (module_spacefilling.f90:389-390):
// CI3 = [0,1,3,2,7,6,4,5], gene exchange = [5,6,0,5,5,0,6,5], reverse =
// [0,0,0,5,0,0,6,5].
constexpr uint32_t pack8(int a0, int a1, int a2, int a3, int a4, int a5, int a6, int a7) {
  return (uint32_t)a0 | (uint32_t)a1 << 3 | (uint32_t)a2 << 6 | (uint32_t)a3 << 9 |
         (uint32_t)a4 << 12 | (uint32_t)a5 << 15 | (uint32_t)a6 << 18 | (uint32_t)a7 << 21;
}
constexpr uint32_t kCI3 = pack8(0, 1, 3, 2, 7, 6, 4, 5);
constexpr uint32_t kGeneX = pack8(5, 6, 0, 5, 5, 0, 6, 5);
constexpr uint32_t kGeneR = pack8(0, 0, 0, 5, 0, 0, 6, 5);

__device__ __forceinline__ unsigned long long hilbert3d(uint32_t c0, uint32_t c1, uint32_t c2) {
  unsigned long long key = 1ull;
#pragma unroll
  for (int i = kMaxLevel - 1; i >= 0; --i) {
    const uint32_t bit = 1u << i;
    uint32_t cval = ((c2 >> i) & 1u) << 2 | ((c1 >> i) & 1u) << 1 | ((c0 >> i) & 1u);
    c0 &= ~bit; c1 &= ~bit; c2 &= ~bit;
    const uint32_t h = (kCI3 >> (3 * cval)) & 7u;
    const uint32_t ex = (kGeneX >> (3 * h)) & 7u;
    const uint32_t rv = (kGeneR >> (3 * h)) & 7u;
    key = (key << 3) | h;
    uint32_t t;
    if (ex == 5u) { t = c0; c0 = c2; c2 = t; }
    else if (ex == 6u) { t = c1; c1 = c2; c2 = t; }
    const uint32_t mask = bit - 1u;
    if (rv & 1u) c0 = ~c0 & mask;
    if (rv & 2u) c1 = ~c1 & mask;
    if (rv & 4u) c2 = ~c2 & mask;
  }
  return key;
}

__global__ __launch_bounds__(256) void keys_kernel(int64_t n, const double* __restrict__ x,
                                                   const double* __restrict__ y,
                                                   const double* __restrict__ z,
                                                   const double* __restrict__ q,
                                                   const DevBox* __restrict__ boxes,
                                                   const uint32_t* __restrict__ sys,
                                                   unsigned long long* __restrict__ keys,
                                                   uint32_t* __restrict__ idx,
                                                   double4* __restrict__ rec) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  // one 32-byte record per particle, so the gather after the sort fetches a
  // particle with one random access instead of four
  rec[i] = make_double4(x[i], y[i], z[i], q[i]);
  const DevBox* __restrict__ box = boxes + (sys ? sys[i] : 0u);
  // int((x - boxmin) / s): the FP64 quotient is correctly rounded and
  // truncated toward zero, exactly as the Fortran int() conversion.
  const long long i0 = (long long)((x[i] - box->boxmin[0]) / box->cell[0]);
  const long long i1 = (long long)((y[i] - box->boxmin[1]) / box->cell[1]);
  const long long i2 = (long long)((z[i] - box->boxmin[2]) / box->cell[2]);
  keys[i] = hilbert3d((uint32_t)i0, (uint32_t)i1, (uint32_t)i2);
  idx[i] = (uint32_t)i;
}

// Points that need not lie inside the box (sinks): each coordinate clamped to
// [boxmin, boxmax] and the cell index to 2^20 - 1 before the same key map.
__global__ __launch_bounds__(256) void sink_keys_kernel(int64_t m, const double* __restrict__ x,
                                                        const double* __restrict__ y,
                                                        const double* __restrict__ z,
                                                        const DevBox* __restrict__ box,
                                                        unsigned long long* __restrict__ keys,
                                                        uint32_t* __restrict__ idx) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= m) return;
  const double p[3] = {x[i], y[i], z[i]};
  uint32_t c[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double v = fmin(fmax(p[d], box->boxmin[d]), box->boxmax[d]);
    long long k = (long long)((v - box->boxmin[d]) / box->cell[d]);
    k = k < 0 ? 0 : (k > (1ll << kMaxLevel) - 1 ? (1ll << kMaxLevel) - 1 : k);
    c[d] = (uint32_t)k;
  }
  keys[i] = hilbert3d(c[0], c[1], c[2]);
  idx[i] = (uint32_t)i;
}

__global__ void sink_groups_kernel(int64_t m, int g, uint32_t* __restrict__ gfirst,
                                   uint32_t* __restrict__ gcount) {
  const int64_t ng = (m + g - 1) / g;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < ng;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t f = k * g;
    gfirst[k] = (uint32_t)f;
    gcount[k] = (uint32_t)(m - f < g ? m - f : g);
  }
}

}  // namespace

void launch_sink_keys(hipStream_t s, int64_t m, const double* x, const double* y, const double* z,
                      const DevBox* box, unsigned long long* keys, uint32_t* idx) {
  if (m <= 0) return;
  sink_keys_kernel<<<(unsigned)((m + 255) / 256), 256, 0, s>>>(m, x, y, z, box, keys, idx);
}

void launch_sink_groups(hipStream_t s, int64_t m, int g, uint32_t* gfirst, uint32_t* gcount) {
  const int64_t ng = (m + g - 1) / g;
  int64_t b = (ng + 255) / 256;
  b = b < 1 ? 1 : (b > 4096 ? 4096 : b);
  sink_groups_kernel<<<(unsigned)b, 256, 0, s>>>(m, g, gfirst, gcount);
}

int box_partial_blocks(int64_t n) {
  int64_t b = (n + kBoxThreads * 16 - 1) / (kBoxThreads * 16);
  return (int)(b < 1 ? 1 : (b > 1024 ? 1024 : b));
}

void launch_box(hipStream_t s, int64_t n, const double* x, const double* y, const double* z,
                int force_cubic, double* partial, DevBox* box, DevScalars* sc) {
  const int nb = box_partial_blocks(n);
  box_partial_kernel<<<nb, kBoxThreads, 0, s>>>(n, x, y, z, partial, sc);
  box_final_kernel<<<1, kWave, 0, s>>>(nb, partial, force_cubic, box, sc);
}

void launch_box_batch(hipStream_t s, int nsys, const int64_t* soff, const double* x,
                      const double* y, const double* z, int force_cubic, DevBox* boxes,
                      uint32_t* sys, DevScalars* sc) {
  box_segment_kernel<<<nsys, kBoxThreads, 0, s>>>(nsys, soff, x, y, z, force_cubic, boxes, sys, sc);
}

void launch_keys(hipStream_t s, int64_t n, const double* x, const double* y, const double* z,
                 const double* q, const DevBox* boxes, const uint32_t* sys,
                 unsigned long long* keys, uint32_t* idx, double4* rec) {
  keys_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, x, y, z, q, boxes, sys, keys, idx, rec);
}

}  // namespace pepc
