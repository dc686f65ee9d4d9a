// common.h -- shared device helpers and data layouts of the MI355X tree solve.
//
// Wave64 primitives (ballot, lane masks, butterfly reductions), the device
// box record and the per-node records the walk reads.  Everything here is
// written for gfx950 (CDNA4): 64-lane wavefronts, 160 KiB LDS per CU.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pepc {

constexpr int kWave = 64;
constexpr int kMaxLevel = 20;          // (64 - 2) / 3, treevars.f90:371

// Status bits raised by kernels (read back by the host after a phase).
enum : uint32_t {
  kFlagDupKeys = 1u << 0,       // equal adjacent keys after the sort
  kFlagNodeOverflow = 1u << 1,  // node pool too small
  kFlagStackOverflow = 1u << 2, // walk stack overflow (never expected)
  kFlagDepth = 1u << 3,         // more than 20 levels needed
  kFlagBadBox = 1u << 4,        // non-finite or degenerate bounding box
  kFlagGroupOverflow = 1u << 5, // group larger than a wavefront
};

// Bounding box and derived per-level tables, module_box.f90:33-37 plus
// module_tree.f90:196-201 (boxlength2).
struct DevBox {
  double boxmin[3];
  double boxmax[3];
  double boxsize[3];
  double cell[3];               // boxsize / 2^20, the refinement length
  double b2[kMaxLevel + 1];     // maxval(boxsize)^2 / 4^l
};

// Device-wide scalars the host reads once per solve (one small copy).
struct DevScalars {
  uint32_t flags;
  uint32_t nnodes;
  uint32_t ntwig;
  uint32_t ngroups;
  uint32_t dup_scan_n;          // particles the duplicate fix-up scans (0 without duplicates)
  uint32_t maxlevel_used;
  uint32_t level_offset[kMaxLevel + 2];   // first node of every level; [kMaxLevel + 1] = nnodes
  uint32_t twig_offset[kMaxLevel + 2];    // first twig rank of every level; [kMaxLevel + 1] = ntwig
  unsigned long long interactions;
  unsigned long long mac_evals;
  unsigned long long p2p;
  unsigned long long m2p;
  unsigned long long amb_nodes;   // twigs the group box could not decide
  unsigned long long walk_steps;  // traversal steps (batches of up to 64 nodes)
  uint32_t dup_count;
  uint32_t pad;
};

// Node record read by the walk for the acceptance test: centre of charge
// and the squared box length of the node's level (for a leaf, whose centre
// is its particle, the last slot holds the particle's charge instead).
struct __align__(32) NodeGeom {
  double cx, cy, cz, b2;
};

// Topology: first child (twig) or particle index (leaf), child count
// (0 for a leaf), and the number of particles below the node.
struct __align__(16) NodeLink {
  uint32_t first;
  uint32_t nchild;
  uint32_t count;
  uint32_t level;
};

// Moments in the form the M2P kernel evaluates (see walk.hip): charge,
// dipole, doubled diagonal quadrupole, off-diagonal quadrupole, trace.
// Slots: 0 q | 1..3 p | 4..6 2*quad | 7 xy | 8 yz | 9 zx | 10 trace
template <typename R>
struct __align__(16) NodeMoments {
  R v[16];
};

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ unsigned long long lanemask_lt() {
  const int l = __lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}

__device__ __forceinline__ int popc64(unsigned long long m) { return __popcll(m); }

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, kWave));
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
  return v;
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t t = (uint32_t)__shfl_xor((int)v, o, kWave);
    v = t < v ? t : v;
  }
  return v;
}

__device__ __forceinline__ int wave_max_i32(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const int t = __shfl_xor(v, o, kWave);
    v = t > v ? t : v;
  }
  return v;
}

// Inclusive prefix sum across the wave (Hillis-Steele over 6 steps).
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
  const int l = __lane_id();
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    uint32_t t = __shfl_up(v, o, kWave);
    if (l >= o) v += t;
  }
  return v;
}

// Number of set bits of m in the lanes below this one (v_mbcnt_lo/hi).
__device__ __forceinline__ uint32_t mask_rank(unsigned long long m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Exclusive prefix sum across the wave of values below 16, built from four
// ballots (bit planes) instead of lane shuffles; *total gets the wave's sum.
__device__ __forceinline__ uint32_t wave_exclusive_scan16(uint32_t v, uint32_t* total) {
  uint32_t ex = 0, tot = 0;
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const unsigned long long m = __ballot((v >> b) & 1u);
    ex += mask_rank(m) << b;
    tot += (uint32_t)__popcll(m) << b;
  }
  *total = tot;
  return ex;
}

// Level of a key: position of the placeholder bit, module_spacefilling.f90:47-57.
__device__ __forceinline__ int level_of_key(unsigned long long k) {
  return (63 - __clzll((long long)k)) / 3;
}

// First level at which two distinct particle keys fall into different cells.
__device__ __forceinline__ int split_level(unsigned long long a, unsigned long long b) {
  unsigned long long x = a ^ b;
  int msb = 63 - __clzll((long long)x);
  return kMaxLevel - msb / 3;
}

}  // namespace pepc
