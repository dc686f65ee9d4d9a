// sort.hip -- device-wide exclusive scan and stable LSD radix sort.
//
// The sort orders (64-bit key, 32-bit value) pairs.  Each pass handles one
// digit of up to 8 bits in three kernels: per-block digit histogram, one
// exclusive scan of the digit-major histogram table, and a stable scatter.
// The scatter ranks equal digits inside a wavefront with ballots (one ballot
// per digit bit gives every lane the mask of lanes holding its digit) and
// keeps per-wave running counts in LDS, so ranks follow input order and the
// sort is stable.  Particle keys need 60 bits (the placeholder bit 60 is the
// same for every key): seven 8-bit passes and one 4-bit pass.
//
// Replaces the single-rank presort of slsort_keys (sl_pepc mpi_pepckeys.c:167,
// sort_radix.c:33-151): for distinct keys both give the same permutation.
#include "kernels.h"

namespace pepc {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;   // 4096

// ---------------------------------------------------------------- scan
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* total) {
  __shared__ uint32_t wsum[kScanThreads / kWave];
  const int w = threadIdx.x / kWave;
  uint32_t inc = wave_inclusive_scan(v);
  if (lane_id() == kWave - 1) wsum[w] = inc;
  __syncthreads();
  uint32_t off = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < kScanThreads / kWave; ++k) {
    uint32_t t = wsum[k];
    off += (k < w) ? t : 0u;
    tot += t;
  }
  __syncthreads();
  *total = tot;
  return off + inc - v;
}

// With ndev != nullptr the scanned length is min(n, *ndev), read on the
// device (n only sizes the grid); entries past it are neither read nor written.
__device__ __forceinline__ int64_t scan_len(int64_t n, const uint32_t* ndev) {
  return (ndev && (int64_t)*ndev < n) ? (int64_t)*ndev : n;
}

__global__ __launch_bounds__(kScanThreads) void scan_reduce_kernel(const uint32_t* __restrict__ in,
                                                                    int64_t n,
                                                                    uint32_t* __restrict__ bsum,
                                                                    const uint32_t* __restrict__ ndev,
                                                                    bool flags) {
  n = scan_len(n, ndev);
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  uint32_t s = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    int64_t e = base + j * kScanThreads + threadIdx.x;
    if (e < n) s += flags ? (in[e] != 0u ? 1u : 0u) : in[e];
  }
  uint32_t tot;
  (void)block_exclusive_scan(s, &tot);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// Exclusive scan of up to kScanTile * 4 block sums in one block; writes the
// grand total to *total when total != nullptr.
__global__ __launch_bounds__(kScanThreads) void scan_small_kernel(uint32_t* __restrict__ v, int nb,
                                                                   uint32_t* __restrict__ total) {
  const int per = (nb + kScanThreads - 1) / kScanThreads;
  const int b0 = threadIdx.x * per;
  uint32_t s = 0;
  for (int k = 0; k < per; ++k)
    if (b0 + k < nb) s += v[b0 + k];
  uint32_t tot;
  uint32_t run = block_exclusive_scan(s, &tot);
  for (int k = 0; k < per; ++k) {
    if (b0 + k < nb) {
      uint32_t t = v[b0 + k];
      v[b0 + k] = run;
      run += t;
    }
  }
  if (threadIdx.x == 0 && total) *total = tot;
}

__global__ __launch_bounds__(kScanThreads) void scan_apply_kernel(const uint32_t* __restrict__ in,
                                                                   uint32_t* __restrict__ out,
                                                                   int64_t n,
                                                                   const uint32_t* __restrict__ bsum,
                                                                   const uint32_t* __restrict__ ndev,
                                                                   bool flags) {
  __shared__ uint32_t tile[kScanTile];
  n = scan_len(n, ndev);
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  if (base >= n) return;   // uniform over the block
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    int64_t e = base + j * kScanThreads + threadIdx.x;
    const uint32_t v = e < n ? in[e] : 0u;
    tile[j * kScanThreads + threadIdx.x] = flags ? (v != 0u ? 1u : 0u) : v;
  }
  __syncthreads();
  uint32_t loc[kScanItems];
  uint32_t s = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    loc[j] = s;
    s += tile[threadIdx.x * kScanItems + j];
  }
  uint32_t tot;
  uint32_t off = block_exclusive_scan(s, &tot) + bsum[blockIdx.x];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) tile[threadIdx.x * kScanItems + j] = off + loc[j];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    int64_t e = base + j * kScanThreads + threadIdx.x;
    if (e < n) out[e] = tile[j * kScanThreads + threadIdx.x];
  }
}

// ---------------------------------------------------------------- radix sort
constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;
constexpr int kSortWaves = kSortThreads / kWave;
constexpr int kSortTile = kSortThreads * kSortItems;   // 2048
constexpr int kRadix = 256;
constexpr int kScanHereBlocks = 64;   // up to 131,072 elements: no separate digit scan

// The digit of element e is bits [shift, shift+nbits) of its key, or, when
// dsrc is given, of dsrc[value] (used to order by system after the keys).
__device__ __forceinline__ uint32_t radix_digit(unsigned long long key, uint32_t val,
                                                const uint32_t* dsrc, int shift, uint32_t mask) {
  return dsrc ? (dsrc[val] >> shift) & mask : (uint32_t)(key >> shift) & mask;
}

__global__ __launch_bounds__(kSortThreads) void radix_hist_kernel(
    const unsigned long long* __restrict__ keys, const uint32_t* __restrict__ vals,
    const uint32_t* __restrict__ dsrc, int64_t n, const uint32_t* __restrict__ ndev,
    int shift, uint32_t mask, uint32_t* __restrict__ hist, int nblocks) {
  if (ndev && (int64_t)*ndev < n) n = *ndev;
  __shared__ uint32_t cnt[kRadix];
  cnt[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kSortTile;
#pragma unroll 4
  for (int j = 0; j < kSortItems; ++j) {
    int64_t e = base + j * kSortThreads + threadIdx.x;
    if (e < n) atomicAdd(&cnt[radix_digit(keys[e], dsrc ? vals[e] : 0u, dsrc, shift, mask)], 1u);
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * nblocks + blockIdx.x] = cnt[threadIdx.x];
}

// Ranks each element within its block (stable), stages the block's elements
// in LDS in digit order, then writes them out so that consecutive threads
// store consecutive positions of one digit's run (coalesced stores).
// hscan is the digit-major table of per-block counts: exclusive-scanned along
// each digit's row by digit_scan_kernel, with the row totals in dtot, or,
// with scan_here, the histogram itself (dtot unused).
__global__ __launch_bounds__(kSortThreads) void radix_scatter_kernel(
    const unsigned long long* __restrict__ kin, const uint32_t* __restrict__ vin,
    unsigned long long* __restrict__ kout, uint32_t* __restrict__ vout,
    const uint32_t* __restrict__ dsrc, int64_t n,
    const uint32_t* __restrict__ ndev, int shift, int nbits, const uint32_t* __restrict__ hscan,
    const uint32_t* __restrict__ dtot, int nblocks, bool scan_here) {
  if (ndev && (int64_t)*ndev < n) n = *ndev;
  const int64_t tile0 = (int64_t)blockIdx.x * kSortTile;
  if (tile0 >= n) return;
  __shared__ uint32_t wcnt[kSortWaves][kRadix];
  __shared__ uint32_t lbase[kRadix];
  __shared__ uint32_t gbase[kRadix];
  __shared__ uint32_t bbase[kRadix];   // digit d's elements in the blocks before this one
  __shared__ unsigned long long skey[kSortTile];
  __shared__ uint32_t sval[kSortTile];
  __shared__ uint8_t sdig[kSortTile];
  const int w = threadIdx.x / kWave;
  const uint32_t mask = (1u << nbits) - 1u;
#pragma unroll
  for (int k = 0; k < kSortWaves; ++k) wcnt[k][threadIdx.x] = 0;
  __syncthreads();

  const int64_t base = tile0 + (int64_t)w * (kSortItems * kWave);
  unsigned long long kk[kSortItems];
  uint32_t vv[kSortItems];
  uint32_t dig[kSortItems];
  uint32_t loc[kSortItems];
  const unsigned long long lt = lanemask_lt();
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const int64_t e = base + j * kWave + lane_id();
    const bool valid = e < n;
    kk[j] = valid ? kin[e] : 0ull;
    vv[j] = valid ? vin[e] : 0u;
    const uint32_t d = valid ? radix_digit(kk[j], vv[j], dsrc, shift, mask) : 0u;
    dig[j] = d;
    unsigned long long peers = __ballot(valid);
    for (int b = 0; b < nbits; ++b) {
      const bool bit = (d >> b) & 1u;
      const unsigned long long m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const uint32_t rank = (uint32_t)popc64(peers & lt);
    uint32_t c = 0;
    if (valid) c = wcnt[w][d];
    __builtin_amdgcn_wave_barrier();
    if (valid && rank == 0) wcnt[w][d] = c + (uint32_t)popc64(peers);
    __builtin_amdgcn_wave_barrier();
    loc[j] = c + rank;
  }
  __syncthreads();
  {
    // thread d: offsets of digit d per wave, and the block-local start of d
    const uint32_t d = threadIdx.x;
    uint32_t run = 0;
#pragma unroll
    for (int k = 0; k < kSortWaves; ++k) {
      const uint32_t t = wcnt[k][d];
      wcnt[k][d] = run;
      run += t;
    }
    uint32_t tot;
    lbase[d] = block_exclusive_scan(run, &tot);
    // digit d's total and its count in earlier blocks: from the digit scan,
    // or, for a few blocks (scan_here), from the histogram row of digit d
    uint32_t all, before;
    if (scan_here) {
      const uint32_t* row = hscan + (int64_t)d * nblocks;
      all = 0;
      before = 0;
      for (int b = 0; b < nblocks; ++b) {
        const uint32_t v = row[b];
        before += b < (int)blockIdx.x ? v : 0u;
        all += v;
      }
    } else {
      all = dtot[d];
      before = hscan[(int64_t)d * nblocks + blockIdx.x];
    }
    bbase[d] = before;
    // global start of digit d: digits below d over all blocks
    gbase[d] = block_exclusive_scan(all, &tot);
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const int64_t e = base + j * kWave + lane_id();
    if (e < n) {
      const uint32_t lp = lbase[dig[j]] + wcnt[w][dig[j]] + loc[j];
      skey[lp] = kk[j];
      sval[lp] = vv[j];
      sdig[lp] = (uint8_t)dig[j];
    }
  }
  __syncthreads();
  const int64_t rem = n - tile0;
  const int cnt = rem < kSortTile ? (int)rem : kSortTile;
  for (int i = threadIdx.x; i < cnt; i += kSortThreads) {
    const uint32_t d = sdig[i];
    const uint32_t pos = gbase[d] + bbase[d] + ((uint32_t)i - lbase[d]);
    kout[pos] = skey[i];
    vout[pos] = sval[i];
  }
}

// Exclusive scan of one digit's row of the block histogram (block b's count
// of digit d -> number of digit-d elements in blocks before b), row total to
// dtot[d].  One block per digit.
__global__ __launch_bounds__(kSortThreads) void digit_scan_kernel(uint32_t* __restrict__ hist,
                                                                 int nblocks,
                                                                 uint32_t* __restrict__ dtot) {
  uint32_t* row = hist + (int64_t)blockIdx.x * nblocks;
  const int per = (nblocks + kSortThreads - 1) / kSortThreads;
  const int b0 = threadIdx.x * per;
  uint32_t sum = 0;
  for (int k = 0; k < per; ++k)
    if (b0 + k < nblocks) sum += row[b0 + k];
  uint32_t tot;
  uint32_t run = block_exclusive_scan(sum, &tot);
  for (int k = 0; k < per; ++k) {
    if (b0 + k < nblocks) {
      const uint32_t t = row[b0 + k];
      row[b0 + k] = run;
      run += t;
    }
  }
  if (threadIdx.x == 0) dtot[blockIdx.x] = tot;
}

__global__ __launch_bounds__(256) void gather4_kernel(int64_t n, const uint32_t* __restrict__ perm,
                                                      const double* __restrict__ x,
                                                      const double* __restrict__ y,
                                                      const double* __restrict__ z,
                                                      const double* __restrict__ q,
                                                      double* __restrict__ xs, double* __restrict__ ys,
                                                      double* __restrict__ zs, double* __restrict__ qs) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t p = perm[i];
  xs[i] = x[p];
  ys[i] = y[p];
  zs[i] = z[p];
  qs[i] = q[p];
}

__global__ __launch_bounds__(256) void copy_pairs_kernel(int64_t n,
                                                         const unsigned long long* __restrict__ kin,
                                                         const uint32_t* __restrict__ vin,
                                                         unsigned long long* __restrict__ kout,
                                                         uint32_t* __restrict__ vout) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  kout[i] = kin[i];
  vout[i] = vin[i];
}

void launch_copy_pairs(hipStream_t s, const unsigned long long* kin, const uint32_t* vin,
                       unsigned long long* kout, uint32_t* vout, int64_t n) {
  copy_pairs_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, kin, vin, kout, vout);
}

__global__ __launch_bounds__(256) void gather_rec_kernel(int64_t n, const uint32_t* __restrict__ perm,
                                                        const double4* __restrict__ rec,
                                                        double* __restrict__ xs, double* __restrict__ ys,
                                                        double* __restrict__ zs, double* __restrict__ qs) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 r = rec[perm[i]];
  xs[i] = r.x;
  ys[i] = r.y;
  zs[i] = r.z;
  qs[i] = r.w;
}

// Inverse of the grow permutation: out[perm[i]] = in[i].
__global__ __launch_bounds__(256) void scatter_back_kernel(int64_t n, const uint32_t* __restrict__ perm,
                                                          const double* __restrict__ in,
                                                          double* __restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[perm[i]] = in[i];
}

// ---------------------------------------------------------------- tie order
// Order of equal keys as the reference's local presort leaves it.
//
// sort_radix_ip / rs_rec_ip (sl_pepc sort_radix.c:33-151) is an in-place MSD
// radix over 10-bit digits from bit 63 down.  A pass over a segment moves
// every element whose digit class k differs from the class of the region it
// sits in along permutation cycles; buckets above 256 elements recurse on
// the next digit, smaller ones get a stable insertion sort (sort_insert.c:
// 14-37).  Equal keys share every bucket, so their final order is the order
// in which the passes over the buckets that contain them leave them.
//
// In one pass an element already in its class's region stays where it is;
// the other slots of region k are filled, in increasing position, by the
// class-k elements in the order the cycles carry them there; and the cycles
// of the pass visit regions as a rotor walk: leaving region r always takes
// r's next not-yet-taken foreign slot (positions in increasing order), and
// the cycles started from region i (phase i, i = 0, 1, ...) end on their
// first return to i.  The walk is replayed here literally.  It is serial by
// nature (every step depends on the class read by the previous one), so the
// block prepares in parallel what the walk reads: the classes, the region
// bounds, and each region's foreign slots in increasing position as a list
// of (position, element, class) entries; one thread then walks, taking one
// list entry per step.
//
// Buckets without equal keys need nothing: for distinct keys the stable LSD
// order is the only sorted order.  So the replay descends only into buckets
// that hold a tie, and writes the final order of each such bucket into the
// permutation.  Runs only when the sort left equal adjacent keys (dup_count
// from dup_detect); one block per system.
constexpr int kTieThreads = 256;
constexpr int kTieWidth = 10;           // sl_tune_intern.h:22 (sort_radix_width_default)
constexpr int kTieThreshold = 256;      // sl_tune_intern.h:30 (sort_radix_threshold_rec)
constexpr int kTieClasses = 1 << kTieWidth;
constexpr int kTieDepth = 8;            // 64 bits in 10-bit digits: at most 7 passes

__global__ __launch_bounds__(kTieThreads) void tie_order_kernel(
    int64_t n, const int64_t* __restrict__ soff, const unsigned long long* __restrict__ ks,
    uint32_t* __restrict__ perm, unsigned long long* __restrict__ kin, uint32_t* __restrict__ ord,
    uint32_t* __restrict__ ord2, uint16_t* __restrict__ cls, uint32_t* __restrict__ fp,
    unsigned long long* __restrict__ fec, DevScalars* __restrict__ sc) {
  if (sc->dup_count == 0) return;
  const int tid = threadIdx.x;
  const int64_t b = soff ? soff[blockIdx.x] : 0;
  const int64_t e = soff ? soff[blockIdx.x + 1] : n;
  if (e - b < 2) return;

  __shared__ uint32_t reg[kTieClasses + 1];   // region starts of the current pass
  __shared__ uint32_t nxt[kTieClasses];       // next unused entry of each region's slot list
  __shared__ uint32_t run_base;
  __shared__ unsigned long long skey[kTieThreshold];
  __shared__ int64_t fs[kTieDepth], fe[kTieDepth], fcur[kTieDepth];
  __shared__ int frh[kTieDepth], fnext[kTieDepth];
  __shared__ int depth, any;
  __shared__ int64_t found;

  if (tid == 0) any = 0;
  __syncthreads();
  for (int64_t i = b + tid; i + 1 < e; i += kTieThreads)
    if (ks[i + 1] == ks[i]) any = 1;
  __syncthreads();
  if (!any) return;

  // keys in input order (the sort's input of this system is [b, e) too) and
  // the initial order of the top segment
  for (int64_t i = b + tid; i < e; i += kTieThreads) {
    kin[perm[i]] = ks[i];
    ord[i] = (uint32_t)i;
  }
  if (tid == 0) {
    depth = 1;
    fs[0] = b;
    fe[0] = e;
    frh[0] = 63;
    fnext[0] = -2;   // pass not run yet
  }
  __syncthreads();

  while (depth > 0) {
    const int d = depth - 1;
    const int64_t s0 = fs[d], e0 = fe[d];
    const int64_t m = e0 - s0;
    if (fnext[d] == -2) {
      // ---- one pass of rs_rec_ip over [s0, e0) on bits rhigh..sh
      const int rhigh = frh[d];
      const int cw = rhigh + 1 < kTieWidth ? rhigh + 1 : kTieWidth;
      const int sh = rhigh - cw + 1;
      const int K = 1 << cw;
      const unsigned long long mask = (unsigned long long)(K - 1);
      for (int c = tid; c <= K; c += kTieThreads) reg[c] = 0;
      for (int c = tid; c < K; c += kTieThreads) nxt[c] = 0;
      __syncthreads();
      for (int64_t t = tid; t < m; t += kTieThreads) {
        const uint32_t el = ord[s0 + t];
        const int c = (int)((kin[el] >> sh) & mask);
        cls[s0 + t] = (uint16_t)c;
        ord2[s0 + t] = el;
        atomicAdd(&reg[c + 1], 1u);
      }
      __syncthreads();
      if (tid == 0)
        for (int c = 0; c < K; ++c) reg[c + 1] += reg[c];
      __syncthreads();
      // elements already inside their class's region stay; count them
      for (int64_t t = tid; t < m; t += kTieThreads) {
        const int c = cls[s0 + t];
        if ((uint32_t)t >= reg[c] && (uint32_t)t < reg[c + 1]) atomicAdd(&nxt[c], 1u);
      }
      __syncthreads();
      // region r's foreign slots, in increasing position, become entries
      // [start_r, start_r+1) of the slot list (fp: position, fec: element and
      // its class); nxt[r] = start_r, and reg[r] is reused for start_r+1
      if (tid == 0) {
        uint32_t acc = 0;
        for (int c = 0; c < K; ++c) {
          const uint32_t f = reg[c + 1] - reg[c] - nxt[c];
          nxt[c] = acc;
          acc += f;
        }
        run_base = 0;
      }
      __syncthreads();
      {
        constexpr int kPer = 4;   // consecutive positions per thread per chunk
        for (int64_t base = 0; base < m; base += (int64_t)kTieThreads * kPer) {
          const int64_t t0 = base + (int64_t)tid * kPer;
          uint32_t flag[kPer], nf = 0;
#pragma unroll
          for (int j = 0; j < kPer; ++j) {
            const int64_t t = t0 + j;
            flag[j] = 0u;
            if (t < m) {
              const int c = cls[s0 + t];
              flag[j] = ((uint32_t)t >= reg[c] && (uint32_t)t < reg[c + 1]) ? 0u : 1u;
            }
            nf += flag[j];
          }
          uint32_t tot;
          uint32_t off = run_base + block_exclusive_scan(nf, &tot);
#pragma unroll
          for (int j = 0; j < kPer; ++j) {
            if (flag[j]) {
              const int64_t t = t0 + j;
              fp[s0 + off] = (uint32_t)t;
              fec[s0 + off] = (unsigned long long)ord[s0 + t] |
                              ((unsigned long long)cls[s0 + t] << 32);
              ++off;
            }
          }
          __syncthreads();
          if (tid == 0) run_base += tot;
          __syncthreads();
        }
      }
      if (tid == 0) {
        for (int c = 0; c < K; ++c) reg[c] = c + 1 < K ? nxt[c + 1] : run_base;
        // The walk: phase i starts a cycle at each of region i's remaining
        // foreign slots (increasing); the carried element goes to the next
        // unused foreign slot of its class's region, whose element is
        // carried on, until an element of class i closes the cycle in the
        // starting slot.  One list entry per step.
        const uint32_t* fpp = fp + s0;
        const unsigned long long* fe = fec + s0;
        uint32_t* o2 = ord2 + s0;
        for (int i = 0; i < K; ++i) {
          while (nxt[i] < reg[i]) {
            const uint32_t k0 = nxt[i]++;
            const uint32_t p0 = fpp[k0];
            const unsigned long long e0 = fe[k0];
            uint32_t el = (uint32_t)e0;
            int c = (int)(e0 >> 32);
            for (;;) {
              if (nxt[c] >= reg[c]) {   // cannot happen: every region receives what it sends
                atomicOr(&sc->flags, kFlagDupKeys);
                break;
              }
              const uint32_t k = nxt[c]++;
              const uint32_t p = fpp[k];
              const unsigned long long e2 = fe[k];
              o2[p] = el;
              el = (uint32_t)e2;
              const int c2 = (int)(e2 >> 32);
              if (c2 == i) {
                o2[p0] = el;
                break;
              }
              c = c2;
            }
          }
        }
        fnext[d] = sh - 1;
        fcur[d] = s0;
      }
      __syncthreads();
      for (int64_t t = tid; t < m; t += kTieThreads) ord[s0 + t] = ord2[s0 + t];
      __syncthreads();
      if (fnext[d] < 0) {
        // all 64 bits consumed: the segment holds one key, its order is final
        for (int64_t t = tid; t < m; t += kTieThreads) perm[s0 + t] = ord[s0 + t];
        __syncthreads();
        if (tid == 0) depth = d;
        __syncthreads();
        continue;
      }
    }
    // ---- next bucket of this pass that holds a tie
    const int sh = fnext[d] + 1;   // low bit of this pass's digit
    if (tid == 0) found = -1;
    __syncthreads();
    for (int64_t base = fcur[d]; base + 1 < e0; base += kTieThreads) {
      const int64_t i = base + tid;
      if (i + 1 < e0 && ks[i + 1] == ks[i]) atomicMin((unsigned long long*)&found, (unsigned long long)i);
      __syncthreads();
      if (found >= 0) break;
      __syncthreads();
    }
    __syncthreads();
    const int64_t t0 = found;
    if (t0 < 0) {
      if (tid == 0) depth = d;
      __syncthreads();
      continue;
    }
    // bucket = positions of [s0, e0) whose key shares ks[t0]'s bits above sh
    __shared__ int64_t cs, ce;
    if (tid == 0) {
      const unsigned long long pre = ks[t0] >> sh;
      int64_t lo = s0, hi = t0;   // first position with the prefix
      while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if ((ks[mid] >> sh) < pre) lo = mid + 1; else hi = mid;
      }
      cs = lo;
      lo = t0 + 1;
      hi = e0;                    // first position past the prefix
      while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if ((ks[mid] >> sh) <= pre) lo = mid + 1; else hi = mid;
      }
      ce = lo;
      fcur[d] = ce;
    }
    __syncthreads();
    const int64_t c0 = cs, c1 = ce, cm = c1 - c0;
    if (cm > kTieThreshold) {
      __syncthreads();
      if (tid == 0) {
        fs[d + 1] = c0;
        fe[d + 1] = c1;
        frh[d + 1] = fnext[d];
        fnext[d + 1] = -2;
        depth = d + 2;
      }
      __syncthreads();
      continue;
    }
    // stable insertion sort of a small bucket (sort_insert.c:14-37): rank
    // by key, equal keys in their current order
    uint32_t el = 0;
    unsigned long long k = 0;
    if (tid < cm) {
      el = ord[c0 + tid];
      k = kin[el];
      skey[tid] = k;
    }
    __syncthreads();
    if (tid < cm) {
      int r = 0;
      for (int u = 0; u < (int)cm; ++u) {
        const unsigned long long ku = skey[u];
        r += (ku < k || (ku == k && u < tid)) ? 1 : 0;
      }
      perm[c0 + r] = el;
    }
    __syncthreads();
  }
}

}  // namespace

void launch_tie_order(hipStream_t s, int64_t n, int64_t nsys, const int64_t* soff,
                      const unsigned long long* ks, uint32_t* perm, unsigned long long* kin,
                      uint32_t* ord, uint32_t* ord2, uint16_t* cls, uint32_t* fp,
                      unsigned long long* fec, DevScalars* sc) {
  if (n < 2) return;
  tie_order_kernel<<<(unsigned)(nsys > 1 ? nsys : 1), kTieThreads, 0, s>>>(
      n, nsys > 1 ? soff : nullptr, ks, perm, kin, ord, ord2, cls, fp, fec, sc);
}

int64_t scan_blocks(int64_t n) { return (n + kScanTile - 1) / kScanTile; }
int64_t sort_blocks(int64_t n) { return (n + kSortTile - 1) / kSortTile; }

void launch_scan(hipStream_t s, const uint32_t* in, uint32_t* out, int64_t n, uint32_t* bsum,
                 uint32_t* total, const uint32_t* ndev, bool flags) {
  if (n <= 0) return;
  const int64_t nb = scan_blocks(n);
  scan_reduce_kernel<<<(unsigned)nb, kScanThreads, 0, s>>>(in, n, bsum, ndev, flags);
  scan_small_kernel<<<1, kScanThreads, 0, s>>>(bsum, (int)nb, total);
  scan_apply_kernel<<<(unsigned)nb, kScanThreads, 0, s>>>(in, out, n, bsum, ndev, flags);
}

// One stable pass over bits [shift, shift + nbits), output in (kout, vout).
// With ndev != nullptr the element count is min(n, *ndev), read on the device
// (n then only sizes the grid), so no host round trip is needed.
void launch_radix_pass_dev(hipStream_t s, const unsigned long long* kin, const uint32_t* vin,
                           unsigned long long* kout, uint32_t* vout, int64_t n, const uint32_t* ndev,
                           int shift, int nbits, uint32_t* hist, uint32_t* bsum,
                           const uint32_t* dsrc) {
  const int nb = (int)sort_blocks(n);
  radix_hist_kernel<<<nb, kSortThreads, 0, s>>>(kin, vin, dsrc, n, ndev, shift, (1u << nbits) - 1u,
                                                hist, nb);
  // Few blocks: each scatter block sums its digits' rows itself, one launch
  // fewer per pass.  Otherwise bsum[0..255] holds the digit totals.
  const bool scan_here = nb <= kScanHereBlocks;
  if (!scan_here) digit_scan_kernel<<<kRadix, kSortThreads, 0, s>>>(hist, nb, bsum);
  radix_scatter_kernel<<<nb, kSortThreads, 0, s>>>(kin, vin, kout, vout, dsrc, n, ndev, shift, nbits,
                                                   hist, bsum, nb, scan_here);
}

void launch_radix_pass(hipStream_t s, const unsigned long long* kin, const uint32_t* vin,
                       unsigned long long* kout, uint32_t* vout, int64_t n, int shift, int nbits,
                       uint32_t* hist, uint32_t* bsum) {
  launch_radix_pass_dev(s, kin, vin, kout, vout, n, nullptr, shift, nbits, hist, bsum, nullptr);
}

// Full sort of particle keys over bits [0, 60).  Input in (k0, v0); the sorted
// result ends in (k0, v0) again after an even number of passes.  With a system
// index per input particle (sysid != nullptr, values < nsys <= 65536), two more
// stable passes over its 16 bits order the particles by (system, key); the
// upper pass only runs when there are more than 256 systems, and then the
// lower one goes to (k1, v1) and the upper one back.
void launch_sort_keys(hipStream_t s, unsigned long long* k0, uint32_t* v0, unsigned long long* k1,
                      uint32_t* v1, int64_t n, uint32_t* hist, uint32_t* bsum,
                      const uint32_t* sysid, int64_t nsys) {
  static const int shifts[8] = {0, 8, 16, 24, 32, 40, 48, 56};
  static const int widths[8] = {8, 8, 8, 8, 8, 8, 8, 4};
  for (int p = 0; p < 8; ++p) {
    if (p % 2 == 0)
      launch_radix_pass(s, k0, v0, k1, v1, n, shifts[p], widths[p], hist, bsum);
    else
      launch_radix_pass(s, k1, v1, k0, v0, n, shifts[p], widths[p], hist, bsum);
  }
  if (sysid) {
    launch_radix_pass_dev(s, k0, v0, k1, v1, n, nullptr, 0, 8, hist, bsum, sysid);
    if (nsys > 256) {
      launch_radix_pass_dev(s, k1, v1, k0, v0, n, nullptr, 8, 8, hist, bsum, sysid);
    } else {
      launch_copy_pairs(s, k1, v1, k0, v0, n);
    }
  }
}

void launch_scatter_back(hipStream_t s, int64_t n, const uint32_t* perm, const double* in,
                         double* out) {
  scatter_back_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, perm, in, out);
}

void launch_gather_rec(hipStream_t s, int64_t n, const uint32_t* perm, const double4* rec,
                       double* xs, double* ys, double* zs, double* qs) {
  gather_rec_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, perm, rec, xs, ys, zs, qs);
}

void launch_gather4(hipStream_t s, int64_t n, const uint32_t* perm, const double* x, const double* y,
                    const double* z, const double* q, double* xs, double* ys, double* zs,
                    double* qs) {
  gather4_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, perm, x, y, z, q, xs, ys, zs, qs);
}

}  // namespace pepc
