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

__global__ __launch_bounds__(kScanThreads) void scan_reduce_kernel(const uint32_t* __restrict__ in,
                                                                    int64_t n,
                                                                    uint32_t* __restrict__ bsum) {
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  uint32_t s = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    int64_t e = base + j * kScanThreads + threadIdx.x;
    if (e < n) s += in[e];
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
                                                                   const uint32_t* __restrict__ bsum) {
  __shared__ uint32_t tile[kScanTile];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    int64_t e = base + j * kScanThreads + threadIdx.x;
    tile[j * kScanThreads + threadIdx.x] = e < n ? in[e] : 0u;
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
__global__ __launch_bounds__(kSortThreads) void radix_scatter_kernel(
    const unsigned long long* __restrict__ kin, const uint32_t* __restrict__ vin,
    unsigned long long* __restrict__ kout, uint32_t* __restrict__ vout,
    const uint32_t* __restrict__ dsrc, int64_t n,
    const uint32_t* __restrict__ ndev, int shift, int nbits, const uint32_t* __restrict__ hscan,
    const uint32_t* __restrict__ dtot, int nblocks) {
  if (ndev && (int64_t)*ndev < n) n = *ndev;
  const int64_t tile0 = (int64_t)blockIdx.x * kSortTile;
  if (tile0 >= n) return;
  __shared__ uint32_t wcnt[kSortWaves][kRadix];
  __shared__ uint32_t lbase[kRadix];
  __shared__ uint32_t gbase[kRadix];
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
    // global start of digit d: digits below d over all blocks
    gbase[d] = block_exclusive_scan(dtot[d], &tot);
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
    const uint32_t pos =
        gbase[d] + hscan[(int64_t)d * nblocks + blockIdx.x] + ((uint32_t)i - lbase[d]);
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

}  // namespace

int64_t scan_blocks(int64_t n) { return (n + kScanTile - 1) / kScanTile; }
int64_t sort_blocks(int64_t n) { return (n + kSortTile - 1) / kSortTile; }

void launch_scan(hipStream_t s, const uint32_t* in, uint32_t* out, int64_t n, uint32_t* bsum,
                 uint32_t* total) {
  if (n <= 0) return;
  const int64_t nb = scan_blocks(n);
  scan_reduce_kernel<<<(unsigned)nb, kScanThreads, 0, s>>>(in, n, bsum);
  scan_small_kernel<<<1, kScanThreads, 0, s>>>(bsum, (int)nb, total);
  scan_apply_kernel<<<(unsigned)nb, kScanThreads, 0, s>>>(in, out, n, bsum);
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
  // bsum[0..255] holds the digit totals
  digit_scan_kernel<<<kRadix, kSortThreads, 0, s>>>(hist, nb, bsum);
  radix_scatter_kernel<<<nb, kSortThreads, 0, s>>>(kin, vin, kout, vout, dsrc, n, ndev, shift, nbits,
                                                   hist, bsum, nb);
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
