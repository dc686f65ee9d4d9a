// tree.hip -- duplicate-key fix-up, level-synchronous octree build, FP64
// upward pass and walk groups.
//
// Build (replaces tree_build_from_particles / insert_helper,
// module_tree_grow.f90:552-697).  For sorted distinct keys let L(a) be the
// first level at which particles a-1 and a fall into different cells (0 at
// both ends).  Particle a then owns the nodes of levels L(a)..max(L(a),L(a+1)):
// twigs down to the level where it separates from its right neighbour, and
// its leaf at the deepest of those levels (the reference's insert_helper
// creates exactly these, a twig before its children in key order).  Nodes
// are stored level by level (BFS), within a level in particle order, so the
// children of every twig are contiguous and the k-th run of children belongs
// to the k-th twig.  A node's BFS index is the number of nodes at lower
// levels plus the number of particles before its owner that own a node at
// its level: per-block counts per level, one scan over the (level, block)
// table, and an emit pass that writes every node to its final place.
//
// Upward pass (tree_node_update_from_children, :762-821, with
// shift_multipoles_up, coulomb/module_interaction_specific.f90:84-140): one
// launch per level, deepest first, FP64, the reference's operation order with
// contraction off, so moments are bit-identical to the reference formula.
//
// Groups (walk tiles, module_tree_walk_simple.f90:135-212): the maximal key
// cells holding at most G particles = children with count <= G of twigs with
// count > G (or the root).
//
// Batched systems: particles sorted by (system, key) form one forest with a
// root per system at level 0; the level-synchronous steps above need no other
// change, every system keeps its own box (keys, boxlength2) and groups never
// cross a system boundary.
#include "kernels.h"

namespace pepc {

namespace {

constexpr int kT = 256;

// ------------------------------------------------------------ duplicate keys
// Batched systems (sys != nullptr): keys are ordered within each system's
// range only, and every check and fix-up stays inside one range.
__device__ __forceinline__ bool same_sys(const uint32_t* sys, int64_t i, int64_t j) {
  return sys == nullptr || sys[i] == sys[j];
}

__global__ void dup_detect_kernel(int64_t n, const unsigned long long* __restrict__ keys,
                                  const uint32_t* __restrict__ sys, DevScalars* __restrict__ sc) {
  uint32_t c = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i + 1 < n;
       i += (int64_t)gridDim.x * blockDim.x)
    c += (keys[i + 1] <= keys[i] && same_sys(sys, i, i + 1)) ? 1u : 0u;
  c = wave_sum(c);
  if (lane_id() == 0 && c) atomicAdd(&sc->dup_count, c);
}

__global__ void dup_prepare_kernel(int64_t n, const unsigned long long* __restrict__ keys,
                                   const uint32_t* __restrict__ sys,
                                   long long* __restrict__ kd, unsigned long long* __restrict__ old,
                                   uint32_t* __restrict__ flag, DevScalars* __restrict__ sc) {
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->dup_scan_n = sc->dup_count ? (uint32_t)n : 0u;
  if (sc->dup_count == 0) return;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    // the last particle of a range behaves as the last particle of a rank
    long long d = (i + 1 < n && same_sys(sys, i, i + 1)) ? (long long)(keys[i + 1] - keys[i]) : 1ll;
    kd[i] = d;
    old[i] = keys[i];
    flag[i] = d < 1 ? 1u : 0u;
  }
}

__global__ void dup_compact_kernel(int64_t n, const long long* __restrict__ kd,
                                   const uint32_t* __restrict__ rank, uint32_t* __restrict__ list,
                                   const DevScalars* __restrict__ sc) {
  if (sc->dup_count == 0) return;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    if (kd[i] < 1) list[rank[i]] = (uint32_t)i;
}

// module_domains.f90:248-297, one thread, in the reference's loop order.
// Only positions that start as duplicates are ever modified to a smaller
// key difference, so visiting the compacted list in order equals visiting
// every position.  With batched systems the loop runs on each system's range
// (base pointer and length of that range), systems in increasing order.
__global__ void dup_serial_kernel(int64_t n_all, unsigned long long* __restrict__ keys_all,
                                  long long* __restrict__ kd_all, const uint32_t* __restrict__ list,
                                  const uint32_t* __restrict__ sys, const int64_t* __restrict__ soff,
                                  DevScalars* __restrict__ sc) {
  const uint32_t nd = sc->dup_count;
  if (nd == 0 || threadIdx.x != 0) return;
  int64_t cur = -1;
  long long n = n_all;
  unsigned long long* keys = keys_all;
  long long* kd = kd_all;
  for (uint32_t r = 0; r < nd; ++r) {
    const int64_t gi = list[r];
    const int64_t sid = sys ? (int64_t)sys[gi] : 0;
    if (sid != cur) {
      cur = sid;
      const int64_t b = sys ? soff[sid] : 0;
      n = sys ? (long long)(soff[sid + 1] - b) : n_all;
      keys = keys_all + b;
      kd = kd_all + b;
      if ((long long)(keys[n - 1] - keys[0]) + 1 < n) {   // :254-262
        atomicOr(&sc->flags, kFlagDupKeys);
        return;
      }
    }
    const long long i1 = (long long)(gi - (keys - keys_all)) + 1;   // 1-based as in the source
    if (kd[i1 - 1] >= 1) continue;
    const long long jmax = (n - i1) > i1 ? (n - i1) : i1;
    for (long long j = 1; j <= jmax; ++j) {
      const long long dn = i1 - j < 1 ? 1 : i1 - j;
      const long long up = i1 + j < n - 1 ? i1 + j : n - 1;
      if (kd[dn - 1] > 1 && i1 - j >= 1) {
        for (long long t = i1 - j + 1; t <= i1; ++t) keys[t - 1] -= 1ull;
        kd[i1 - 1] += 1;
        kd[i1 - j - 1] -= 1;
        break;
      } else if (kd[up - 1] > 1 && i1 + j <= n - 1) {
        for (long long t = i1 + 1; t <= i1 + j; ++t) keys[t - 1] += 1ull;
        kd[i1 - 1] += 1;
        kd[i1 + j - 1] -= 1;
        break;
      }
    }
  }
}

// key_to_intcoord_hilbert / key_to_coord, module_spacefilling.f90:445-565.
constexpr int kC3[8] = {0, 1, 3, 2, 6, 7, 5, 4};
constexpr int kGX[8] = {5, 6, 0, 5, 5, 0, 6, 5};
constexpr int kGR[8] = {0, 0, 0, 5, 0, 0, 6, 5};

__device__ void key_to_intcoord(unsigned long long key, long long ic[3]) {
  const int lev = level_of_key(key);
  ic[0] = ic[1] = ic[2] = 0;
  for (int i = 0; i < lev; ++i) {
    const int h = (int)((key >> (3 * i)) & 7ull);
    const int cval = kC3[h], rev = kGR[h], ex = kGX[h];
    if (i > 0) {
      for (int j = 0; j < 3; ++j)
        if ((rev >> j) & 1) ic[j] = (~ic[j]) & ((1ll << i) - 1);
      for (int j = 0; j < 2; ++j) {
        if (!((ex >> j) & 1)) continue;
        for (int k = j + 1; k < 3; ++k)
          if ((ex >> k) & 1) { long long t = ic[j]; ic[j] = ic[k]; ic[k] = t; }
      }
    }
    for (int j = 0; j < 3; ++j) ic[j] |= (long long)((cval >> j) & 1) << i;
  }
  for (int j = 0; j < 3; ++j) ic[j] <<= (kMaxLevel - lev);
}

__global__ void dup_recoord_kernel(int64_t n, const unsigned long long* __restrict__ keys,
                                   const unsigned long long* __restrict__ old,
                                   double* __restrict__ x, double* __restrict__ y,
                                   double* __restrict__ z, const DevBox* __restrict__ boxes,
                                   const uint32_t* __restrict__ sys,
                                   const DevScalars* __restrict__ sc) {
  if (sc->dup_count == 0) return;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (keys[i] == old[i]) continue;
    const DevBox* __restrict__ box = boxes + (sys ? sys[i] : 0u);
    long long ic[3];
    key_to_intcoord(keys[i], ic);
    x[i] = ((double)ic[0] + 0.5) * box->cell[0] + box->boxmin[0];
    y[i] = ((double)ic[1] + 0.5) * box->cell[1] + box->boxmin[1];
    z[i] = ((double)ic[2] + 0.5) * box->cell[2] + box->boxmin[2];
  }
}

// ------------------------------------------------------------ build
// Every kernel after the node count is known bails out once an earlier stage
// raised a flag that makes the node arrays inconsistent, so no index read
// from those arrays is ever used.
__device__ __forceinline__ bool tree_broken(const DevScalars* sc) {
  return (sc->flags & (kFlagNodeOverflow | kFlagDupKeys | kFlagBadBox)) != 0u;
}

// A system boundary acts like the ends of the array: the first particle of a
// system owns that system's root (level 0).
__global__ void levels_kernel(int64_t n, const unsigned long long* __restrict__ keys,
                              const uint32_t* __restrict__ sys,
                              uint8_t* __restrict__ lo, uint8_t* __restrict__ hi,
                              uint32_t* __restrict__ gmark, DevScalars* __restrict__ sc) {
  for (int64_t a = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; a < n;
       a += (int64_t)gridDim.x * blockDim.x) {
    int la = 0, lb = 0;
    bool dup = false;
    if (a > 0 && same_sys(sys, a - 1, a)) {
      if (keys[a - 1] == keys[a]) dup = true; else la = split_level(keys[a - 1], keys[a]);
    }
    if (a + 1 < n && same_sys(sys, a, a + 1)) {
      if (keys[a + 1] == keys[a]) dup = true; else lb = split_level(keys[a], keys[a + 1]);
    }
    if (dup) {
      atomicOr(&sc->flags, kFlagDupKeys);
      la = lb = 0;
    }
    const int lam = la > lb ? la : lb;
    lo[a] = (uint8_t)la;
    hi[a] = (uint8_t)lam;
    gmark[a] = 0;   // group marks start cleared (group_mark_kernel sets them)
  }
}

// ---- BFS node numbering
// Tiles of kBfsTile sorted particles; wave w of a tile owns the kBfsWave
// consecutive particles from base + w kBfsWave and takes them 64 at a time,
// one per lane, so the nodes one pass writes for a level are consecutive.
// tab[l * nb + b] counts the nodes of level l owned by tile b's particles,
// tab[kLevels * nb + l * nb + b] the twigs among them; after the exclusive
// scan (level-major) the entries are the BFS index of the tile's first node
// of that level and the twig rank of its first twig there.
constexpr int kBfsTile = 4096;
constexpr int kBfsWaves = kT / kWave;           // 4
constexpr int kBfsWave = kBfsTile / kBfsWaves;  // 1024 particles per wave
constexpr int kLevels = kMaxLevel + 1;

// The wave's node and twig counts per level over its particles: lane l
// accumulates level l (levels < 21 <= 64 lanes).
__device__ __forceinline__ void bfs_wave_counts(int64_t n, int64_t base,
                                                const uint8_t* __restrict__ lo,
                                                const uint8_t* __restrict__ hi, uint32_t& cn,
                                                uint32_t& ct) {
  const int lane = lane_id();
  cn = 0;
  ct = 0;
  for (int it = 0; it < kBfsWave / kWave; ++it) {
    const int64_t a = base + (int64_t)it * kWave + lane;
    const bool valid = a < n;
    const int l0 = valid ? lo[a] : kLevels, l1 = valid ? hi[a] : -1;
    const int lmin = (int)wave_min_u32((uint32_t)l0), lmax = (int)wave_max_i32(l1);
    for (int l = lmin; l <= lmax; ++l) {
      const uint32_t c = (uint32_t)popc64(__ballot(l0 <= l && l <= l1));
      const uint32_t d = (uint32_t)popc64(__ballot(l0 <= l && l < l1));
      if (lane == l) { cn += c; ct += d; }
    }
  }
}

__global__ __launch_bounds__(kT) void bfs_count_kernel(int64_t n, const uint8_t* __restrict__ lo,
                                                      const uint8_t* __restrict__ hi,
                                                      uint32_t* __restrict__ tab, int64_t nb,
                                                      const DevScalars* __restrict__ sc) {
  if (tree_broken(sc)) return;
  __shared__ uint32_t wn[kBfsWaves][kLevels], wt[kBfsWaves][kLevels];
  const int w = threadIdx.x / kWave, lane = lane_id();
  uint32_t cn, ct;
  bfs_wave_counts(n, (int64_t)blockIdx.x * kBfsTile + (int64_t)w * kBfsWave, lo, hi, cn, ct);
  if (lane < kLevels) { wn[w][lane] = cn; wt[w][lane] = ct; }
  __syncthreads();
  if (threadIdx.x < kLevels) {
    const int l = threadIdx.x;
    uint32_t a = 0, b = 0;
    for (int k = 0; k < kBfsWaves; ++k) { a += wn[k][l]; b += wt[k][l]; }
    tab[(int64_t)l * nb + blockIdx.x] = a;
    tab[(int64_t)(kLevels + l) * nb + blockIdx.x] = b;
  }
}

// Exclusive scans of the node table and of the twig table, level-major:
// block (k, l) scans row l of table k (one count per tile) into the output
// half of tab and starts it at the sum of rows 0..l-1, which it reads
// itself with coalesced loads, so the 2 x 21 rows are scanned in one launch
// with no wait between blocks.  The totals and per-level starts go to the
// scalars (node count, first node of every level; twig count, first twig
// rank of every level).  A node count above the pool raises the overflow
// flag before any node is written.
constexpr int kBfsScanT = 1024;

__global__ __launch_bounds__(kBfsScanT) void bfs_scan_kernel(uint32_t* __restrict__ tab, int64_t nb,
                                                            int64_t cap, DevScalars* __restrict__ sc) {
  if (tree_broken(sc)) return;
  const int kind = blockIdx.x / kLevels, l = blockIdx.x % kLevels;
  const uint32_t* t = tab + (int64_t)kind * kLevels * nb;
  uint32_t* out = tab + (int64_t)(2 + kind) * kLevels * nb + (int64_t)l * nb;
  constexpr int kBatch = 8;   // loads in flight together per thread
  // sum of the earlier rows
  const int64_t plen = (int64_t)l * nb;
  uint32_t prev = 0;
  for (int64_t e0 = threadIdx.x; e0 < plen; e0 += (int64_t)kBfsScanT * kBatch) {
    uint32_t v[kBatch];
#pragma unroll
    for (int j = 0; j < kBatch; ++j) {
      const int64_t e = e0 + (int64_t)j * kBfsScanT;
      v[j] = e < plen ? t[e] : 0u;
    }
#pragma unroll
    for (int j = 0; j < kBatch; ++j) prev += v[j];
  }
  // this row: thread i holds entries [i * per, (i + 1) * per)
  const uint32_t* row = t + plen;
  const int64_t per = (nb + kBfsScanT - 1) / kBfsScanT;
  const int64_t b0 = (int64_t)threadIdx.x * per;
  uint32_t sum = 0;
  for (int64_t k0 = 0; k0 < per; k0 += kBatch) {
    uint32_t v[kBatch];
#pragma unroll
    for (int j = 0; j < kBatch; ++j) {
      const int64_t e = b0 + k0 + j;
      v[j] = (k0 + j < per && e < nb) ? row[e] : 0u;
    }
#pragma unroll
    for (int j = 0; j < kBatch; ++j) sum += v[j];
  }
  __shared__ uint32_t ws[kBfsScanT / kWave], wp[kBfsScanT / kWave];
  const int w = threadIdx.x / kWave;
  const uint32_t inc = wave_inclusive_scan(sum);
  const uint32_t pw = wave_sum(prev);
  if (lane_id() == kWave - 1) ws[w] = inc;
  if (lane_id() == 0) wp[w] = pw;
  __syncthreads();
  uint32_t off = 0, tot = 0, start = 0;
  for (int k = 0; k < kBfsScanT / kWave; ++k) {
    off += k < w ? ws[k] : 0u;
    tot += ws[k];
    start += wp[k];
  }
  uint32_t run = start + off + inc - sum;
  for (int64_t k0 = 0; k0 < per; k0 += kBatch) {
    uint32_t v[kBatch];
#pragma unroll
    for (int j = 0; j < kBatch; ++j) {
      const int64_t e = b0 + k0 + j;
      v[j] = (k0 + j < per && e < nb) ? row[e] : 0u;
    }
#pragma unroll
    for (int j = 0; j < kBatch; ++j) {
      const int64_t e = b0 + k0 + j;
      if (k0 + j < per && e < nb) {
        out[e] = run;
        run += v[j];
      }
    }
  }
  if (threadIdx.x == 0) {
    if (kind) sc->twig_offset[l] = start;
    else sc->level_offset[l] = start;
    if (l == kLevels - 1) {
      const uint32_t total = start + tot;
      if (kind) {
        sc->twig_offset[kLevels] = total;
        sc->ntwig = total;
      } else {
        sc->level_offset[kLevels] = total;
        sc->nnodes = total;
        if ((int64_t)total > cap) atomicOr(&sc->flags, kFlagNodeOverflow);
      }
    }
  }
}

// Every node to its BFS place.  The wave recounts its particles per level
// for its start within the tile, then passes over them 64 at a time, levels
// deepest first so that a twig's first child (its owner's node one level
// down) was numbered just before it; the lanes owning a level take
// consecutive indices (ballot ranks).  Leaves get their geometry record
// (centre = particle, last slot = charge); twigs their rank-indexed first
// child and node index.
__global__ __launch_bounds__(kT) void bfs_emit_kernel(
    int64_t n, const uint8_t* __restrict__ lo, const uint8_t* __restrict__ hi,
    const uint32_t* __restrict__ tab, int64_t nb, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ z, const double* __restrict__ q,
    NodeGeom* __restrict__ geom, NodeLink* __restrict__ link, uint32_t* __restrict__ nfirst,
    uint8_t* __restrict__ ntype, uint32_t* __restrict__ fc_rank, uint32_t* __restrict__ node_rank,
    uint32_t* __restrict__ parent, const DevScalars* __restrict__ sc) {
  if (tree_broken(sc)) return;
  // children of the last twig end at the last node
  if (blockIdx.x == 0 && threadIdx.x == 0) fc_rank[sc->ntwig] = sc->nnodes;
  __shared__ uint32_t wn[kBfsWaves][kLevels], wt[kBfsWaves][kLevels];
  const int w = threadIdx.x / kWave, lane = lane_id();
  const int64_t base = (int64_t)blockIdx.x * kBfsTile + (int64_t)w * kBfsWave;
  uint32_t rn, rt;   // lane l: next node index / twig rank of level l for this wave
  bfs_wave_counts(n, base, lo, hi, rn, rt);
  if (lane < kLevels) { wn[w][lane] = rn; wt[w][lane] = rt; }
  __syncthreads();
  if (lane < kLevels) {
    rn = tab[(int64_t)(2 * kLevels + lane) * nb + blockIdx.x];
    rt = tab[(int64_t)(3 * kLevels + lane) * nb + blockIdx.x];
    for (int k = 0; k < w; ++k) { rn += wn[k][lane]; rt += wt[k][lane]; }
  }
  for (int it = 0; it < kBfsWave / kWave; ++it) {
    const int64_t a = base + (int64_t)it * kWave + lane;
    const bool valid = a < n;
    const int l0 = valid ? lo[a] : kLevels, l1 = valid ? hi[a] : -1;
    NodeGeom g = {0.0, 0.0, 0.0, 0.0};
    if (valid) { g.cx = x[a]; g.cy = y[a]; g.cz = z[a]; g.b2 = q[a]; }
    const int lmin = (int)wave_min_u32((uint32_t)l0), lmax = (int)wave_max_i32(l1);
    uint32_t child = 0;
    for (int l = lmax; l >= lmin; --l) {
      const bool own = l0 <= l && l <= l1;
      const bool twig = own && l < l1;
      const unsigned long long mn = __ballot(own), mt = __ballot(twig);
      const uint32_t idx = (uint32_t)__builtin_amdgcn_readlane((int)rn, l) + mask_rank(mn);
      const uint32_t r = (uint32_t)__builtin_amdgcn_readlane((int)rt, l) + mask_rank(mt);
      if (lane == l) { rn += (uint32_t)popc64(mn); rt += (uint32_t)popc64(mt); }
      if (own) {
        NodeLink lk;
        lk.level = (uint32_t)l;
        nfirst[idx] = (uint32_t)a;
        if (!twig) {                           // leaf (a leaf has no acceptance test:
          lk.first = (uint32_t)a;              // its b2 slot holds the charge)
          lk.nchild = 0;
          lk.count = 1;
          geom[idx] = g;
        } else {
          lk.first = child;
          lk.nchild = 0xFFFFFFFFu;
          lk.count = 0;
          fc_rank[r] = child;
          node_rank[r] = idx;
        }
        link[idx] = lk;
        ntype[idx] = twig ? 1 : 0;
        if (l == 0) parent[idx] = 0xFFFFFFFFu;   // a root (every other node's parent is
                                                // written by the upward pass)
        child = idx;                           // first child of this particle's twig one level up
      }
    }
  }
}


// Moment record slots, reference layout (t_tree_node_interaction_data order).
enum { M_CX = 0, M_CY, M_CZ, M_Q, M_AQ, M_PX, M_PY, M_PZ, M_QX, M_QY, M_QZ, M_XY, M_YZ, M_ZX, M_BMAX };

// The walk accepts a twig for a target iff theta2 * dist2 > b2 with the b2
// written here, so the MAC selection (coulomb/module_interaction_specific.f90:
// 256-275) is fixed per node: 0 Barnes-Hut, boxlength2(level); 1 Bmax,
// min(bmax**2, 3.0 * boxlength2); any other value never accepts (N^2).
__device__ __forceinline__ double walk_b2(const double* m, double boxlength2, int mac_select) {
  if (mac_select == 0) return boxlength2;
  if (mac_select == 1) {
    const double bm = m[M_BMAX] * m[M_BMAX], c = 3.0 * boxlength2;
    return bm < c ? bm : c;
  }
  return INFINITY;
}

__device__ __forceinline__ void write_walk_forms(int64_t i, const double* m, double b2, int prec,
                                                 NodeGeom* geom, NodeMoments<float>* wm32,
                                                 NodeMoments<double>* wm64) {
  NodeGeom g;
  g.cx = m[M_CX]; g.cy = m[M_CY]; g.cz = m[M_CZ]; g.b2 = b2;
  geom[i] = g;
  double w[12];
  w[0] = m[M_Q];
  w[1] = m[M_PX]; w[2] = m[M_PY]; w[3] = m[M_PZ];
  w[4] = 2.0 * m[M_QX]; w[5] = 2.0 * m[M_QY]; w[6] = 2.0 * m[M_QZ];
  w[7] = m[M_XY]; w[8] = m[M_YZ]; w[9] = m[M_ZX];
  w[10] = (m[M_QX] + m[M_QY]) + m[M_QZ];
  w[11] = 0.0;
  if (prec == kWalkFP64) {
    double4* o = reinterpret_cast<double4*>(&wm64[i]);
    o[0] = make_double4(w[0], w[1], w[2], w[3]);
    o[1] = make_double4(w[4], w[5], w[6], w[7]);
    o[2] = make_double4(w[8], w[9], w[10], w[11]);
  } else {
    float4* o = reinterpret_cast<float4*>(&wm32[i]);
    o[0] = make_float4((float)w[0], (float)w[1], (float)w[2], (float)w[3]);
    o[1] = make_float4((float)w[4], (float)w[5], (float)w[6], (float)w[7]);
    o[2] = make_float4((float)w[8], (float)w[9], (float)w[10], 0.0f);
  }
}

// Moments of twig i from its children (contiguous from lk.first).  A leaf
// child's moments are its particle's (coc = x, charge = q, |charge| = |q|,
// higher moments and bmax 0), a twig child's come from its record.
struct UpArgs {
  int prec;
  int mac_select;
  NodeLink* link;
  const double *x, *y, *z, *q;
  double* mom;
  NodeGeom* geom;
  NodeMoments<float>* wm32;
  NodeMoments<double>* wm64;
  const DevBox* boxes;
  const uint32_t* sys;
  const uint32_t* nfirst;
  const uint8_t* ntype;       // 0 leaf, 1 twig
  uint8_t* kidtwig;           // bit j: child j is a twig (read by the walk's pushes)
  const uint32_t* node_rank;  // node index of each twig rank
  const uint32_t* fc_rank;    // first child of each twig rank (+1 sentinel)
  uint32_t* parent;           // parent node of each node
  const DevScalars* sc;
};

__device__ void twig_moments(int64_t i, uint32_t c0, uint32_t nc, int level, const UpArgs& u) {
  const double b2 = u.boxes[u.sys ? u.sys[u.nfirst[i]] : 0u].b2[level];
  // The children's kinds come first (one byte each, one line); then every
  // child's charge, |charge| and centre are requested together, a leaf
  // child's from its geometry record (its particle's position and charge), a
  // twig child's from the first 48 bytes of its moment record with its
  // particle count from its link, so the loads of all children overlap.
  uint32_t leafmask = 0, count = 0;
#pragma unroll
  for (uint32_t j = 0; j < 8; ++j)
    if (j < nc && u.ntype[c0 + j] == 0) leafmask |= 1u << j;
  double cq[8], ca[8], cx[8], cy[8], cz[8];
#pragma unroll
  for (uint32_t j = 0; j < 8; ++j) {
    cq[j] = ca[j] = cx[j] = cy[j] = cz[j] = 0.0;
    if (j < nc) {
      if (leafmask & (1u << j)) {
        const NodeGeom g = u.geom[c0 + j];
        count += 1u;
        cx[j] = g.cx; cy[j] = g.cy; cz[j] = g.cz; cq[j] = g.b2;
        ca[j] = fabs(cq[j]);
      } else {
        count += u.link[c0 + j].count;
        const double2* r = reinterpret_cast<const double2*>(u.mom + (int64_t)(c0 + j) * 16);
        const double2 r0 = r[0], r1 = r[1], r2 = r[2];
        cx[j] = r0.x; cy[j] = r0.y; cz[j] = r1.x; cq[j] = r1.y; ca[j] = r2.x;
      }
    }
  }
  double charge = 0.0, absq = 0.0;
#pragma unroll
  for (uint32_t j = 0; j < 8; ++j)
    if (j < nc) charge = charge + cq[j];
#pragma unroll
  for (uint32_t j = 0; j < 8; ++j)
    if (j < nc) absq = absq + ca[j];
  double coc[3] = {0.0, 0.0, 0.0};
  if (absq != 0.0) {
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j) {
      if (j < nc) {
        coc[0] = coc[0] + (cx[j] * ca[j]);
        coc[1] = coc[1] + (cy[j] * ca[j]);
        coc[2] = coc[2] + (cz[j] * ca[j]);
      }
    }
    for (int d = 0; d < 3; ++d) coc[d] = coc[d] / absq;
  } else {
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j) {
      if (j < nc) {
        coc[0] = coc[0] + cx[j];
        coc[1] = coc[1] + cy[j];
        coc[2] = coc[2] + cz[j];
      }
    }
    for (int d = 0; d < 3; ++d) coc[d] = coc[d] / (double)nc;
  }
  double dip[3] = {0.0, 0.0, 0.0}, quad[3] = {0.0, 0.0, 0.0}, xy = 0.0, yz = 0.0, zx = 0.0;
  double bmax = 0.0;
#pragma unroll
  for (uint32_t j = 0; j < 8; ++j) {
    if (j >= nc) continue;
    // higher moments of the child: zero for a leaf, record slots 5..14 else
    double c[15];
#pragma unroll
    for (int k = 0; k < 15; ++k) c[k] = 0.0;
    c[M_CX] = cx[j]; c[M_CY] = cy[j]; c[M_CZ] = cz[j]; c[M_Q] = cq[j]; c[M_AQ] = ca[j];
    if (!(leafmask & (1u << j))) {
      const double2* r = reinterpret_cast<const double2*>(u.mom + (int64_t)(c0 + j) * 16);
      const double2 r2 = r[2], r3 = r[3], r4 = r[4], r5 = r[5], r6 = r[6], r7 = r[7];
      c[M_PX] = r2.y; c[M_PY] = r3.x; c[M_PZ] = r3.y; c[M_QX] = r4.x; c[M_QY] = r4.y;
      c[M_QZ] = r5.x; c[M_XY] = r5.y; c[M_YZ] = r6.x; c[M_ZX] = r6.y; c[M_BMAX] = r7.x;
    }
    double s[3];
    for (int d = 0; d < 3; ++d) s[d] = coc[d] - c[M_CX + d];
    for (int d = 0; d < 3; ++d) dip[d] = dip[d] + c[M_PX + d] + c[M_Q] * s[d];
    for (int d = 0; d < 3; ++d)
      quad[d] = quad[d] + c[M_QX + d] + c[M_PX + d] * s[d] + 0.5 * c[M_Q] * (s[d] * s[d]);
    xy = xy + c[M_XY] + c[M_PX] * s[1] + c[M_PY] * s[0] + c[M_Q] * s[0] * s[1];
    yz = yz + c[M_YZ] + c[M_PY] * s[2] + c[M_PZ] * s[1] + c[M_Q] * s[1] * s[2];
    zx = zx + c[M_ZX] + c[M_PZ] * s[0] + c[M_PX] * s[2] + c[M_Q] * s[2] * s[0];
    const double v = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]) + c[M_BMAX];
    if (j == 0 || v > bmax) bmax = v;
  }
  double m[16];
  m[M_CX] = coc[0]; m[M_CY] = coc[1]; m[M_CZ] = coc[2];
  m[M_Q] = charge; m[M_AQ] = absq;
  for (int d = 0; d < 3; ++d) { m[M_PX + d] = dip[d]; m[M_QX + d] = quad[d]; }
  m[M_XY] = xy; m[M_YZ] = yz; m[M_ZX] = zx; m[M_BMAX] = bmax; m[15] = 0.0;
  u.link[i].nchild = nc;
  u.link[i].count = count;
  u.kidtwig[i] = (uint8_t)(~leafmask & ((1u << nc) - 1u));
  for (uint32_t j = 0; j < nc; ++j) u.parent[c0 + j] = (uint32_t)i;
  double2* out = reinterpret_cast<double2*>(u.mom + i * 16);
#pragma unroll
  for (int k = 0; k < 8; ++k) out[k] = make_double2(m[2 * k], m[2 * k + 1]);
  write_walk_forms(i, m, walk_b2(m, b2, u.mac_select), u.prec, u.geom, u.wm32, u.wm64);
}

// One launch per level, deepest first, one thread per twig of the level:
// twig ranks follow the BFS order, so a level's twigs are one contiguous run
// of ranks, [twig_offset[level], twig_offset[level + 1]).
// A twig's children come from the rank-indexed first-child array (children
// of consecutive twigs are consecutive), read together with its node index.
// The launch boundary orders a level's writes before the next level's reads
// on every XCD; a single-launch scheme would need an agent-scope release per
// node, which writes back the L2 of its XCD.
__global__ __launch_bounds__(kT) void upward_level_kernel(int level, UpArgs u) {
  if (tree_broken(u.sc)) return;
  const int64_t tb = u.sc->twig_offset[level], te = u.sc->twig_offset[level + 1];
  for (int64_t r = tb + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < te;
       r += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = u.node_rank[r];
    const uint32_t c0 = u.fc_rank[r];
    twig_moments(i, c0, u.fc_rank[r + 1] - c0, level, u);
  }
}

// Consistency of the grown tree (tree_check, module_tree.f90:438-521, counts
// the leaves and twigs reachable from the root against the recorded totals).
// Every node is checked locally: a leaf holds one particle; a twig's child
// range lies inside the node array and holds 1..8 nodes one level deeper
// whose parent is the twig and whose particle counts add up to the twig's;
// a node with a parent lies in that parent's child range.  res[0..6] =
// {leaves, twigs, roots, violations, particles below roots, sum and sum of
// squares of the leaves' particle indices}; the host compares them with n,
// the twig count, the system count and the sums of 0..n-1.
__global__ __launch_bounds__(kT) void tree_check_kernel(const NodeLink* __restrict__ link,
                                                       const uint32_t* __restrict__ parent,
                                                       int64_t n, const DevScalars* __restrict__ sc,
                                                       unsigned long long* __restrict__ res) {
  const int64_t m = sc->nnodes;
  unsigned long long v[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * blockDim.x) {
    const NodeLink lk = link[i];
    const uint32_t p = parent[i];
    if (p == 0xFFFFFFFFu) {
      v[2] += 1;
      v[4] += lk.count;
    } else if ((int64_t)p >= m) {
      v[3] += 1;
    } else {
      const NodeLink pl = link[p];
      if (!(pl.nchild > 0 && (uint32_t)i >= pl.first && (uint32_t)i - pl.first < pl.nchild)) v[3] += 1;
    }
    if (lk.nchild == 0) {
      v[0] += 1;
      if ((int64_t)lk.first >= n || lk.count != 1u) v[3] += 1;
      v[5] += lk.first;
      v[6] += (unsigned long long)lk.first * lk.first;
    } else {
      v[1] += 1;
      if (lk.nchild > 8u || (int64_t)lk.first + lk.nchild > m) {
        v[3] += 1;
      } else {
        uint32_t sum = 0;
        for (uint32_t j = 0; j < lk.nchild; ++j) {
          const NodeLink c = link[lk.first + j];
          sum += c.count;
          if (c.level != lk.level + 1 || parent[lk.first + j] != (uint32_t)i) v[3] += 1;
        }
        if (sum != lk.count) v[3] += 1;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    const unsigned long long t = wave_sum(v[k]);
    if (lane_id() == 0 && t) atomicAdd(&res[k], t);
  }
}

// Leaf records in the reference layout, for read-back only (not part of a
// solve): coc = x, charge = q, abs_charge = |q|, the rest 0.
__global__ void leaf_mom_kernel(const NodeLink* __restrict__ link, const double* __restrict__ x,
                                const double* __restrict__ y, const double* __restrict__ z,
                                const double* __restrict__ q, double* __restrict__ mom,
                                const DevScalars* __restrict__ sc) {
  if (tree_broken(sc)) return;
  const int64_t m = sc->nnodes;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * blockDim.x) {
    const NodeLink lk = link[i];
    if (lk.nchild != 0) continue;
    double* o = mom + i * 16;
    for (int k = 0; k < 16; ++k) o[k] = 0.0;
    o[M_CX] = x[lk.first]; o[M_CY] = y[lk.first]; o[M_CZ] = z[lk.first];
    o[M_Q] = q[lk.first]; o[M_AQ] = fabs(q[lk.first]);
  }
}

// ------------------------------------------------------------ groups
// A root holding at most G particles is one group.  A twig holding more than
// G particles marks each child with at most G as a group start (the child's first particle gets the group size).  With
// pack > 0, runs of consecutive such children are merged while the merged
// size stays <= pack; a child that is split further ends a run.
// Items: the roots (0 .. nroot-1), then the twigs by rank.
__global__ void group_mark_kernel(int group_size, int pack, const NodeLink* __restrict__ link,
                                  const uint32_t* __restrict__ nfirst,
                                  const uint32_t* __restrict__ node_rank, uint32_t* __restrict__ gmark,
                                  const DevScalars* __restrict__ sc) {
  if (tree_broken(sc)) return;
  const int64_t nroot = sc->level_offset[1];
  const int64_t items = nroot + (int64_t)sc->ntwig;
  const uint32_t G = (uint32_t)group_size, P = (uint32_t)pack;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < items;
       t += (int64_t)gridDim.x * blockDim.x) {
    if (t < nroot) {
      const uint32_t cnt = link[t].count;
      if (cnt <= G) gmark[nfirst[t]] = cnt;   // a root that is one group
      continue;
    }
    const NodeLink lk = link[node_rank[t - nroot]];
    if (lk.count <= G) continue;
    uint32_t run_first = 0xFFFFFFFFu, run_cnt = 0;
    for (uint32_t j = 0; j < lk.nchild; ++j) {
      const uint32_t c = lk.first + j;
      const uint32_t cc = link[c].count;
      if (cc <= G) {
        if (P > 0 && run_first != 0xFFFFFFFFu && run_cnt + cc <= P) {
          run_cnt += cc;
        } else {
          if (run_first != 0xFFFFFFFFu) gmark[run_first] = run_cnt;
          run_first = nfirst[c];
          run_cnt = cc;
        }
      } else {
        if (run_first != 0xFFFFFFFFu) gmark[run_first] = run_cnt;
        run_first = 0xFFFFFFFFu;
        run_cnt = 0;
      }
    }
    if (run_first != 0xFFFFFFFFu) gmark[run_first] = run_cnt;
  }
}

__global__ void group_emit_kernel(int64_t n, const uint32_t* __restrict__ gmark,
                                  const uint32_t* __restrict__ rank, uint32_t* __restrict__ gfirst,
                                  uint32_t* __restrict__ gcount) {
  for (int64_t a = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; a < n;
       a += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t c = gmark[a];
    if (c) {
      const uint32_t g = rank[a];
      gfirst[g] = (uint32_t)a;
      gcount[g] = c;
    }
  }
}

// Root node of each system (level-0 nodes, one per non-empty system).
__global__ void sysroot_kernel(const uint32_t* __restrict__ sys, const uint32_t* __restrict__ nfirst,
                               uint32_t* __restrict__ sysroot, const DevScalars* __restrict__ sc) {
  if (tree_broken(sc)) return;
  const int64_t nroot = sc->level_offset[1];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nroot;
       i += (int64_t)gridDim.x * blockDim.x)
    sysroot[sys[nfirst[i]]] = (uint32_t)i;
}

}  // namespace

static unsigned grid_for(int64_t n) {
  int64_t b = (n + kT - 1) / kT;
  if (b < 1) b = 1;
  if (b > 8192) b = 8192;
  return (unsigned)b;
}

void launch_dup_detect(hipStream_t s, int64_t n, const unsigned long long* keys, const uint32_t* sys,
                       DevScalars* sc) {
  dup_detect_kernel<<<grid_for(n), kT, 0, s>>>(n, keys, sys, sc);
}

void launch_dupfix(hipStream_t s, int64_t n, unsigned long long* keys, double* x, double* y,
                   double* z, const DevBox* boxes, const uint32_t* sys, const int64_t* soff,
                   long long* kd, unsigned long long* oldkeys, uint32_t* dflag, uint32_t* dlist,
                   uint32_t* bsum, DevScalars* sc) {
  dup_prepare_kernel<<<grid_for(n), kT, 0, s>>>(n, keys, sys, kd, oldkeys, dflag, sc);
  launch_scan(s, dflag, dflag, n, bsum, nullptr, &sc->dup_scan_n);
  dup_compact_kernel<<<grid_for(n), kT, 0, s>>>(n, kd, dflag, dlist, sc);
  dup_serial_kernel<<<1, 1, 0, s>>>(n, keys, kd, dlist, sys, soff, sc);
  dup_recoord_kernel<<<grid_for(n), kT, 0, s>>>(n, keys, oldkeys, x, y, z, boxes, sys, sc);
}

int64_t bfs_tiles(int64_t n) {
  const int64_t b = (n + kBfsTile - 1) / kBfsTile;
  return b < 1 ? 1 : b;
}

int64_t bfs_table_words(int64_t n) { return 4 * (int64_t)kLevels * bfs_tiles(n); }

void launch_tree_build(hipStream_t s, const TreeBuffers& t, int group_size, int group_pack) {
  const int64_t n = t.n, cap = t.cap, nb = bfs_tiles(n);
  levels_kernel<<<grid_for(n), kT, 0, s>>>(n, t.keys, t.sys, t.lo, t.hi, t.gmark, t.sc);
  bfs_count_kernel<<<(unsigned)nb, kT, 0, s>>>(n, t.lo, t.hi, t.tab, nb, t.sc);
  bfs_scan_kernel<<<2 * kLevels, kBfsScanT, 0, s>>>(t.tab, nb, cap, t.sc);
  bfs_emit_kernel<<<(unsigned)nb, kT, 0, s>>>(n, t.lo, t.hi, t.tab, nb, t.x, t.y, t.z, t.q, t.geom,
                                              t.link, t.nfirst, t.ntype, t.fc_rank, t.node_rank,
                                              t.parent, t.sc);
  launch_upward(s, t);
  group_mark_kernel<<<grid_for(n), kT, 0, s>>>(group_size, group_pack, t.link, t.nfirst, t.node_rank,
                                               t.gmark, t.sc);
  launch_scan(s, t.gmark, t.cnt, n, t.bsum, &t.sc->ngroups, nullptr, true);
  group_emit_kernel<<<grid_for(n), kT, 0, s>>>(n, t.gmark, t.cnt, t.gfirst, t.gcount);
  if (t.sys) sysroot_kernel<<<grid_for(t.nsys), kT, 0, s>>>(t.sys, t.nfirst, t.sysroot, t.sc);
}

void launch_upward(hipStream_t s, const TreeBuffers& t) {
  UpArgs u;
  u.prec = t.prec; u.mac_select = t.mac_select; u.link = t.link;
  u.x = t.x; u.y = t.y; u.z = t.z; u.q = t.q;
  u.mom = t.mom64; u.geom = t.geom; u.wm32 = t.wm32; u.wm64 = t.wm64;
  u.boxes = t.box; u.sys = t.sys; u.nfirst = t.nfirst;
  u.node_rank = t.node_rank; u.fc_rank = t.fc_rank; u.ntype = t.ntype; u.parent = t.parent;
  u.kidtwig = t.kidtwig;
  u.sc = t.sc;
  // A fixed grid near the resident capacity (six 256-thread blocks per CU of
  // MI355X) with a twig-stride loop: the large levels keep every CU busy and
  // the many small levels launch few idle blocks.
  constexpr unsigned kUpBlocks = 1536;
  for (int l = kMaxLevel; l >= 0; --l)
    upward_level_kernel<<<kUpBlocks, kT, 0, s>>>(l, u);
}

void launch_tree_check(hipStream_t s, const NodeLink* link, const uint32_t* parent, int64_t n,
                       const DevScalars* sc, unsigned long long* res) {
  tree_check_kernel<<<256, kT, 0, s>>>(link, parent, n, sc, res);
}

void launch_leaf_moments(hipStream_t s, const TreeBuffers& t) {
  leaf_mom_kernel<<<grid_for(t.cap), kT, 0, s>>>(t.link, t.x, t.y, t.z, t.q, t.mom64, t.sc);
}

}  // namespace pepc
