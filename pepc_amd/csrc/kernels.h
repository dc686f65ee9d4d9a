// kernels.h -- host-side launch entry points of the device kernels.
#pragma once

#include "common.h"

namespace pepc {

// keys.hip
int box_partial_blocks(int64_t n);
void launch_box(hipStream_t s, int64_t n, const double* x, const double* y, const double* z,
                int force_cubic, double* partial, DevBox* box, DevScalars* sc);
// Batched systems: per-system boxes over the ranges [soff[s], soff[s+1]) and
// the system index of every particle.
void launch_box_batch(hipStream_t s, int nsys, const int64_t* soff, const double* x,
                      const double* y, const double* z, int force_cubic, DevBox* boxes,
                      uint32_t* sys, DevScalars* sc);
// sys == nullptr: one system with boxes[0].
// Also writes rec[i] = {x, y, z, q} (input order) for launch_gather_rec.
void launch_keys(hipStream_t s, int64_t n, const double* x, const double* y, const double* z,
                 const double* q, const DevBox* boxes, const uint32_t* sys,
                 unsigned long long* keys, uint32_t* idx, double4* rec);
// Keys of points that need not lie in the box (sinks): coordinates clamped
// into the box first; idx[i] = i.
void launch_sink_keys(hipStream_t s, int64_t m, const double* x, const double* y, const double* z,
                      const DevBox* box, unsigned long long* keys, uint32_t* idx);
// Runs of g consecutive sinks: gfirst[k] = k g, gcount[k] = min(g, m - k g).
void launch_sink_groups(hipStream_t s, int64_t m, int g, uint32_t* gfirst, uint32_t* gcount);

// sort.hip
int64_t scan_blocks(int64_t n);
int64_t sort_blocks(int64_t n);
// Exclusive scan of n values (in place allowed); *total gets the sum when
// total != nullptr.  With ndev != nullptr only the first min(n, *ndev) values
// are scanned (the count is read on the device; n sizes the grid).  With
// flags, each value counts as 1 when nonzero and 0 otherwise.
void launch_scan(hipStream_t s, const uint32_t* in, uint32_t* out, int64_t n, uint32_t* bsum,
                 uint32_t* total, const uint32_t* ndev = nullptr, bool flags = false);
void launch_radix_pass(hipStream_t s, const unsigned long long* kin, const uint32_t* vin,
                       unsigned long long* kout, uint32_t* vout, int64_t n, int shift, int nbits,
                       uint32_t* hist, uint32_t* bsum);
// One stable pass; afterwards bsum[0..255] holds the element count of each
// digit value.
void launch_radix_pass_dev(hipStream_t s, const unsigned long long* kin, const uint32_t* vin,
                           unsigned long long* kout, uint32_t* vout, int64_t n, const uint32_t* ndev,
                           int shift, int nbits, uint32_t* hist, uint32_t* bsum,
                           const uint32_t* dsrc);
void launch_sort_keys(hipStream_t s, unsigned long long* k0, uint32_t* v0, unsigned long long* k1,
                      uint32_t* v1, int64_t n, uint32_t* hist, uint32_t* bsum,
                      const uint32_t* sysid, int64_t nsys);
// Reorders runs of equal keys in perm (after launch_sort_keys and
// launch_dup_detect) into the order the reference's in-place MSD presort
// gives them; nothing to do (one uniform load per block) when dup_count is 0.
// Scratch: kin[n] (keys by input index), ord/ord2[n], cls[n], and the
// per-pass slot lists fp[n] (positions) and fec[n] (element | class << 32).
// One block per system (soff: nsys + 1 offsets when nsys > 1).
void launch_tie_order(hipStream_t s, int64_t n, int64_t nsys, const int64_t* soff,
                      const unsigned long long* ks, uint32_t* perm, unsigned long long* kin,
                      uint32_t* ord, uint32_t* ord2, uint16_t* cls, uint32_t* fp,
                      unsigned long long* fec, DevScalars* sc);
// out[perm[i]] = in[i] (back to input order)
void launch_scatter_back(hipStream_t s, int64_t n, const uint32_t* perm, const double* in,
                         double* out);
// Sorted SoA particles from the 32-byte records: xs[i] = rec[perm[i]].x, ...
void launch_gather_rec(hipStream_t s, int64_t n, const uint32_t* perm, const double4* rec,
                       double* xs, double* ys, double* zs, double* qs);
void launch_gather4(hipStream_t s, int64_t n, const uint32_t* perm, const double* x, const double* y,
                    const double* z, const double* q, double* xs, double* ys, double* zs,
                    double* qs);

enum WalkPrecision { kWalkFP32Kahan = 0, kWalkFP64 = 1 };

// tree.hip
struct TreeBuffers {
  int64_t n;
  int64_t cap;                       // node capacity
  const unsigned long long* keys;    // sorted particle keys
  const double *x, *y, *z, *q;       // sorted particles
  uint8_t* lo;                       // L(a): level of the first node of particle a
  uint8_t* hi;                       // leaf level of particle a
  uint32_t* cnt;                     // per-particle scratch (group marks -> group ranks)
  uint32_t* tab;                     // (level, tile) node and twig counts and their scans, bfs_table_words(n)
  uint32_t* fc_rank;                 // first child by twig rank (+1 sentinel)
  uint32_t* node_rank;               // node index by twig rank
  uint32_t* nfirst;                  // first particle of each node
  uint8_t* ntype;                    // node kind: 0 leaf, 1 twig
  uint8_t* kidtwig;                  // per twig: bit j set when child j is a twig
  uint32_t* parent;                  // parent node (roots: 0xFFFFFFFF)
  NodeLink* link;
  NodeGeom* geom;
  double* mom64;                     // 16 doubles per node, reference layout
  NodeMoments<float>* wm32;          // walk form, FP32
  NodeMoments<double>* wm64;         // walk form, FP64
  uint32_t* gmark;                   // group marks per particle -> ranks
  uint32_t* gfirst;
  uint32_t* gcount;
  uint32_t* hist;
  uint32_t* bsum;
  DevBox* box;                       // one box per system
  DevScalars* sc;
  int prec;                          // walk record form to write (WalkPrecision)
  int mac_select;                    // 0 Barnes-Hut, 1 Bmax, other: never accept
  // batched systems (sys == nullptr for a single system)
  const uint32_t* sys;               // system index per sorted particle
  const int64_t* soff;               // system ranges, nsys + 1 entries
  uint32_t* sysroot;                 // root node per system
  int64_t nsys;
};

void launch_dup_detect(hipStream_t s, int64_t n, const unsigned long long* keys, const uint32_t* sys,
                       DevScalars* sc);
void launch_dupfix(hipStream_t s, int64_t n, unsigned long long* keys, double* x, double* y,
                   double* z, const DevBox* boxes, const uint32_t* sys, const int64_t* soff,
                   long long* kd, unsigned long long* oldkeys, uint32_t* dflag, uint32_t* dlist,
                   uint32_t* bsum, DevScalars* sc);
void launch_tree_build(hipStream_t s, const TreeBuffers& t, int group_size, int group_pack);
int64_t bfs_table_words(int64_t n);
void launch_leaf_moments(hipStream_t s, const TreeBuffers& t);
// Tree consistency counters res[7] (zeroed by the caller), see tree.hip.
void launch_tree_check(hipStream_t s, const NodeLink* link, const uint32_t* parent, int64_t n,
                       const DevScalars* sc, unsigned long long* res);
// Upward pass alone (all levels) over an existing topology.
void launch_upward(hipStream_t s, const TreeBuffers& t);

// walk.hip
struct WalkArgs {
  int64_t n;
  const double *x, *y, *z, *q;
  const NodeGeom* geom;
  const NodeLink* link;
  const uint8_t* kidtwig;            // per twig: bit j set when child j is a twig
  const NodeMoments<float>* wm32;
  const NodeMoments<double>* wm64;
  const uint32_t* gfirst;
  const uint32_t* gcount;
  DevScalars* sc;
  double theta2, eps2, sign;
  double* ex;
  double* ey;
  double* ez;
  double* pot;
  double* work;
  uint32_t* np2p;
  uint32_t* nm2p;
  uint32_t* gstat;                   // 8 words per group: mac, p2p, m2p, amb, steps, count
  uint32_t* queue;                   // work-queue heads, kQueueStride words apart
  int nqueue = 1;                    // queues: one contiguous range of groups per XCD
  const uint32_t* sys;               // batched systems: system per particle (else nullptr)
  const uint32_t* sysroot;           // batched systems: root node per system
  int nbox;                          // periodic: lattice vectors walked per group (else 1)
  const double* vbox;                // periodic: 3 * nbox lattice vectors, (0,0,0) last
  // sinks (targets other than the grown particles; tx == nullptr otherwise):
  // target coordinates, group count, and gfirst == nullptr for one sink per
  // group.  work = interactions (no +1); gstat is not written.
  const double *tx = nullptr, *ty = nullptr, *tz = nullptr;
  uint32_t nsink_groups = 0;
};
void launch_walk_reset(hipStream_t s, const WalkArgs& a);
// Work queues of the persistent walk: up to kMaxQueues heads, each on a
// 64-byte line of its own (kQueueWords words in all).
constexpr int kMaxQueues = 8;
constexpr int kQueueStride = 16;
constexpr int kQueueWords = kMaxQueues * kQueueStride;
void launch_queue_reset(hipStream_t s, uint32_t* queue);   // work-queue heads only
// list_tile: M2P list capacity in LDS (0 = default: 64 FP32, 32 FP64;
// FP32 takes 64 or 128, FP64 32 or 64).
bool walk_list_tile_ok(int precision, int list_tile);
void launch_walk(hipStream_t s, const WalkArgs& a, int precision, int blocks, int list_tile);
void launch_walk_stats(hipStream_t s, const WalkArgs& a);
int walk_default_blocks(int precision, int list_tile);

// lattice.hip: periodic boundaries (far field).  Tables of complex values
// (double2) hold l = 0..lmax, m = 0..l at l(l+1)/2 + m.
constexpr int kLatticeLM = 20;                  // multipole order of the box
constexpr int kLatticeNM = 231;                 // entries for l <= 20
constexpr int kLatticeNT = 861;                 // entries for l <= 40 (Taylor)
constexpr int kLatticeMaxBlocks = 1024;         // partial-moment blocks
constexpr int kLatticeFact = 81;                // factorials 0..80
struct LatticeGeom {
  double center[3];                             // LatticeCenter
  double origin[3];                             // LatticeOrigin
  double lat[9];                                // rows t_lattice_1..3
  double latinv[9];                             // inverse of lat
  int per[3];                                   // periodic axes
  int simple;                                   // principal axes
  int correction;                               // extrinsic correction (kExtrinsic*)
  double volume;                                // |(t1 x t2) . t3|
  // Fictitious-charge correction: the neighbour boxes' fictitious charges
  // merged by lattice point, the points where all of them meet (and cancel)
  // left out.  fpt[4k..4k+2] = integer lattice coordinates of point k,
  // fpt[4k+3] = mask of the charges there; lattice_mu_kernel writes
  // fpos[4k..4k+3] = {x, y, z, merged charge}.  nfpt = 0 otherwise.
  const int* fpt;
  double* fpos;
  int nfpt;
};
// Extrinsic-to-intrinsic corrections, module_fmm_framework.f90:42-46.
enum { kExtrinsicNone = 0, kExtrinsicRedlack = 1, kExtrinsicFictCharge = 2 };
// fict block of the far field: 4 x {x, y, z, q} fictitious charges, their
// count, then the box dipole (3) and quadrupole trace of the last timestep.
constexpr int kLatticeFict = 21;
// Lattice coefficients for neighbour vectors vec[3*nb] (central box last),
// UL scale 2*layers+1.  Writes mstar, lstar (scratch) and ml.
void launch_lattice_coefficients(hipStream_t s, const double* vec, int nb, double scale,
                                 const double* fact, double2* mstar, double2* lstar,
                                 int maxiter, double2* ml);
int lattice_moment_blocks(int64_t n);
// Recurrence constants: inv[l(l+1)/2 + m] = 1 / ((l-m)(l+m)) for l > m,
// inv[kLatticeNM + m] = 1 / (2m); kLatticeNM + 21 doubles.
void launch_lattice_inv(hipStream_t s, double* inv);
// Per-step moments, mu and the field added to ex..pot (sign applied).
// part: lattice_moment_blocks(n) * kLatticeNM, dpart: 4 per block, om and
// mu: kLatticeNM, dip: 4, fict: kLatticeFict doubles, coef: 8 * kLatticeNM.
void launch_lattice_far(hipStream_t s, int64_t n, const double* x, const double* y, const double* z,
                        const double* q, const LatticeGeom& g, const double2* ml, const double* fact,
                        const double* inv, double sign, double2* part,
                        double* dpart, double2* om, double* dip, double2* mu, double* fict,
                        double* coef, double* ex, double* ey, double* ez, double* pot);
// The field part alone at m other points (mu, fict and coef of the sources
// already computed by launch_lattice_far or launch_lattice_sources).
void launch_lattice_sources(hipStream_t s, int64_t n, const double* x, const double* y,
                            const double* z, const double* q, const LatticeGeom& g,
                            const double2* ml, const double* fact, double2* part, double* dpart,
                            double2* om, double* dip, double2* mu, double* fict, double* coef);
// Positions x[idx[i]] (idx == nullptr: x[i]) for outputs i < m.
void launch_lattice_field(hipStream_t s, int64_t m, const double* x, const double* y, const double* z,
                          const LatticeGeom& g, const double* coef, const double* inv,
                          const double* fict, double sign, double* ex, double* ey, double* ez,
                          double* pot, const int64_t* idx = nullptr);
void launch_constrain_periodic(hipStream_t s, int64_t n, double* x, double* y, double* z,
                               const LatticeGeom& g);

// step.hip: particle state of the on-device timestep (state order = key
// order of the last grow) and the gather targets of the reorder.
struct StepState {
  double *x, *y, *z, *vx, *vy, *vz, *q, *m, *ax, *ay, *az;
  long long* label;
  double *tvx, *tvy, *tvz, *tm;
  long long* tlabel;
};
void launch_kick_drift(hipStream_t s, int64_t n, double dt, const StepState& st);
void launch_reorder_kick(hipStream_t s, int64_t n, double dt, const uint32_t* perm, const double* xs,
                         const double* ys, const double* zs, const double* qs, const double* ex,
                         const double* ey, const double* ez, const StepState& st);
// Chunks: [cbeg[c], cbeg[c+1]) lies in one system; system s owns chunks
// [sys_chunk[s], sys_chunk[s+1]).
constexpr int64_t kEnergyChunk = 4096;
void launch_energy(hipStream_t s, int nsys, int nchunk, const int64_t* cbeg, const int* sys_chunk,
                   const double* pot, const StepState& st, double* partial, double* hist,
                   int64_t cap, uint32_t* counter);
void launch_counter_set(hipStream_t s, uint32_t* counter, uint32_t v);

// adjoint.hip: gradient of the frozen-list solve (sorted order in and out)
struct AdjointArgs {
  int64_t n, cap;
  const double *x, *y, *z, *q;
  const NodeGeom* geom;
  const NodeLink* link;
  const double* mom64;
  const uint32_t *gfirst, *gcount;
  const uint32_t *sys, *sysroot;
  DevScalars* sc;
  double theta2, eps2, sign;
  const double *gex, *gey, *gez, *gpot;   // dL/dE, dL/dpot
  double *gx, *gy, *gz, *gq;              // dL/dx, dL/dq
  double* nadj;                           // 16 doubles per node
  uint32_t *parent, *leafof, *queue;
};
void launch_adjoint(hipStream_t s, const AdjointArgs& a, int blocks);

// Frozen interaction lists (adjoint.hip): per group, the nodes of its leaf
// and accepted-twig interactions in traversal order.
struct ListArgs {
  const double *x, *y, *z;
  const NodeGeom* geom;
  const NodeLink* link;
  const uint32_t *gfirst, *gcount;
  const uint32_t *sys, *sysroot;
  DevScalars* sc;
  double theta2;
  uint32_t* count;      // per group (count pass)
  const uint32_t* off;  // per group (record pass)
  uint32_t* entries;    // record pass
  uint32_t* queue;
};
// record = false: counts and their 64-bit total; record = true: entries.
void launch_lists(hipStream_t s, const ListArgs& a, bool record, int blocks,
                  unsigned long long* total);
void launch_leaf_geom(hipStream_t s, const NodeLink* link, const double* x, const double* y,
                      const double* z, const double* q, NodeGeom* geom, int64_t cap,
                      const DevScalars* sc);
struct ReplayArgs {
  const double *x, *y, *z;
  const NodeGeom* geom;
  const NodeLink* link;
  const double* mom64;
  const uint32_t *gfirst, *gcount;
  const uint32_t *count, *off, *entries;
  DevScalars* sc;
  double eps2, sign;
  double *ex, *ey, *ez, *pot;
  uint32_t* queue;
};
void launch_replay(hipStream_t s, const ReplayArgs& a, int blocks);

// walk.hip: FP64 direct sum
// nbox mirror-box vectors vbox (nullptr: open boundaries).
void launch_direct(hipStream_t s, int64_t n, const double* x, const double* y, const double* z,
                   const double* q, int64_t ntest, const int64_t* testidx, double eps2, double sign,
                   double* ex, double* ey, double* ez, double* pot, int nbox = 1,
                   const double* vbox = nullptr);

}  // namespace pepc
